// beamform.hpp -- C++ drop-in for the reference entry point
//
//     void form_beams(ucomplex4 *J, const ucomplex4 *E, const ucomplex4 *A,
//                     const float *G);                         (cpu.cxx:15-16)
//
// The reference fixes the shape at compile time (arraysizes.hxx:22-27); here
// the shape is set once with bf::set_shape() and form_beams() then has the
// reference's four-pointer signature, with host arrays in the reference
// layouts (arraysizes.hxx:74-142). It runs on the GPU through the C ABI of
// beamform.h. Where the reference exits the process on an error
// (cuda.cu:255-262), this throws bf::error.
//
// Header-only; link with -lbeamform_amd.
#pragma once

#include "beamform.h"

#include <cstdint>
#include <stdexcept>
#include <string>

namespace bf {

// One 4-bit complex value, byte-compatible with the reference's ucomplex4
// (icomplex4.hxx:22-40): real part in the high nibble, imaginary part in the
// low nibble, each stored as value + 8.
struct ucomplex4 {
  unsigned char data;
  ucomplex4() = default;
  constexpr ucomplex4(signed char re, signed char im)
      : data((unsigned char)((((unsigned char)re << 4) | (im & 0x0f)) ^ 0x88)) {}
  constexpr signed char real() const { return (signed char)(data ^ 0x88) >> 4; }
  constexpr signed char imag() const { return (signed char)((unsigned char)(data ^ 0x88) << 4) >> 4; }
  constexpr bool operator==(const ucomplex4 &o) const { return data == o.data; }
  constexpr bool operator!=(const ucomplex4 &o) const { return data != o.data; }
};
static_assert(sizeof(ucomplex4) == 1, "one byte per complex value");
static_assert(ucomplex4(1, 2).data == 0x9a && ucomplex4(-1, -2).data == 0x76, "reference encoding");

struct error : std::runtime_error {
  int status;
  error(int s, const std::string &what) : std::runtime_error(what), status(s) {}
};

inline void check(int status) {
  if (status != BF_OK) throw error(status, std::string(bf_status_string(status)) + ": " + bf_last_error());
}

// The shape used by form_beams(); defaults to the reference's shipped
// benchmark configuration (arraysizes.hxx:22-27).
inline bf_shape &current_shape() {
  static bf_shape s{8192, 68, 512, 2, 96, BF_J_BFPT, 0};
  return s;
}

inline void set_shape(int64_t ntimes, int64_t nfrequencies, int64_t ndishes, int64_t nbeams,
                      int64_t npolarizations = 2, int32_t j_layout = BF_J_BFPT) {
  current_shape() = bf_shape{ntimes, nfrequencies, ndishes, npolarizations, nbeams, j_layout, 0};
}

// Sizes in complex values (= bytes), as Esize/2, Asize/2, Jsize/2 and Gsize in
// arraysizes.hxx:79-142.
inline size_t e_count() {
  const bf_shape &s = current_shape();
  return (size_t)(s.ntimes * s.nfrequencies * s.ndishes * s.npolarizations);
}
inline size_t a_count() {
  const bf_shape &s = current_shape();
  return (size_t)(s.nfrequencies * s.nbeams * s.ndishes);
}
inline size_t g_count() {
  const bf_shape &s = current_shape();
  return (size_t)(s.nfrequencies * s.nbeams);
}
inline size_t j_count() {
  const bf_shape &s = current_shape();
  return (size_t)(s.nbeams * s.nfrequencies * s.npolarizations * s.ntimes);
}

// The drop-in: host arrays, current shape, J fully overwritten.
inline void form_beams(ucomplex4 *J, const ucomplex4 *E, const ucomplex4 *A, const float *G) {
  check(bf_form_beams(reinterpret_cast<uint8_t *>(J), reinterpret_cast<const uint8_t *>(E),
                      reinterpret_cast<const uint8_t *>(A), G, &current_shape()));
}

// RAII plan for repeated calls with resident weights (device buffers).
class plan {
public:
  plan(const bf_shape &shape, const ucomplex4 *A, const float *G) {
    check(bf_plan_create(&shape, reinterpret_cast<const uint8_t *>(A), G, &p_));
  }
  ~plan() { bf_plan_destroy(p_); }
  plan(const plan &) = delete;
  plan &operator=(const plan &) = delete;
  void run_device(ucomplex4 *J_dev, const ucomplex4 *E_dev, void *stream = nullptr) const {
    check(bf_form_beams_device(p_, reinterpret_cast<uint8_t *>(J_dev), reinterpret_cast<const uint8_t *>(E_dev),
                               stream));
  }
  void run_host(ucomplex4 *J, const ucomplex4 *E) const {
    check(bf_form_beams_host(p_, reinterpret_cast<uint8_t *>(J), reinterpret_cast<const uint8_t *>(E)));
  }
  bf_plan *get() const { return p_; }

private:
  bf_plan *p_ = nullptr;
};

} // namespace bf
