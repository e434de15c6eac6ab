// Host check of the gfx950 kernels' memory accesses: runs the tile
// pipeline's address arithmetic (the shared helpers e_row_index,
// e_col_offset, j_row_bfpt, j_row_bftp and chunk_tiles of bf_device.hpp, the
// work items of bf_items.hpp) for every wave and lane of every workgroup of
// plain launches (form_beams_mfma) and batched launches (form_beams_batch),
// over random shapes and chunk lists, and checks that
//  * every 16-byte E load lies inside the E buffer of its launch or chunk,
//  * every J store lies inside the J buffer of its launch or chunk,
//  * every J byte is written exactly once (nothing missed, nothing twice),
//  * with the LDS staging area (either J layout), every staging write and
//    read stays in its slot's row, and every J byte stored comes from the
//    staging byte that holds that time of that job (the staging area is
//    modelled as per-byte tags of job, row and time, written as the epilogue
//    writes it, and kept from job to job as the batch kernel's LDS is),
// for both J layouts, with and without the LDS staging area, with and
// without a dish tail, all beam tilings (NT = 1..3) and grids. The loop
// structure restates run_tiles in bf_mfma_kernel.hpp: its prologue loads the
// first RP 32-dish K steps of the wave's first tile (even for a wave with no
// tile), each active tile loads its own steps RP..KQ-1 and the first RP steps
// of the wave's next tile (tile + 8, active or not), and each round either
// stores the wave's words directly (4 bytes per lane, polarization and
// 16-beam group of [b][f][p][t]; 8 bytes per lane and group of [b][f][t][p])
// or, when staged, stores the whole round's rows cooperatively.
// Prints "ok" and exits 0 when every check passes. Built with -DPLANT=n
// (1..3), the model makes one planted mistake of its own (drops the row
// clamp at T - 1; reads the other staging slot; writes [b][f][p][t] staging
// words one row off), which the checks must catch.
#ifndef PLANT
#define PLANT 0
#endif
#include "bf_items.hpp"
#include "bf_mfma_kernel.hpp"

#include <cstdio>
#include <cstring>
#include <vector>

using namespace bf;

namespace {

int failures = 0;
void expect(bool ok, const char *what, long long a = 0, long long b = 0) {
  if (!ok && failures++ < 20) std::printf("FAIL %s (%lld, %lld)\n", what, a, b);
}

uint64_t state = 0x9e3779b97f4a7c15ull;
uint64_t next() {
  state ^= state << 13;
  state ^= state >> 7;
  state ^= state << 17;
  return state;
}
int64_t pick(int64_t lo, int64_t hi) { return lo + (int64_t)(next() % (uint64_t)(hi - lo + 1)); }

// The buffers of one job: E of the launch or chunk, J with a write count per byte.
struct Buffers {
  int64_t e_bytes = 0;
  std::vector<uint8_t> jcount;
};

// A workgroup's J staging area: per byte, the tag (job serial, staging row,
// position in the J row) of what the epilogue last wrote there.
struct Staging {
  std::vector<int64_t> tag;
  int64_t serial = 0;
};

struct Tiling {
  int nt, ks, nbg;
  bool dt, staged;
  int jl;
};

// run_tiles for one job (a frequency of a launch or chunk, one beam group,
// tiles [tile_first, tile_end)); e_base is the job's E offset in the buffer.
void run_job(const Tiling &c, Buffers &buf, Staging &lds, int64_t e_base, int64_t e_tstride, int64_t T, int F, int f,
             int D, int B, int bg, int64_t tile_first, int64_t tile_end) {
  const int KQ = c.ks / 2;            // 32-dish K steps
  const int RP0 = c.nt >= 3 ? 4 : 8;  // E prefetch distance of run_tiles
  const int RP = RP0 < KQ ? RP0 : KQ;
  const int NQ = 2 * c.nt;             // 16-beam groups
  const int64_t step = kWaves;
  const int64_t rounds = tile_end > tile_first ? (tile_end - tile_first + step - 1) / step : 0;
  const bool fast_store = (T & 15) == 0;
  const int64_t jbytes = (int64_t)buf.jcount.size();

  auto load = [&](int64_t tile, int lane, int j) {
    const int ld_row = lane & 15, ld_piece = lane >> 4;
    const int64_t row = PLANT == 1 ? tile * kTileT + ld_row : e_row_index(tile, ld_row, T);
    const int64_t off = e_base + 16 * ld_piece + row * e_tstride + e_col_offset(j, ld_piece, D, c.dt);
    expect(off >= 0 && off + 16 <= buf.e_bytes, "E load inside the buffer", off, buf.e_bytes);
  };
  // n bytes written at off (a piece of n <= 0 bytes writes nothing: its
  // byte stores are all predicated off, so its address is never accessed)
  auto store = [&](int64_t off, int n) {
    if (n > 0) expect(off >= 0 && off + n <= jbytes, "J store inside the buffer", off, jbytes);
    for (int i = 0; i < n; ++i)
      if (off + i >= 0 && off + i < jbytes) ++buf.jcount[(size_t)(off + i)];
  };
  // store_piece / store_row16: 16 bytes when whole, else bytes i < n
  auto piece16 = [&](int64_t off, int n) { store(off, fast_store || n >= 16 ? 16 : (n > 0 ? n : 0)); };

  const int64_t serial = ++lds.serial;
  // job, staging row (beam and polarization) and position in the J row
  auto tag_of = [&](int srow, int64_t t) { return serial << 40 | (int64_t)srow << 24 | t; };
  const int64_t stage_bytes = (int64_t)mfma_lds_staging(c.nt);
  if ((int64_t)lds.tag.size() < stage_bytes) lds.tag.assign((size_t)stage_bytes, -1);
  // a row's bytes in one slot: [lo, lo + 128)
  auto row_lo = [&](int slot, int srow) { return ((int64_t)slot * c.nt * 64 + srow) * 128; };
  // a [b][f][t][p] beam row's bytes in one slot: [lo, lo + 256)
  auto beam_lo = [&](int slot, int srow) { return ((int64_t)slot * c.nt * 32 + srow) * 256; };
  // the epilogue's staging writes of one tile
  auto stage_tile = [&](int64_t r, int wave, int64_t tile) {
    const int slot = (int)(r & 1);
    for (int lane = 0; lane < 64; ++lane) {
      const int kg = lane >> 4, col = lane & 15;
      if (c.jl == 1) { // 8 bytes per beam: times 4 kg .. 4 kg + 3, polarizations interleaved
        for (int qq = 0; qq < NQ; ++qq) {
          const int srow = qq * 16 + col;
          const int64_t at = stage_pair_bftp(c.nt, slot, srow, wave, kg);
          expect(at >= beam_lo(slot, srow) && at + 8 <= beam_lo(slot, srow) + 256 && at + 8 <= stage_bytes,
                 "staging write inside its row", at, srow);
          for (int i = 0; i < 8; ++i) lds.tag[(size_t)(at + i)] = tag_of(srow, 2 * (tile * kTileT + 4 * kg) + i);
        }
        continue;
      }
      for (int qq = 0; qq < NQ; ++qq)
        for (int q = 0; q < 2; ++q) {
          const int srow = (qq * 16 + col) * 2 + q;
          const int64_t at = 4 * (int64_t)stage_word(c.nt, slot, PLANT == 3 ? srow ^ 1 : srow, wave, kg);
          expect(at >= 0 && at + 4 <= stage_bytes, "staging write inside the area", at, srow);
          for (int i = 0; i < 4; ++i) lds.tag[(size_t)(at + i)] = tag_of(srow, tile * kTileT + 4 * kg + i);
        }
    }
  };

  for (int wave = 0; wave < kWaves; ++wave)
    for (int lane = 0; lane < 64; ++lane)
      for (int q = 0; q < RP; ++q) load(tile_first + wave, lane, q); // prologue

  for (int64_t r = 0; r < rounds; ++r) {
    for (int wave = 0; wave < kWaves; ++wave) {
      const int64_t tile = tile_first + wave + step * r;
      if (tile >= tile_end) continue;
      for (int lane = 0; lane < 64; ++lane)
        for (int j = 0; j < KQ; ++j) {
          const int j2 = j + RP;
          if (j2 >= KQ)
            load(tile + step, lane, j2 - KQ);
          else
            load(tile, lane, j2);
        }
      if (c.staged) {
        stage_tile(r, wave, tile);
        continue;
      }
      const int64_t t0 = tile * kTileT;
      const int nvalid = T - t0 < kTileT ? (int)(T - t0) : kTileT;
      for (int lane = 0; lane < 64; ++lane) {
        const int kg = lane >> 4, col = lane & 15;
        for (int qq = 0; qq < NQ; ++qq) {
          const int b = bg * c.nt * 32 + qq * 16 + col;
          if (b >= B) continue;
          if (c.jl == 0) {
            for (int q = 0; q < 2; ++q) { // store_row4
              const int n = nvalid - 4 * kg;
              store(j_row_bfpt(b, F, f, q, T) + t0 + 4 * kg, fast_store || n >= 4 ? 4 : (n > 0 ? n : 0));
            }
          } else {
            const int n = 2 * (nvalid - 4 * kg); // store_row8
            store(j_row_bftp(b, F, f, T) + 2 * (t0 + 4 * kg), n >= 8 ? 8 : (n > 0 ? n : 0));
          }
        }
      }
    }
    if (!c.staged) continue;
    const int64_t t_round = (tile_first + r * step) * kTileT;
    const int64_t t_end = tile_end * kTileT < T ? tile_end * kTileT : T;
    for (int k = 0; k < c.nt; ++k)
      for (int thread = 0; thread < kThreads; ++thread) {
        const int q = thread + kThreads * k;
        if (c.jl == 0) {
          const int srow = q >> 3, seg = q & 7;
          const int slot = (int)(r & 1) ^ (PLANT == 2 ? 1 : 0);
          const int64_t at = stage_piece_byte(c.nt, slot, srow, seg);
          expect(at >= row_lo(slot, srow) && at + 16 <= row_lo(slot, srow) + 128 && at + 16 <= stage_bytes,
                 "staging read inside its row", at, srow);
          const int b = bg * c.nt * 32 + (srow >> 1);
          const int64_t t = t_round + 16 * seg;
          if (b < B && t < t_end) {
            const int n = t_end - t < 16 ? (int)(t_end - t) : 16;
            for (int i = 0; i < n; ++i)
              expect(lds.tag[(size_t)(at + i)] == tag_of(srow, t + i), "J byte from the staging byte of its time", t + i,
                     lds.tag[(size_t)(at + i)] & 0xffffffll);
            piece16(j_row_bfpt(b, F, f, srow & 1, T) + t, n);
          }
        } else {
          const int srow = q >> 4, seg = q & 15;
          const int slot = (int)(r & 1);
          const int64_t at = stage_piece_bftp(c.nt, slot, srow, seg);
          expect(at >= beam_lo(slot, srow) && at + 16 <= beam_lo(slot, srow) + 256 && at + 16 <= stage_bytes,
                 "staging read inside its row", at, srow);
          const int b = bg * c.nt * 32 + srow;
          const int64_t t = t_round + 8 * seg;
          if (b < B && t < t_end) {
            const int n = t_end - t < 8 ? (int)(2 * (t_end - t)) : 16;
            for (int i = 0; i < n; ++i)
              expect(lds.tag[(size_t)(at + i)] == tag_of(srow, 2 * t + i), "J byte from the staging byte of its time",
                     2 * t + i, lds.tag[(size_t)(at + i)] & 0xffffffll);
            piece16(j_row_bftp(b, F, f, T) + 2 * t, n);
          }
        }
      }
  }
}

Tiling random_tiling(int D, int B, int jl) {
  Tiling c;
  const int kneed = (D + kDishesPerStep - 1) / kDishesPerStep;
  c.ks = 2;
  while (c.ks < kneed) c.ks *= 2;
  c.dt = c.ks * kDishesPerStep != D;
  const int max_nt = c.ks >= 64 ? 2 : kMaxBeamTiles;
  c.nt = (int)pick(1, max_nt);
  const int nt_tot = (B + kBeamTile - 1) / kBeamTile;
  c.nbg = (nt_tot + c.nt - 1) / c.nt;
  c.jl = jl;
  c.staged = mfma_stages_j(c.nt, c.ks) && (next() & 1);
  return c;
}

void check_written_once(const Buffers &buf, const char *what, long long id) {
  for (size_t i = 0; i < buf.jcount.size(); ++i)
    if (buf.jcount[i] != 1) {
      expect(false, what, id, (long long)i);
      return;
    }
}

// A plain launch: work items (f, bg, time chunk) as in form_beams_mfma.
void check_plain(int id) {
  const int64_t T = next() % 4 ? pick(1, 3000) : pick(3000, 40000);
  const int F = (int)pick(1, 4);
  const int D = 8 * (int)pick(1, 128);
  const int B = (int)pick(1, 300);
  const Tiling c = random_tiling(D, B, (int)(next() & 1));
  const int64_t ntiles = (T + kTileT - 1) / kTileT;
  const int tpc = kWaves * (int)pick(1, (int)((ntiles + kWaves - 1) / kWaves));
  const int nchunks = (int)((ntiles + tpc - 1) / tpc);
  Buffers buf;
  buf.e_bytes = T * F * D * 2;
  buf.jcount.assign((size_t)(B * F * 2 * T), 0);
  Staging lds;
  for (int f = 0; f < F; ++f)
    for (int bg = 0; bg < c.nbg; ++bg)
      for (int ch = 0; ch < nchunks; ++ch) {
        int64_t tb, te;
        chunk_tiles(ch, tpc, ntiles, &tb, &te);
        run_job(c, buf, lds, (int64_t)f * D * 2, (int64_t)F * D * 2, T, F, f, D, B, bg, tb, te);
      }
  check_written_once(buf, "plain launch: every J byte written once", id);
}

// A batched launch: segments of balanced or fixed-size items, each chunk with
// its own E and J buffers of exactly its ntimes samples.
void check_batch(int id) {
  const int n = (int)pick(1, 12);
  const int D = 8 * (int)pick(1, 128);
  const int B = (int)pick(1, 300);
  const Tiling c = random_tiling(D, B, (int)(next() & 1));
  std::vector<bf_chunk> chunks((size_t)n);
  for (bf_chunk &k : chunks) {
    std::memset(&k, 0, sizeof k);
    k.ntimes = next() % 3 ? pick(1, 2500) : pick(2500, 20000);
    k.nfreq = (int)pick(0, 3); // 0 means 1
  }
  std::vector<BatchItem> segs;
  std::vector<int> work;
  if (next() & 1) {
    balanced_items(chunks.data(), n, c.nbg, (int)pick(1, 600), segs, work);
  } else {
    fixed_items(chunks.data(), n, c.nbg, kWaves * (int)pick(1, 32), segs);
  }
  std::vector<Buffers> bufs((size_t)n);
  for (int i = 0; i < n; ++i) {
    const int nf = chunk_nfreq(chunks[i]);
    bufs[i].e_bytes = chunks[i].ntimes * nf * D * 2;
    bufs[i].jcount.assign((size_t)(B * nf * 2 * chunks[i].ntimes), 0);
  }
  Staging lds;
  // through the segment records the kernel reads (E addresses as offsets
  // from a chunk base that is never dereferenced)
  for (const BatchItem &s : segs) {
    bf_chunk k = chunks[s.chunk];
    k.E = reinterpret_cast<const uint8_t *>(uintptr_t(1) << 40);
    const SegRec r = seg_record(k, s, D);
    expect(r.nf == chunk_nfreq(chunks[s.chunk]) && r.T == k.ntimes && r.channel == k.channel + s.f,
           "segment record fields", s.chunk, s.f);
    run_job(c, bufs[s.chunk], lds, (int64_t)(r.E - k.E), (int64_t)r.nf * D * 2, r.T, r.nf, r.f_bg & 0xffff, D, B,
            r.f_bg >> 16, r.tile_begin, r.tile_end);
  }
  for (int i = 0; i < n; ++i) check_written_once(bufs[i], "batched launch: every chunk's J byte written once", id);
}

} // namespace

int main() {
  for (int i = 0; i < 60; ++i) check_plain(i);
  for (int i = 0; i < 60; ++i) check_batch(i);
  if (failures) {
    std::printf("%d failures\n", failures);
    return 1;
  }
  std::printf("ok\n");
  return 0;
}
