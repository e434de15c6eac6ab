"""ctypes binding of libbeamform_amd.so (the C ABI in include/beamform.h).

This is the Python face of the library, used by the tests and the benchmark.
It exposes the same operation as the reference's entry point
``form_beams(J, E, A, G)`` (reference cpu.cxx:15-16): 4-bit complex voltages E
and weights A in ucomplex4 byte encoding, float32 gains G, 4-bit complex beams
J. Device buffers are passed as raw pointers (for example
``torch.Tensor.data_ptr()``), streams as ``hipStream_t`` integers (for example
``torch.cuda.current_stream().cuda_stream``).

The library must be built (``make -C beamforming_amd``); importing this
module raises if it is missing. There is no CPU fallback.
"""
import ctypes
import os

import numpy as np

__all__ = [
    "BF_J_BFPT", "BF_J_BFTP", "BeamformError", "Shape", "Plan", "lib", "library_path",
    "form_beams", "adler32", "adler32_device", "adler32_device_noalloc", "fill_uniform_device", "version", "Chunk", "Batch",
    "Server", "ServerConfig", "Request", "Result", "ServerCounters", "Ingest", "IngestConfig",
    "BF_ERR_INVALID_ARGUMENT", "BF_ERR_UNSUPPORTED", "BF_ERR_OVER_BUDGET", "BF_ERR_NOT_READY", "BF_ERR_CLOSED",
]

BF_ERR_INVALID_ARGUMENT = 1
BF_ERR_UNSUPPORTED = 2
BF_ERR_OVER_BUDGET = 5
BF_ERR_NOT_READY = 6
BF_ERR_CLOSED = 7

BF_J_BFPT = 0  # J[b][f][p][t] (reference Jlinear, arraysizes.hxx:91-101)
BF_J_BFTP = 1  # J[b][f][t][p] (reference J2linear, arraysizes.hxx:103-112)

_HERE = os.path.dirname(os.path.abspath(__file__))
library_path = os.path.join(_HERE, "libbeamform_amd.so")


class BeamformError(RuntimeError):
    """A non-zero status returned by the library."""

    def __init__(self, status, message):
        super().__init__(message)
        self.status = status


class Shape(ctypes.Structure):
    """bf_shape: T, F, D, P, B and the J layout."""

    _fields_ = [
        ("ntimes", ctypes.c_int64),
        ("nfrequencies", ctypes.c_int64),
        ("ndishes", ctypes.c_int64),
        ("npolarizations", ctypes.c_int64),
        ("nbeams", ctypes.c_int64),
        ("j_layout", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
    ]

    @classmethod
    def make(cls, ntimes, nfrequencies, ndishes, nbeams, npolarizations=2, j_layout=BF_J_BFPT):
        return cls(ntimes, nfrequencies, ndishes, npolarizations, nbeams, j_layout, 0)

    @property
    def e_bytes(self):
        return self.ntimes * self.nfrequencies * self.ndishes * self.npolarizations

    @property
    def a_bytes(self):
        return self.nfrequencies * self.nbeams * self.ndishes

    @property
    def g_count(self):
        return self.nfrequencies * self.nbeams

    @property
    def j_bytes(self):
        return self.nbeams * self.nfrequencies * self.npolarizations * self.ntimes

    def as_tuple(self):
        return (self.ntimes, self.nfrequencies, self.ndishes, self.npolarizations, self.nbeams, self.j_layout)


class Chunk(ctypes.Structure):
    """bf_chunk: one chunk of a batch (device E and J, its ntimes, first
    channel and number of consecutive channels)."""

    _fields_ = [
        ("E", ctypes.c_void_p),
        ("J", ctypes.c_void_p),
        ("ntimes", ctypes.c_int64),
        ("channel", ctypes.c_int32),
        ("nfreq", ctypes.c_int32),
    ]


class ServerConfig(ctypes.Structure):
    """bf_server_config."""

    _fields_ = [
        ("hbm_budget", ctypes.c_size_t),
        ("max_inflight", ctypes.c_int32),
        ("max_batch", ctypes.c_int32),
        ("max_launches", ctypes.c_int32),
        ("use_graph", ctypes.c_int32),
        ("poll_us", ctypes.c_int32),
    ]


class Request(ctypes.Structure):
    """bf_request: one chunk for the server."""

    _fields_ = [
        ("E", ctypes.c_void_p),
        ("ntimes", ctypes.c_int64),
        ("channel", ctypes.c_int32),
        ("nfreq", ctypes.c_int32),
        ("e_on_host", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("tag", ctypes.c_uint64),
    ]


class Result(ctypes.Structure):
    """bf_result: a finished request."""

    _fields_ = [
        ("ticket", ctypes.c_uint64),
        ("tag", ctypes.c_uint64),
        ("J", ctypes.c_void_p),
        ("j_bytes", ctypes.c_size_t),
        ("status", ctypes.c_int32),
        ("batch_size", ctypes.c_int32),
        ("t_submit", ctypes.c_double),
        ("t_admit", ctypes.c_double),
        ("t_launch", ctypes.c_double),
        ("t_done", ctypes.c_double),
    ]

    @property
    def latency(self):
        return self.t_done - self.t_submit


class ServerCounters(ctypes.Structure):
    """bf_server_counters."""

    _fields_ = [(n, ctypes.c_uint64) for n in (
        "submitted", "admitted", "completed", "released", "rejected", "failed",
        "launches", "max_batch", "waiting", "inflight")] + [
        ("arena_bytes", ctypes.c_size_t),
        ("bytes_in_use", ctypes.c_size_t),
        ("peak_bytes", ctypes.c_size_t),
        ("mean_batch", ctypes.c_double),
        ("clock", ctypes.c_double),
    ]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class IngestConfig(ctypes.Structure):
    """bf_ingest_config."""

    _fields_ = [
        ("nfrequencies", ctypes.c_int64),
        ("ndishes", ctypes.c_int64),
        ("npolarizations", ctypes.c_int64),
        ("chunk_ntimes", ctypes.c_int64),
        ("group_channels", ctypes.c_int32),
        ("ring_slots", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("reserved2", ctypes.c_int32),
    ]


def _load():
    if not os.path.exists(library_path):
        raise ImportError(f"{library_path} is not built; run `make -C {_HERE}` (hipcc, gfx950)")
    lib = ctypes.CDLL(library_path)
    vp, u8p, fp = ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p
    sig = {
        "bf_status_string": (ctypes.c_char_p, [ctypes.c_int]),
        "bf_version": (ctypes.c_char_p, []),
        "bf_last_error": (ctypes.c_char_p, []),
        "bf_plan_create": (ctypes.c_int, [ctypes.POINTER(Shape), u8p, fp, ctypes.POINTER(vp)]),
        "bf_plan_create_device": (ctypes.c_int, [ctypes.POINTER(Shape), u8p, fp, vp, ctypes.POINTER(vp)]),
        "bf_plan_update_weights": (ctypes.c_int, [vp, u8p, fp]),
        "bf_plan_destroy": (ctypes.c_int, [vp]),
        "bf_plan_shape": (ctypes.c_int, [vp, ctypes.POINTER(Shape)]),
        "bf_plan_describe": (ctypes.c_int, [vp, ctypes.c_char_p, ctypes.c_size_t]),
        "bf_form_beams_device": (ctypes.c_int, [vp, u8p, u8p, vp]),
        "bf_form_beams_host": (ctypes.c_int, [vp, u8p, u8p]),
        "bf_form_beams": (ctypes.c_int, [u8p, u8p, u8p, fp, ctypes.POINTER(Shape)]),
        "bf_adler32_device": (ctypes.c_int, [u8p, ctypes.c_size_t, ctypes.POINTER(ctypes.c_uint32), vp]),
        "bf_adler32_device_noalloc": (ctypes.c_int, [u8p, ctypes.c_size_t, ctypes.POINTER(ctypes.c_uint32), vp]),
        "bf_adler32": (ctypes.c_uint32, [u8p, ctypes.c_size_t]),
        "bf_fill_uniform_device": (ctypes.c_int, [u8p, ctypes.c_size_t, ctypes.c_uint64, ctypes.c_uint64, vp]),
        "bf_form_beams_channel": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_int64, u8p, u8p, vp]),
        "bf_form_beams_channels": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int64, u8p, u8p, vp]),
        "bf_batch_create": (ctypes.c_int, [vp, ctypes.c_int, ctypes.POINTER(vp)]),
        "bf_batch_destroy": (ctypes.c_int, [vp]),
        "bf_batch_set_item_tiles": (ctypes.c_int, [vp, ctypes.c_int]),
        "bf_batch_run": (ctypes.c_int, [vp, ctypes.POINTER(Chunk), ctypes.c_int, vp]),
        "bf_batch_run_graph": (ctypes.c_int, [vp, ctypes.POINTER(Chunk), ctypes.c_int, vp]),
        "bf_batch_describe": (ctypes.c_int, [vp, ctypes.c_char_p, ctypes.c_size_t]),
        "bf_plan_autotune": (ctypes.c_int, [vp, u8p, u8p, vp, ctypes.c_int, ctypes.c_char_p, ctypes.c_size_t]),
        "bf_plan_configure": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_int, ctypes.c_double]),
        "bf_batch_autotune": (ctypes.c_int, [vp, ctypes.POINTER(Chunk), ctypes.c_int, vp, ctypes.c_int,
                                             ctypes.c_char_p, ctypes.c_size_t]),
        "bf_server_create": (ctypes.c_int, [vp, ctypes.POINTER(ServerConfig), ctypes.POINTER(vp)]),
        "bf_server_destroy": (ctypes.c_int, [vp]),
        "bf_server_submit": (ctypes.c_int, [vp, ctypes.POINTER(Request), ctypes.POINTER(ctypes.c_uint64)]),
        "bf_server_submit_many": (ctypes.c_int, [vp, ctypes.POINTER(Request), ctypes.c_int,
                                                 ctypes.POINTER(ctypes.c_uint64)]),
        "bf_server_poll": (ctypes.c_int, [vp, ctypes.POINTER(Result)]),
        "bf_server_wait": (ctypes.c_int, [vp, ctypes.c_double, ctypes.POINTER(Result)]),
        "bf_server_read": (ctypes.c_int, [vp, ctypes.c_uint64, u8p, ctypes.c_size_t]),
        "bf_server_release": (ctypes.c_int, [vp, ctypes.c_uint64]),
        "bf_server_stats": (ctypes.c_int, [vp, ctypes.POINTER(ServerCounters)]),
        "bf_server_clock": (ctypes.c_double, [vp]),
        "bf_server_update_weights": (ctypes.c_int, [vp, u8p, fp]),
        "bf_ingest_create": (ctypes.c_int, [ctypes.POINTER(IngestConfig), ctypes.POINTER(vp)]),
        "bf_ingest_destroy": (ctypes.c_int, [vp]),
        "bf_ingest_push": (ctypes.c_int, [vp, u8p, ctypes.c_int64, ctypes.c_int]),
        "bf_ingest_push_many": (ctypes.c_int, [vp, ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(ctypes.c_int64),
                                               ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]),
        "bf_ingest_pop": (ctypes.c_int, [vp, vp, ctypes.POINTER(Chunk), ctypes.POINTER(ctypes.c_uint64)]),
        "bf_ingest_release": (ctypes.c_int, [vp, ctypes.c_uint64, vp]),
        "bf_ingest_pending": (ctypes.c_int, [vp, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int)]),
        "bf_ingest_sync": (ctypes.c_int, [vp]),
        "bf_ingest_wait_release": (ctypes.c_int, [vp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


lib = _load()


def _check(status):
    if status != 0:
        msg = lib.bf_last_error().decode() or lib.bf_status_string(status).decode()
        raise BeamformError(status, f"{lib.bf_status_string(status).decode()}: {msg}")


def _host_ptr(arr, dtype):
    a = np.ascontiguousarray(arr, dtype=dtype)
    return a, a.ctypes.data_as(ctypes.c_void_p)


def version():
    return lib.bf_version().decode()


class Plan:
    """A form-beams plan: shape + resident weights/gains + kernel choice."""

    def __init__(self, shape, A=None, G=None, *, A_dev=None, G_dev=None, stream=0):
        self.shape = shape
        self._h = ctypes.c_void_p()
        if A_dev is not None:
            _check(lib.bf_plan_create_device(ctypes.byref(shape), ctypes.c_void_p(A_dev), ctypes.c_void_p(G_dev),
                                             ctypes.c_void_p(stream), ctypes.byref(self._h)))
        else:
            A_arr, A_p = _host_ptr(A, np.uint8)
            G_arr, G_p = _host_ptr(G, np.float32)
            if A_arr.size != shape.a_bytes or G_arr.size != shape.g_count:
                raise ValueError("A or G size does not match the shape")
            _check(lib.bf_plan_create(ctypes.byref(shape), A_p, G_p, ctypes.byref(self._h)))

    def close(self):
        if self._h:
            lib.bf_plan_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def handle(self):
        return self._h

    def describe(self):
        buf = ctypes.create_string_buffer(512)
        _check(lib.bf_plan_describe(self._h, buf, 512))
        return buf.value.decode()

    def update_weights(self, A, G):
        A_arr, A_p = _host_ptr(A, np.uint8)
        G_arr, G_p = _host_ptr(G, np.float32)
        if A_arr.size != self.shape.a_bytes or G_arr.size != self.shape.g_count:
            raise ValueError("A or G size does not match the shape")
        _check(lib.bf_plan_update_weights(self._h, A_p, G_p))

    def form_beams_device(self, J_ptr, E_ptr, stream=0):
        """Asynchronous launch on device buffers (raw pointers)."""
        _check(lib.bf_form_beams_device(self._h, ctypes.c_void_p(J_ptr), ctypes.c_void_p(E_ptr),
                                        ctypes.c_void_p(stream)))

    def form_beams_channel(self, channel, ntimes, J_ptr, E_ptr, stream=0):
        """One channel of a channel-table plan, with its own ntimes (device pointers)."""
        _check(lib.bf_form_beams_channel(self._h, channel, ntimes, ctypes.c_void_p(J_ptr), ctypes.c_void_p(E_ptr),
                                         ctypes.c_void_p(stream)))

    def form_beams_channels(self, channel, nfreq, ntimes, J_ptr, E_ptr, stream=0):
        """nfreq consecutive channels of a channel-table plan in one launch (device pointers)."""
        _check(lib.bf_form_beams_channels(self._h, channel, nfreq, ntimes, ctypes.c_void_p(J_ptr),
                                          ctypes.c_void_p(E_ptr), ctypes.c_void_p(stream)))

    def autotune(self, J_ptr, E_ptr, stream=0, iters=0):
        """Times candidate configurations on the given device buffers, keeps the best; returns the report."""
        buf = ctypes.create_string_buffer(4096)
        _check(lib.bf_plan_autotune(self._h, ctypes.c_void_p(J_ptr), ctypes.c_void_p(E_ptr), ctypes.c_void_p(stream),
                                    iters, buf, 4096))
        return buf.value.decode()

    def configure(self, beam_tiles=0, stage_j=True, wg_scale=1.0):
        """Sets the kernel configuration explicitly (see bf_plan_configure)."""
        _check(lib.bf_plan_configure(self._h, int(beam_tiles), int(bool(stage_j)), float(wg_scale)))

    def form_beams_host_ptr(self, J_ptr, E_ptr):
        """Synchronous copy-in / run / copy-out on host buffers given as raw
        pointers (E: e_bytes, J: j_bytes)."""
        _check(lib.bf_form_beams_host(self._h, ctypes.c_void_p(J_ptr), ctypes.c_void_p(E_ptr)))

    def form_beams_host(self, E):
        """Synchronous: host E (numpy uint8) -> host J (numpy uint8)."""
        E_arr, E_p = _host_ptr(E, np.uint8)
        if E_arr.size != self.shape.e_bytes:
            raise ValueError("E size does not match the shape")
        J = np.empty(self.shape.j_bytes, dtype=np.uint8)
        _check(lib.bf_form_beams_host(self._h, J.ctypes.data_as(ctypes.c_void_p), E_p))
        return J


def form_beams(E, A, G, shape):
    """One-shot host call: returns J (numpy uint8) for host E, A, G."""
    E_arr, E_p = _host_ptr(E, np.uint8)
    A_arr, A_p = _host_ptr(A, np.uint8)
    G_arr, G_p = _host_ptr(G, np.float32)
    J = np.empty(shape.j_bytes, dtype=np.uint8)
    _check(lib.bf_form_beams(J.ctypes.data_as(ctypes.c_void_p), E_p, A_p, G_p, ctypes.byref(shape)))
    return J


def adler32(data):
    """Adler-32 of host bytes (reference adler32.c:7-20)."""
    arr, p = _host_ptr(data, np.uint8)
    return int(lib.bf_adler32(p, arr.size))


def adler32_device(ptr, nbytes, stream=0):
    """Adler-32 of device bytes, computed on the device."""
    out = ctypes.c_uint32()
    _check(lib.bf_adler32_device(ctypes.c_void_p(ptr), nbytes, ctypes.byref(out), ctypes.c_void_p(stream)))
    return int(out.value)


def adler32_device_noalloc(ptr, nbytes, stream=0):
    """Adler-32 of device bytes, computed on the device with the library's
    static scratch (no allocation; calls overlapping on one GPU race)."""
    out = ctypes.c_uint32()
    _check(lib.bf_adler32_device_noalloc(ctypes.c_void_p(ptr), nbytes, ctypes.byref(out), ctypes.c_void_p(stream)))
    return int(out.value)


def fill_uniform_device(ptr, nbytes, seed, offset=0, stream=0):
    """Seeded uniform ucomplex4 bytes into device memory (same stream as host oracle)."""
    _check(lib.bf_fill_uniform_device(ctypes.c_void_p(ptr), nbytes, seed, offset, ctypes.c_void_p(stream)))


def _chunk_array(chunks):
    arr = (Chunk * max(1, len(chunks)))()
    for i, c in enumerate(chunks):
        if isinstance(c, Chunk):
            arr[i] = c
        else:
            E, J, ntimes, channel = c[:4]
            arr[i] = Chunk(E, J, ntimes, channel, c[4] if len(c) > 4 else 1)
    return arr


class Batch:
    """A batch context over a channel-table plan: many chunks (each with its
    own ntimes and channels) per persistent launch (balanced work items, or
    fixed-size items through a device-side queue)."""

    def __init__(self, plan, max_chunks):
        self.plan = plan  # keeps the plan alive
        self._h = ctypes.c_void_p()
        _check(lib.bf_batch_create(plan.handle, max_chunks, ctypes.byref(self._h)))

    def close(self):
        if self._h:
            lib.bf_batch_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def set_item_tiles(self, n):
        _check(lib.bf_batch_set_item_tiles(self._h, n))

    def run(self, chunks, stream=0, graph=False):
        """chunks: list of Chunk or (E_ptr, J_ptr, ntimes, channel[, nfreq])."""
        arr = _chunk_array(chunks)
        fn = lib.bf_batch_run_graph if graph else lib.bf_batch_run
        _check(fn(self._h, arr, len(chunks), ctypes.c_void_p(stream)))

    def autotune(self, chunks, stream=0, iters=0):
        arr = _chunk_array(chunks)
        buf = ctypes.create_string_buffer(4096)
        _check(lib.bf_batch_autotune(self._h, arr, len(chunks), ctypes.c_void_p(stream), iters, buf, 4096))
        return buf.value.decode()

    def describe(self):
        buf = ctypes.create_string_buffer(512)
        _check(lib.bf_batch_describe(self._h, buf, 512))
        return buf.value.decode()


class Server:
    """The streaming request loop over a channel-table plan (continuous
    batching with HBM-budget admission)."""

    def __init__(self, plan, hbm_budget, max_inflight=0, max_batch=0, max_launches=0, use_graph=False, poll_us=0):
        self.plan = plan  # the plan must outlive the server
        cfg = ServerConfig(hbm_budget, max_inflight, max_batch, max_launches, int(bool(use_graph)), int(poll_us))
        self._h = ctypes.c_void_p()
        _check(lib.bf_server_create(plan.handle, ctypes.byref(cfg), ctypes.byref(self._h)))

    def close(self):
        if self._h:
            lib.bf_server_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def submit(self, E_ptr, ntimes, channel, nfreq=1, tag=0, e_on_host=False):
        """Queues one chunk; returns its ticket."""
        req = Request(E_ptr, ntimes, channel, nfreq, int(bool(e_on_host)), 0, tag)
        t = ctypes.c_uint64()
        _check(lib.bf_server_submit(self._h, ctypes.byref(req), ctypes.byref(t)))
        return t.value

    def submit_many(self, requests):
        """Queues several chunks atomically: requests is a list of tuples
        (E_ptr, ntimes, channel, nfreq, tag[, e_on_host]). Returns tickets."""
        n = len(requests)
        arr = (Request * max(1, n))()
        for i, q in enumerate(requests):
            arr[i] = Request(q[0], q[1], q[2], q[3], int(bool(q[5])) if len(q) > 5 else 0, 0, q[4])
        tickets = (ctypes.c_uint64 * max(1, n))()
        _check(lib.bf_server_submit_many(self._h, arr, n, tickets))
        return list(tickets[:n])

    def poll(self):
        """A finished request (Result) or None."""
        r = Result()
        st = lib.bf_server_poll(self._h, ctypes.byref(r))
        if st == BF_ERR_NOT_READY:
            return None
        _check(st)
        return r

    def wait(self, timeout=-1.0):
        """Waits for a finished request; None on timeout."""
        r = Result()
        st = lib.bf_server_wait(self._h, float(timeout), ctypes.byref(r))
        if st == BF_ERR_NOT_READY:
            return None
        _check(st)
        return r

    def read(self, ticket, nbytes):
        """Copies a finished request's J to a new host array."""
        out = np.empty(nbytes, dtype=np.uint8)
        _check(lib.bf_server_read(self._h, ticket, out.ctypes.data_as(ctypes.c_void_p), nbytes))
        return out

    def release(self, ticket):
        _check(lib.bf_server_release(self._h, ticket))

    def stats(self):
        c = ServerCounters()
        _check(lib.bf_server_stats(self._h, ctypes.byref(c)))
        return c.as_dict()

    def clock(self):
        return lib.bf_server_clock(self._h)

    def update_weights(self, A, G):
        """New weights and gains (host arrays of the table's shape) for the
        requests submitted from now on."""
        A_arr, A_p = _host_ptr(A, np.uint8)
        G_arr, G_p = _host_ptr(G, np.float32)
        if A_arr.size != self.plan.shape.a_bytes or G_arr.size != self.plan.shape.g_count:
            raise ValueError("A or G size does not match the channel table")
        _check(lib.bf_server_update_weights(self._h, A_p, G_p))


class Ingest:
    """E ingest: packets E[t][F][D][P] into per-channel-group device chunks."""

    def __init__(self, nfrequencies, ndishes, chunk_ntimes, group_channels, ring_slots=4, npolarizations=2):
        cfg = IngestConfig(nfrequencies, ndishes, npolarizations, chunk_ntimes, group_channels, ring_slots, 0, 0)
        self._h = ctypes.c_void_p()
        _check(lib.bf_ingest_create(ctypes.byref(cfg), ctypes.byref(self._h)))

    def close(self):
        if self._h:
            lib.bf_ingest_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def push(self, packet_ptr, ntimes, on_host):
        """Appends a packet; returns False (nothing copied) when no ring slot is free.
        The packet is read asynchronously: keep it unchanged until sync()."""
        st = lib.bf_ingest_push(self._h, ctypes.c_void_p(packet_ptr), ntimes, int(bool(on_host)))
        if st == BF_ERR_NOT_READY:
            return False
        _check(st)
        return True

    def push_many(self, packet_ptrs, ntimes, on_host):
        """Pushes packets in order until one is refused; returns how many were
        taken (fewer than given when a ring slot is missing)."""
        n = len(packet_ptrs)
        ptrs = (ctypes.c_void_p * max(1, n))(*packet_ptrs)
        nts = (ctypes.c_int64 * max(1, n))(*ntimes)
        pushed = ctypes.c_int()
        st = lib.bf_ingest_push_many(self._h, ptrs, nts, n, int(bool(on_host)), ctypes.byref(pushed))
        if st not in (0, BF_ERR_NOT_READY):
            _check(st)
        return pushed.value

    def pop(self, stream=0):
        """Next complete chunk as (Chunk, slot_id), or None."""
        c = Chunk()
        sid = ctypes.c_uint64()
        st = lib.bf_ingest_pop(self._h, ctypes.c_void_p(stream), ctypes.byref(c), ctypes.byref(sid))
        if st == BF_ERR_NOT_READY:
            return None
        _check(st)
        return c, sid.value

    def release(self, slot_id, stream=0):
        _check(lib.bf_ingest_release(self._h, slot_id, ctypes.c_void_p(stream)))

    def sync(self):
        """Waits until every pushed packet has been read."""
        _check(lib.bf_ingest_sync(self._h))

    def wait_release(self):
        """Blocks until the oldest pending release has finished and its slot
        is free; False (at once) when no release is pending."""
        st = lib.bf_ingest_wait_release(self._h)
        if st == BF_ERR_NOT_READY:
            return False
        _check(st)
        return True

    def pending(self):
        n = ctypes.c_int64()
        r = ctypes.c_int()
        _check(lib.bf_ingest_pending(self._h, ctypes.byref(n), ctypes.byref(r)))
        return n.value, r.value
