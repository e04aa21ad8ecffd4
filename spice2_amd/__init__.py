"""spice2_amd: MI355X-native Spice simulation engine.

The product is native: libspice_hip.so (C ABI in include/spice_hip.h) and the
C++20 user API in include/spice/.  This module only loads the in-tree native
libraries with ctypes for tests and benchmarks; it never falls back to a
Python or CPU implementation - a missing library raises.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
LIB = PKG / "lib"

SPICE_OK = 0
SPICE_EPRECONDITION = -3
SPICE_DELIVER_PULL = 0
SPICE_DELIVER_PUSH = 1
SPICE_MAX_BLOCK = 16

_vp = C.c_void_p
_i64 = C.c_int64
_u64 = C.c_uint64


class SpiceStats(C.Structure):
    _fields_ = [("steps", C.c_int64), ("events", C.c_uint64), ("rows", C.c_uint64),
                ("deliver_ms", C.c_double),
                ("deliver_launches", C.c_int64), ("update_ms", C.c_double),
                ("update_launches", C.c_int64), ("compact_ms", C.c_double),
                ("csr_rows_host_resolved", C.c_int64), ("csr_ambiguous_draws", C.c_int64),
                ("csr_ms", C.c_double)]


class SpiceError(RuntimeError):
    pass


class SpicePrecondition(SpiceError):
    """A reference precondition (SPICE_PRE / SPICE_INV) failed."""


def _require(path: Path) -> Path:
    if not path.exists():
        raise FileNotFoundError(
            f"{path} is missing: build the native libraries first (python -m spice2_amd.build)")
    return path


_rt = None
_wl = None


def runtime() -> C.CDLL:
    """libspice_hip.so with argument types set for the C ABI."""
    global _rt
    if _rt is None:
        L = C.CDLL(str(_require(LIB / "libspice_hip.so")))
        L.spice_last_error.restype = C.c_char_p
        L.spice_ctx_create.argtypes = [C.c_int, _i64, C.POINTER(_vp)]
        L.spice_ctx_destroy.argtypes = [_vp]
        L.spice_ctx_set_delivery.argtypes = [_vp, C.c_int]
        L.spice_ctx_set_timing.argtypes = [_vp, C.c_int]
        L.spice_ctx_set_tile.argtypes = [_vp, _i64]
        L.spice_ctx_set_push_chunks.argtypes = [_vp, C.c_int32]
        L.spice_ctx_tile.restype = _i64
        L.spice_ctx_tile.argtypes = [_vp]
        L.spice_ctx_stream.restype = _vp
        L.spice_ctx_stream.argtypes = [_vp]
        L.spice_ctx_sync.argtypes = [_vp]
        L.spice_ctx_stats.argtypes = [_vp, C.POINTER(SpiceStats)]
        L.spice_pop_create.argtypes = [_vp, _i64, C.c_int32, C.c_int32, C.POINTER(_i64)]
        L.spice_pop_upload.argtypes = [_vp, _i64, _vp, C.c_int32]
        L.spice_pop_download.argtypes = [_vp, _i64, _vp, C.c_int32]
        L.spice_pop_size.restype = _i64
        L.spice_pop_size.argtypes = [_vp, _i64]
        L.spice_proj_create_fixed_probability.argtypes = [_vp, _i64, _i64, C.c_double, _u64, _u64,
                                                          _i64, C.POINTER(_i64)]
        L.spice_proj_create_csr.argtypes = [_vp, _i64, _i64, _vp, _vp, _i64, C.POINTER(_i64)]
        L.spice_proj_synapses.argtypes = [_vp, _i64, C.POINTER(_i64)]
        L.spice_proj_download.argtypes = [_vp, _i64, _vp, _vp]
        L.spice_spikes.argtypes = [_vp, _i64, _i64, _vp, _i64, C.POINTER(_i64)]
        L.spice_history_len.restype = _i64
        L.spice_history_len.argtypes = [_vp, _i64]
        L.spice_instance_create.argtypes = [_vp, _i64, C.POINTER(_i64)]
        L.spice_instance_steps.restype = _i64
        L.spice_instance_steps.argtypes = [_vp, _i64]
        L.spice_instance_count.restype = _i64
        L.spice_instance_count.argtypes = [_vp]
        L.spice_pop_create_in.argtypes = [_vp, _i64, _i64, C.c_int32, C.c_int32, C.POINTER(_i64)]
        L.spice_ctx_set_graphs.argtypes = [_vp, C.c_int]
        L.spice_run.argtypes = [_vp, C.c_int32, _vp]
        L.spice_pop_set_tile.argtypes = [_vp, _i64, _i64]
        L.spice_pop_tile.restype = _i64
        L.spice_pop_tile.argtypes = [_vp, _i64]
        L.spice_ctx_autotune.argtypes = [_vp, C.POINTER(_i64), C.c_int32, C.c_int32, C.POINTER(_i64)]
        L.spice_rng_jump.argtypes = [_u64, _u64, _u64, C.POINTER(_u64), C.POINTER(_u64)]
        L.spice_rng_jump.restype = None
        L.spice_rng_jump_poly.argtypes = [_u64, C.POINTER(_u64), C.POINTER(_u64)]
        L.spice_rng_jump_poly.restype = None
        L.spice_exp_draw_interval.argtypes = [C.c_double, C.c_double, C.POINTER(C.c_double),
                                              C.POINTER(C.c_double)]
        L.spice_exp_draw_interval.restype = C.c_int
        L.spice_log_cr.argtypes = [C.c_double]
        L.spice_log_cr.restype = C.c_double
        _rt = L
    return _rt


def check(status: int) -> None:
    if status == SPICE_OK:
        return
    msg = runtime().spice_last_error().decode()
    if status == SPICE_EPRECONDITION:
        raise SpicePrecondition(msg)
    raise SpiceError(f"spice status {status}: {msg}")


class Context:
    """A device context driven directly through the C ABI (no user models)."""

    def __init__(self, max_delay: int = 1, device: int = 0):
        self.L = runtime()
        h = _vp()
        check(self.L.spice_ctx_create(device, max_delay, C.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            self.L.spice_ctx_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def population(self, size: int, words: int = 1, draws: int = 0) -> int:
        out = _i64()
        check(self.L.spice_pop_create(self.h, size, words, draws, C.byref(out)))
        return out.value

    def fixed_probability(self, src: int, dst: int, p: float, seed, delay: int = 1) -> int:
        out = _i64()
        check(self.L.spice_proj_create_fixed_probability(self.h, src, dst, p, int(seed[0]),
                                                         int(seed[1]), delay, C.byref(out)))
        return out.value

    def csr(self, src: int, dst: int, offsets, neighbors, delay: int = 1) -> int:
        off = np.ascontiguousarray(offsets, np.int64)
        nb = np.ascontiguousarray(neighbors, np.int32)
        out = _i64()
        check(self.L.spice_proj_create_csr(self.h, src, dst, off.ctypes.data_as(_vp),
                                           nb.ctypes.data_as(_vp), delay, C.byref(out)))
        return out.value

    def download(self, proj: int, src_size: int):
        n = _i64()
        check(self.L.spice_proj_synapses(self.h, proj, C.byref(n)))
        off = np.zeros(src_size + 1, np.int64)
        nb = np.zeros(max(n.value, 1), np.int32)
        check(self.L.spice_proj_download(self.h, proj, off.ctypes.data_as(_vp), nb.ctypes.data_as(_vp)))
        return off, nb[: n.value]

    def stats(self) -> SpiceStats:
        s = SpiceStats()
        check(self.L.spice_ctx_stats(self.h, C.byref(s)))
        return s


def workloads() -> C.CDLL:
    """libspice_workloads.so: Brunel / conductance networks over the C++ API."""
    global _wl
    if _wl is None:
        runtime()
        L = C.CDLL(str(_require(LIB / "libspice_workloads.so")))
        L.wl_error.restype = C.c_char_p
        L.wl_brunel_create.restype = _vp
        L.wl_brunel_create.argtypes = [_i64, _i64, _i64, C.c_double, C.c_float, C.c_float, C.c_float,
                                       C.c_float, C.c_float, C.POINTER(C.c_uint32), C.c_int, C.c_int]
        L.wl_random_weight_create.restype = _vp
        L.wl_random_weight_create.argtypes = [_i64, _i64, C.c_double, C.c_float, C.c_float, C.c_float, C.c_float,
                                              C.POINTER(C.c_uint32), C.c_int]
        L.wl_coba_create.restype = _vp
        L.wl_coba_create.argtypes = [_i64, _i64, C.c_double, C.c_float, C.c_float, C.c_float,
                                     C.c_float, C.c_int, C.POINTER(C.c_uint32), C.c_int, C.c_int]
        L.wl_brunel_plastic_create.restype = _vp
        L.wl_brunel_plastic_create.argtypes = [_i64, _i64, _i64, C.c_double, C.c_float, C.c_float, C.c_float,
                                               C.c_float, C.c_float, C.POINTER(C.c_uint32), C.c_int, C.c_int,
                                               C.c_float, C.c_int]
        L.wl_destroy.argtypes = [_vp]
        L.wl_step.argtypes = [_vp, _i64]
        L.wl_sync.argtypes = [_vp]
        L.wl_spikes.argtypes = [_vp, C.c_int, _i64, _vp, _i64, C.POINTER(_i64)]
        L.wl_neurons.argtypes = [_vp, C.c_int, _vp, _i64]
        L.wl_context.restype = _vp
        L.wl_context.argtypes = [_vp]
        L.wl_set_graphs.argtypes = [_vp, C.c_int]
        L.wl_batch_brunel_create.restype = _vp
        L.wl_batch_brunel_create.argtypes = [C.c_int, C.POINTER(_i64), C.POINTER(C.c_double),
                                             C.POINTER(C.c_float), C.POINTER(C.c_uint32), C.c_float,
                                             C.c_float, C.c_int]
        L.wl_batch_destroy.argtypes = [_vp]
        L.wl_batch_step.argtypes = [_vp, _i64]
        L.wl_batch_sync.argtypes = [_vp]
        L.wl_batch_set_graphs.argtypes = [_vp, C.c_int]
        L.wl_batch_spikes.argtypes = [_vp, C.c_int, C.c_int, _i64, _vp, _i64, C.POINTER(_i64)]
        L.wl_batch_context.restype = _vp
        L.wl_batch_context.argtypes = [_vp]
        L.wl_batch_set_window.argtypes = [_vp, _i64]
        _dp = C.POINTER(C.c_double)
        _ip = C.POINTER(_i64)
        L.wl_serve_run.argtypes = [C.c_int, C.POINTER(C.c_int), _ip, _dp, _ip, _dp, C.POINTER(C.c_uint32), _ip,
                                   _i64, _i64, C.c_int, _dp, _dp, C.POINTER(_u64), _ip, _dp]
        L.wl_job_alone.argtypes = [C.c_int, _i64, C.c_double, _i64, C.POINTER(C.c_uint32), C.POINTER(_u64)]
        _wl = L
    return _wl


class Network:
    """A Brunel or conductance network built by libspice_workloads.so."""

    def __init__(self, handle, sizes, state_bytes):
        self.L = workloads()
        if not handle:
            msg = self.L.wl_error().decode()
            if msg.startswith("Assertion failed"):
                raise SpicePrecondition(msg)
            raise SpiceError(msg)
        self.h = handle
        self.sizes = sizes
        self.state_bytes = state_bytes

    @classmethod
    def brunel(cls, nP, nE, nI, p, rate, w_ex, w_in, dt=1e-4, delay=15e-4, seeds=(1337,),
               mode=SPICE_DELIVER_PULL):
        L = workloads()
        s = (C.c_uint32 * len(seeds))(*seeds)
        h = L.wl_brunel_create(nP, nE, nI, p, rate, w_ex, w_in, dt, delay, s, len(seeds), mode)
        return cls(h, (nP, nE, nI), 8)

    @classmethod
    def brunel_plastic(cls, nP, nE, nI, p, rate, w_ex, w_in, rule, w0=0.0, dt=1e-4, delay=15e-4,
                       seeds=(1337,), mode=SPICE_DELIVER_PULL):
        """Brunel+ : E->E plastic (rule 1: the sample's synapse, 2: pair STDP)."""
        L = workloads()
        s = (C.c_uint32 * len(seeds))(*seeds)
        h = L.wl_brunel_plastic_create(nP, nE, nI, p, rate, w_ex, w_in, dt, delay, s, len(seeds), rule, w0,
                                       mode)
        return cls(h, (nP, nE, nI), 8)

    @classmethod
    def random_weight(cls, nP, nL, p, rate, scale, dt=1e-4, delay=15e-4, seeds=(1337,)):
        """Poisson -> LIF over fixed_probability(p), one synapse weight per synapse
        drawn by the synapse's init() (uniform in [0, scale))."""
        L = workloads()
        s = (C.c_uint32 * len(seeds))(*seeds)
        h = L.wl_random_weight_create(nP, nL, p, rate, scale, dt, delay, s, len(seeds))
        return cls(h, (nP, nL), 8)

    @classmethod
    def coba(cls, nE, nI, p, w_ex, w_in, dt=1e-4, delay=1e-4, init=1, seeds=(1337,),
             mode=SPICE_DELIVER_PULL):
        L = workloads()
        s = (C.c_uint32 * len(seeds))(*seeds)
        h = L.wl_coba_create(nE, nI, p, w_ex, w_in, dt, delay, init, s, len(seeds), mode)
        return cls(h, (nE, nI), 16)

    def _check(self, r):
        if r != 0:
            msg = self.L.wl_error().decode()
            raise (SpicePrecondition if r == -3 else SpiceError)(msg)

    def close(self):
        if self.h:
            self.L.wl_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def step(self, n: int = 1):
        self._check(self.L.wl_step(self.h, n))

    def sync(self):
        self._check(self.L.wl_sync(self.h))

    def spikes(self, pop: int, age: int = 0) -> np.ndarray:
        out = np.zeros(max(self.sizes[pop], 1), np.int32)
        n = _i64()
        self._check(self.L.wl_spikes(self.h, pop, age, out.ctypes.data_as(_vp), len(out), C.byref(n)))
        return out[: n.value].copy()

    def neurons(self, pop: int) -> np.ndarray:
        """Raw AoS bytes of a stateful population, viewed as uint32 words."""
        out = np.zeros(self.sizes[pop] * self.state_bytes // 4, np.uint32)
        self._check(self.L.wl_neurons(self.h, pop, out.ctypes.data_as(_vp), out.nbytes))
        return out

    def context(self):
        return self.L.wl_context(self.h)

    def stats(self) -> SpiceStats:
        s = SpiceStats()
        check(runtime().spice_ctx_stats(self.context(), C.byref(s)))
        return s

    def set_timing(self, on: bool):
        check(runtime().spice_ctx_set_timing(self.context(), int(on)))

    def set_graphs(self, on: bool):
        self._check(self.L.wl_set_graphs(self.h, int(on)))


def job_footprint(kind: int, N: int, p: float, tile: int = 8192) -> int:
    """Device bytes a served job holds (admission estimate): per synapse 4 B
    target + 2 B tile-local target (+ 12 B state and 4 B event slot for the
    plastic E->E projection); per source and projection 8 B offset and
    2 B per tile run start; per neuron its state, a 31-slot spike ring and
    31-slot event counters per incoming projection."""
    ring = 15 + 16
    if kind == 1:
        E, I = N * 4 // 5, N // 5
        pops = [(E, 16, 2), (I, 16, 2)]
        projs = [(E, E), (E, I), (I, E), (I, I)]
    else:
        P_, E, I = N // 2, N * 4 // 10, N // 10
        pops = [(P_, 0, 0), (E, 8, 3), (I, 8, 3)]
        projs = [(P_, E), (P_, I), (E, E), (E, I), (I, E), (I, I)]
    total = 0
    for a, b in projs:
        nnz = a * b * p
        total += nnz * 6 + a * (8 + 2 * (b // tile + 2))
    if kind == 2:
        total += (N * 4 // 10) ** 2 * p * 16
    for n, sb, nin in pops:
        total += n * (sb + 4 * ring + 4 * ring * nin + 8 * 16 // 64 + 8)
    return int(total * 1.1) + (1 << 20)


def serve(kind, N, p, steps, arrival, seeds, footprint, budget, window=100, graphs=True):
    """Runs jobs through spice::server (libspice_workloads.so, wl_serve_run).
    Returns (admitted, finished, result hashes, synaptic events, host seconds
    spent building / running rounds / retiring / generating connectivity
    (part of the rounds))."""
    L = workloads()
    n = len(kind)
    arr = lambda t, xs: (t * n)(*xs)
    adm = (C.c_double * n)()
    fin = (C.c_double * n)()
    res = (_u64 * n)()
    ev = _i64()
    ph = (C.c_double * 4)()
    sw = (C.c_uint32 * (2 * n))(*[int(w) for s in seeds for w in s])
    r = L.wl_serve_run(n, arr(C.c_int, [int(x) for x in kind]), arr(_i64, [int(x) for x in N]),
                       arr(C.c_double, [float(x) for x in p]), arr(_i64, [int(x) for x in steps]),
                       arr(C.c_double, [float(x) for x in arrival]), sw, arr(_i64, [int(x) for x in footprint]),
                       int(budget), int(window), int(bool(graphs)), adm, fin, res, C.byref(ev), ph)
    if r != 0:
        raise SpiceError(L.wl_error().decode())
    return np.array(adm[:]), np.array(fin[:]), [int(x) for x in res], ev.value, list(ph)


def job_alone(kind, N, p, steps, seed):
    """The result hash of one served job run on a network of its own."""
    L = workloads()
    out = _u64()
    s = (C.c_uint32 * 2)(*seed)
    if L.wl_job_alone(int(kind), int(N), float(p), int(steps), s, C.byref(out)) != 0:
        raise SpiceError(L.wl_error().decode())
    return out.value


class BrunelBatch:
    """Many Brunel networks stepped together in one spice::batch (one context).

    Instance i: N[i] neurons (P/E/I = N/2, N*4/10, N/10), probability p[i],
    Poisson rate rate[i], weights +0.2/(N p) and -1/(N p), seed seeds[i] (two
    32-bit words)."""

    def __init__(self, N, p, rate, seeds, dt=1e-4, delay=15e-4, mode=SPICE_DELIVER_PULL):
        self.L = workloads()
        n = len(N)
        self.n = n
        self.N = [int(x) for x in N]
        Na = (_i64 * n)(*self.N)
        pa = (C.c_double * n)(*[float(x) for x in p])
        ra = (C.c_float * n)(*[float(x) for x in rate])
        sa = (C.c_uint32 * (2 * n))(*[int(w) for s in seeds for w in s])
        h = self.L.wl_batch_brunel_create(n, Na, pa, ra, sa, dt, delay, mode)
        if not h:
            msg = self.L.wl_error().decode()
            raise (SpicePrecondition if msg.startswith("Assertion failed") else SpiceError)(msg)
        self.h = h

    def _check(self, r):
        if r != 0:
            msg = self.L.wl_error().decode()
            raise (SpicePrecondition if r == -3 else SpiceError)(msg)

    def close(self):
        if self.h:
            self.L.wl_batch_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def step(self, n: int = 1):
        self._check(self.L.wl_batch_step(self.h, n))

    def sync(self):
        self._check(self.L.wl_batch_sync(self.h))

    def set_graphs(self, on: bool):
        self._check(self.L.wl_batch_set_graphs(self.h, int(on)))

    def set_window(self, steps: int):
        self._check(self.L.wl_batch_set_window(self.h, steps))

    def autotune(self, candidates=(1024, 2048, 4096, 8192, 16384), reps: int = 3):
        """Pull tile per target-size bucket (< 4096, < 16384, < 65536, larger),
        timed by replaying the last block's delivery; returns the tiles chosen."""
        c = (_i64 * len(candidates))(*candidates)
        out = (_i64 * 4)()
        check(runtime().spice_ctx_autotune(self.context(), c, len(candidates), reps, out))
        return list(out)

    def sizes(self, inst: int):
        N = self.N[inst]
        return (N // 2, N * 4 // 10, N // 10)

    def spikes(self, inst: int, pop: int, age: int = 0) -> np.ndarray:
        out = np.zeros(max(self.sizes(inst)[pop], 1), np.int32)
        n = _i64()
        self._check(self.L.wl_batch_spikes(self.h, inst, pop, age, out.ctypes.data_as(_vp), len(out),
                                           C.byref(n)))
        return out[: n.value].copy()

    def context(self):
        return self.L.wl_batch_context(self.h)

    def stats(self) -> SpiceStats:
        s = SpiceStats()
        check(runtime().spice_ctx_stats(self.context(), C.byref(s)))
        return s

    def set_timing(self, on: bool):
        check(runtime().spice_ctx_set_timing(self.context(), int(on)))
