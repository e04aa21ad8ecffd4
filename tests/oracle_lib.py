"""ctypes bindings to the CPU oracle (oracle/liboracle.so) and, when it has been
built in this tree, the reference library (oracle/_ref/libspice_ref.so).

Test infrastructure only: imported by tests/, by __graft_entry__.smoke() as the
checker, and by bench.py for the CPU-baseline leg.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[1]
ORACLE_SO = ROOT / "oracle" / "liboracle.so"
REF_SO = ROOT / "oracle" / "_ref" / "libspice_ref.so"
REF_STRICT_SO = ROOT / "oracle" / "_ref" / "libspice_ref_strict.so"

ORC_POISSON, ORC_LIF, ORC_COBA = 0, 1, 2
ORC_V, ORC_GEX, ORC_GIN = 0, 1, 2

_u32p = C.POINTER(C.c_uint32)
_vp = C.c_void_p


class U128(C.Structure):
    _fields_ = [("lo", C.c_uint64), ("hi", C.c_uint64)]


class Rng(C.Structure):
    _fields_ = [("s0", C.c_uint64), ("s1", C.c_uint64)]


class Kahan(C.Structure):
    _fields_ = [("c", C.c_float), ("sum", C.c_float)]


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(_vp)


def _seeds(seeds):
    arr = (C.c_uint32 * len(seeds))(*seeds)
    return arr, len(seeds)


_oracle = None


def oracle() -> C.CDLL:
    """Loads oracle/liboracle.so (built by `make -C oracle oracle`)."""
    global _oracle
    if _oracle is None:
        if not ORACLE_SO.exists():
            raise FileNotFoundError(f"{ORACLE_SO} missing: run __graft_entry__.build()")
        L = C.CDLL(str(ORACLE_SO))
        L.orc_murmur3_bytes.restype = U128
        L.orc_murmur3_bytes.argtypes = [_vp, C.c_uint64]
        L.orc_murmur3_u128.restype = U128
        L.orc_murmur3_u128.argtypes = [U128]
        L.orc_seed_chain.argtypes = [_u32p, C.c_int, C.c_int, _vp]
        L.orc_rng_next.restype = C.c_uint64
        L.orc_rng_next.argtypes = [C.POINTER(Rng)]
        L.orc_canonical_f32.restype = C.c_float
        L.orc_canonical_f32.argtypes = [C.POINTER(Rng)]
        L.orc_canonical_f64_leftopen.restype = C.c_double
        L.orc_canonical_f64_leftopen.argtypes = [C.POINTER(Rng)]
        L.orc_exponential.restype = C.c_double
        L.orc_exponential.argtypes = [C.POINTER(Rng), C.c_double]
        L.orc_uniform_f32.restype = C.c_float
        L.orc_uniform_f32.argtypes = [C.POINTER(Rng), C.c_float, C.c_float]
        L.orc_kahan_add.restype = C.c_float
        L.orc_kahan_add.argtypes = [C.POINTER(Kahan), C.c_float]
        L.orc_fp_max_degree.restype = C.c_int64
        L.orc_fp_max_degree.argtypes = [C.c_int64, C.c_double]
        L.orc_fp_size.restype = C.c_int64
        L.orc_fp_size.argtypes = [C.c_int64, C.c_int64, C.c_double]
        L.orc_fp_generate.restype = C.c_int64
        L.orc_fp_generate.argtypes = [C.c_int64, C.c_int64, C.c_double, U128, _vp, _vp]
        L.orc_csr_hash.restype = C.c_uint64
        L.orc_csr_hash.argtypes = [C.c_int64, _vp, _vp]
        L.orc_net_create.restype = _vp
        L.orc_net_create.argtypes = [C.c_float, C.c_float, _u32p, C.c_int]
        L.orc_net_destroy.argtypes = [_vp]
        L.orc_add_population.argtypes = [_vp, C.c_int, C.c_int64, C.c_float, C.c_int]
        L.orc_connect.argtypes = [_vp, C.c_int, C.c_int, C.c_double, C.c_float, C.c_float, C.c_int]
        L.orc_connect_csr.argtypes = [_vp, C.c_int, C.c_int, _vp, _vp, C.c_float, C.c_float, C.c_int]
        L.orc_connect_plastic.argtypes = [_vp, C.c_int, C.c_int, C.c_double, C.c_float, C.c_int, C.c_float]
        L.orc_get_synapses.argtypes = [_vp, C.c_int, _vp, _vp, _vp]
        L.orc_step.argtypes = [_vp]
        L.orc_spikes.restype = C.c_int64
        L.orc_spikes.argtypes = [_vp, C.c_int, C.c_int64, C.POINTER(C.POINTER(C.c_int32))]
        L.orc_get_field.argtypes = [_vp, C.c_int, C.c_int, _vp]
        L.orc_get_twait.argtypes = [_vp, C.c_int, _vp]
        L.orc_events.restype = C.c_int64
        L.orc_events.argtypes = [_vp]
        L.orc_state_hash.restype = C.c_uint64
        L.orc_state_hash.argtypes = [_vp]
        L.orc_csr_synapses.restype = C.c_int64
        L.orc_csr_synapses.argtypes = [_vp, C.c_int]
        L.orc_proj_hash.restype = C.c_uint64
        L.orc_proj_hash.argtypes = [_vp, C.c_int]
        L.orc_fnv1a.restype = C.c_uint64
        L.orc_fnv1a.argtypes = [C.c_uint64, _vp, C.c_uint64]
        L.orc_last_error.restype = C.c_char_p
        _oracle = L
    return _oracle


def seed_chain(seeds, count):
    L = oracle()
    s, n = _seeds(seeds)
    out = np.zeros(2 * count, np.uint64)
    if L.orc_seed_chain(s, n, count, _ptr(out)) != 0:
        raise ValueError(L.orc_last_error().decode())
    return out.reshape(count, 2)


def rng_from(seed_lohi) -> Rng:
    return Rng(int(seed_lohi[0]), int(seed_lohi[1]))


def fp_generate(src, dst, p, seed_lohi):
    """fixed_probability CSR (offsets int64[src+1], neighbors int32[nnz])."""
    L = oracle()
    size = L.orc_fp_size(src, dst, p) if src > 0 and dst > 0 else 0
    off = np.zeros(src + 1, np.int64)
    nb = np.zeros(max(size, 1), np.int32)
    n = L.orc_fp_generate(src, dst, p, U128(int(seed_lohi[0]), int(seed_lohi[1])), _ptr(off), _ptr(nb))
    return off, nb[:n].copy()


def csr_hash(off, nb):
    L = oracle()
    src = len(off) - 1
    return int(L.orc_csr_hash(src, _ptr(np.ascontiguousarray(off, np.int64)),
                              _ptr(np.ascontiguousarray(nb, np.int32))))


class Net:
    """Oracle network (reference step semantics, direct in-order delivery)."""

    def __init__(self, dt, max_delay, seeds):
        self.L = oracle()
        s, n = _seeds(seeds)
        self.h = self.L.orc_net_create(dt, max_delay, s, n)
        if not self.h:
            raise ValueError(self.L.orc_last_error().decode())
        self.sizes = []
        self._keep = []

    def close(self):
        if self.h:
            self.L.orc_net_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def add(self, model, size, rate=0.0, init=0):
        i = self.L.orc_add_population(self.h, model, size, rate, init)
        self.sizes.append(size)
        return i

    def connect(self, src, dst, p, delay, weight, field=ORC_V):
        r = self.L.orc_connect(self.h, src, dst, p, delay, weight, field)
        if r < 0:
            raise ValueError(self.L.orc_last_error().decode())
        return r

    def connect_csr(self, src, dst, off, nb, delay, weight, field=ORC_V):
        off = np.ascontiguousarray(off, np.int64)
        nb = np.ascontiguousarray(nb, np.int32)
        self._keep += [off, nb]
        r = self.L.orc_connect_csr(self.h, src, dst, _ptr(off), _ptr(nb), delay, weight, field)
        if r < 0:
            raise ValueError(self.L.orc_last_error().decode())
        return r

    def connect_plastic(self, src, dst, p, delay, rule, w0=0.0):
        r = self.L.orc_connect_plastic(self.h, src, dst, p, delay, rule, w0)
        if r < 0:
            raise ValueError(self.L.orc_last_error().decode())
        return r

    def synapses(self, c):
        """(W, Zpre, Zpost) of a plastic projection in CSR order."""
        n = self.proj_synapses(c)
        out = [np.zeros(max(n, 1), np.float32) for _ in range(3)]
        if self.L.orc_get_synapses(self.h, c, *[_ptr(a) for a in out]) != 0:
            raise ValueError(self.L.orc_last_error().decode())
        return tuple(a[:n] for a in out)

    def step(self):
        if self.L.orc_step(self.h) != 0:
            raise ValueError(self.L.orc_last_error().decode())

    def spikes(self, pop, age=0):
        p = C.POINTER(C.c_int32)()
        n = self.L.orc_spikes(self.h, pop, age, C.byref(p))
        if n < 0:
            raise ValueError(self.L.orc_last_error().decode())
        return np.ctypeslib.as_array(p, (n,)).copy() if n else np.zeros(0, np.int32)

    def field(self, pop, field=ORC_V):
        out = np.zeros(self.sizes[pop], np.float32)
        self.L.orc_get_field(self.h, pop, field, _ptr(out))
        return out

    def twait(self, pop):
        out = np.zeros(self.sizes[pop], np.int32)
        self.L.orc_get_twait(self.h, pop, _ptr(out))
        return out

    def events(self):
        return int(self.L.orc_events(self.h))

    def state_hash(self):
        return int(self.L.orc_state_hash(self.h))

    def proj_hash(self, c):
        return int(self.L.orc_proj_hash(self.h, c))

    def proj_synapses(self, c):
        return int(self.L.orc_csr_synapses(self.h, c))


ORC_RULE_BRUNELPLUS = 1
ORC_RULE_STDP_PAIR = 2
ORC_RULE_COUNT = 3


def brunel_plastic_net(nP, nE, nI, p, rate, wex, win, rule, w0=0.0, dt=1e-4, delay=15e-4, seeds=(1337,)):
    """Brunel+ (samples/brunel+.cpp:101-117): E->E plastic."""
    n = Net(dt, delay, list(seeds))
    P = n.add(ORC_POISSON, nP, rate)
    E = n.add(ORC_LIF, nE)
    I = n.add(ORC_LIF, nI)
    n.connect(P, E, p, delay, wex)
    n.connect(P, I, p, delay, wex)
    n.connect_plastic(E, E, p, delay, rule, w0)
    n.connect(E, I, p, delay, wex)
    n.connect(I, E, p, delay, win)
    n.connect(I, I, p, delay, win)
    return n, (P, E, I)


def brunel_net(nP, nE, nI, p, rate, wex, win, dt=1e-4, delay=15e-4, seeds=(1337,)):
    """Brunel network of samples/brunel.cpp:75-103, parameterised."""
    n = Net(dt, delay, list(seeds))
    P = n.add(ORC_POISSON, nP, rate)
    E = n.add(ORC_LIF, nE)
    I = n.add(ORC_LIF, nI)
    for s, d, w in ((P, E, wex), (P, I, wex), (E, E, wex), (E, I, wex), (I, E, win), (I, I, win)):
        n.connect(s, d, p, delay, w)
    return n, (P, E, I)


# ----------------------------------------------------------------- reference

_ref = {}


def ref_available() -> bool:
    return REF_SO.exists() and REF_STRICT_SO.exists()


def ref(strict: bool = False) -> C.CDLL:
    """Loads the reference compiled from its own sources: with its release flags
    (fast-math family), or with strict IEEE semantics when strict=True."""
    if strict not in _ref:
        L = C.CDLL(str(REF_STRICT_SO if strict else REF_SO))
        L.ref_last_error.restype = C.c_char_p
        L.ref_murmur3.argtypes = [_vp, C.c_uint64, _vp]
        L.ref_seed_chain.argtypes = [_u32p, C.c_int, C.c_int, _vp]
        for f in ("ref_xoroshiro", "ref_canonical_f32", "ref_canonical_f64_leftopen"):
            getattr(L, f).argtypes = [_u32p, C.c_int, C.c_int, C.c_int64, _vp]
        L.ref_exponential.argtypes = [_u32p, C.c_int, C.c_int, C.c_double, C.c_int64, _vp]
        L.ref_uniform_f32.argtypes = [_u32p, C.c_int, C.c_int, C.c_float, C.c_float, C.c_int64, _vp]
        L.ref_kahan.argtypes = [C.c_float, C.c_int64, _vp, _vp]
        L.ref_fp_size.restype = C.c_int64
        L.ref_fp_size.argtypes = [C.c_int64, C.c_int64, C.c_double]
        L.ref_fp_generate.restype = C.c_int64
        L.ref_fp_generate.argtypes = [C.c_int64, C.c_int64, C.c_double, _u32p, C.c_int, C.c_int,
                                      _vp, _vp, C.c_int64]
        L.ref_fp_hash.argtypes = [C.c_int64, C.c_int64, C.c_double, _u32p, C.c_int, C.c_int,
                                  _vp, _vp, _vp, C.c_int]
        L.ref_brunel_run.argtypes = [C.c_int64, C.c_int64, C.c_int64, C.c_double, C.c_float,
                                     C.c_float, C.c_float, C.c_float, C.c_float, _u32p, C.c_int,
                                     C.c_int64, _vp, _vp, C.c_int64, _vp, _vp, _vp, _vp]
        L.ref_vogels_run.argtypes = [C.c_int64, C.c_int64, C.c_double, C.c_float, C.c_float,
                                     C.c_float, C.c_float, C.c_int, _u32p, C.c_int, C.c_int64,
                                     _vp, _vp, C.c_int64, _vp, _vp, _vp]
        L.ref_brunel_plastic_run.argtypes = [C.c_int64, C.c_int64, C.c_int64, C.c_double, C.c_float,
                                             C.c_float, C.c_float, C.c_float, C.c_float, _u32p, C.c_int,
                                             C.c_int, C.c_float, C.c_int64, _vp, _vp, C.c_int64, _vp,
                                             _vp, _vp]
        _ref[strict] = L
    return _ref[strict]
