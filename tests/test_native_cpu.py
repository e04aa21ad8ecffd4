"""CPU checks of the native library: it loads, exports every entry point the C ABI
header declares, and its host-side pieces (jump-ahead, logarithm, draw interval,
the C++ API restated from the reference's unit tests) are right.  Nothing here
launches device work."""
import ctypes as C
import math
import random
import re
import subprocess
from decimal import Decimal, getcontext

import numpy as np
import pytest

import oracle_lib as O
import spice2_amd as S

HEADER = S.ROOT / "include" / "spice_hip.h"


def declared_functions():
    text = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    return sorted(set(re.findall(r"^\s*[\w\s\*]+?\b(spice_\w+)\s*\(", text, flags=re.M)))


def test_runtime_exports_every_declared_symbol():
    L = S.runtime()
    names = declared_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_workloads_library_loads():
    L = S.workloads()
    for n in ("wl_brunel_create", "wl_coba_create", "wl_step", "wl_spikes", "wl_neurons"):
        assert hasattr(L, n)


def test_api_host_program():
    """The C++ API restated against the reference's own unit tests (CSR, Kahan,
    UInt128, RNG bounds, seed chain, concepts)."""
    exe = S.ROOT / "tests" / "native" / "bin" / "api_host_tests"
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "all passed" in r.stdout


def _jump(s0, s1, n):
    a, b = C.c_uint64(), C.c_uint64()
    S.runtime().spice_rng_jump(s0, s1, n, C.byref(a), C.byref(b))
    return a.value, b.value


def test_jump_ahead_equals_stepping():
    L = O.oracle()
    rs = np.random.default_rng(3)
    for n in (0, 1, 2, 63, 64, 65, 127, 128, 1000, 4097, 65537, 250000):
        s0, s1 = (int(x) for x in rs.integers(1, 2**63, 2))
        r = O.Rng(s0, s1)
        for _ in range(n):
            L.orc_rng_next(r)
        assert _jump(s0, s1, n) == (r.s0, r.s1), n


def test_jump_composes_to_published_2_64_jump():
    """x^(2^63) applied twice equals the jump polynomial for 2^64 steps that the
    xoroshiro128+ authors publish (0xdf900294d8f554a5, 0x170865df4b3201fc)."""
    lo, hi = C.c_uint64(), C.c_uint64()
    S.runtime().spice_rng_jump_poly(2**63, C.byref(lo), C.byref(hi))
    L = O.oracle()

    def apply(poly, s0, s1):
        acc0 = acc1 = 0
        r = O.Rng(s0, s1)
        for i in range(128):
            if (poly >> i) & 1:
                acc0 ^= r.s0
                acc1 ^= r.s1
            L.orc_rng_next(r)
        return acc0, acc1

    p63 = lo.value | (hi.value << 64)
    s = (0x0123456789ABCDEF, 0x55AA55AA55AA55AA)
    twice = apply(p63, *apply(p63, *s))
    published = apply(0xDF900294D8F554A5 | (0x170865DF4B3201FC << 64), *s)
    assert twice == published


def _cr_log(x: float) -> float:
    getcontext().prec = 70
    return float(Decimal(x).ln())  # Decimal -> float rounds correctly


def test_log_is_correctly_rounded():
    L = S.runtime()
    rnd = random.Random(11)
    xs = [1.0, 0.5, 2.0**-53, 1 - 2.0**-53, 2.0**-1074, 1e-300, 0.7071067811865476,
          1.4142135623730951, 3.0, 1e300]
    xs += [((rnd.getrandbits(53) >> 0) + 1) * 2.0**-53 for _ in range(3000)]
    for x in xs:
        assert L.spice_log_cr(x) == _cr_log(x), x


def test_draw_interval_contains_the_c_library_value():
    """glibc's log (through Python's math.log) is either the correctly rounded
    value or, for draws flagged ambiguous, the other neighbour."""
    L = S.runtime()
    rnd = random.Random(12)
    lo, hi = C.c_double(), C.c_double()
    flagged = 0
    for _ in range(200000):
        u = (rnd.getrandbits(53) + 1) * 2.0**-53
        for scale in (9.0, 49.0):
            amb = L.spice_exp_draw_interval(u, scale, C.byref(lo), C.byref(hi))
            g = -scale * math.log(u)
            flagged += amb
            if amb:
                assert g in (lo.value, hi.value) and lo.value < hi.value
            else:
                assert lo.value == hi.value == g
    assert 0.05 < flagged / 400000 < 0.15  # about 10% of draws are near a midpoint


def test_oracle_and_library_are_separate():
    """The product library does not link the oracle (the checker)."""
    out = subprocess.run(["ldd", str(S.LIB / "libspice_hip.so")], capture_output=True, text=True).stdout
    assert "oracle" not in out
    out = subprocess.run(["ldd", str(S.LIB / "libspice_workloads.so")], capture_output=True, text=True).stdout
    assert "oracle" not in out
