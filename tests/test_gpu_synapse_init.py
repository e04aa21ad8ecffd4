"""Per-synapse init() (synapse_population.h:34-41): the reference calls init(syn,
src, dst, rng) for every synapse in CSR order with a generator seeded from the
network's seed chain (the stateful target population takes seed 0, the
topology seed 1, the init generator seed 2).  An init() that draws runs on the
host; one that takes no draws runs on the device (the STDP rule's, covered by
the plastic tests).  Here the weights of a drawing init() read back from the
device must equal the oracle's generator outputs in CSR order, bit for bit, and
the CSR must equal the oracle's fixed_probability."""
import ctypes as C

import numpy as np
import pytest

import oracle_lib as O
import spice2_amd as S

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("nP,nL,p", [(500, 300, 0.1), (3000, 2000, 0.02)])
def test_drawing_init_weights_follow_the_seed_chain(nP, nL, p):
    scale, seeds = np.float32(2e-4), (1337, 7)
    print("random_weight", nP, nL, p, flush=True)
    with S.Network.random_weight(nP, nL, p, 20.0, float(scale), seeds=seeds) as net:
        L = S.runtime()
        n = C.c_int64()
        S.check(L.spice_proj_synapses(net.context(), 0, C.byref(n)))
        off = np.zeros(nP + 1, np.int64)
        nb = np.zeros(max(n.value, 1), np.int32)
        S.check(L.spice_proj_download(net.context(), 0, off.ctypes.data_as(C.c_void_p), nb.ctypes.data_as(C.c_void_p)))
        got = np.zeros(max(n.value, 1), np.float32)
        L.spice_proj_get_synapses.argtypes = [C.c_void_p, C.c_int64, C.c_void_p, C.c_int32]
        S.check(L.spice_proj_get_synapses(net.context(), 0, got.ctypes.data_as(C.c_void_p), 4))
    chain = O.seed_chain(list(seeds), 3)
    ooff, onb = O.fp_generate(nP, nL, p, chain[1])
    assert np.array_equal(off, ooff) and np.array_equal(nb[: n.value], onb)
    rng = O.rng_from(chain[2])
    lib = O.oracle()
    want = np.array([lib.orc_canonical_f32(C.byref(rng)) for _ in range(n.value)], np.float32) * scale
    assert n.value > 0
    assert np.array_equal(got[: n.value].view(np.uint32), want.view(np.uint32))
