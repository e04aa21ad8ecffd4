"""Plastic synapses on the device: Brunel+ (samples/brunel+.cpp) with the
sample's plastic E->E synapse and with the pair-STDP rule.  The lazy catch-up
(synapse_population.h:87-115), the sweep every 64 steps (snn.cpp:17-19) and the
source-ordered delivery are compared with the reference-built fixtures
(tests/golden/brunelplus_c0.npz, stdp_small.npz) step by step, and the synapse
state with the oracle."""
import ctypes as C
import json

import numpy as np
import pytest

import oracle_lib as O
import spice2_amd as S

pytestmark = pytest.mark.gpu

G = O.ROOT / "tests" / "golden"
NETS = json.loads((G / "networks.json").read_text())
FNV = 0xCBF29CE484222325
MODES = {"pull": S.SPICE_DELIVER_PULL, "push": S.SPICE_DELIVER_PUSH}


def fnv(h, arr):
    arr = np.ascontiguousarray(arr)
    return int(O.oracle().orc_fnv1a(h, O._ptr(arr), arr.nbytes))


def network(rec, mode):
    N = rec["N"]
    w0 = float(np.uint32(rec["w0_bits"]).view(np.float32))
    return S.Network.brunel_plastic(N // 2, N * 4 // 10, N // 10, rec["p"], 20.0, float(np.float32(2.0 / N)),
                                    float(np.float32(-10.0 / N)), rec["rule"], w0, mode=mode)


@pytest.mark.parametrize("mode", ["pull", "push"])
@pytest.mark.parametrize("name", ["brunelplus_c0", "stdp_small"])
def test_plastic_network_matches_reference(name, mode):
    rec = NETS[name]
    z = np.load(G / f"{name}.npz")
    print("network", name, mode, flush=True)
    mismatch = None
    with network(rec, MODES[mode]) as net:
        for t in range(rec["steps"]):
            net.step()
            lists = [net.spikes(p) for p in range(3)]
            h = FNV
            for x in lists:
                h = fnv(h, x)
            if [len(x) for x in lists] != z["counts"][t].tolist() or h != int(z["raster_hash"][t]):
                mismatch = t
                break
            if t % 50 == 49:
                sh = fnv(fnv(FNV, net.neurons(1)), net.neurons(2))
                assert sh == int(z["state_hash"][t]), (name, t)
    assert mismatch is None, f"{name}: first differing step {mismatch}"


def test_plastic_synapse_state_matches_oracle():
    """W, Zpre and Zpost of every E->E synapse after 200 steps (three sweeps and
    the deliveries in between) against the oracle's."""
    rec = NETS["stdp_small"]
    N = rec["N"]
    w0 = np.float32(np.uint32(rec["w0_bits"]).view(np.float32))
    steps = 200
    onet, _ = O.brunel_plastic_net(N // 2, N * 4 // 10, N // 10, rec["p"], 20.0, np.float32(2.0 / N),
                                   np.float32(-10.0 / N), rec["rule"], w0)
    for _ in range(steps):
        onet.step()
    W, Zpre, Zpost = onet.synapses(2)
    onet.close()
    with network(rec, S.SPICE_DELIVER_PULL) as net:
        net.step(steps)
        net.sync()
        L = S.runtime()
        L.spice_proj_get_synapses.argtypes = [C.c_void_p, C.c_int64, C.c_void_p, C.c_int32]
        got = np.zeros((len(W), 3), np.float32)
        S.check(L.spice_proj_get_synapses(net.context(), 2, got.ctypes.data_as(C.c_void_p), 12))
    for k, want in enumerate((W, Zpre, Zpost)):
        assert np.array_equal(got[:, k], want), (k, int(np.sum(got[:, k] != want)))
