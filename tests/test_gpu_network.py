"""End-to-end parity of the device engine with the reference: spike rasters and
neuron state of the Brunel networks (tests/golden/brunel_*.npz, recorded from
the reference) for every step, through the C++ user API and the C ABI."""
import hashlib
import json
import subprocess

import numpy as np
import pytest

import oracle_lib as O
import spice2_amd as S

pytestmark = pytest.mark.gpu

G = O.ROOT / "tests" / "golden"
NETS = json.loads((G / "networks.json").read_text())
FNV = 0xCBF29CE484222325


def fnv(h, arr):
    arr = np.ascontiguousarray(arr)
    return int(O.oracle().orc_fnv1a(h, O._ptr(arr), arr.nbytes))


MODES = {"pull": S.SPICE_DELIVER_PULL, "push": S.SPICE_DELIVER_PUSH}


def brunel(rec, mode=S.SPICE_DELIVER_PULL):
    N = rec["N"]
    wex = float(np.uint32(rec["w_ex_bits"]).view(np.float32))
    win = float(np.uint32(rec["w_in_bits"]).view(np.float32))
    return S.Network.brunel(N // 2, N * 4 // 10, N // 10, rec["p"], rec["rate"], wex, win,
                            seeds=tuple(rec["seeds"]), mode=mode)


def step_record(net, pops=(0, 1, 2)):
    lists = [net.spikes(p) for p in pops]
    h = FNV
    for x in lists:
        h = fnv(h, x)
    return [len(x) for x in lists], h, lists


@pytest.mark.parametrize("mode", ["pull", "push"])
@pytest.mark.parametrize("name", ["c0", "sweep0", "sweep1", "sweep2"])
def test_brunel_raster_every_step(name, mode):
    rec = NETS[name]
    z = np.load(G / f"brunel_{name}.npz")
    print("network", name, mode, flush=True)
    with brunel(rec, MODES[mode]) as net:
        for t in range(rec["steps"]):
            net.step()
            cnt, h, _ = step_record(net)
            assert cnt == z["counts"][t].tolist(), (name, t)
            assert h == int(z["raster_hash"][t]), (name, t)
            if t % 50 == 49 or t == rec["steps"] - 1:
                sh = fnv(fnv(FNV, net.neurons(1)), net.neurons(2))
                assert sh == int(z["state_hash"][t]), (name, t)
        v = np.concatenate([net.neurons(1), net.neurons(2)]).reshape(-1, 2)[:, 0]
        assert (v == z["final_V"].view(np.uint32)).all()


def test_brunel_c0_full_raster_and_events():
    rec = NETS["c0"]
    z = np.load(G / "brunel_c0.npz")
    with brunel(rec) as net:
        got = []
        for _ in range(rec["steps"]):
            net.step()
            got += [net.spikes(p) for p in range(3)]
        assert (np.concatenate(got) == z["raster"]).all()
        ev = net.stats().events
    # the engine delivers (counts) each step's spikes right after the step; the
    # reference does so d - 1 = 14 steps later, so it has delivered the spikes
    # of steps 0..299 after 314 steps
    onet, _ = O.brunel_net(10000, 8000, 2000, 0.1, 20.0, np.float32(1e-4), np.float32(-5e-4))
    for _ in range(rec["steps"] + 14):
        onet.step()
    assert ev == onet.events()


@pytest.mark.parametrize("mode", ["pull", "push"])
def test_blocks_of_steps_equal_single_steps(mode):
    """300 steps queued and run in blocks of 15 reach the same state and history
    as steps run one by one."""
    rec = NETS["c0"]
    z = np.load(G / "brunel_c0.npz")
    with brunel(rec, MODES[mode]) as net:
        net.step(300)
        for age in range(15):
            t = 299 - age
            lists = [net.spikes(p, age) for p in range(3)]
            assert [len(x) for x in lists] == z["counts"][t].tolist(), age
            h = FNV
            for x in lists:
                h = fnv(h, x)
            assert h == int(z["raster_hash"][t]), age
        with pytest.raises(S.SpicePrecondition, match="age < _spike_counts.size"):
            net.spikes(0, 15)
        sh = fnv(fnv(FNV, net.neurons(1)), net.neurons(2))
        assert sh == int(z["state_hash"][299])


def test_irregular_reads():
    rec = NETS["sweep1"]
    z = np.load(G / "brunel_sweep1.npz")
    rs = np.random.default_rng(1)
    t = 0
    with brunel(rec) as net:
        while t < rec["steps"]:
            k = int(rs.integers(1, 23))
            k = min(k, rec["steps"] - t)
            net.step(k)
            t += k
            cnt, h, _ = step_record(net)
            assert cnt == z["counts"][t - 1].tolist() and h == int(z["raster_hash"][t - 1]), t


@pytest.mark.parametrize("mode", ["pull", "push"])
def test_coba_network_delay_one(mode):
    vg = NETS["coba_small"]
    z = np.load(G / "coba_small.npz")
    with S.Network.coba(vg["nE"], vg["nI"], vg["p"], vg["w_ex"], vg["w_in"], delay=vg["delay"],
                        init=vg["init"], seeds=tuple(vg["seeds"]), mode=MODES[mode]) as net:
        for t in range(vg["steps"]):
            net.step()
            cnt = [len(net.spikes(0)), len(net.spikes(1))]
            assert cnt == z["counts"][t][1:].tolist(), t
            if t % 100 == 99:
                sh = fnv(fnv(FNV, net.neurons(0)), net.neurons(1))
                assert sh == int(z["state_hash"][t]), t


@pytest.mark.slow
def test_coba_benchmark_size_against_oracle():
    """The conductance network at its benchmark size (N=250k, 80% E / 20% I,
    p=0.004, delay 1 step; 0.27 nS and 4.5 nS over a 10 nS leak, scaled by
    80/(N p)): spike counts of every one of the first 200 steps and the neuron
    state every 50 steps equal the oracle's on the same seed, in both delivery
    modes."""
    nE, nI, p, steps = 200_000, 50_000, 0.004, 200
    scale = 80.0 / ((nE + nI) * p)
    wex, win = np.float32(0.027 * scale), np.float32(0.45 * scale)
    n = O.Net(1e-4, 1e-4, [1337])
    E = n.add(O.ORC_COBA, nE, init=1)
    I = n.add(O.ORC_COBA, nI, init=1)
    for s, d, w, f in ((E, E, wex, O.ORC_GEX), (E, I, wex, O.ORC_GEX), (I, E, win, O.ORC_GIN),
                       (I, I, win, O.ORC_GIN)):
        n.connect(s, d, p, 1e-4, w, f)
    want, want_state = [], {}
    for t in range(steps):
        n.step()
        want.append([len(n.spikes(E)), len(n.spikes(I))])
        if t % 50 == 49:
            want_state[t] = n.state_hash()
    n.close()
    assert sum(map(sum, want)) > 0
    for mode in ("pull", "push"):
        print("coba N=250k", mode, flush=True)
        with S.Network.coba(nE, nI, p, float(wex), float(win), delay=1e-4, init=1, seeds=(1337,),
                            mode=MODES[mode]) as net:
            for t in range(steps):
                net.step()
                assert [len(net.spikes(0)), len(net.spikes(1))] == want[t], (mode, t)
                if t in want_state:
                    assert fnv(fnv(FNV, net.neurons(0)), net.neurons(1)) == want_state[t], (mode, t)


def test_sample_program_output_is_byte_identical():
    """spice2_amd/samples/brunel.cpp (user code over the C++ API) prints the same
    JSON raster as the reference's samples/brunel.cpp."""
    exe = S.LIB / "brunel_sample"
    r = subprocess.run([str(exe)], capture_output=True, timeout=300)
    assert r.returncode == 0, r.stderr.decode()
    assert hashlib.md5(r.stdout).hexdigest() == "dc05acddc0e3fb6da4e490641121fbe9"


@pytest.mark.parametrize("tile", [64, 640, 4096, 32768])
def test_pull_tile_sizes(tile):
    """Pull delivery gives the reference raster for any tile size (tiles that do
    not divide the population, a single tile, tiles of 64 targets)."""
    import ctypes as C
    rec = NETS["sweep1"]
    z = np.load(G / "brunel_sweep1.npz")
    with brunel(rec) as net:
        S.check(S.runtime().spice_ctx_set_tile(net.context(), tile))
        for t in range(120):
            net.step()
            cnt, h, _ = step_record(net)
            assert cnt == z["counts"][t].tolist() and h == int(z["raster_hash"][t]), (tile, t)


@pytest.mark.parametrize("chunks,tile", [(1, 8192), (3, 640), (64, 64)])
def test_push_chunks(chunks, tile):
    """Push delivery gives the reference raster for any number of spike chunks
    per (tile, step), including chunks left without spikes (64 chunks)."""
    rec = NETS["sweep1"]
    z = np.load(G / "brunel_sweep1.npz")
    print("push chunks", chunks, "tile", tile, flush=True)
    with brunel(rec, S.SPICE_DELIVER_PUSH) as net:
        S.check(S.runtime().spice_ctx_set_tile(net.context(), tile))
        S.check(S.runtime().spice_ctx_set_push_chunks(net.context(), chunks))
        for t in range(120):
            net.step()
            cnt, h, _ = step_record(net)
            assert cnt == z["counts"][t].tolist() and h == int(z["raster_hash"][t]), (chunks, t)


def test_push_chunks_range():
    rec = NETS["sweep0"]
    with brunel(rec, S.SPICE_DELIVER_PUSH) as net:
        with pytest.raises(S.SpiceError, match="push chunks"):
            S.check(S.runtime().spice_ctx_set_push_chunks(net.context(), 65))


def test_mode_is_fixed_after_first_step():
    rec = NETS["sweep0"]
    with brunel(rec) as net:
        net.step(3)
        net.sync()
        with pytest.raises(S.SpiceError, match="before the first step"):
            S.check(S.runtime().spice_ctx_set_delivery(net.context(), S.SPICE_DELIVER_PUSH))


def test_device_errors_surface():
    with pytest.raises(S.SpicePrecondition, match="at least 1dt"):
        S.Network.brunel(100, 80, 20, 0.1, 20.0, 1e-3, -1e-3, delay=0.0)
    with pytest.raises(S.SpicePrecondition, match="maximum delay"):
        # delay above max_delay cannot be expressed through this workload; use the C ABI
        with S.Context(max_delay=2) as c:
            a = c.population(10, 0)
            b = c.population(10, 1)
            c.fixed_probability(a, b, 0.5, (1, 2), delay=3)


@pytest.mark.slow
@pytest.mark.parametrize("mode", ["pull", "push"])
def test_brunel_n500k_rasters(mode):
    """N=500k, p=0.02, 1000 steps: per-step spike counts and raster hashes, and the
    neuron state every 100 steps, against the reference."""
    if "c1" not in NETS:
        pytest.skip("N=500k golden vectors not recorded")
    rec = NETS["c1"]
    z = np.load(G / "brunel_c1.npz")
    with brunel(rec, MODES[mode]) as net:
        for t in range(rec["steps"]):
            net.step()
            cnt, h, _ = step_record(net)
            assert cnt == z["counts"][t].tolist(), t
            assert h == int(z["raster_hash"][t]), t
            if t % 100 == 99:
                sh = fnv(fnv(FNV, net.neurons(1)), net.neurons(2))
                assert sh == int(z["state_hash"][t]), t
