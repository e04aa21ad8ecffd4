"""Pins the CPU oracle (oracle/spice_oracle.c) before anything is checked against it.

Each oracle function is compared with golden vectors that the reference itself
produced (tests/golden/, written by tests/golden/make_golden.py from the
reference compiled from its own sources), with the reference's own unit-test
known answers, and with closed forms where the mathematics fixes the result.
When the reference build is present in this tree (oracle/_ref), random cases are
additionally compared live.
"""
import json
import math
import struct

import numpy as np
import pytest

import oracle_lib as O

G = O.ROOT / "tests" / "golden"
RNG = json.loads((G / "rng.json").read_text())
CSR = json.loads((G / "csr.json").read_text())
NETS = json.loads((G / "networks.json").read_text())


def h(x):
    return int(x, 16)


# ------------------------------------------------------------------ RNG

def test_seed_chains_match_reference():
    """random.h:149-152,163-167 (murmur3 seed chain)."""
    for rec in RNG["seed_chains"]:
        got = O.seed_chain(rec["seeds"], 12).ravel()
        assert [int(v) for v in got] == [h(v) for v in rec["chain"]], rec["seeds"]


def test_seed_chain_1337_known_values():
    """S0 and S8 of {1337} as recorded when the reference sample was first run."""
    ch = O.seed_chain([1337], 9)
    assert int(ch[0, 0]) == 0x6748B9668C35B07B and int(ch[0, 1]) == 0x3C011C497F4B0865
    assert int(ch[8, 0]) == 0xD64E4EF75FCD078F and int(ch[8, 1]) == 0xB3D44233FF13C9B5


def test_empty_seed_list_is_rejected():
    """random.h:151: SPICE_PRE(il.size() > 0)."""
    with pytest.raises(ValueError, match="at least 1 seed"):
        O.seed_chain([], 1)


def test_murmur3_every_tail_length():
    """random.h:32-106, all 16 tail cases and multi-block inputs."""
    L = O.oracle()
    data = np.array(RNG["murmur3_bytes"]["data"], np.uint8)
    for ln, want in enumerate(RNG["murmur3_bytes"]["by_len"]):
        r = L.orc_murmur3_bytes(O._ptr(data), ln)
        assert (r.lo, r.hi) == (h(want[0]), h(want[1])), ln


def test_murmur3_single_block_equals_bytes_form():
    """random.h:108-139 is the 16-byte special case of random.h:32-106."""
    L = O.oracle()
    rs = np.random.default_rng(5)
    for _ in range(50):
        lo, hi = (int(v) for v in rs.integers(0, 2**63, 2, dtype=np.int64))
        b = np.frombuffer(struct.pack("<QQ", lo, hi), np.uint8).copy()
        a = L.orc_murmur3_bytes(O._ptr(b), 16)
        c = L.orc_murmur3_u128(O.U128(lo, hi))
        assert (a.lo, a.hi) == (c.lo, c.hi)


def _draw(fn, seed_lohi, n, *args):
    r = O.rng_from(seed_lohi)
    return [fn(r, *args) for _ in range(n)]


def test_xoroshiro_outputs():
    """random.h:222-234."""
    L = O.oracle()
    s8 = O.seed_chain([1337], 9)[8]
    assert _draw(L.orc_rng_next, s8, 64) == [h(v) for v in RNG["xoroshiro_S8"]]
    s0 = O.seed_chain([1], 1)[0]
    assert _draw(L.orc_rng_next, s0, 64) == [h(v) for v in RNG["xoroshiro_1_S0"]]


def test_xoroshiro_closed_form_first_output():
    """Output of xoroshiro128+ is s0 + s1 of the state before the step."""
    L = O.oracle()
    r = O.Rng(1, 2)
    assert L.orc_rng_next(r) == 3
    assert L.orc_rng_next(O.Rng(2**64 - 1, 5)) == 4


def test_canonical_and_exponential_match_reference():
    L = O.oracle()
    s8 = O.seed_chain([1337], 9)[8]
    f = np.array(_draw(L.orc_canonical_f32, s8, 64), np.float32).view(np.uint32)
    assert f.tolist() == RNG["canonical_f32_S8"]
    d = np.array(_draw(L.orc_canonical_f64_leftopen, s8, 64), np.float64).view(np.uint64)
    assert [int(x) for x in d] == [h(v) for v in RNG["canonical_f64_leftopen_S8"]]
    s2 = O.seed_chain([1337], 3)[2]
    for scale, want in RNG["exponential"].items():
        got = np.array(_draw(L.orc_exponential, s2, 64, float(scale)), np.float64).view(np.uint64)
        assert [int(x) for x in got] == [h(v) for v in want], scale
    s1 = O.seed_chain([1337], 2)[1]
    u = np.array(_draw(L.orc_uniform_f32, s1, 64, -0.06, -0.05), np.float32).view(np.uint32)
    assert u.tolist() == RNG["uniform_f32_S1"]


def test_canonical_bounds_with_constant_generators():
    """test/util/random.cpp:124-157: [0,1) right-open, (0,1] left-open for the
    constant generators 0 and 2^64-1 (closed form of random.h:236-247)."""
    lo = 0
    hi = 2**64 - 1
    f_lo, f_hi = (lo >> 40) / 2**24, (hi >> 40) / 2**24
    d_lo, d_hi = ((lo >> 11) + 1) / 2**53, ((hi >> 11) + 1) / 2**53
    assert 0 <= f_lo < 1 and 0 <= f_hi < 1 and 0 < d_lo <= 1 and 0 < d_hi <= 1
    # the oracle realises these through generators whose next output is lo/hi:
    # (s0, s1) = (x, 0) gives output x.
    L = O.oracle()
    for x, fw, dw in ((lo, f_lo, d_lo), (hi, f_hi, d_hi)):
        assert L.orc_canonical_f32(O.Rng(x, 0)) == np.float32(fw)
        assert L.orc_canonical_f64_leftopen(O.Rng(x, 0)) == dw


def test_exponential_is_minus_scale_log_u():
    """random.h:270-272 restated through Python's math.log (the same libm)."""
    L = O.oracle()
    s = O.seed_chain([99], 1)[0]
    r1, r2 = O.rng_from(s), O.rng_from(s)
    for _ in range(1000):
        u = L.orc_canonical_f64_leftopen(r1)
        assert L.orc_exponential(r2, 49.0) == -49.0 * math.log(u)


def test_kahan_matches_reference_sequence():
    """numeric.h:9-15 plus the reset at >= 1 s (snn.cpp:8-10) over 25000 steps."""
    z = np.load(G / "kahan.npz")
    L = O.oracle()
    k = O.Kahan(0.0, 0.0)
    dts, sums = [], []
    for _ in range(len(z["dt_bits"])):
        d = L.orc_kahan_add(k, np.float32(1e-4))
        if k.sum >= 1:
            k.sum = 0.0
        dts.append(d)
        sums.append(k.sum)
    assert (np.array(dts, np.float32).view(np.uint32) == z["dt_bits"]).all()
    assert (np.array(sums, np.float32).view(np.uint32) == z["sum_bits"]).all()


def test_kahan_property_from_reference_tests():
    """test/util/numeric.cpp:9-29: 1000 x 0.001f sums to exactly 1 and the
    returned deltas also sum to exactly 1, while a naive float sum does not."""
    L = O.oracle()
    k = O.Kahan(0.0, 0.0)
    naive = np.float32(0)
    deltas = np.float32(0)
    for i in range(1000):
        assert abs(k.sum - i * 0.001) <= 1e-5 * max(1, i * 0.001)
        naive = np.float32(naive + np.float32(0.001))
        d = L.orc_kahan_add(k, np.float32(0.001))
        deltas = np.float32(deltas + np.float32(d))
        assert abs(d - 0.001) <= 0.001 / 100
    assert naive != 1 and np.float32(k.sum) == 1 and deltas == 1


# ------------------------------------------------------ fixed_probability

def test_fixed_probability_small_cases_in_full():
    """topology.cpp:73-112 against complete reference CSRs."""
    z = np.load(G / "csr_small.npz")
    for i, rec in enumerate(CSR["full"]):
        if "error" in rec:
            continue
        seed = O.seed_chain(rec["seeds"], rec["advance"] + 1)[rec["advance"]]
        off, nb = O.fp_generate(rec["src"], rec["dst"], rec["p"], seed)
        assert O.oracle().orc_fp_size(rec["src"], rec["dst"], rec["p"]) == rec["size"]
        assert (off == z[f"off{i}"]).all(), rec
        assert (nb == z[f"nb{i}"]).all(), rec


def test_fixed_probability_hashes():
    """Per-projection FNV-1a of the Brunel-family CSRs (and the 500k ones when recorded)."""
    for rec in CSR["hashed"]:
        if rec["src"] * rec["dst"] > 3e8:
            continue  # the N=500k projections are checked by test_fixed_probability_big
        seed = O.seed_chain(rec["seeds"], rec["advance"] + 1)[rec["advance"]]
        off, nb = O.fp_generate(rec["src"], rec["dst"], rec["p"], seed)
        assert len(nb) == rec["count"], rec["case"]
        assert f"{O.csr_hash(off, nb):016x}" == rec["fnv"], rec["case"]
        assert nb[: off[1]][:8].tolist() == [v for v in rec["row0"] if v >= 0], rec["case"]


def test_fixed_probability_c0_survey_hashes():
    """The six config-0 projections, hashes recorded from the reference sample."""
    want = ["12d319e63b90e715", "1341a36fe801da59", "c164bd9c2123f634", "646a02217199087e",
            "843d6282c6643c86", "cd70cd601474402f"]
    net, _ = O.brunel_net(10000, 8000, 2000, 0.1, 20.0, np.float32(1e-4), np.float32(-5e-4))
    assert [f"{net.proj_hash(c):016x}" for c in range(6)] == want
    assert sum(net.proj_synapses(c) for c in range(6)) == 19996472


def test_fixed_probability_invariants():
    """Rows strictly ascending, targets in range, degree <= max_degree, mean degree ~ D p."""
    L = O.oracle()
    for src, dst, p in ((500, 3000, 0.05), (200, 50, 0.7), (1000, 100000, 0.001)):
        off, nb = O.fp_generate(src, dst, p, (12345, 678))
        md = L.orc_fp_max_degree(dst, p)
        deg = np.diff(off)
        assert (deg <= md).all() and (deg >= 0).all()
        assert nb.min(initial=0) >= 0 and nb.max(initial=0) < dst
        for r in range(src):
            row = nb[off[r]:off[r + 1]]
            assert (np.diff(row) > 0).all()
        mean = deg.mean()
        assert abs(mean - dst * p) < 5 * math.sqrt(dst * p * (1 - p) / src) + 1


def test_fixed_probability_edges():
    """p == 0 and dst == 0 give empty rows (topology.cpp:84-85); p == 1 is complete."""
    off, nb = O.fp_generate(7, 5, 0.0, (1, 2))
    assert (off == 0).all() and len(nb) == 0
    off, nb = O.fp_generate(5, 0, 0.5, (1, 2))
    assert (off == 0).all() and len(nb) == 0
    off, nb = O.fp_generate(4, 6, 1.0, (1, 2))
    assert (np.diff(off) == 6).all() and (nb.reshape(4, 6) == np.arange(6)).all()


def test_fixed_probability_second_restatement_tiny():
    """A second, independent restatement in Python (geometric skips from the same
    stream, std::round = half away from zero) agrees on tiny sizes.  The pin is
    the reference's own output above; this guards the C code against typos in
    paths the golden cases might not reach."""
    L = O.oracle()
    for src, dst, p, seed in ((3, 9, 0.3, (5, 6)), (4, 20, 0.5, (7, 8)), (2, 40, 0.1, (9, 1))):
        off, nb = O.fp_generate(src, dst, p, seed)
        r = O.Rng(*seed)
        scale = 1 / p - 1
        md = int(dst * p + 3 * math.sqrt(dst * p * (1 - p)))
        rows = []
        for _ in range(src):
            row, noise = [], 0.0
            while True:
                u = ((L.orc_rng_next(r) >> 11) + 1) / 2**53
                noise += -scale * math.log(u)
                fl = math.floor(noise)
                d = len(row) + int(fl) + (1 if noise - fl >= 0.5 else 0)
                if d >= dst or len(row) >= md:
                    break
                row.append(d)
            rows.append(row)
        assert np.diff(off).tolist() == [len(x) for x in rows]
        assert nb.tolist() == [v for x in rows for v in x]


# ----------------------------------------------------------------- network

def _brunel(rec, steps=None):
    N = rec["N"]
    wex = np.uint32(rec["w_ex_bits"]).view(np.float32)
    win = np.uint32(rec["w_in_bits"]).view(np.float32)
    return O.brunel_net(N // 2, N * 4 // 10, N // 10, rec["p"], rec["rate"], wex, win,
                        seeds=rec["seeds"])


@pytest.mark.parametrize("name", ["c0", "sweep0", "sweep1", "sweep2"])
def test_brunel_rasters_match_reference(name):
    """snn.cpp:7-28 end to end: per-step spike lists and neuron-state hashes."""
    rec = NETS[name]
    z = np.load(G / f"brunel_{name}.npz")
    net, pops = _brunel(rec)
    for t in range(rec["steps"]):
        net.step()
        lists = [net.spikes(p) for p in pops]
        assert [len(x) for x in lists] == z["counts"][t].tolist(), (name, t)
        hv = 0xCBF29CE484222325
        for x in lists:
            hv = O.oracle().orc_fnv1a(hv, O._ptr(x), 4 * len(x))
        assert hv == int(z["raster_hash"][t]), (name, t)
        assert net.state_hash() == int(z["state_hash"][t]), (name, t)
    v = np.concatenate([net.field(pops[1]), net.field(pops[2])])
    assert (v.view(np.uint32) == z["final_V"].view(np.uint32)).all()


def test_brunel_c0_full_raster():
    z = np.load(G / "brunel_c0.npz")
    assert len(z["raster"]) == 14389
    assert (z["raster"] == z["fast_raster"]).all()  # the release (fast-math) build agrees
    rec = NETS["c0"]
    net, pops = _brunel(rec)
    got = []
    for _ in range(rec["steps"]):
        net.step()
        got += [net.spikes(p) for p in pops]
    assert (np.concatenate(got) == z["raster"]).all()


def test_coba_small_matches_reference():
    """vogels.cpp:18-45 neuron, per-neuron init, delay of one step."""
    vg = NETS["coba_small"]
    z = np.load(G / "coba_small.npz")
    n = O.Net(1e-4, vg["delay"], vg["seeds"])
    E = n.add(O.ORC_COBA, vg["nE"], init=vg["init"])
    I = n.add(O.ORC_COBA, vg["nI"], init=vg["init"])
    wex, win = np.float32(vg["w_ex"]), np.float32(vg["w_in"])
    for s, d, w, f in ((E, E, wex, O.ORC_GEX), (E, I, wex, O.ORC_GEX), (I, E, win, O.ORC_GIN),
                       (I, I, win, O.ORC_GIN)):
        n.connect(s, d, vg["p"], vg["delay"], w, f)
    for t in range(vg["steps"]):
        n.step()
        assert [len(n.spikes(E)), len(n.spikes(I))] == z["counts"][t][1:].tolist(), t
        assert n.state_hash() == int(z["state_hash"][t]), t


def test_connect_preconditions():
    """snn.h:34-38 and the empty-source precondition of topology.cpp:106."""
    n = O.Net(1e-4, 15e-4, [1])
    P = n.add(O.ORC_POISSON, 10, 20.0)
    E = n.add(O.ORC_LIF, 10)
    with pytest.raises(ValueError, match="at least 1dt"):
        n.connect(P, E, 0.1, 0.0, 1.0)
    with pytest.raises(ValueError, match="_max_delay"):
        n.connect(P, E, 0.1, 16e-4, 1.0)
    Z = n.add(O.ORC_LIF, 0)
    with pytest.raises(ValueError, match="offsets.size"):
        n.connect(Z, E, 0.1, 1e-4, 1.0)


# ------------------------------------------------- live, against oracle/_ref

ref = pytest.mark.skipif(not O.ref_available(), reason="reference build absent (oracle/_ref)")


@ref
def test_random_csrs_against_reference_build():
    R = O.ref()
    rs = np.random.default_rng(2024)
    for _ in range(25):
        src, dst = int(rs.integers(1, 400)), int(rs.integers(1, 5000))
        p = float(rs.choice([0.001, 0.01, 0.05, 0.1, 0.3, 0.5, 0.9, 1.0]))
        sd = [int(x) for x in rs.integers(0, 2**32, int(rs.integers(1, 4)))]
        adv = int(rs.integers(0, 10))
        s, n = O._seeds(sd)
        hv = np.zeros(1, np.uint64)
        cnt = np.zeros(1, np.int64)
        row0 = np.zeros(4, np.int32)
        assert R.ref_fp_hash(src, dst, p, s, n, adv, O._ptr(hv), O._ptr(cnt), O._ptr(row0), 4) == 0
        seed = O.seed_chain(sd, adv + 1)[adv]
        off, nb = O.fp_generate(src, dst, p, seed)
        assert O.csr_hash(off, nb) == int(hv[0]) and len(nb) == cnt[0]


@ref
def test_random_brunel_networks_against_reference_build():
    R = O.ref(strict=True)
    rs = np.random.default_rng(7)
    for _ in range(4):
        N = int(rs.integers(500, 6000))
        p = float(rs.choice([0.05, 0.1, 0.2]))
        rate = float(rs.uniform(10, 40))
        wex, win = np.float32(3.0 / N), np.float32(-12.0 / N)
        steps = 120
        sd = [int(rs.integers(0, 2**31))]
        s, n = O._seeds(sd)
        cnt = np.zeros(steps * 3, np.int32)
        rh = np.zeros(steps, np.uint64)
        sh = np.zeros(steps, np.uint64)
        used = np.zeros(1, np.int64)
        assert R.ref_brunel_run(N // 2, N * 4 // 10, N // 10, p, rate, wex, win, np.float32(1e-4),
                                np.float32(15e-4), s, n, steps, O._ptr(cnt), None, 0, O._ptr(used),
                                O._ptr(rh), O._ptr(sh), None) == 0
        net, pops = O.brunel_net(N // 2, N * 4 // 10, N // 10, p, rate, wex, win, seeds=sd)
        for t in range(steps):
            net.step()
            assert [len(net.spikes(q)) for q in pops] == cnt.reshape(-1, 3)[t].tolist()
            assert net.state_hash() == int(sh[t])


# ---------------------------------------------------------------- plastic path

def _kat_syn():
    """The projection of test/detail/synapse_population.cpp:30-44: 3 sources, 5
    targets, edges (0,0) (0,1) (0,3) (1,3) (2,4)."""
    off = np.array([0, 3, 4, 5], np.int64)
    nb = np.array([0, 1, 3, 3, 4], np.int32)
    L = O.oracle()
    L.orc_syn_create.restype = O.C.c_void_p
    L.orc_syn_create.argtypes = [O.C.c_int64, O.C.c_int64, O._vp, O._vp, O.C.c_int, O.C.c_float]
    L.orc_syn_update.argtypes = [O._vp, O.C.c_int64, O.C.c_float, O._vp]
    L.orc_syn_deliver.argtypes = [O._vp, O.C.c_int64, O.C.c_float, O._vp, O.C.c_int64, O._vp, O._vp]
    L.orc_syn_destroy.argtypes = [O._vp]
    return L, L.orc_syn_create(3, 5, O._ptr(off), O._ptr(nb), O.ORC_RULE_COUNT, 0.0)


@pytest.mark.parametrize("ops,expected", [
    ([("d", 0)], 1),                          # deliver
    ([("u", 0), ("d", 0)], 1),                # update then deliver in one step: one update
    ([("u", 0), ("u", 0), ("d", 0)], 1),      # repeated updates in one step: one update
    ([("u", 0), ("d", 1)], 2),                # update and deliver in different steps: two
    ([("d", 9)], 10),                         # skip ahead to step 10: ten updates
    ([("u", 4), ("d", 9)], 10),
])
def test_plastic_catch_up_matches_reference_unit_tests(ops, expected):
    """test/detail/synapse_population.cpp:96-167 (DeliverPlastic): after the
    operations, targets 3 and 4 hold the update count of their synapses."""
    L, s = _kat_syn()
    hist = np.zeros(5, np.uint64)
    dst = np.zeros(5, np.float32)
    spikes = np.array([1, 2], np.int32)
    try:
        for op, t in ops:
            if op == "u":
                assert L.orc_syn_update(s, t, 1.0, O._ptr(hist)) == 0
            else:
                assert L.orc_syn_deliver(s, t, 1.0, O._ptr(spikes), 2, O._ptr(hist), O._ptr(dst)) == 0
    finally:
        L.orc_syn_destroy(s)
    assert dst[3] == expected and dst[4] == expected


@pytest.mark.parametrize("name", ["brunelplus_c0", "stdp_small"])
def test_plastic_network_matches_reference(name):
    """Brunel+ with the sample's plastic E->E synapse (N=20000, first 150 steps)
    and with the pair-STDP rule (N=4000, 600 steps): spike counts, raster hashes
    and neuron-state hashes of every step equal the reference's."""
    rec = NETS[name]
    z = np.load(G / f"{name}.npz")
    N, p = rec["N"], rec["p"]
    w0 = float(np.uint32(rec["w0_bits"]).view(np.float32))
    steps = min(rec["steps"], 150 if N > 10000 else rec["steps"])
    net, _ = O.brunel_plastic_net(N // 2, N * 4 // 10, N // 10, p, 20.0, np.float32(2.0 / N),
                                  np.float32(-10.0 / N), rec["rule"], w0)
    try:
        for t in range(steps):
            net.step()
            lists = [net.spikes(k) for k in range(3)]
            assert [len(x) for x in lists] == z["counts"][t].tolist(), t
            h = 0xCBF29CE484222325
            for x in lists:
                h = int(O.oracle().orc_fnv1a(h, O._ptr(np.ascontiguousarray(x)), x.nbytes))
            assert h == int(z["raster_hash"][t]), t
            assert net.state_hash() == int(z["state_hash"][t]), t
    finally:
        net.close()
