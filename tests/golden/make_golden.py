"""Writes the golden vectors under tests/golden/ from the reference itself.

The reference (denniskb/spice2) is compiled from its own sources by
oracle/Makefile into oracle/_ref/ in two builds: with its release flags
(libspice_ref.so: -O2 -ffast-math family -march=haswell, asserts on) and with
strict IEEE semantics (libspice_ref_strict.so).  This script calls both through
oracle/ref_driver.cpp and stores inputs and outputs as small data files.

Run from the repo root after `make -C oracle`:  python tests/golden/make_golden.py [--big]
--big also writes the N=500k Brunel and N=250k conductance-network vectors
(about 3 minutes and 11 GB of RAM).
"""
from __future__ import annotations

import argparse
import json
import sys
import time
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent))
import oracle_lib as O  # noqa: E402

P = O._ptr


def hexs(a):
    return [f"{int(x):016x}" for x in np.asarray(a, np.uint64).ravel()]


def seeds_arg(seeds):
    return O._seeds(list(seeds))


def rng_vectors(R):
    out = {}
    lists = [[1337], [1], [0], [0xFFFFFFFF], [1337, 5], [1, 2, 3], [7, 8, 9, 10, 11],
             [1, 2, 3, 4, 5, 6]]
    out["seed_chains"] = []
    for sd in lists:
        s, n = seeds_arg(sd)
        o = np.zeros(2 * 12, np.uint64)
        assert R.ref_seed_chain(s, n, 12, P(o)) == 0
        out["seed_chains"].append({"seeds": sd, "chain": hexs(o)})
    # murmur3 over every tail length 0..40 of a fixed byte pattern
    data = np.array([(i * 37 + 11) & 0xFF for i in range(48)], np.uint8)
    mm = []
    for ln in range(0, 41):
        o = np.zeros(2, np.uint64)
        R.ref_murmur3(P(data), ln, P(o))
        mm.append(hexs(o))
    out["murmur3_bytes"] = {"data": data.tolist(), "by_len": mm}
    s, n = seeds_arg([1337])
    x = np.zeros(64, np.uint64)
    R.ref_xoroshiro(s, n, 8, 64, P(x))
    out["xoroshiro_S8"] = hexs(x)
    s1, n1 = seeds_arg([1])
    R.ref_xoroshiro(s1, n1, 0, 64, P(x))
    out["xoroshiro_1_S0"] = hexs(x)
    f = np.zeros(64, np.float32)
    R.ref_canonical_f32(s, n, 8, 64, P(f))
    out["canonical_f32_S8"] = [int(v) for v in f.view(np.uint32)]
    d = np.zeros(64, np.float64)
    R.ref_canonical_f64_leftopen(s, n, 8, 64, P(d))
    out["canonical_f64_leftopen_S8"] = hexs(d.view(np.uint64))
    out["exponential"] = {}
    for scale in (9.0, 49.0, 249.0, 0.2):
        R.ref_exponential(s, n, 2, scale, 64, P(d))
        out["exponential"][repr(scale)] = hexs(d.view(np.uint64))
    R.ref_uniform_f32(s, n, 1, np.float32(-0.06), np.float32(-0.05), 64, P(f))
    out["uniform_f32_S1"] = [int(v) for v in f.view(np.uint32)]
    return out


def kahan_vectors(R):
    n = 25000
    dt = np.zeros(n, np.float32)
    sm = np.zeros(n, np.float32)
    R.ref_kahan(np.float32(1e-4), n, P(dt), P(sm))
    return {"dt_bits": dt.view(np.uint32), "sum_bits": sm.view(np.uint32)}


def fp_case(R, src, dst, p, seeds, adv, full):
    s, n = seeds_arg(seeds)
    size = R.ref_fp_size(src, dst, p)
    rec = {"src": src, "dst": dst, "p": p, "seeds": list(seeds), "advance": adv, "size": int(size)}
    if full:
        off = np.zeros(max(src + 1, 1), np.int64)
        nb = np.zeros(max(size, 1), np.int32)
        cnt = R.ref_fp_generate(src, dst, p, s, n, adv, P(off), P(nb), max(size, 1))
        if cnt < 0:
            rec["error"] = R.ref_last_error().decode()
            return rec, None
        rec["count"] = int(cnt)
        return rec, (off[: src + 1].copy(), nb[:cnt].copy())
    h = np.zeros(1, np.uint64)
    c = np.zeros(1, np.int64)
    row0 = np.zeros(8, np.int32)
    r = R.ref_fp_hash(src, dst, p, s, n, adv, P(h), P(c), P(row0), 8)
    if r < 0:
        rec["error"] = R.ref_last_error().decode()
        return rec, None
    rec.update(count=int(c[0]), fnv=f"{int(h[0]):016x}", row0=row0.tolist())
    return rec, None


BRUNEL_PROJ = [(0, 1), (0, 2), (1, 1), (1, 2), (2, 1), (2, 2)]


def brunel_cases():
    """(name, N, p, rate, w_ex, w_in, steps, seeds) of the Brunel family."""
    N0 = 20000
    cases = [("c0", N0, 0.1, 20.0, 2.0 / N0, -10.0 / N0, 300, [1337])]
    # small instances of the swept family: N, p = min(0.1, 2000/N), w = 0.2/(N p), -1/(N p)
    for i, (N, rate) in enumerate([(5000, 10.0), (12000, 17.5), (31000, 29.0)]):
        p = min(0.1, 2000.0 / N)
        cases.append((f"sweep{i}", N, p, rate, 0.2 / (N * p), -1.0 / (N * p), 400, [1337, i]))
    return cases


def run_brunel(R, N, p, rate, wex, win, steps, seeds, with_raster):
    nP, nE, nI = N // 2, N * 4 // 10, N // 10
    s, n = seeds_arg(seeds)
    cnt = np.zeros(steps * 3, np.int32)
    cap = 4_000_000 if with_raster else 0
    ras = np.zeros(max(cap, 1), np.int32)
    used = np.zeros(1, np.int64)
    rh = np.zeros(steps, np.uint64)
    sh = np.zeros(steps, np.uint64)
    fv = np.zeros(nE + nI, np.float32)
    r = R.ref_brunel_run(nP, nE, nI, p, rate, np.float32(wex), np.float32(win), np.float32(1e-4),
                         np.float32(15e-4), s, n, steps, P(cnt), P(ras) if with_raster else None,
                         cap, P(used), P(rh), P(sh), P(fv))
    assert r == 0, R.ref_last_error()
    out = {"counts": cnt.reshape(steps, 3), "raster_hash": rh, "state_hash": sh, "final_V": fv}
    if with_raster:
        assert used[0] <= cap
        out["raster"] = ras[: used[0]].copy()
    return out


def run_vogels(R, nE, nI, p, wex, win, delay, steps, seeds, init):
    s, n = seeds_arg(seeds)
    cnt = np.zeros(steps * 3, np.int32)
    used = np.zeros(1, np.int64)
    rh = np.zeros(steps, np.uint64)
    sh = np.zeros(steps, np.uint64)
    r = R.ref_vogels_run(nE, nI, p, np.float32(wex), np.float32(win), np.float32(1e-4),
                         np.float32(delay), init, s, n, steps, P(cnt), None, 0, P(used), P(rh), P(sh))
    assert r == 0, R.ref_last_error()
    return {"counts": cnt.reshape(steps, 3), "raster_hash": rh, "state_hash": sh}


def run_plastic(R, N, p, rule, w0, steps, seeds):
    """Brunel+ (samples/brunel+.cpp:101-117): E->E plastic with `rule`."""
    nP, nE, nI = N // 2, N * 4 // 10, N // 10
    s, n = seeds_arg(seeds)
    cnt = np.zeros(steps * 3, np.int32)
    used = np.zeros(1, np.int64)
    rh = np.zeros(steps, np.uint64)
    sh = np.zeros(steps, np.uint64)
    r = R.ref_brunel_plastic_run(nP, nE, nI, p, 20.0, np.float32(2.0 / N), np.float32(-10.0 / N),
                                 np.float32(1e-4), np.float32(15e-4), s, n, rule, np.float32(w0), steps,
                                 P(cnt), None, 0, P(used), P(rh), P(sh))
    assert r == 0, R.ref_last_error()
    return {"counts": cnt.reshape(steps, 3), "raster_hash": rh, "state_hash": sh}


PLASTIC_CASES = [  # name, N, p, rule, w0, steps
    ("brunelplus_c0", 20000, 0.1, 1, 0.0, 300),
    ("stdp_small", 4000, 0.1, 2, 2.0 / 4000, 600),
]


def plastic_fixtures(Rs, Rf):
    """Adds the plastic-synapse fixtures to networks.json (other entries kept)."""
    nets = json.loads((HERE / "networks.json").read_text())
    for name, N, p, rule, w0, steps in PLASTIC_CASES:
        b = run_plastic(Rs, N, p, rule, w0, steps, [1337])
        a = run_plastic(Rf, N, p, rule, w0, steps, [1337])
        np.savez_compressed(HERE / f"{name}.npz", counts=b["counts"], raster_hash=b["raster_hash"],
                            state_hash=b["state_hash"], fast_counts=a["counts"])
        nets[name] = {"N": N, "p": p, "rule": rule, "w0_bits": int(np.float32(w0).view(np.uint32)),
                      "steps": steps, "seeds": [1337],
                      "fast_equals_strict_raster": bool((a["raster_hash"] == b["raster_hash"]).all())}
        print(name, b["counts"].sum(0), nets[name]["fast_equals_strict_raster"])
    (HERE / "networks.json").write_text(json.dumps(nets, indent=1))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--big", action="store_true")
    ap.add_argument("--plastic-only", action="store_true",
                    help="only (re)write the plastic-synapse fixtures")
    args = ap.parse_args()
    if args.plastic_only:
        plastic_fixtures(O.ref(strict=True), O.ref(strict=False))
        return
    Rf, Rs = O.ref(False), O.ref(True)
    t0 = time.time()

    rv = rng_vectors(Rf)
    assert rng_vectors(Rs) == rv
    (HERE / "rng.json").write_text(json.dumps(rv, indent=1))
    kv = kahan_vectors(Rf)
    np.savez_compressed(HERE / "kahan.npz", **kv)
    print("rng/kahan", time.time() - t0)

    # fixed_probability: small cases in full, edge cases, and hashes of larger ones
    small = [(50, 40, 0.1, [1337], 2), (30, 1000, 0.5, [1337], 3), (10, 10, 1.0, [7], 0),
             (100, 3, 0.9, [7], 1), (1, 100000, 0.001, [9], 0), (64, 64, 0.25, [1, 2], 5),
             (300, 200, 0.02, [1337], 4), (7, 5, 0.0, [1], 0), (5, 0, 0.5, [1], 0),
             (2000, 2000, 0.05, [42], 1), (1, 1, 1.0, [3], 0), (3, 1, 0.5, [3], 0)]
    recs, arrays = [], {}
    for i, (src, dst, p, sd, adv) in enumerate(small):
        rec, csr = fp_case(Rf, src, dst, p, sd, adv, True)
        rec2, csr2 = fp_case(Rs, src, dst, p, sd, adv, True)
        assert rec == rec2 and (csr is None) == (csr2 is None)
        if csr is not None:
            assert (csr[0] == csr2[0]).all() and (csr[1] == csr2[1]).all()
            arrays[f"off{i}"], arrays[f"nb{i}"] = csr
        recs.append(rec)
    # empty source population: the reference's precondition fires
    rec, _ = fp_case(Rf, 0, 10, 0.1, [1], 0, True)
    recs.append(rec)
    np.savez_compressed(HERE / "csr_small.npz", **arrays)

    hashed = []
    for name, N, p, *_ in brunel_cases():
        sizes = (N // 2, N * 4 // 10, N // 10)
        seeds = [1337] if name == "c0" else [1337, int(name[-1])]
        for k, (a, b) in enumerate(BRUNEL_PROJ):
            rec, _ = fp_case(Rf, sizes[a], sizes[b], p, seeds, 2 + k, False)
            rec["case"] = f"{name}.{k}"
            hashed.append(rec)
    print("csr", time.time() - t0)

    # Brunel network rasters (both builds)
    nets = {}
    for name, N, p, rate, wex, win, steps, seeds in brunel_cases():
        full = name == "c0"
        a = run_brunel(Rf, N, p, rate, wex, win, steps, seeds, full)
        b = run_brunel(Rs, N, p, rate, wex, win, steps, seeds, full)
        same_raster = bool((a["raster_hash"] == b["raster_hash"]).all())
        print(name, "fast==strict raster:", same_raster, "spikes", a["counts"].sum(0))
        rec = {"counts": b["counts"], "raster_hash": b["raster_hash"],
               "state_hash": b["state_hash"], "final_V": b["final_V"],
               "fast_counts": a["counts"], "fast_raster_hash": a["raster_hash"]}
        if full:
            rec["raster"] = b["raster"]
            rec["fast_raster"] = a["raster"]
        np.savez_compressed(HERE / f"brunel_{name}.npz", **rec)
        nets[name] = {"N": N, "p": p, "rate": rate, "w_ex_bits": int(np.float32(wex).view(np.uint32)),
                      "w_in_bits": int(np.float32(win).view(np.uint32)), "steps": steps,
                      "seeds": seeds, "fast_equals_strict_raster": same_raster}
    print("brunel", time.time() - t0)

    # small conductance network (delay 1 step, V0 ~ U[-60,-50] mV)
    vg = {"nE": 3200, "nI": 800, "p": 0.02, "w_ex": 0.027 * 80 / (4000 * 0.02),
          "w_in": 0.45 * 80 / (4000 * 0.02), "delay": 1e-4, "steps": 1000, "seeds": [1337], "init": 1}
    a = run_vogels(Rf, vg["nE"], vg["nI"], vg["p"], vg["w_ex"], vg["w_in"], vg["delay"], vg["steps"],
                   vg["seeds"], vg["init"])
    b = run_vogels(Rs, vg["nE"], vg["nI"], vg["p"], vg["w_ex"], vg["w_in"], vg["delay"], vg["steps"],
                   vg["seeds"], vg["init"])
    vg["fast_equals_strict_raster"] = bool((a["raster_hash"] == b["raster_hash"]).all())
    np.savez_compressed(HERE / "coba_small.npz", counts=b["counts"], raster_hash=b["raster_hash"],
                        state_hash=b["state_hash"], fast_counts=a["counts"])
    nets["coba_small"] = vg
    print("coba", time.time() - t0, vg["fast_equals_strict_raster"], b["counts"].sum(0))

    if args.big:
        N, p = 500000, 0.02
        sizes = (N // 2, N * 4 // 10, N // 10)
        for k, (a_, b_) in enumerate(BRUNEL_PROJ):
            rec, _ = fp_case(Rf, sizes[a_], sizes[b_], p, [1337], 2 + k, False)
            rec["case"] = f"c1.{k}"
            hashed.append(rec)
            print("c1 csr", k, time.time() - t0)
        wex, win = 0.2 / (N * p), -1.0 / (N * p)
        b = run_brunel(Rs, N, p, 20.0, wex, win, 1000, [1337], False)
        np.savez_compressed(HERE / "brunel_c1.npz", counts=b["counts"], raster_hash=b["raster_hash"],
                            state_hash=b["state_hash"])
        nets["c1"] = {"N": N, "p": p, "rate": 20.0, "w_ex_bits": int(np.float32(wex).view(np.uint32)),
                      "w_in_bits": int(np.float32(win).view(np.uint32)), "steps": 1000, "seeds": [1337]}
        print("c1 run", time.time() - t0, b["counts"].sum(0))
        nE, nI, p2 = 200000, 50000, 0.004
        for k, (a_, b_) in enumerate([(0, 0), (0, 1), (1, 0), (1, 1)]):
            sz = (nE, nI)
            rec, _ = fp_case(Rf, sz[a_], sz[b_], p2, [1337], 2 + k, False)
            rec["case"] = f"c2.{k}"
            hashed.append(rec)

    (HERE / "csr.json").write_text(json.dumps({"full": recs, "hashed": hashed}, indent=1))
    (HERE / "networks.json").write_text(json.dumps(nets, indent=1))
    print("done", time.time() - t0)


if __name__ == "__main__":
    main()
