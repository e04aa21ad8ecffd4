"""Benchmark: synaptic events/s of the Brunel N=500k network on one MI355X.

Workload (one GPU): Brunel with N = 500000 (250k Poisson at 20 Hz, 200k
excitatory and 50k inhibitory LIF), all six projections fixed_probability
p = 0.02 (about 2.5e9 synapses, int32 targets), weights +0.2/(N p) and
-1.0/(N p), dt = 1e-4 s, delay 15 steps, seed {1337}.  A step is one simulation
step (neuron updates of all populations plus delivery of the spikes due).  The
connectivity is generated on the device before the timed region (reported as
csr_generation_s).

`value` = synaptic events delivered in the K timed steps / wall time of those
steps (inputs resident in HBM).  `roofline` is for the delivery kernel, timed
with HIP events on the context stream: algorithmic bytes = 4 B per synaptic
event (int32 target) + 8 B per delivered spike (row offsets), over 8 TB/s.
`cpu_baseline` times the CPU oracle (single core) on the same network on a
bounded sample of steps.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--mode pull|push] [--no-cpu]
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import subprocess
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
N, P_CONN = 500_000, 0.02
NP, NE, NI = N // 2, N * 4 // 10, N // 10
W_EX, W_IN = 0.2 / (N * P_CONN), -1.0 / (N * P_CONN)


def cpu_model() -> str:
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def traffic_from_profiles():
    """HBM bytes per delivery launch from the committed rocprofv3 PMC summary of
    this workload (profiles/deliver_traffic.json), if present."""
    f = ROOT / "profiles" / "deliver_traffic.json"
    if not f.exists():
        return None
    try:
        return json.loads(f.read_text()).get("bytes_per_launch")
    except (OSError, ValueError):
        return None


COPY_BYTES = 1 << 30  # bytes read (and written) by one calibration copy


def copy_calibration(reps: int = 5):
    """A plain device copy of COPY_BYTES, `reps` times: the known-size kernel that
    scales the FETCH_SIZE / WRITE_SIZE counters of a rocprofv3 --pmc run."""
    import torch
    x = torch.empty(COPY_BYTES // 4, dtype=torch.float32, device="cuda").fill_(1.0)
    y = torch.empty_like(x)
    for _ in range(reps):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": COPY_BYTES, "reps": reps}), flush=True)


def _pmc_rows(folder: Path, counter: str):
    """(kernel name, value) per dispatch from a rocprofv3 --pmc output folder
    (rocpd SQLite databases, or counter_collection CSVs)."""
    import csv
    import sqlite3
    rows = []
    for f in sorted(Path(folder).rglob("*results.db")):
        con = sqlite3.connect(str(f))
        try:
            for k, n, v in con.execute("select kernel_name, counter_name, value from counters_collection"):
                if n == counter:
                    rows.append((k, float(v)))
        finally:
            con.close()
    for f in sorted(Path(folder).rglob("*counter_collection.csv")):
        with open(f, newline="") as fh:
            for r in csv.DictReader(fh):
                if r.get("Counter_Name") == counter:
                    rows.append((r.get("Kernel_Name", ""), float(r.get("Counter_Value", "nan"))))
    return rows


def traffic_summary(fetch_run: Path, write_run: Path, calib_fetch: Path, calib_write: Path, out: Path):
    """Average HBM bytes per delivery launch from two rocprofv3 --pmc runs of this
    benchmark (FETCH_SIZE, then WRITE_SIZE; KiB per dispatch), each scaled by what
    the same counter reads for the calibration copy of COPY_BYTES."""
    def avg(folder, counter, match):
        v = [x for k, x in _pmc_rows(folder, counter) if match in k]
        return (sum(v) / len(v), len(v)) if v else (None, 0)
    cf, _ = avg(calib_fetch, "FETCH_SIZE", "copy")
    cw, _ = avg(calib_write, "WRITE_SIZE", "copy")
    f, nf = avg(fetch_run, "FETCH_SIZE", "deliver_pull_kernel")
    w, nw = avg(write_run, "WRITE_SIZE", "deliver_pull_kernel")
    kib = COPY_BYTES / 1024.0
    sf = kib / cf if cf else None
    sw = kib / cw if cw else None
    res = {"kernel": "deliver_pull_kernel", "dispatches": [nf, nw],
           "fetch_kib_raw": f, "write_kib_raw": w,
           "copy_fetch_kib_raw": cf, "copy_write_kib_raw": cw, "copy_kib_expected": kib,
           "fetch_scale": sf, "write_scale": sw,
           "fetch_bytes": f * 1024 * sf if f and sf else None,
           "write_bytes": w * 1024 * sw if w and sw else None}
    res["bytes_per_launch"] = (res["fetch_bytes"] + res["write_bytes"]
                               if res["fetch_bytes"] and res["write_bytes"] else None)
    Path(out).write_text(json.dumps(res, indent=1) + "\n")
    print(json.dumps(res), flush=True)


def cpu_baseline(net, steps_warm: int, steps_timed: int):
    """The CPU oracle (one core) on the same network, using the device-generated
    CSR (downloaded, checked identical in the tests), timed over a bounded sample."""
    sys.path.insert(0, str(ROOT / "tests"))
    import oracle_lib as O
    import spice2_amd as S

    L = S.runtime()
    ctx = net.context()
    onet = O.Net(1e-4, 15e-4, [1337])
    P = onet.add(O.ORC_POISSON, NP, 20.0)
    E = onet.add(O.ORC_LIF, NE)
    I = onet.add(O.ORC_LIF, NI)
    sizes = {P: NP, E: NE, I: NI}
    projs = [(P, E, W_EX), (P, I, W_EX), (E, E, W_EX), (E, I, W_EX), (I, E, W_IN), (I, I, W_IN)]
    import ctypes as C
    for k, (s, d, w) in enumerate(projs):
        n = C.c_int64()
        S.check(L.spice_proj_synapses(ctx, k, C.byref(n)))
        off = np.zeros(sizes[s] + 1, np.int64)
        nb = np.zeros(max(n.value, 1), np.int32)
        S.check(L.spice_proj_download(ctx, k, off.ctypes.data_as(C.c_void_p), nb.ctypes.data_as(C.c_void_p)))
        onet.connect_csr(s, d, off, nb, 15e-4, np.float32(w))
    for _ in range(steps_warm):
        onet.step()
    e0 = onet.events()
    t0 = time.perf_counter()
    for _ in range(steps_timed):
        onet.step()
    dt = time.perf_counter() - t0
    ev = onet.events() - e0
    onet.close()
    return {"value": ev / dt, "unit": "synaptic events/s", "cores": 1, "kind": "port",
            "sample": f"Brunel N=500k p=0.02, steps {steps_warm}..{steps_warm + steps_timed - 1} "
                      f"({steps_timed} steps, {ev} events, {dt:.2f} s) on one core of "
                      f"'{cpu_model()}' (nproc {os.cpu_count()})",
            "ms_per_step": 1e3 * dt / steps_timed}


def sweep_instances(n: int, seed: int = 2024):
    """The batched sweep: N log-uniform in [5000, 100000], p = min(0.1, 2000/N),
    Poisson rate uniform in [10, 30] Hz, seeds {1337, i} (numpy PCG64(seed))."""
    rng = np.random.Generator(np.random.PCG64(seed))
    N = np.round(np.exp(rng.uniform(np.log(5e3), np.log(1e5), n))).astype(np.int64)
    p = np.minimum(0.1, 2000.0 / N)
    rate = rng.uniform(10.0, 30.0, n)
    return N, p, rate, [(1337, i) for i in range(n)]


def run_batch(n: int, steps: int, window: int, plain_sample: int):
    """`n` Brunel instances in one batch (one context; windows of `window` steps
    replayed from HIP graphs; pull tiles autotuned per target-size bucket) against
    a stratified sample of the same instances run one after another through the
    plain single-network path.  Returns the batched and plain aggregate synaptic
    events/s and whether the sampled instances' final spike histories agree."""
    import spice2_amd as S
    N, p, rate, seeds = sweep_instances(n)
    t0 = time.perf_counter()
    b = S.BrunelBatch(N, p, rate, seeds)
    setup_s = time.perf_counter() - t0
    b.set_window(window)
    b.set_graphs(True)
    b.step(window)            # warm: one window (captures its graph)
    b.sync()
    tiles = b.autotune()
    b.step(window)            # re-capture with the tuned tiles
    b.sync()
    e0 = b.stats().events
    t1 = time.perf_counter()
    b.step(steps)
    b.sync()
    dt_b = time.perf_counter() - t1
    ev_b = b.stats().events - e0
    total_steps = steps + 2 * window
    order = np.argsort(N)
    sample = [int(order[k]) for k in np.linspace(0, n - 1, plain_sample).round().astype(int)]
    ev_p = 0
    dt_p = 0.0
    same = True
    for i in sample:
        wex = float(np.float32(0.2 / (N[i] * p[i])))
        win = float(np.float32(-1.0 / (N[i] * p[i])))
        with S.Network.brunel(int(N[i]) // 2, int(N[i]) * 4 // 10, int(N[i]) // 10, float(p[i]),
                              float(rate[i]), wex, win, seeds=seeds[i]) as net:
            net.step(2 * window)
            net.sync()
            a = net.stats().events
            t2 = time.perf_counter()
            net.step(steps)
            net.sync()
            dt_p += time.perf_counter() - t2
            ev_p += net.stats().events - a
            for pop in range(3):
                for age in range(15):
                    same = same and np.array_equal(net.spikes(pop, age), b.spikes(i, pop, age))
    b.close()
    return {"instances": n, "steps": steps, "window": window, "graphs": True,
            "tiles_by_bucket": tiles, "setup_s": setup_s, "events": ev_b, "seconds": dt_b,
            "value": ev_b / dt_b, "neurons_total": int(N.sum()),
            "plain_sample": len(sample), "plain_events": ev_p, "plain_seconds": dt_p,
            "plain_value": ev_p / dt_p, "speedup_vs_plain": (ev_b / dt_b) / (ev_p / dt_p),
            "sample_histories_identical": bool(same), "steps_before_timing": total_steps - steps,
            "sweep": "N log-uniform [5000, 100000], p = min(0.1, 2000/N), rate U[10, 30] Hz, "
                     "seeds {1337, i}, numpy PCG64(2024)"}


def serve_requests(n: int, rate: float, seed: int = 2025):
    """The served job stream: kinds uniform over Brunel (N log-uniform in
    [10k, 200k], p = min(0.1, 2000/N)), conductance E/I (N in [10k, 100k],
    p = min(0.1, 1000/N), delay 1) and Brunel with pair-STDP E->E (N in
    [10k, 200k], p = min(0.1, 1000/N)); 500-5000 steps; Poisson arrivals at
    `rate` per second; seeds {2025, i} (numpy PCG64(seed))."""
    rng = np.random.Generator(np.random.PCG64(seed))
    kind = rng.integers(0, 3, n)
    hi = np.where(kind == 1, 1e5, 2e5)
    N = np.round(np.exp(rng.uniform(np.log(1e4), np.log(hi)))).astype(np.int64)
    p = np.minimum(0.1, np.where(kind == 0, 2000.0, 1000.0) / N)
    steps = rng.integers(500, 5001, n)
    arrival = np.cumsum(rng.exponential(1.0 / rate, n))
    return kind, N, p, steps, arrival, [(2025, i) for i in range(n)]


def run_serve(n: int, rate: float, budget_gb: float, window: int, alone_sample: int, graphs: bool = True):
    """The job stream through spice::server; p50/p99 latency (arrival to
    finish), aggregate synaptic events/s over the run, and a sample of jobs
    re-run alone whose results must be identical."""
    import spice2_amd as S
    kind, N, p, steps, arr, seeds = serve_requests(n, rate)
    fp = [S.job_footprint(int(k), int(x), float(q)) for k, x, q in zip(kind, N, p)]
    t0 = time.perf_counter()
    adm, fin, res, ev, ph = S.serve(kind, N, p, steps, arr, seeds, fp, int(budget_gb * 2**30), window,
                                    graphs=graphs)
    wall = time.perf_counter() - t0
    lat = fin - arr
    sample = np.linspace(0, n - 1, alone_sample).round().astype(int)
    same = all(res[i] == S.job_alone(int(kind[i]), int(N[i]), float(p[i]), int(steps[i]), seeds[i])
               for i in sample)
    return {"requests": n, "arrival_rate_per_s": rate, "budget_gb": budget_gb, "window": window,
            "latency_p50_s": float(np.percentile(lat, 50)), "latency_p99_s": float(np.percentile(lat, 99)),
            "queue_wait_p50_s": float(np.percentile(adm - arr, 50)),
            "events": int(ev), "makespan_s": float(fin.max()), "wall_s": wall,
            "host_seconds": {"build": ph[0], "rounds": ph[1], "retire": ph[2], "generation_in_rounds": ph[3]},
            "value": ev / float(fin.max()), "unit": "synaptic events/s",
            "kinds": [int((kind == k).sum()) for k in range(3)],
            "sample_alone": len(sample), "sample_results_identical": bool(same),
            "mix": "Brunel N 10k-200k p=min(0.1,2000/N) | conductance E/I N 10k-100k p=min(0.1,1000/N) "
                   "delay 1 | Brunel + pair-STDP E->E N 10k-200k p=min(0.1,1000/N); steps U[500,5000]; "
                   "Poisson arrivals; numpy PCG64(2025)"}


def step_roofline(events: int, rows: int, steps: int, value: float) -> dict:
    """Whole-step roofline of the Brunel workload: algorithmic bytes per step =
    4 B per synaptic event (int32 target) + 8 B per delivered row (CSR offsets)
    + the LIF state read and written once (2 x 8 B per E/I neuron; Poisson
    neurons have no state).  At 8 TB/s this bounds events/s; `frac` is the
    measured rate over that bound."""
    per_step = (4.0 * events + 8.0 * rows) / steps + 2 * 8.0 * (NE + NI)
    bound = (events / steps) / (per_step / (HBM_PEAK_GBS * 1e9))
    return {"bytes_per_step": per_step, "roofline_events_per_s": bound, "frac": value / bound,
            "note": "whole step (update + compaction + delivery) against HBM peak"}


def run_mode(mode: int, steps: int, warmup: int, tile: int | None = None, graphs: bool = True,
             timing_steps: int = 300, push_chunks: int = 0):
    """The headline run: warmup, then `steps` timed steps (windows replayed from
    HIP graphs unless graphs=False; host clock around the steps and a device
    sync), then `timing_steps` more steps launched one by one with HIP events
    around every kernel (per-kernel times for the roofline)."""
    import spice2_amd as S

    t0 = time.perf_counter()
    net = S.Network.brunel(NP, NE, NI, P_CONN, 20.0, W_EX, W_IN, seeds=(1337,), mode=mode)
    st0 = net.stats()
    csr_s = time.perf_counter() - t0
    if tile:
        S.check(S.runtime().spice_ctx_set_tile(net.context(), tile))
    S.check(S.runtime().spice_ctx_set_push_chunks(net.context(), push_chunks))
    net.set_graphs(graphs)
    net.step(warmup)
    net.sync()
    a = net.stats()
    t1 = time.perf_counter()
    net.step(steps)
    net.sync()
    t2 = time.perf_counter()
    b = net.stats()
    net.set_graphs(False)
    net.set_timing(True)
    c = net.stats()
    net.step(timing_steps)
    net.sync()
    d = net.stats()
    return net, {"events": b.events - a.events, "rows": b.rows - a.rows, "seconds": t2 - t1,
                 "graphs": graphs, "timing_steps": timing_steps,
                 "timing_events": d.events - c.events, "timing_rows": d.rows - c.rows,
                 "deliver_ms": d.deliver_ms - c.deliver_ms,
                 "deliver_launches": d.deliver_launches - c.deliver_launches,
                 "update_ms": d.update_ms - c.update_ms, "compact_ms": d.compact_ms - c.compact_ms,
                 "csr_s": csr_s, "csr_device_ms": st0.csr_ms,
                 "csr_host_rows": st0.csr_rows_host_resolved}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--mode", choices=["push", "pull"], default="pull")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=300)
    ap.add_argument("--tile", type=int, default=None, help="targets per pull tile")
    ap.add_argument("--no-graphs", action="store_true", help="launch the timed steps one by one")
    ap.add_argument("--no-other-mode", action="store_true", help="skip timing the other delivery mode")
    ap.add_argument("--push-chunks", type=int, default=0, help="push: spike chunks per (tile, step); 0 = auto")
    ap.add_argument("--other-steps", type=int, default=300, help="timed steps of the other delivery mode")
    ap.add_argument("--timing-steps", type=int, default=300,
                    help="steps after the timed region with per-kernel HIP event timing")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-sweep leg")
    ap.add_argument("--batch-instances", type=int, default=256)
    ap.add_argument("--batch-steps", type=int, default=1000)
    ap.add_argument("--batch-window", type=int, default=100)
    ap.add_argument("--plain-sample", type=int, default=16)
    ap.add_argument("--no-serve", action="store_true", help="skip the served-stream leg")
    ap.add_argument("--serve-requests", type=int, default=1000, help="jobs in the served stream")
    ap.add_argument("--serve-rate", type=float, default=40.0, help="arrivals per second")
    ap.add_argument("--serve-budget-gb", type=float, default=160.0)
    ap.add_argument("--serve-window", type=int, default=100)
    ap.add_argument("--serve-alone", type=int, default=6)
    ap.add_argument("--serve-no-graphs", action="store_true", help="server rounds launch kernels one by one")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run the known-size copy kernel (PMC counter calibration)")
    ap.add_argument("--traffic-summary", nargs=5, metavar=("FETCH_RUN", "WRITE_RUN", "CALIB_FETCH",
                                                            "CALIB_WRITE", "OUT"),
                    help="summarise rocprofv3 --pmc runs into HBM bytes per delivery launch")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.traffic_summary:
        traffic_summary(*map(Path, args.traffic_summary))
        return
    if args.gpus != 1:
        print(f"bench.py: only one GPU is supported; --gpus {args.gpus} is not (sharding by "
              "target-neuron slice with a spike all-gather is designed, not built)", file=sys.stderr)
        sys.exit(2)

    import spice2_amd as S
    mode = S.SPICE_DELIVER_PUSH if args.mode == "push" else S.SPICE_DELIVER_PULL
    net, r = run_mode(mode, args.steps, args.warmup, args.tile, graphs=not args.no_graphs,
                      timing_steps=args.timing_steps, push_chunks=args.push_chunks)
    tile_used = S.runtime().spice_ctx_tile(net.context())

    ev = r["events"]
    value = ev / r["seconds"]
    # algorithmic bytes of the delivery kernel: 4 B per event (int32 target) and
    # 8 B per delivered row (the spike's CSR offsets), counted exactly on the device,
    # over the per-kernel timing pass
    alg_bytes = 4.0 * r["timing_events"] + 8.0 * r["timing_rows"]
    launches = max(r["deliver_launches"], 1)
    avg_ms = r["deliver_ms"] / launches
    achieved = (alg_bytes / launches) / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else None
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(net, 100, args.cpu_steps)
    net.close()
    other = None
    if not args.no_other_mode:  # the same workload with the other delivery mode
        omode = S.SPICE_DELIVER_PULL if args.mode == "push" else S.SPICE_DELIVER_PUSH
        onet, o = run_mode(omode, args.other_steps, args.warmup, args.tile, graphs=not args.no_graphs,
                           timing_steps=min(args.timing_steps, args.other_steps), push_chunks=args.push_chunks)
        onet.close()
        o_launches = max(o["deliver_launches"], 1)
        o_ms = o["deliver_ms"] / o_launches
        o_bytes = (4.0 * o["timing_events"] + 8.0 * o["timing_rows"]) / o_launches
        o_gbs = o_bytes / (o_ms * 1e-3) / 1e9 if o_ms > 0 else None
        other = {"delivery": "pull" if args.mode == "push" else "push", "steps": args.other_steps,
                 "value": o["events"] / o["seconds"], "ms_per_step": 1e3 * o["seconds"] / args.other_steps,
                 "deliver_ms_per_step": o["deliver_ms"] / o["timing_steps"],
                 "roofline": {"bound": "hbm", "achieved": o_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                              "frac": (o_gbs / HBM_PEAK_GBS) if o_gbs else None,
                              "algorithmic_bytes_per_launch": o_bytes, "avg_launch_ms": o_ms},
                 "note": "same network and seed; spike rasters are identical in both modes (event counts are integers)"}
    batch = None
    if not args.no_batch:
        batch = run_batch(args.batch_instances, args.batch_steps, args.batch_window, args.plain_sample)
    served = None
    if not args.no_serve:
        served = run_serve(args.serve_requests, args.serve_rate, args.serve_budget_gb, args.serve_window,
                           args.serve_alone, graphs=not args.serve_no_graphs)

    line = {
        "metric": "synaptic events/s (Brunel N=500k, p=0.02, deterministic rasters)",
        "value": value,
        "unit": "synaptic events/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * r["seconds"] / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 state, int32 targets, u32 event counts",
        "data": "synthetic (fixed_probability connectivity generated on the device from seed 1337)",
        "config": {"workload": "Brunel N=500000 (P 250k / E 200k / I 50k), p=0.02, "
                               "w=+0.2/(N p), -1/(N p), dt=1e-4, delay 15 steps, seed 1337",
                   "delivery": args.mode, "pull_tile": tile_used if args.mode == "pull" else None,
                   "events_timed": ev, "rows_timed": r["rows"],
                   "csr_generation_s": r["csr_s"], "csr_device_ms": r["csr_device_ms"],
                   "csr_rows_resolved_on_host": r["csr_host_rows"],
                   "graphs": r["graphs"], "timing_steps": r["timing_steps"],
                   "update_ms_per_step": r["update_ms"] / r["timing_steps"],
                   "compact_ms_per_step": r["compact_ms"] / r["timing_steps"],
                   "deliver_ms_per_step": r["deliver_ms"] / r["timing_steps"]},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": (achieved / HBM_PEAK_GBS) if achieved else None,
                     "traffic": traffic_from_profiles(),
                     "kernel": "deliver (per launch = one block of up to 15 steps)",
                     "algorithmic_bytes_per_launch": alg_bytes / launches,
                     "avg_launch_ms": avg_ms},
        "step_roofline": step_roofline(ev, r["rows"], args.steps, value),
        "other_mode": other,
        "cpu_baseline": cpu,
        "batched": batch,
        "served": served,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
