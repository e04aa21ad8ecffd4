"""Served job stream (spice::server): jobs of three kinds (Brunel, conductance
E/I with delay 1, Brunel with a pair-STDP E->E projection) arrive over time,
are admitted by device-memory budget and step together with the jobs already
running.  Every job's result (neuron state after its last step and that step's
spike counts) must equal the result of the same network run alone."""
import numpy as np
import pytest

import spice2_amd as S

pytestmark = pytest.mark.gpu

# kind, N, p, steps, arrival (s), seed
JOBS = [(0, 6000, 0.1, 300, 0.00, (7, 0)), (1, 4000, 0.05, 200, 0.00, (7, 1)),
        (2, 5000, 0.1, 250, 0.01, (7, 2)), (0, 3000, 0.1, 150, 0.02, (7, 3)),
        (1, 2500, 0.08, 400, 0.03, (7, 4)), (2, 3000, 0.1, 110, 0.05, (7, 5)),
        (0, 8000, 0.05, 120, 0.05, (7, 6)), (1, 6000, 0.03, 90, 0.20, (7, 7)),
        (0, 2000, 0.1, 333, 0.25, (7, 8))]


def test_served_jobs_equal_jobs_alone():
    kind, N, p, steps, arr, seeds = map(list, zip(*JOBS))
    fp = [S.job_footprint(k, n, q) for k, n, q in zip(kind, N, p)]
    budget = 2 * max(fp)  # forces later jobs to wait for memory
    print("serving", len(JOBS), "jobs, budget", budget, flush=True)
    adm, fin, res, ev, _ = S.serve(kind, N, p, steps, arr, seeds, fp, budget, window=50)
    assert ev > 0
    assert (adm >= np.array(arr) - 1e-9).all() and (fin > adm).all()
    for i, (k, n, q, st, _, sd) in enumerate(JOBS):
        alone = S.job_alone(k, n, q, st, sd)
        assert res[i] == alone, (i, JOBS[i])
