"""Batched networks (spice::batch): many independent Brunel networks of uneven
sizes, rates and seeds stepped together in one context.  Every instance must
produce exactly the raster it produces alone, and the raster of the oracle
(the CPU restatement of the reference) on the same parameters; windows replayed
from HIP graphs must equal windows launched one by one."""
import numpy as np
import pytest

import oracle_lib as O
import spice2_amd as S

pytestmark = pytest.mark.gpu

# (N, p, rate, seed) - uneven sizes, connection probability min(0.1, 200/N)
INSTANCES = [(2000, 0.1, 20.0, (1337, 0)), (5200, 200 / 5200, 13.5, (1337, 1)),
             (3100, 200 / 3100, 27.0, (1337, 2)), (8000, 200 / 8000, 10.0, (1337, 3)),
             (640, 0.1, 30.0, (1337, 4))]


def make_batch(mode=S.SPICE_DELIVER_PULL, graphs=False):
    b = S.BrunelBatch([x[0] for x in INSTANCES], [x[1] for x in INSTANCES], [x[2] for x in INSTANCES],
                      [x[3] for x in INSTANCES], mode=mode)
    b.set_graphs(graphs)
    return b


def weights(N, p):
    return np.float32(0.2 / (N * p)), np.float32(-1.0 / (N * p))


def test_batch_equals_standalone_every_step():
    steps = 45
    print("batch of", len(INSTANCES), flush=True)
    with make_batch() as b:
        alone = []
        for N, p, rate, seed in INSTANCES:
            wex, win = weights(N, p)
            alone.append(S.Network.brunel(N // 2, N * 4 // 10, N // 10, p, rate, float(wex), float(win),
                                          seeds=seed))
        try:
            for t in range(steps):
                b.step()
                for i, net in enumerate(alone):
                    net.step()
                    for pop in range(3):
                        got, want = b.spikes(i, pop), net.spikes(pop)
                        assert np.array_equal(got, want), (i, pop, t)
        finally:
            for net in alone:
                net.close()


@pytest.mark.parametrize("inst", [0, 2, 4])
def test_batch_instance_matches_oracle(inst):
    N, p, rate, seed = INSTANCES[inst]
    wex, win = weights(N, p)
    steps = 120
    onet, _ = O.brunel_net(N // 2, N * 4 // 10, N // 10, p, rate, wex, win, seeds=seed)
    with make_batch() as b:
        for t in range(steps):
            b.step()
            onet.step()
            for pop in range(3):
                assert np.array_equal(b.spikes(inst, pop), onet.spikes(pop)), (inst, pop, t)
    onet.close()


@pytest.mark.parametrize("mode", ["pull", "push"])
def test_batch_graph_replay_equals_eager(mode):
    m = S.SPICE_DELIVER_PULL if mode == "pull" else S.SPICE_DELIVER_PUSH
    with make_batch(m, graphs=False) as a, make_batch(m, graphs=True) as g:
        # windows of 100 steps: the first records the graph, later ones replay it
        for w in range(4):
            a.step(100)
            g.step(100)
            for i in range(len(INSTANCES)):
                for pop in range(3):
                    for age in range(15):
                        assert np.array_equal(a.spikes(i, pop, age), g.spikes(i, pop, age)), (w, i, pop, age)
        assert a.stats().events == g.stats().events


def test_autotuned_tiles_keep_rasters():
    """Re-tiling target populations (the autotuner replays the last block's
    delivery with each candidate tile) changes neither the counts it leaves
    behind nor the rasters that follow."""
    with make_batch() as a, make_batch() as t:
        a.step(30)
        t.step(30)
        a.sync()
        t.sync()
        ev_before = t.stats().events
        tiles = t.autotune((64, 1024, 4096), reps=1)
        assert t.stats().events == ev_before
        assert all(x in (0, 64, 1024, 4096) for x in tiles), tiles
        for step in range(40):
            a.step()
            t.step()
            for i in range(len(INSTANCES)):
                for pop in range(3):
                    assert np.array_equal(a.spikes(i, pop), t.spikes(i, pop)), (step, i, pop)
