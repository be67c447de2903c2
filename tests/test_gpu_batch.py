"""Batched stepping of problems of different sizes: one launch per time level
covers every problem, and the result must be bitwise what each context
computes alone (same kernels, same shapes), eager and graph-replayed, with
shot passes and checkpoint segments of different lengths in play."""
import numpy as np
import pytest

from _cases import unit_config, unit_phantom
from _pkg import tomo

pytestmark = pytest.mark.gpu

SIZES = [(64, 16, 512), (96, 24, 768), (128, 32, 1024)]  # n, shots (n/4), steps (8n)


def problem(n, shots, steps, seed, coverage, passes=12):
    c = unit_config(n, steps + 2)
    c.checkpoint_every = 100
    c.shots_per_pass = passes
    ii, jj = tomo.transducers_ring(n, n, shots, 0.39 * n)
    fo = unit_phantom(c, seed=seed)
    e = tomo.Engine(c, device=0)
    e.set_transducers(ii, jj, one_based=False)
    e.set_receivers(ii, jj, one_based=False)
    e.set_groups(1)
    if coverage:
        e.enable_coverage(True)
    e.set_field(fo)
    meas = e.forward(fp64=True)
    e.set_measured(meas)
    e.set_field((0.4 * fo).astype(np.float32))
    return e


@pytest.mark.parametrize("graph,coverage", [(False, False), (True, True)])
def test_batch_matches_single_contexts(graph, coverage):
    print(f"[case] batch graph={graph} coverage={coverage}", flush=True)
    alone = [problem(n, s, t, 10 + i, coverage) for i, (n, s, t) in enumerate(SIZES)]
    batched = [problem(n, s, t, 10 + i, coverage) for i, (n, s, t) in enumerate(SIZES)]
    try:
        want = [e.gradient() for e in alone]
        with tomo.Batch(batched, graph=graph) as b:
            for rep in range(2):  # the second call replays the program (or graph)
                mis = b.gradient()
                assert list(mis) == want, (rep, list(mis), want)
                for ea, eb, (n, s, t) in zip(alone, batched, SIZES):
                    for g in (0, s // 2, s - 1):
                        np.testing.assert_array_equal(eb.get_gradient(g), ea.get_gradient(g))
                    if coverage:
                        np.testing.assert_array_equal(eb.get_coverage(), ea.get_coverage())
            t = b.timing()
            cells = sum(3 * t_ * s * n * n for n, s, t_ in SIZES)
            assert t["calls"] == 2 and t["ms"] > 0
            assert abs(t["cells"] / 2 - cells) / cells < 0.01
    finally:
        for e in alone + batched:
            e.close()


def test_batch_rejects_mixed_checkpoints():
    a = problem(64, 16, 512, 1, False)
    c = unit_config(64, 514)
    c.checkpoint_every = 50
    ii, jj = tomo.transducers_ring(64, 64, 16, 25.0)
    e = tomo.Engine(c, device=0)
    try:
        e.set_transducers(ii, jj, one_based=False)
        e.set_receivers(ii, jj, one_based=False)
        e.set_groups(1)
        e.set_measured(np.zeros((16, 514, 16), np.float32))
        with pytest.raises(tomo.TomoError):
            tomo.Batch([a, e])
    finally:
        a.close()
        e.close()


def test_batch_refuses_changed_contexts():
    """The program holds the contexts' device pointers and shapes as they were
    at creation: after a context changes them, the batch must refuse to run."""
    print("[case] batch after a context changed", flush=True)
    a = problem(64, 16, 512, 3, False)
    b = problem(96, 24, 768, 4, False)
    try:
        with tomo.Batch([a, b]) as bt:
            bt.gradient()
            a.enable_coverage(True)
            with pytest.raises(tomo.TomoError, match="changed"):
                bt.gradient()
        with tomo.Batch([a, b]) as bt:   # built for the new state: runs
            bt.gradient()
            b.set_groups(2)              # fewer shots: per-shot buffers and tables rebuilt
            with pytest.raises(tomo.TomoError, match="changed"):
                bt.gradient()
    finally:
        a.close()
        b.close()


def test_batch_refuses_oversized_program():
    """A program above the accepted launch count is refused at creation,
    before any launch (one context of 6000 steps: about 18k launches)."""
    print("[case] batch program size limit", flush=True)
    c = unit_config(64, 6002)
    c.checkpoint_every = 100
    ii, jj = tomo.transducers_ring(64, 64, 4, 25.0)
    e = tomo.Engine(c, device=0)
    try:
        e.set_transducers(ii, jj, one_based=False)
        e.set_receivers(ii, jj, one_based=False)
        e.set_groups(1)
        e.set_measured(np.zeros((4, 6002, 4), np.float32))
        with pytest.raises(tomo.TomoError, match="launch program"):
            tomo.Batch([e])
    finally:
        e.close()


def test_shared_workspace_matches_own_storage():
    """Three problems of different sizes share one wavefield workspace (every
    shot in one pass) and are stepped one after another, twice, the middle
    one by graph replay: the pool grows to the largest need and every
    gradient equals the same problem's on storage of its own, bit for bit.
    A batch refuses contexts with a workspace, and the workspace refuses to
    be destroyed while contexts use it."""
    print("[case] shared workspace", flush=True)
    sizes = [(64, 16, 512), (128, 32, 1024), (96, 24, 768)]
    alone = [problem(n, s, t, 20 + i, False, passes=0) for i, (n, s, t) in enumerate(sizes)]
    shared = [problem(n, s, t, 20 + i, False, passes=0) for i, (n, s, t) in enumerate(sizes)]
    ws = tomo.Workspace(0)
    try:
        shared[1].set_graphs(True)
        alone[1].set_graphs(True)
        for e in shared:
            ws.attach(e)
        want = [e.gradient() for e in alone]
        ref = [[e.get_gradient(g) for g in range(s)] for e, (n, s, t) in zip(alone, sizes)]
        for rep in range(2):
            for i, e in enumerate(shared):
                assert e.gradient() == want[i], (rep, i)
                for g in range(sizes[i][1]):
                    np.testing.assert_array_equal(e.get_gradient(g), ref[i][g])
        with pytest.raises(tomo.TomoError, match="workspace"):
            tomo.Batch(shared)
        with pytest.raises(tomo.TomoError, match="still use"):
            ws.close()
        for e in shared:
            e.detach_workspace()
        assert shared[0].gradient() == want[0]   # back on storage of its own
        ws.close()
    finally:
        for e in alone + shared:
            e.close()
