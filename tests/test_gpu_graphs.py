"""Graph replay of gradient passes: every replay must give bitwise the
gradients and misfit of the eager launches, with eager work on the same
context between replays, across field changes, and after a change that
forces a new capture (coverage on, shots per pass)."""
import numpy as np
import pytest

from _cases import unit_config, unit_geometry, unit_phantom
from _pkg import tomo

pytestmark = pytest.mark.gpu


def make(c, ii, jj, meas, graphs):
    e = tomo.Engine(c, device=0)
    e.set_transducers(ii, jj, one_based=False)
    e.set_receivers(ii, jj, one_based=False)
    e.set_groups(1)
    e.set_measured(meas)
    if graphs:
        e.set_graphs(True)
    return e


@pytest.mark.parametrize("n,ckpt", [(96, 100), (128, 0)])
def test_graph_replays_match_eager(n, ckpt):
    print(f"[case] graphs n={n} ckpt={ckpt}", flush=True)
    c = unit_config(n, 400)
    c.checkpoint_every = ckpt
    ii, jj, _ = unit_geometry(c, 12)
    fo = unit_phantom(c)
    with tomo.Engine(c, device=0) as g:
        g.set_transducers(ii, jj, one_based=False)
        g.set_receivers(ii, jj, one_based=False)
        g.set_groups(1)
        g.set_field(fo)
        meas = g.forward(fp64=True)
    eager = make(c, ii, jj, meas, False)
    graph = make(c, ii, jj, meas, True)
    rng = np.random.default_rng(4)
    try:
        for rnd in range(4):
            if rnd == 2:  # forces a new capture
                eager.enable_coverage(True)
                graph.enable_coverage(True)
            f = (fo * rng.uniform(0.2, 0.9)).astype(np.float32)
            eager.set_field(f)
            graph.set_field(f)
            m_e = eager.gradient()
            m_g = graph.gradient()
            assert m_g == m_e, (rnd, m_g, m_e)
            for s in range(0, 12, 5):
                np.testing.assert_array_equal(graph.get_gradient(s), eager.get_gradient(s))
            if rnd >= 2:
                np.testing.assert_array_equal(graph.get_coverage(), eager.get_coverage())
            # eager work on the graph context between replays
            graph.misfit()
            graph.forward()
        t = graph.timing_graphs()
        assert t["replays"] > 0 and t["ms"] > 0
    finally:
        eager.close()
        graph.close()


def test_graph_replays_time_their_stencil_loops():
    """The replayed graph records its loops' event pairs: the forward and
    adjoint launch and cell counts equal the eager run's, and their times
    fall inside the replays' own time."""
    print("[case] graph loop timings", flush=True)
    c = unit_config(128, 600)
    c.checkpoint_every = 100
    ii, jj, _ = unit_geometry(c, 16)
    fo = unit_phantom(c)
    with tomo.Engine(c, device=0) as g:
        g.set_transducers(ii, jj, one_based=False)
        g.set_receivers(ii, jj, one_based=False)
        g.set_groups(1)
        g.set_field(fo)
        meas = g.forward(fp64=True)
    eager = make(c, ii, jj, meas, False)
    graph = make(c, ii, jj, meas, True)
    try:
        for e in (eager, graph):
            e.set_field((0.5 * fo).astype(np.float32))
            e.gradient()  # the graph is captured here
            e.timing_reset()
            for _ in range(3):
                e.gradient()
        te, tg, tr = eager.timing(), graph.timing(), graph.timing_graphs()
        for k in ("fwd_launches", "adj_launches", "fwd_cells", "adj_cells"):
            assert tg[k] == te[k], (k, tg[k], te[k])
        assert tr["replays"] == 3 and tr["untimed_cells"] == 0
        inner = tg["fwd_ms"] + tg["adj_ms"]
        print(f"[case] graph replays {tr['ms']:.3f} ms, stencil loops inside {inner:.3f} ms, "
              f"eager loops {te['fwd_ms'] + te['adj_ms']:.3f} ms", flush=True)
        assert 0.3 * tr["ms"] < inner <= 1.001 * tr["ms"]
    finally:
        eager.close()
        graph.close()


def test_graph_replays_with_passes():
    print("[case] graphs with passes", flush=True)
    c = unit_config(64, 300)
    c.shots_per_pass = 5
    ii, jj, _ = unit_geometry(c, 12)
    fo = unit_phantom(c)
    with tomo.Engine(c, device=0) as g:
        g.set_transducers(ii, jj, one_based=False)
        g.set_receivers(ii, jj, one_based=False)
        g.set_groups(1)
        g.set_field(fo)
        meas = g.forward(fp64=True)
    eager = make(c, ii, jj, meas, False)
    graph = make(c, ii, jj, meas, True)
    try:
        for scale in (0.3, 0.6, 0.6):
            f = (fo * scale).astype(np.float32)
            eager.set_field(f)
            graph.set_field(f)
            assert graph.gradient() == eager.gradient()
            for s in range(12):
                np.testing.assert_array_equal(graph.get_gradient(s), eager.get_gradient(s))
    finally:
        eager.close()
        graph.close()


def test_autotune_picks_a_valid_shape_and_keeps_results():
    print("[case] autotune", flush=True)
    c = unit_config(128, 300)
    ii, jj, _ = unit_geometry(c, 8)
    fo = unit_phantom(c)
    with tomo.Engine(c, device=0) as e:
        e.set_transducers(ii, jj, one_based=False)
        e.set_receivers(ii, jj, one_based=False)
        e.set_groups(1)
        e.set_field(fo)
        meas = e.forward(fp64=True)
        e.set_measured(meas)
        e.set_field((0.5 * fo).astype(np.float32))
        e.gradient()
        g_default = np.stack([e.get_gradient(s) for s in range(8)])
        shape, (tf, ta) = e.autotune()
        assert shape[0] in (2, 4) and shape[1] in (8, 16, 32) and shape[2] in (8, 16, 32)
        assert shape[3] in (1, 2, 4) and shape[4] in (1, 2)
        assert 0 < tf < 1e3 and 0 < ta < 1e3
        assert e.get_shape() == shape
        # cached for the size class: same answer without re-timing
        assert e.autotune()[0] == shape
        e.gradient()
        g_tuned = np.stack([e.get_gradient(s) for s in range(8)])
        # shapes of one lane width agree bitwise, across lane widths to rounding
        scale = np.abs(g_default).max()
        assert np.abs(g_tuned - g_default).max() <= 1e-4 * scale
