"""Two and three stencil steps per launch (the intermediate levels held in
registers; two or three for the adjoint), against one step per launch: traces, misfits and
gradients must be bitwise equal -- the per-cell arithmetic is the same -- for
each full-row lane width (nx = 128, 192, 256, 384, 512 -> 2, 3, 4, 6, 8 cells
per lane; 3 and 6 cells move as single floats and float pairs), with
an odd number of steps (a single-step tail), checkpoint segments of odd and
even length (adjoint pairs that stop short of a segment's first level), the
edge/receiver passes around each pair, and receivers on strip and halo
rows.  Where the two-step adjoint does not apply (two receivers in one lane
segment) the engine falls back to one-step adjoint launches with
the same results; with coverage on, a pass per checkpoint segment sums the
coverage map in the one-step sweep's order."""
import numpy as np
import pytest

from _cases import unit_config, unit_phantom
from _pkg import tomo

pytestmark = pytest.mark.gpu


def adj_steps(n):
    """Adjoint steps per launch by default: three at widths 256 and 512, else two."""
    return 3 if n in (256, 512) else 2


def engine(n, nt, ckpt, steps, shots=6, coverage=False, receivers=None):
    c = unit_config(n, nt)
    c.checkpoint_every = ckpt
    ii, jj = tomo.transducers_ring(n, n, shots, 0.39 * n)
    e = tomo.Engine(c, device=0)
    if coverage:
        e.enable_coverage(True)
    e.set_transducers(ii, jj, one_based=False)
    ri, rj = (ii, jj) if receivers is None else receivers
    e.set_receivers(ri, rj, one_based=False)
    e.set_groups(1)
    e.set_time_blocking(steps)
    return e, c


def compare(e1, e2, c, shots, coverage=False):
    fo = unit_phantom(c, seed=7)
    e1.set_field(fo)
    meas = e1.forward(fp64=True)
    f = (0.5 * fo).astype(np.float32)
    for e in (e1, e2):
        e.set_measured(meas)
        e.set_field(f)
    m1, m2 = e1.gradient(), e2.gradient()
    print(f"  misfit {m1!r} {m2!r}", flush=True)
    assert m1 == m2
    for s in range(shots):
        np.testing.assert_array_equal(e2.get_gradient(s), e1.get_gradient(s))
    if coverage:
        np.testing.assert_array_equal(e2.get_coverage(), e1.get_coverage())


@pytest.mark.parametrize("steps", [2, 3])
@pytest.mark.parametrize("n,nt,ckpt", [(128, 301, 0), (256, 301, 49), (512, 203, 50), (512, 160, 7),
                                       (192, 241, 40), (384, 203, 50)])
def test_blocked_launches_match_single_steps(n, nt, ckpt, steps):
    print(f"[case] time blocking {steps} n={n} nt={nt} ckpt={ckpt}", flush=True)
    e1, c = engine(n, nt, ckpt, 1)
    e2, _ = engine(n, nt, ckpt, steps)
    try:
        assert e1.time_blocking() == 1 and e2.time_blocking() == steps
        assert e1.adjoint_blocking() == 1 and e2.adjoint_blocking() == adj_steps(n)
        compare(e1, e2, c, 6)
        # the FP32 trace forward (its own rolling ring) agrees too
        np.testing.assert_array_equal(e2.forward(fp64=False), e1.forward(fp64=False))
    finally:
        e1.close()
        e2.close()


@pytest.mark.parametrize("steps", [2, 3])
def test_many_receivers_on_strip_and_halo_rows(steps):
    """128 sources and receivers on the ring at 256^2 (receivers on most
    rows, so on strip boundaries and halo rows of every level)."""
    print(f"[case] time blocking {steps}, 256 with 128 receivers", flush=True)
    e1, c = engine(256, 241, 60, 1, shots=128)
    e2, _ = engine(256, 241, 60, steps, shots=128)
    try:
        assert e2.time_blocking() == steps and e2.adjoint_blocking() == 3
        compare(e1, e2, c, 128)
    finally:
        e1.close()
        e2.close()


@pytest.mark.parametrize("ny,rows", [(96, 32), (160, 32), (96, 8)])
def test_odd_strip_counts_on_rectangular_grids(ny, rows):
    """Rectangular grids (nx = 128, ny = 96 or 160) with 3, 5 or 12 strips.
    The three-step forward's last strip forms the bottom edge rows at the end
    of its walk.  The two-step adjoint's strips alternate their walking
    direction (even strips down, odd up): with an odd number of strips the
    last one walks down, with an even number it walks up and writes the
    bottom edge row at the start."""
    print(f"[case] rectangular 128x{ny}, rows {rows}", flush=True)
    engines = []
    for steps in (1, 3):
        c = unit_config(128, 241)
        c.ny = ny
        c.mask_y0, c.mask_y1 = ny // 4, ny - ny // 4
        c.checkpoint_every = 40
        ii, jj = tomo.transducers_ring(128, ny, 8, 0.39 * min(128, ny))
        e = tomo.Engine(c, device=0)
        e.set_transducers(ii, jj, one_based=False)
        e.set_receivers(ii, jj, one_based=False)
        e.set_groups(1)
        e.set_shape(4, rows, rows, 1, 1)
        e.set_time_blocking(steps)
        engines.append(e)
    e1, e3 = engines
    try:
        assert e1.time_blocking() == 1 and e3.time_blocking() == 3 and e3.adjoint_blocking() == 2
        compare(e1, e3, c, 8)
        np.testing.assert_array_equal(e3.forward(fp64=False), e1.forward(fp64=False))
    finally:
        e1.close()
        e3.close()


@pytest.mark.parametrize("n,nt,ckpt,steps", [(256, 201, 40, 2), (256, 201, 40, 3), (128, 160, 7, 3),
                                             (128, 121, 0, 3)])
def test_coverage_with_blocked_adjoint(n, nt, ckpt, steps):
    """Coverage on: the two-step adjoint carries none, a pass per checkpoint
    segment (or over the full history) sums u(k)^2 newest level first, corner
    cells as the mean of their edge neighbours -- the values and order of the
    one-step sweep, so the coverage maps agree bit for bit."""
    print(f"[case] coverage, blocking {steps}, n={n} nt={nt} ckpt={ckpt}", flush=True)
    e1, c = engine(n, nt, ckpt, 1, coverage=True)
    e2, _ = engine(n, nt, ckpt, steps, coverage=True)
    try:
        assert e1.adjoint_blocking() == 1
        assert e2.time_blocking() == steps and e2.adjoint_blocking() == adj_steps(n)
        compare(e1, e2, c, 6, coverage=True)
        assert e2.get_coverage().max() > 0
    finally:
        e1.close()
        e2.close()


@pytest.mark.parametrize("n,nt,ckpt,adj,rows,depth", [(512, 160, 7, 2, 0, 1), (256, 301, 49, 2, 0, 1),
                                                       (512, 203, 50, 3, 16, 1), (256, 241, 40, 3, 32, 1),
                                                       (512, 98, 0, 3, 32, 1), (512, 203, 50, 3, 32, 2),
                                                       (256, 160, 7, 3, 16, 2), (512, 203, 50, 3, 64, 1),
                                                       (256, 241, 40, 3, 64, 1), (512, 98, 0, 3, 64, 1)])
def test_adjoint_steps_and_strip_heights(n, nt, ckpt, adj, rows, depth):
    """The two-step adjoint where three would apply (tomo_set_adjoint_blocking),
    and the three-step adjoint with 16-, 32- and 64-row strips, full history
    and one or two rows of read-ahead, against one step per launch."""
    print(f"[case] adjoint steps {adj} rows {rows} depth {depth} n={n} nt={nt} ckpt={ckpt}", flush=True)
    e1, c = engine(n, nt, ckpt, 1)
    e2, _ = engine(n, nt, ckpt, 3)
    try:
        e2.set_adjoint_blocking(adj)
        if rows:
            e2.set_shape(4, 32 if n == 512 else 16, rows, 1, depth)
        assert e2.adjoint_blocking() == adj
        compare(e1, e2, c, 6)
    finally:
        e1.close()
        e2.close()


def test_fallback_to_one_step_adjoint():
    print("[case] adjoint fallback: crowded receivers", flush=True)
    # two receivers in one 4-cell lane segment of a row: records in passes
    # after the blocked forward launches, one-step adjoint
    ri = np.array([100, 101, 150], np.int32)
    rj = np.array([120, 120, 60], np.int32)
    for steps in (2, 3):
        e1, c = engine(256, 201, 40, 1, receivers=(ri, rj))
        e2, _ = engine(256, 201, 40, steps, receivers=(ri, rj))
        try:
            assert e2.time_blocking() == steps and e2.adjoint_blocking() == 1
            compare(e1, e2, c, 6)
        finally:
            e1.close()
            e2.close()


@pytest.mark.parametrize("steps", [2, 3])
def test_reference_preset_blocked_forward_matches_single_steps(steps):
    """The reference preset (no sponge, sensor lines with a receiver in
    every cell, records from the small passes after each launch) runs the
    blocked forward kernels; its traces and gradients equal one step per
    launch bitwise (the adjoint stays at one step: overwrite injection on
    the lines)."""
    from _cases import ref_config, ref_geometry, ref_phantom
    print(f"[case] reference preset, forward steps per launch {steps}", flush=True)
    c = ref_config(192, 401)
    ii, jj, _ = ref_geometry(c, 32)
    fo = ref_phantom(c)
    es = []
    try:
        for tb in (1, steps):
            e = tomo.Engine(c, device=0)
            e.set_transducers(ii, jj, one_based=True)
            e.set_groups(32)
            e.set_time_blocking(tb)
            es.append(e)
        e1, e2 = es
        assert e1.time_blocking() == 1 and e2.time_blocking() == steps
        assert e2.adjoint_blocking() == 1
        e1.set_field(fo)
        e2.set_field(fo)
        meas = e1.forward(fp64=True)
        np.testing.assert_array_equal(e2.forward(fp64=False), e1.forward(fp64=False))
        f = (0.5 * fo).astype(np.float32)
        for e in es:
            e.set_measured(meas)
            e.set_field(f)
        m1, m2 = e1.gradient(), e2.gradient()
        assert m1 == m2
        for s in range(e1.num_shots):
            np.testing.assert_array_equal(e2.get_gradient(s), e1.get_gradient(s))
    finally:
        for e in es:
            e.close()
