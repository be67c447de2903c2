"""Two forward steps per launch (the intermediate level held in registers)
against one step per launch: traces, misfits and gradients must be bitwise
equal -- the per-cell arithmetic is the same -- for each full-row lane width
(nx = 128, 256, 512 -> 2, 4, 8 cells per lane), with an odd number of steps
(a single-step tail), checkpoint segments of odd and even length, and the
edge/receiver passes around each pair."""
import numpy as np
import pytest

from _cases import unit_config, unit_phantom
from _pkg import tomo

pytestmark = pytest.mark.gpu


def engine(n, nt, ckpt, steps, shots=6):
    c = unit_config(n, nt)
    c.checkpoint_every = ckpt
    ii, jj = tomo.transducers_ring(n, n, shots, 0.39 * n)
    e = tomo.Engine(c, device=0)
    e.set_transducers(ii, jj, one_based=False)
    e.set_receivers(ii, jj, one_based=False)
    e.set_groups(1)
    e.set_time_blocking(steps)
    return e, c


@pytest.mark.parametrize("n,nt,ckpt", [(128, 301, 0), (256, 301, 49), (512, 203, 50)])
def test_two_step_launches_match_single_steps(n, nt, ckpt):
    print(f"[case] time blocking n={n} nt={nt} ckpt={ckpt}", flush=True)
    e1, c = engine(n, nt, ckpt, 1)
    e2, _ = engine(n, nt, ckpt, 2)
    try:
        assert e1.time_blocking() == 1 and e2.time_blocking() == 2
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
        for s in range(6):
            g1, g2 = e1.get_gradient(s), e2.get_gradient(s)
            np.testing.assert_array_equal(g2, g1)
    finally:
        e1.close()
        e2.close()
