"""The 128x128 plain-SIRT configuration end to end against the FP64 oracle:
32 ring transducers (radius 50) fired one at a time, 1000 time steps,
2-4 inclusions in the central 64x64 region, data from the FP64 solver, 10
uniform SIRT iterations at a fixed relaxation.  The engine (FP32 bulk
arithmetic, FP64 reductions) and the oracle (FP64 throughout) run the same
iterations; reported per iteration and bounded:

  relative error eps (%)        |eps_gpu - eps_oracle| <= 0.05 (points)
  field after each iteration    relative L2 <= 2e-3
  traces of the final field     relative L2 <= 2e-4

Also the per-pixel coverage weights of the update: cbar / (c + eps cmax) on
the engine's own coverage (relative L2 of the update <= 1e-5).
"""
import numpy as np
import pytest

import oracle as orc
from _cases import rel_l2
from _pkg import tomo

pytestmark = pytest.mark.gpu

EPS_TOL = 0.05
FIELD_TOL = 2e-3
TRACE_TOL = 2e-4


def c1_problem(seed=1000, steps=1000):
    n = 128
    c = tomo.config_unit(n, steps + 2)
    ii, jj = tomo.transducers_ring(n, n, 32, 50.0)
    region = (c.mask_x0, c.mask_x1, c.mask_y0, c.mask_y1)
    fo = np.zeros(n * n, np.float32)
    tomo.phantom_inclusions(fo, n, n, seed, 2 + seed % 3, 4.0, 12.0, 0.05, 0.10, region)
    return c, ii, jj, fo


def test_c1_ten_sirt_iterations_match_oracle():
    print("[case] C1 10 iterations", flush=True)
    c, ii, jj, fo = c1_problem()
    n2 = c.nx * c.ny
    cells = (ii + c.nx * jj).astype(np.int32)
    bbox = np.stack([ii, ii, jj, jj], axis=1).astype(np.int32).ravel()
    with tomo.Engine(c, device=0) as e:
        e.set_transducers(ii, jj, one_based=False)
        e.set_receivers(ii, jj, one_based=False)
        e.set_groups(1)
        e.set_truth(fo)
        e.set_field(fo)
        meas = e.forward(fp64=True)
        e.set_field(np.zeros(n2, np.float32))
        e.set_measured(meas)
        e.gradient()
        g = np.stack([e.get_gradient(s) for s in range(32)]).astype(np.float64).sum(0) / 32
        scale = float(0.02 / np.abs(g).max())
        f_orc = np.zeros(n2, np.float64)
        m64 = meas.astype(np.float64)
        rows = []
        for it in range(10):
            _, eps_gpu = e.sirt_iteration(scale, weighted=False, mode=tomo.UPDATE_MEAN)
            G, _ = orc.gradient(c, f_orc, bbox, cells, cells, m64, np.float64)
            f_orc = orc.update_uniform(c, f_orc, G, scale, tomo.UPDATE_MEAN)
            eps_orc = orc.error(f_orc, fo)
            f_gpu = e.get_field()
            err = rel_l2(f_gpu, f_orc)
            rows.append((it, eps_gpu, eps_orc, err))
            print(f"  iter {it}: eps gpu {eps_gpu:.4f} oracle {eps_orc:.4f}  field rel L2 {err:.2e}", flush=True)
        for it, eg, eo, err in rows:
            assert abs(eg - eo) <= EPS_TOL, rows
            assert err <= FIELD_TOL, rows
        assert rows[-1][1] < 100.0  # the iteration converges
        tr_gpu = e.forward()
    tr_orc = orc.forward(c, f_orc, bbox, cells, np.float64)
    err = rel_l2(tr_gpu, tr_orc)
    print(f"  final traces rel L2 {err:.2e}", flush=True)
    assert err <= TRACE_TOL


def test_coverage_pixel_weights():
    print("[case] pixel weights", flush=True)
    c, ii, jj, fo = c1_problem(1001, steps=400)
    with tomo.Engine(c, device=0) as e:
        e.set_transducers(ii, jj, one_based=False)
        e.set_receivers(ii, jj, one_based=False)
        e.set_groups(1)
        e.enable_coverage(True)
        e.set_field(fo)
        meas = e.forward(fp64=True)
        e.set_measured(meas)
        f0 = (0.3 * fo).astype(np.float32)
        e.set_field(f0)
        e.gradient()
        cov = e.get_coverage().astype(np.float64)
        G = np.stack([e.get_gradient(s) for s in range(32)]).astype(np.float64)
        e.set_pixel_weights(1, 0.05)
        e.update_uniform(3.0, tomo.UPDATE_MEAN)
        f1 = e.get_field().astype(np.float64)
    m = np.zeros((c.ny, c.nx), bool)
    m[c.mask_y0:c.mask_y1, c.mask_x0:c.mask_x1] = True
    m = m.ravel()
    w = cov[m].mean() / (cov + 0.05 * cov.max())
    d = 3.0 * G.sum(0) / (1.0 + f0.astype(np.float64)) / 32 * w
    want = np.where(m, f0 + d, f0)
    assert rel_l2(f1 - f0, want - f0) <= 1e-5
