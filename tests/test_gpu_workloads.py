"""The benchmark workloads against the FP64 oracle, through the C ABI, with
the kernels each workload runs by default (three forward steps per launch,
three adjoint steps at widths 256 and 512 and two otherwise, checkpoint
recompute):

  C3  512^2, 4000 steps, checkpoint every 200, smooth +-8 % velocity and four
      inclusions; 8 of the 128 ring shots (every 16th) against all 128 ring
      receivers: traces at the truth, per-shot and summed gradients and the
      misfit at a half-contrast field, and one uniform SIRT update.
  C2  256^2, 64 shots, 2000 steps, 10 %-noise data: one weighted SIRT
      iteration with per-pixel coverage weights, and one ART sweep.
  C4  sides 128, 192, 256 and 384 of the uneven set: n/4 shots, 8n steps,
      checkpoint every 128, passes of 32 shots: one gradient of every shot
      and one weighted update.  (Side 512 runs the kernels of C3.)

Measured traces come from the oracle's FP64 forward.  Every relative L2 is
printed.  The gradients' error grows with the grid and the step count in
FP32 itself: the oracle run in FP32 differs from its FP64 run by 1.0-1.4e-4
per C3 shot, but by 0.7-1.2e-3 per shot at C4's side 384.  So each test also
runs the oracle in FP32 on the same inputs, and the engine may differ from
the FP64 oracle by at most twice the FP32 oracle's own error (plus 1e-5), and
never by more than the caps:

  traces                 2e-4
  per-shot gradient      5e-3 cap, max over shots <= 2 x the FP32 oracle's
  summed gradient        5e-3 cap, <= 2 x the FP32 oracle's
  misfit                 1e-3 cap (relative), <= 2 x the FP32 oracle's
  field update           1e-3 at C3, 3e-3 at C2 / C4 (relative L2 of f_new - f)
  weights                1e-3 (relative to the largest weight)
"""
import time

import numpy as np
import pytest

import oracle as orc
from _cases import rel_l2
from _pkg import tomo

W = __import__("tomo_amd.workloads", fromlist=["problem"])

pytestmark = pytest.mark.gpu

TRACES = 2e-4
G_SHOT = 5e-3
G_SUM = 5e-3
MISFIT = 1e-3
UPDATE = 1e-3


def say(*a):
    print("[case]", *a, flush=True)


def engine(c, si, sj, ri, rj):
    e = tomo.Engine(c, device=0)
    e.set_transducers(si, sj, one_based=False)
    e.set_receivers(ri, rj, one_based=False)
    e.set_groups(1)
    return e


def boxes(si, sj):
    return np.stack([si, si, sj, sj], axis=1).astype(np.int32).ravel()


def pixel_weights(c, cov, eps=0.05):
    m = np.zeros((c.ny, c.nx), bool)
    m[c.mask_y0:c.mask_y1, c.mask_x0:c.mask_x1] = True
    m = m.ravel()
    return cov[m].mean() / (cov + eps * cov.max())


def errors(G, Gw, mis, mis_w):
    per = [rel_l2(G[s], Gw[s]) for s in range(len(Gw))]
    return max(per), float(np.median(per)), rel_l2(G.sum(0), Gw.sum(0)), abs(mis - mis_w.sum()) / mis_w.sum()


def check_gradients(tag, Gg, Gw, mis, mis_w, G32, mis32):
    """Engine vs FP64 oracle, bounded by twice the FP32 oracle's own error."""
    mx, md, tot, dm = errors(Gg, Gw, mis, mis_w)
    mx32, md32, tot32, dm32 = errors(G32.astype(np.float64), Gw, mis32.sum(), mis_w)
    say(f"  {tag}: per-shot gradient rel L2 max {mx:.2e} median {md:.2e}; summed {tot:.2e}; misfit {dm:.2e}")
    say(f"  {tag} FP32 oracle: per-shot max {mx32:.2e} median {md32:.2e}; summed {tot32:.2e}; misfit {dm32:.2e}")
    assert mx < min(G_SHOT, 2 * mx32 + 1e-5), (mx, mx32)
    assert tot < min(G_SUM, 2 * tot32 + 1e-5), (tot, tot32)
    assert dm < min(MISFIT, 2 * dm32 + 1e-5), (dm, dm32)


def oracle_gradients(c, f, bbox, cells, meas, coverage=False):
    """FP64 and FP32 oracle gradients of the same inputs."""
    out64 = orc.gradient(c, np.asarray(f, np.float64), bbox, cells, cells, meas, np.float64, coverage=coverage)
    G32, mis32 = orc.gradient(c, np.asarray(f, np.float32), bbox, cells, cells, meas.astype(np.float32),
                              np.float32)
    return out64, G32, mis32


def test_c3_blocked_kernels_match_oracle_over_4000_steps():
    say("C3 512x512, 8 of 128 shots, 4000 steps, checkpoint 200")
    c, ii, jj, fo = W.problem("c3")
    assert (c.nx, c.nt, c.checkpoint_every) == (512, 4002, 200)
    sel = np.arange(0, 128, 16)
    si, sj = ii[sel], jj[sel]
    cells = (ii + c.nx * jj).astype(np.int32)
    bbox = boxes(si, sj)
    t = time.perf_counter()
    meas = orc.forward(c, fo, bbox, cells, np.float64)
    f1 = (0.5 * fo).astype(np.float32)
    (Gw, mis_w), G32, mis32 = oracle_gradients(c, f1, bbox, cells, meas)
    say(f"  oracle: {time.perf_counter() - t:.1f} s on {orc.num_threads()} threads")
    with engine(c, si, sj, ii, jj) as e:
        assert e.time_blocking() == 3 and e.adjoint_blocking() == 3
        say("  shape", e.get_shape(), "forward steps per launch", e.time_blocking(),
            "adjoint", e.adjoint_blocking())
        e.set_field(fo)
        tr = e.forward()
        r = rel_l2(tr, meas)
        say(f"  traces at the truth: rel L2 {r:.2e}")
        assert r < TRACES
        e.set_field(f1)
        e.set_measured(meas.astype(np.float32))
        mis = e.gradient()
        Gg = np.stack([e.get_gradient(s) for s in range(8)]).astype(np.float64)
        check_gradients("C3", Gg, Gw, mis, mis_w, G32, mis32)
        scale = float(0.01 / np.abs(Gg.sum(0) / 8).max())
        e.update_uniform(scale, tomo.UPDATE_MEAN)
        f_gpu = e.get_field().astype(np.float64)
    f_orc = orc.update_uniform(c, f1.astype(np.float64), Gw, scale, tomo.UPDATE_MEAN)
    r = rel_l2(f_gpu - f1, f_orc - f1)
    say(f"  uniform update: rel L2 {r:.2e}")
    assert r < UPDATE


def test_c2_weighted_coverage_iteration_and_art_sweep_on_noisy_data():
    say("C2 256x256, 64 shots, 2000 steps, 10 % noise")
    c, ii, jj, fo = W.problem("c2")
    cells = (ii + c.nx * jj).astype(np.int32)
    bbox = boxes(ii, jj)
    t = time.perf_counter()
    clean = orc.forward(c, fo, bbox, cells, np.float64).astype(np.float32)
    meas = tomo.add_noise(clean.copy(), 0.10, 77)
    assert rel_l2(meas, clean) > 0.05
    m64 = meas.astype(np.float64)
    f0 = np.zeros(fo.size)
    (Gw, mis_w, cov_w), G32, mis32 = oracle_gradients(c, f0, bbox, cells, m64, coverage=True)
    say(f"  oracle gradient: {time.perf_counter() - t:.1f} s")
    with engine(c, ii, jj, ii, jj) as e:
        e.enable_coverage(True)
        e.set_measured(meas)
        e.set_truth(fo)
        e.set_pixel_weights(1, 0.05)
        mis = e.gradient()
        Gg = np.stack([e.get_gradient(s) for s in range(64)]).astype(np.float64)
        check_gradients("C2", Gg, Gw, mis, mis_w, G32, mis32)
        rc = rel_l2(e.get_coverage(), cov_w)
        say(f"  coverage rel L2 {rc:.2e}")
        assert rc < 1e-4
        scale = float(0.02 / np.abs(Gg.sum(0) / 64).max())
        w = e.update_weighted(scale)
        f_gpu = e.get_field().astype(np.float64)
        fw, ww = orc.update_weighted(c, f0.copy(), Gw, scale)
        f_orc = f0 + pixel_weights(c, cov_w) * (fw - f0)
        rw = float(np.abs(w - ww).max() / np.abs(ww).max())
        ru = rel_l2(f_gpu, f_orc)
        say(f"  weighted update with coverage weights: weights {rw:.2e}, field rel L2 {ru:.2e}")
        assert rw < 1e-3 and ru < 3 * UPDATE
        # one ART sweep from that field (no pixel weights, as in the C2 study)
        e.set_pixel_weights(0)
        f_start = f_gpu.astype(np.float32)
        e.set_field(f_start)
        art_scale = scale / 64
        t = time.perf_counter()
        _, eps = e.art_iteration(art_scale)
        f_gpu = e.get_field().astype(np.float64)
    t = time.perf_counter()
    f = f_start.astype(np.float64)
    for g in range(64):
        Gs, _ = orc.gradient(c, f, bbox[4 * g:4 * g + 4], cells, cells, m64[g:g + 1], np.float64)
        orc.update_group(c, f, Gs[0], art_scale)
    ra = rel_l2(f_gpu - f_start, f - f_start)
    de = abs(eps - orc.error(f, fo))
    say(f"  ART sweep (oracle {time.perf_counter() - t:.1f} s): field change rel L2 {ra:.2e}, eps diff {de:.2e}")
    assert ra < 5e-3 and de < 5e-3


@pytest.mark.parametrize("n", [128, 192, 256, 384])
def test_c4_sides_match_oracle(n):
    say(f"C4 side {n}: {n // 4} shots, {8 * n} steps, checkpoint 128, passes of 32")
    c = tomo.config_unit(n, 8 * n + 2)
    c.checkpoint_every = 128
    c.shots_per_pass = 32
    ii, jj = tomo.transducers_ring(n, n, n // 4, 0.39 * n)
    fo = W.c1_phantom(n, 4000 + n, (c.mask_x0, c.mask_x1, c.mask_y0, c.mask_y1))
    cells = (ii + c.nx * jj).astype(np.int32)
    bbox = boxes(ii, jj)
    t = time.perf_counter()
    meas = orc.forward(c, fo, bbox, cells, np.float64)
    (Gw, mis_w), G32, mis32 = oracle_gradients(c, np.zeros(fo.size), bbox, cells, meas)
    say(f"  oracle: {time.perf_counter() - t:.1f} s")
    with engine(c, ii, jj, ii, jj) as e:
        say("  shape", e.get_shape(), "forward steps per launch", e.time_blocking(),
            "adjoint", e.adjoint_blocking())
        assert e.time_blocking() == 3 and e.adjoint_blocking() == (3 if n in (256, 512) else 2)
        e.set_field(fo)
        r = rel_l2(e.forward(), meas)
        say(f"  traces rel L2 {r:.2e}")
        assert r < TRACES
        e.set_field(np.zeros(fo.size, np.float32))
        e.set_measured(meas.astype(np.float32))
        mis = e.gradient()
        Gg = np.stack([e.get_gradient(s) for s in range(n // 4)]).astype(np.float64)
        check_gradients(f"C4 {n}", Gg, Gw, mis, mis_w, G32, mis32)
        scale = float(0.02 / np.abs(Gg.sum(0) / (n // 4)).max())
        w = e.update_weighted(scale)
        f_gpu = e.get_field().astype(np.float64)
    f_orc, ww = orc.update_weighted(c, np.zeros(fo.size), Gw, scale)
    rw = float(np.abs(w - ww).max() / np.abs(ww).max())
    ru = rel_l2(f_gpu, f_orc)
    say(f"  weighted update: weights {rw:.2e}, field rel L2 {ru:.2e}")
    assert rw < 1e-3 and ru < 3 * UPDATE
