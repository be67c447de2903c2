"""GPU parity of the HIP engine (through the C ABI) against the CPU oracle on
the same seeded inputs.

Tolerances (relative L2 against the FP64 oracle) come from the FP32
arithmetic itself: the oracle run in FP32 differs from its FP64 run by up to
1e-4 on traces, 3.3e-3 on a single shot's gradient (the time sum
of z * Lu cancels heavily) and 3e-4 on the summed gradient; the engine is
allowed twice that.  The FP64 solver must agree to the float rounding of
its returned traces (1e-7).  Checkpointed,
full-history and shot-pass runs of the engine must be bitwise identical."""
import numpy as np
import pytest

import oracle as orc
from _cases import (rel_l2, ref_config, ref_geometry, ref_phantom, unit_config, unit_geometry,
                    unit_phantom)
from _pkg import tomo

pytestmark = pytest.mark.gpu

TR32 = 2e-4      # traces, FP32 engine vs FP64 oracle
TR64 = 1e-7      # traces, FP64 engine vs FP64 oracle: the traces come back as float, so
                  # float rounding (6e-8) bounds it; the FP32 solver lands at ~3e-6
G_SHOT = 8e-3    # per-shot gradient
G_SUM = 1e-3     # summed gradient
MISFIT = 1e-4


def say(*a):
    print("[case]", *a, flush=True)


def ref_engine(n, nt, G, field=None, **cfg_over):
    c = ref_config(n, nt)
    for k, v in cfg_over.items():
        setattr(c, k, v)
    e = tomo.Engine(c)
    ii, jj, bbox = ref_geometry(c, G)
    e.set_transducers(ii, jj, one_based=True)
    ng = e.set_groups(G)
    assert ng == len(bbox) // 4
    if field is not None:
        e.set_field(field)
    return e, c, bbox


def unit_engine(n, nt, count=16, **cfg_over):
    c = unit_config(n, nt)
    for k, v in cfg_over.items():
        setattr(c, k, v)
    e = tomo.Engine(c)
    ii, jj, bbox = unit_geometry(c, count)
    e.set_transducers(ii, jj, one_based=False)
    e.set_receivers(ii, jj, one_based=False)
    e.set_groups(1)
    return e, c, bbox, ii, jj


def unit_receivers(c, ii, jj):
    cells = (ii + c.nx * jj).astype(np.int32)
    return cells, cells.copy()


# ------------------------------------------------------------------ forward

def test_receivers_match_oracle():
    say("receivers ref 192")
    e, c, _ = ref_engine(192, 401, 160)
    ri, rj = e.receivers()
    rc, _ = orc.receivers_lines(c)
    np.testing.assert_array_equal(ri + c.nx * rj, rc)


@pytest.mark.parametrize("n,nt,G", [(64, 201, 8), (192, 401, 160)])
def test_forward_ref_preset(n, nt, G):
    say("forward ref", n, nt, G)
    c = ref_config(n, nt)
    fo = ref_phantom(c)
    e, c, bbox = ref_engine(n, nt, G, field=fo)
    rc, _ = orc.receivers_lines(c)
    want = orc.forward(c, fo, bbox, rc, np.float64)
    got32 = e.forward(fp64=False)
    got64 = e.forward(fp64=True)
    assert rel_l2(got32, want) < TR32, rel_l2(got32, want)
    assert rel_l2(got64, want) < TR64, rel_l2(got64, want)
    # shot by shot, and the zero levels below record_first
    for s in range(want.shape[0]):
        assert rel_l2(got32[s], want[s]) < 2 * TR32
    assert not got32[:, :2].any()


def test_forward_unit_preset_sponge_ricker():
    say("forward unit 96")
    e, c, bbox, ii, jj = unit_engine(96, 400)
    fo = unit_phantom(c)
    e.set_field(fo)
    rc, _ = unit_receivers(c, ii, jj)
    want = orc.forward(c, fo, bbox, rc, np.float64)
    got32 = e.forward(fp64=False)
    got64 = e.forward(fp64=True)
    assert np.abs(want).max() > 0
    assert rel_l2(got32, want) < TR32, rel_l2(got32, want)
    assert rel_l2(got64, want) < TR64, rel_l2(got64, want)


# ----------------------------------------------------------------- gradient

@pytest.mark.parametrize("n,nt,G", [(64, 201, 8), (192, 401, 32)])
def test_gradient_ref_preset(n, nt, G):
    say("gradient ref", n, nt, G)
    c = ref_config(n, nt)
    fo = ref_phantom(c)
    e, c, bbox = ref_engine(n, nt, G)
    rc, ra = orc.receivers_lines(c)
    meas = orc.forward(c, fo, bbox, rc, np.float64)
    f0 = np.zeros(fo.size)
    Gw, mis_w = orc.gradient(c, f0, bbox, rc, ra, meas, np.float64)
    e.set_measured(meas.astype(np.float32))
    mis = e.gradient()
    Gg = np.stack([e.get_gradient(g) for g in range(e.num_shots)])
    per_shot = [rel_l2(Gg[s], Gw[s]) for s in range(len(Gw))]
    assert max(per_shot) < G_SHOT, per_shot
    assert rel_l2(Gg.sum(0), Gw.sum(0)) < G_SUM
    assert abs(mis - mis_w.sum()) < MISFIT * mis_w.sum()


def test_gradient_unit_preset_with_coverage():
    say("gradient unit 96")
    e, c, bbox, ii, jj = unit_engine(96, 400)
    fo = unit_phantom(c)
    rc, ra = unit_receivers(c, ii, jj)
    meas = orc.forward(c, fo, bbox, rc, np.float64)
    f0 = np.zeros(fo.size)
    Gw, mis_w, cov_w = orc.gradient(c, f0, bbox, rc, ra, meas, np.float64, coverage=True)
    e.enable_coverage(True)
    e.set_measured(meas.astype(np.float32))
    mis = e.gradient()
    Gg = np.stack([e.get_gradient(g) for g in range(e.num_shots)])
    assert max(rel_l2(Gg[s], Gw[s]) for s in range(len(Gw))) < G_SHOT
    assert rel_l2(Gg.sum(0), Gw.sum(0)) < G_SUM
    assert abs(mis - mis_w.sum()) < MISFIT * mis_w.sum()
    assert rel_l2(e.get_coverage(), cov_w) < 1e-4


def test_fixed_point_zero_gradient():
    say("fixed point ref 64")
    c = ref_config(64, 161)
    fo = ref_phantom(c)
    e, c, bbox = ref_engine(64, 161, 8, field=fo)
    e.set_measured(e.forward())
    assert e.gradient() == 0.0
    assert all(not e.get_gradient(g).any() for g in range(e.num_shots))


def test_checkpoint_and_shot_passes_are_bitwise_identical():
    say("checkpoint/passes ref 64")
    c = ref_config(64, 201)
    fo = ref_phantom(c)
    rc, _ = orc.receivers_lines(c)
    results = []
    for C, P in [(0, 0), (50, 0), (7, 0), (0, 3), (13, 5)]:
        e, c2, bbox = ref_engine(64, 201, 8, checkpoint_every=C, shots_per_pass=P)
        meas = orc.forward(c2, fo, bbox, rc, np.float64).astype(np.float32)
        e.set_measured(meas)
        m = e.gradient()
        results.append((m, np.stack([e.get_gradient(g) for g in range(e.num_shots)])))
        e.close()
    for m, G in results[1:]:
        assert m == results[0][0]
        np.testing.assert_array_equal(G, results[0][1])


# (vec, forward rows, adjoint rows, forward prefetch depth, adjoint prefetch depth)
SHAPES = [(1, 4, 4, 1, 1), (1, 8, 8, 1, 1), (1, 16, 4, 1, 1), (2, 4, 4, 1, 1), (2, 16, 8, 2, 4),
          (2, 32, 32, 4, 2), (4, 4, 8, 4, 4), (4, 16, 16, 4, 2), (4, 32, 4, 2, 1)]


@pytest.mark.parametrize("n", [64, 192])
def test_stencil_shapes_agree(n):
    """Every kernel shape (scalar or vectorised, rows per lane) gives the
    oracle's traces and gradients; shapes of one vector width agree bitwise."""
    say("shapes ref", n)
    nt = 161 if n == 64 else 401   # the scattered wave needs ~320 levels to come back at 192
    c = ref_config(n, nt)
    fo = ref_phantom(c)
    G = 8 if n == 64 else 32
    rc, ra = orc.receivers_lines(c)
    _, _, bbox = ref_geometry(c, G)
    meas = orc.forward(c, fo, bbox, rc, np.float64)
    Gw, _ = orc.gradient(c, np.zeros(fo.size), bbox, rc, ra, meas, np.float64)
    by_vec = {}
    for shape in SHAPES:
        vec = shape[0]
        e, c2, _ = ref_engine(n, nt, G, field=fo)
        e.set_shape(*shape)
        assert e.get_shape() == shape
        tr = e.forward()
        e.set_field(np.zeros(fo.size, np.float32))
        e.set_measured(meas.astype(np.float32))
        e.gradient()
        Gg = np.stack([e.get_gradient(g) for g in range(e.num_shots)])
        e.close()
        assert rel_l2(tr, meas) < TR32, shape
        assert rel_l2(Gg.sum(0), Gw.sum(0)) < G_SUM, shape
        if vec in by_vec:
            np.testing.assert_array_equal(tr, by_vec[vec][0])
            np.testing.assert_array_equal(Gg, by_vec[vec][1])
        else:
            by_vec[vec] = (tr, Gg)


def test_vectorised_unit_preset():
    say("shapes unit 96")
    e, c, bbox, ii, jj = unit_engine(96, 300)
    fo = unit_phantom(c)
    rc, ra = unit_receivers(c, ii, jj)
    meas = orc.forward(c, fo, bbox, rc, np.float64)
    Gw, _ = orc.gradient(c, np.zeros(fo.size), bbox, rc, ra, meas, np.float64)
    for shape in [(2, 8, 4, 2, 2), (4, 16, 16, 4, 2)]:
        e.set_shape(*shape)
        e.set_field(fo)
        assert rel_l2(e.forward(), meas) < TR32
        e.set_field(np.zeros(fo.size, np.float32))
        e.set_measured(meas.astype(np.float32))
        e.gradient()
        Gg = np.stack([e.get_gradient(g) for g in range(e.num_shots)])
        assert rel_l2(Gg.sum(0), Gw.sum(0)) < G_SUM


# ------------------------------------------------------------------ updates

def _engine_with_gradients(n=64, nt=201, G=8):
    c = ref_config(n, nt)
    fo = ref_phantom(c)
    e, c, bbox = ref_engine(n, nt, G)
    rc, ra = orc.receivers_lines(c)
    meas = orc.forward(c, fo, bbox, rc, np.float64).astype(np.float32)
    e.set_measured(meas)
    e.set_truth(fo)
    return e, c, fo


@pytest.mark.parametrize("mode", [tomo.UPDATE_SUM, tomo.UPDATE_MEAN])
def test_update_uniform_matches_oracle(mode):
    say("update uniform", mode)
    e, c, fo = _engine_with_gradients()
    f0 = (0.01 * np.random.default_rng(2).standard_normal(fo.size)).astype(np.float32)
    e.set_field(f0)
    e.gradient()
    G = np.stack([e.get_gradient(g) for g in range(e.num_shots)]).astype(np.float64)
    scale = 1e3
    e.update_uniform(scale, mode)
    want = orc.update_uniform(c, f0.astype(np.float64), G, scale, mode)
    np.testing.assert_allclose(e.get_field(), want, rtol=2e-6, atol=1e-9)


def test_update_weighted_and_group_match_oracle():
    say("update weighted / group")
    e, c, fo = _engine_with_gradients()
    e.gradient()
    G = np.stack([e.get_gradient(g) for g in range(e.num_shots)]).astype(np.float64)
    w = e.update_weighted(2e3)
    want, ww = orc.update_weighted(c, np.zeros(fo.size), G, 2e3)
    np.testing.assert_allclose(w, ww, rtol=1e-12)
    np.testing.assert_allclose(e.get_field(), want, rtol=2e-6, atol=1e-9)
    f1 = e.get_field()
    e.gradient(3, 1)
    g3 = e.get_gradient(3).astype(np.float64)
    e.update_group(3, 5e3)
    want = orc.update_group(c, f1.astype(np.float64), g3, 5e3)
    np.testing.assert_allclose(e.get_field(), want, rtol=2e-6, atol=1e-9)
    eps = e.error()
    assert abs(eps - orc.error(e.get_field().astype(np.float64), fo)) < 1e-9


def test_update_avg_weight_matches_oracle():
    """sirt-mimo-avg-weight.cu:80-203 on the engine's own gradients: the
    distances w_g, their sum and the field after two updates (the second
    from a non-zero field, which enters the update whole)."""
    say("update avg-weight")
    e, c, fo = _engine_with_gradients()
    f = np.zeros(fo.size, np.float32)
    for s in (2e3, 5e2):
        e.set_field(f)
        e.gradient()
        G = np.stack([e.get_gradient(g) for g in range(e.num_shots)]).astype(np.float32)
        w = e.update_avg_weight(s)
        want, ww = orc.update_avg_weight(c, f.copy(), G, s)
        np.testing.assert_allclose(w, ww, rtol=1e-12)
        np.testing.assert_allclose(e.get_field(), want, rtol=2e-6, atol=1e-9)
        f = e.get_field()
    assert np.isfinite(f).all()


# ----------------------------------------------------- whole iterations

@pytest.mark.parametrize("weighted", [False, True])
def test_sirt_iterations_match_oracle(weighted):
    """Three SIRT iterations at the reference preset (G = 160) with the
    scales of the reference's own convergence log."""
    say("sirt iterations ref 192 G=160 weighted", weighted)
    c = ref_config()
    fo = orc.phantom_reference()
    e, c, bbox = ref_engine(192, 401, 160)
    rc, ra = orc.receivers_lines(c)
    meas = orc.forward(c, fo, bbox, rc, np.float64)
    e.set_measured(meas.astype(np.float32))
    e.set_truth(fo)
    scales = [22500.0, 15500.0, 28200.0]
    f = np.zeros(fo.size)
    for it, s in enumerate(scales):
        Gw, _ = orc.gradient(c, f, bbox, rc, ra, meas, np.float64)
        if weighted:
            orc.update_weighted(c, f, Gw, s)
        else:
            orc.update_uniform(c, f, Gw, s, tomo.UPDATE_MEAN)
        _, eps = e.sirt_iteration(s, weighted=weighted, mode=tomo.UPDATE_MEAN)
        say("  iter", it, "eps engine", eps, "oracle", orc.error(f, fo))
        assert abs(eps - orc.error(f, fo)) < 2e-3
    assert rel_l2(e.get_field(), f) < 2e-3


def test_art_iteration_matches_oracle():
    say("art iteration ref 64")
    c = ref_config(64, 201)
    fo = ref_phantom(c)
    e, c, bbox = ref_engine(64, 201, 16)
    rc, ra = orc.receivers_lines(c)
    meas = orc.forward(c, fo, bbox, rc, np.float64)
    e.set_measured(meas.astype(np.float32))
    e.set_truth(fo)
    scale = 3e4
    f = np.zeros(fo.size)
    for g in range(len(bbox) // 4):
        Gw, _ = orc.gradient(c, f, bbox[4 * g:4 * g + 4], rc, ra, meas[g:g + 1], np.float64)
        orc.update_group(c, f, Gw[0], scale)
    _, eps = e.art_iteration(scale)
    assert rel_l2(e.get_field(), f) < 5e-3
    assert abs(eps - orc.error(f, fo)) < 5e-3


def test_data_dynamic_kat_on_engine():
    """Iteration-1 KAT of the reference log sirt-dyn-10-0.txt / sirt-norm-dyn-10-0.txt
    through the engine (FP32 GPU data, as in the reference's inverse-crime runs)."""
    import json
    import os
    say("kat ref 192 G=10")
    kat = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "data_dynamic_kat.json")))["logs"]
    fo = orc.phantom_reference().astype(np.float64)
    e, c, bbox = ref_engine(192, 401, 10, field=fo.astype(np.float32))
    meas = e.forward()
    e.set_field(np.zeros(fo.size, np.float32))
    e.set_measured(meas)
    e.gradient()
    G = np.stack([e.get_gradient(g) for g in range(e.num_shots)]).astype(np.float64)
    mask = np.zeros((192, 192), bool)
    mask[21:180, 21:180] = True
    for name, d in (("sirt-dyn-10-0.txt", G.sum(0) / len(G) * mask.ravel()),
                    ("sirt-norm-dyn-10-0.txt", orc.update_weighted(c, np.zeros(fo.size), G, 1.0)[0])):
        row = kat[name][0]
        eps = 100 * np.linalg.norm(row[3] * d - fo) / np.linalg.norm(fo)
        say("  ", name, "eps", eps, "log", row[1])
        assert abs(eps - row[1]) < 0.04
        assert abs(d @ fo / (d @ d) - row[3]) < 0.005 * row[3]


@pytest.mark.parametrize("weighted", [False, True])
def test_line_search_reproduces_data_dynamic_logs(weighted):
    """The ground-truth line search (sirt-test-weight.cu / sirt-norm-test-weight.cu)
    produced the data_dynamic logs: six iterations at G = 10 from f = 0 must
    pick the logged scales (grid 0, 100, .., 1e6) and reach the logged errors.
    Tolerances: eps within 0.05 + 0.02 (k - 1) points, scale within 1 % + 1 %
    per iteration (the reference evaluated eps in FP32 and the trajectories
    separate slowly)."""
    import json
    import os
    name = "sirt-norm-dyn-10-0.txt" if weighted else "sirt-dyn-10-0.txt"
    say("line search G=10", name)
    kat = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "data_dynamic_kat.json")))["logs"][name]
    fo = orc.phantom_reference()
    e, c, bbox = ref_engine(192, 401, 10, field=fo)
    meas = e.forward()
    e.set_field(np.zeros(fo.size, np.float32))
    e.set_measured(meas)
    e.set_truth(fo)
    for it, eps_log, _, scale_log in kat:
        e.gradient()
        scale, eps_pred = e.line_search(weighted=weighted)
        if weighted:
            e.update_weighted(scale)
        else:
            e.update_uniform(scale, tomo.UPDATE_MEAN)
        eps = e.error()
        say(f"  iter {it}: scale {scale:.0f} (log {scale_log:.0f}) eps {eps:.4f} (log {eps_log:.4f}, predicted {eps_pred:.4f})")
        assert abs(eps - eps_pred) < 1e-3
        assert abs(eps - eps_log) <= 0.05 + 0.02 * (it - 1)
        assert abs(scale - scale_log) <= (0.01 + 0.01 * it) * scale_log
