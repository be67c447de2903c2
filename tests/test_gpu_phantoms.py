"""Phantoms stacked along the shot axis (tomo_set_phantoms): one context
holds several phantoms of one geometry, and every stencil launch advances
the shots of all of them.  Each phantom must get bitwise what a context of
its own computes: the FP64 data, the per-shot gradients, the weighted
update's weights and field, the coverage maps and the error, at every lane
width of the blocked kernels (nx = 128, 192, 256, 384, 512), eagerly and
through graph replay."""
import numpy as np
import pytest

from _cases import unit_config, unit_phantom
from _pkg import tomo

pytestmark = pytest.mark.gpu

K = 3  # phantoms per stacked context


def geometry(n):
    shots = max(n // 16, 8)
    ii, jj = tomo.transducers_ring(n, n, shots, 0.39 * n)
    return ii, jj


def config(n):
    c = unit_config(n, 2 * n + 2)
    c.checkpoint_every = 100
    return c


def single(n, fo, coverage):
    c = config(n)
    ii, jj = geometry(n)
    e = tomo.Engine(c, device=0)
    e.set_transducers(ii, jj, one_based=False)
    e.set_receivers(ii, jj, one_based=False)
    e.set_groups(1)
    if coverage:
        e.enable_coverage(True)
        e.set_pixel_weights(1, 0.05)
    e.set_truth(fo)
    e.set_field(fo)
    meas = e.forward(fp64=True)
    e.set_measured(meas)
    e.set_field((0.4 * fo).astype(np.float32))
    return e, meas


def stacked(n, fos, meas, coverage, graphs):
    c = config(n)
    ii, jj = geometry(n)
    e = tomo.Engine(c, device=0)
    e.set_phantoms(len(fos))
    e.set_transducers(ii, jj, one_based=False)
    e.set_receivers(ii, jj, one_based=False)
    e.set_groups(1)
    if coverage:
        e.enable_coverage(True)
        e.set_pixel_weights(1, 0.05)
    if graphs:
        e.set_graphs(True)
    truth = np.concatenate(fos)
    e.set_truth(truth)
    e.set_field(truth)
    data = e.forward(fp64=True)
    e.set_measured(np.concatenate(meas))
    e.set_field(np.concatenate([(0.4 * fo).astype(np.float32) for fo in fos]))
    return e, data


@pytest.mark.parametrize("n,coverage,graphs", [(128, False, False), (192, False, True), (256, True, False),
                                               (384, False, False), (512, False, True)])
def test_stacked_phantoms_match_single_contexts(n, coverage, graphs):
    print(f"[case] stacked phantoms n={n} coverage={coverage} graphs={graphs}", flush=True)
    c = config(n)
    fos = [unit_phantom(c, seed=40 + p) for p in range(K)]
    alone = []
    meas = []
    st = None
    try:
        for fo in fos:
            e, m = single(n, fo, coverage)
            alone.append(e)
            meas.append(m)
        st, data = stacked(n, fos, meas, coverage, graphs)
        S = alone[0].num_shots
        assert st.num_shots == K * S
        # the FP64 solver over the stacked coefficient images
        np.testing.assert_array_equal(data, np.concatenate(meas))
        for it in range(2):
            mis_alone = [e.gradient() for e in alone]
            mis = st.gradient()
            assert abs(mis - sum(mis_alone)) <= 1e-12 * sum(mis_alone), (it, mis, mis_alone)
            for p in range(K):
                for g in (0, S // 2, S - 1):
                    np.testing.assert_array_equal(st.get_gradient(p * S + g), alone[p].get_gradient(g))
            if coverage:
                cov = st.get_coverage().reshape(K, -1)
                for p in range(K):
                    np.testing.assert_array_equal(cov[p], alone[p].get_coverage())
            w = st.update_weighted(2e3)
            for p in range(K):
                np.testing.assert_array_equal(w[p * S:(p + 1) * S], alone[p].update_weighted(2e3))
            f = st.get_field().reshape(K, -1)
            eps = st.error()
            for p in range(K):
                np.testing.assert_array_equal(f[p], alone[p].get_field())
                assert eps[p] == alone[p].error()
    finally:
        for e in alone:
            e.close()
        if st is not None:
            st.close()


def test_stacked_uniform_iteration_and_refusals():
    print("[case] stacked phantoms: uniform SIRT iteration, one-phantom calls refused", flush=True)
    n = 128
    c = config(n)
    fos = [unit_phantom(c, seed=60 + p) for p in range(2)]
    alone, meas = [], []
    st = None
    try:
        for fo in fos:
            e, m = single(n, fo, False)
            alone.append(e)
            meas.append(m)
        st, _ = stacked(n, fos, meas, False, False)
        _, eps = st.sirt_iteration(5e3, weighted=False, mode=tomo.UPDATE_MEAN)
        for p in range(2):
            _, ep = alone[p].sirt_iteration(5e3, weighted=False, mode=tomo.UPDATE_MEAN)
            assert eps[p] == ep
            np.testing.assert_array_equal(st.get_field().reshape(2, -1)[p], alone[p].get_field())
        # each phantom at its own relaxation (tomo_update_*_each)
        scales = np.array([3e3, 7e3], np.float32)
        st.gradient()
        st.update_uniform(scales, tomo.UPDATE_MEAN)
        st.gradient()
        w = st.update_weighted(scales)
        for p in range(2):
            alone[p].gradient()
            alone[p].update_uniform(float(scales[p]), tomo.UPDATE_MEAN)
            alone[p].gradient()
            wp = alone[p].update_weighted(float(scales[p]))
            S = alone[p].num_shots
            np.testing.assert_array_equal(w[p * S:(p + 1) * S], wp)
            np.testing.assert_array_equal(st.get_field().reshape(2, -1)[p], alone[p].get_field())
        with pytest.raises(ValueError):
            st.update_uniform(np.ones(3, np.float32))
        st.gradient()
        for call in (lambda: st.update_group(0, 1.0), lambda: st.art_iteration(1.0),
                     lambda: st.line_search(False, 0.0, 100.0, 10)):
            with pytest.raises(tomo.TomoError):
                call()
    finally:
        for e in alone:
            e.close()
        if st is not None:
            st.close()


def test_stacked_reference_preset_and_passes_across_phantoms():
    """The reference preset (sensor lines, one-way edges, the scalar and
    one-step kernels) stacked, with passes of 5 shots that cut across the
    phantoms' 24 shots (groups of 8 transducers): each phantom's gradients,
    misfits and uniform update equal a context of its own, bitwise."""
    from _cases import ref_config, ref_geometry
    print("[case] stacked phantoms, reference preset, passes across phantoms", flush=True)
    n, nt, G = 64, 201, 8
    c = ref_config(n, nt)
    c.shots_per_pass = 5
    ii, jj, _ = ref_geometry(c, G)
    rng = np.random.default_rng(12)
    fos = []
    for p in range(3):
        fo = np.zeros((n, n), np.float32)
        yy, xx = np.mgrid[0:n, 0:n]
        cx, cy = rng.uniform(0.4 * n, 0.6 * n, 2)
        fo[(xx - cx) ** 2 + (yy - cy) ** 2 <= (0.1 * n) ** 2] = rng.uniform(0.02, 0.06)
        fos.append(fo.ravel())
    alone = []
    st = tomo.Engine(c, device=0)
    try:
        meas = []
        for fo in fos:
            e = tomo.Engine(c, device=0)
            e.set_transducers(ii, jj, one_based=True)
            e.set_groups(G)
            e.set_truth(fo)
            e.set_field(fo)
            meas.append(e.forward(fp64=True))
            e.set_measured(meas[-1])
            e.set_field(np.zeros(n * n, np.float32))
            alone.append(e)
        st.set_phantoms(len(fos))
        st.set_transducers(ii, jj, one_based=True)
        S = alone[0].num_shots
        assert S == 24
        assert st.set_groups(G) == len(fos) * S
        st.set_truth(np.concatenate(fos))
        st.set_field(np.concatenate(fos))
        np.testing.assert_array_equal(st.forward(fp64=True), np.concatenate(meas))
        st.set_measured(np.concatenate(meas))
        st.set_field(np.zeros(len(fos) * n * n, np.float32))
        for it in range(2):
            mis = st.gradient()
            mis_alone = [e.gradient() for e in alone]
            assert abs(mis - sum(mis_alone)) <= 1e-12 * sum(mis_alone)
            for p in range(len(fos)):
                for g in range(S):
                    np.testing.assert_array_equal(st.get_gradient(p * S + g), alone[p].get_gradient(g))
            st.update_uniform(2e4, tomo.UPDATE_SUM)
            for e in alone:
                e.update_uniform(2e4, tomo.UPDATE_SUM)
            f = st.get_field().reshape(len(fos), -1)
            eps = st.error()
            for p in range(len(fos)):
                np.testing.assert_array_equal(f[p], alone[p].get_field())
                assert eps[p] == alone[p].error()
    finally:
        st.close()
        for e in alone:
            e.close()


def test_phantom_arguments_are_checked():
    """Bad phantom counts and indices are refused with an error, and a
    context keeps working after them."""
    print("[case] stacked phantoms: argument checks", flush=True)
    n = 128
    c = config(n)
    ii, jj = geometry(n)
    with tomo.Engine(c, device=0) as e:
        for bad in (0, -1, 70000):
            with pytest.raises(tomo.TomoError):
                e.set_phantoms(bad)
        e.set_phantoms(2)
        e.set_transducers(ii, jj, one_based=False)
        e.set_receivers(ii, jj, one_based=False)
        assert e.set_groups(1) == 2 * len(ii)
        fos = [unit_phantom(c, seed=80 + p) for p in range(2)]
        e.set_truth(np.concatenate(fos))
        e.set_field(np.concatenate(fos))
        e.set_measured(e.forward(fp64=True))
        e.set_field(np.zeros(2 * n * n, np.float32))
        e.gradient()
        with tomo.Preconditioner(device=0, seed=1) as net:
            for bad in (-1, 2):
                with pytest.raises(tomo.TomoError):
                    net.add_sample(e, 1.0, phantom=bad)
            with pytest.raises(tomo.TomoError):
                net.add_sample(e, 1.0)  # the one-phantom call on a stack
            net.add_sample(e, 1.0, phantom=1)
            assert net.num_samples() == 1
        # new transducers drop the groups: the shots must be set up again
        e.set_transducers(ii, jj, one_based=False)
        with pytest.raises(tomo.TomoError):
            e.gradient()
