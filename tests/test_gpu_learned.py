"""The learned preconditioner end to end on the GPU: trajectory sampling and
training reduce the loss; the learned iteration never keeps a step that
raises the data misfit; the shipped network (weights/preconditioner.npy,
trained on 256^2 plain-SIRT trajectories) gains over plain SIRT on a
held-out phantom; and training samples taken from stacked phantoms equal
those of one context per phantom."""
import os

import numpy as np
import pytest

from _pkg import tomo
import tomo_amd.learned as L

pytestmark = pytest.mark.gpu

WEIGHTS = os.path.join(tomo.PKG_DIR, "weights", "preconditioner.npy")


def test_sampling_and_training_reduce_loss():
    print("[case] learned training 64^2", flush=True)
    with tomo.Preconditioner(device=0, seed=3) as net:
        rep = L.train(net, 64, range(100, 106), 4, steps=400, frac=0.25, rounds=0, train_steps=150, batch=4,
                      lr=2e-3, target="best")
        assert rep.samples == 24 and net.num_samples() == 24
        first = np.mean(rep.losses[:10])
        last = np.mean(rep.losses[-10:])
        print(f"  loss {first:.4f} -> {last:.4f}", flush=True)
        assert last < 0.7 * first


def test_learned_iteration_keeps_only_misfit_decreasing_steps():
    """An untrained network with a large output: every kept learned step has
    a smaller misfit than the field it started from (the next iteration's
    misfit), rejected ones fall back to the plain step."""
    print("[case] learned fallback 96^2", flush=True)
    p = L.Problem(96, L.phantom(96, 31), steps=500, shots=16)
    try:
        sc = L.calibrate_scale(p, 0.1)
        p.reset()
        with tomo.Preconditioner(device=0, seed=5) as net:
            params = net.get_params() * 3.0
            net.set_params(params)
            tr = L.run_learned(net, p, sc, 8, alpha=1.0)
        print("  misfit", ["%.4g" % m for m in tr.misfit], "accepted", tr.accepted, flush=True)
        for k in range(len(tr.misfit) - 1):
            if tr.accepted[k]:
                assert tr.misfit[k + 1] < tr.misfit[k]
        assert np.isfinite(tr.eps).all()
    finally:
        p.close()


def test_shipped_network_gains_over_plain_sirt():
    """Held-out 256^2 phantom (seed 9002) under the time-to-target study's
    protocol (source peak frequency 0.05, relaxation 0.3 x the
    misfit-optimal first step): 3 learned iterations then plain ones against
    plain SIRT, 8 iterations each; the learned run must end at least 5
    points of relative error lower."""
    import tomo_amd.studies as S
    print("[case] shipped network 256^2", flush=True)
    p = L.Problem(256, L.phantom(256, 9002), fp=S.TTT_FP)
    try:
        sc = L.calibrate_scale(p, S.TTT_FRAC)
        p.reset()
        plain = L.run_plain(p, sc, 8)
        p.reset()
        with tomo.Preconditioner(device=0) as net:
            net.set_params(np.load(WEIGHTS, allow_pickle=False))
            learned = L.run_learned(net, p, sc, 3)
        rest = L.run_plain(p, sc, 5)
        eps_l = learned.eps + rest.eps
        print("  plain  ", ["%.2f" % e for e in plain.eps], flush=True)
        print("  learned", ["%.2f" % e for e in eps_l], flush=True)
        assert eps_l[-1] < plain.eps[-1] - 5.0
    finally:
        p.close()


def test_stacked_samples_equal_single_context_samples():
    """Samples taken from a stack of phantoms (one batched gradient per
    iteration, one sample per phantom) are those of one context per phantom:
    two networks trained on them from the same initial parameters and batch
    order end bitwise equal."""
    print("[case] stacked training samples 128^2", flush=True)
    n, steps, shots, iters, seeds = 128, 256, 8, 3, [7301, 7302]
    stack = L.stacked_problem(n, seeds, steps=steps, shots=shots)
    singles = [L.Problem(n, L.phantom(n, s), steps=steps, shots=shots) for s in seeds]
    try:
        scale = L.calibrate_scale_stacked(stack, len(seeds), 0.25)
        stack.set_field(np.zeros(len(seeds) * n * n, np.float32))
        with tomo.Preconditioner(device=0, seed=9) as a, tomo.Preconditioner(device=0, seed=9) as b:
            a.reserve_samples(len(seeds) * iters, n, n)
            best = L.collect_plain_stacked(a, stack, len(seeds), scale, iters)
            best_f = [None] * len(seeds)
            best_e = [None] * len(seeds)
            for _ in range(iters):
                for p, q in enumerate(singles):
                    q.engine.gradient()
                    b.add_sample(q.engine, scale)
                for p, q in enumerate(singles):
                    q.engine.update_uniform(scale, tomo.UPDATE_MEAN)
                    e = q.engine.error()
                    if best_e[p] is None or e < best_e[p]:
                        best_e[p], best_f[p] = e, q.engine.get_field()
            for it in range(iters):
                for p in range(len(seeds)):
                    b.set_targets(it * len(seeds) + p, 1, best_f[p])
            assert best == best_e
            assert a.num_samples() == b.num_samples() == len(seeds) * iters
            la, lb = a.train(20, 4, 1e-3, seed=2), b.train(20, 4, 1e-3, seed=2)
            np.testing.assert_array_equal(la, lb)
            np.testing.assert_array_equal(a.get_params(), b.get_params())
    finally:
        stack.close()
        for q in singles:
            q.close()
