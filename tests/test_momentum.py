"""Heavy-ball momentum of the plain SIRT loop (learned.run_plain), CPU: the
loop is host logic over the engine's field, so a stand-in engine whose
"iteration" is a gradient step on a diagonal quadratic pins it.  The
expected fields below are worked out by hand, not by the formula the loop
uses."""
from types import SimpleNamespace

import numpy as np
import pytest

from _pkg import tomo  # noqa: F401  (loads the package as tomo_amd)
import tomo_amd.learned as L


class Quadratic:
    """misfit(f) = 1/2 sum lam (f - t)^2; an iteration is the step
    f -= scale * lam * (f - t) and returns (misfit before the step, error
    after it).  `misfits`, when given, replaces the returned misfits."""

    def __init__(self, lam, target, misfits=None):
        self.lam = np.asarray(lam, np.float64)
        self.t = np.asarray(target, np.float64)
        self.f = np.zeros_like(self.t)
        self.misfits = list(misfits) if misfits is not None else None
        self.history = []

    def get_field(self):
        return self.f.astype(np.float32)

    def set_field(self, f):
        self.f = np.asarray(f, np.float64)

    def error(self):
        return float(100.0 * np.linalg.norm(self.f - self.t) / np.linalg.norm(self.t))

    def sirt_iteration(self, scale, weighted=False, mode=None):
        m = 0.5 * float(np.sum(self.lam * (self.f - self.t) ** 2))
        if self.misfits is not None:
            m = self.misfits[len(self.history)]
        self.history.append(self.f.copy())
        self.f = self.f - scale * self.lam * (self.f - self.t)
        return m, self.error()


def run(engine, scale, iters, momentum):
    return L.run_plain(SimpleNamespace(engine=engine), scale, iters, momentum=momentum)


def eps_of(f, t):
    return 100.0 * np.linalg.norm(np.asarray(f) - t) / np.linalg.norm(t)


def test_first_steps_by_hand():
    t = np.array([1.0, 1.0])
    q = Quadratic([1.0, 0.1], t)
    tr = run(q, 0.5, 3, 0.6)
    # f0 = 0; step 1 takes no momentum: f1 = 0.5 * lam = (0.5, 0.05)
    # step 2: f1 + 0.5 lam (t - f1) = (0.75, 0.0975), plus 0.6 (f1 - f0) = (0.3, 0.03)
    # step 3 from f2 = (1.05, 0.1275): the step gives (1.025, 0.171125), plus 0.6 (f2 - f1) = (0.33, 0.0465)
    f1, f2, f3 = (0.5, 0.05), (1.05, 0.1275), (1.355, 0.217625)
    np.testing.assert_allclose(tr.eps, [eps_of(f, t) for f in (f1, f2, f3)], rtol=1e-6)
    np.testing.assert_allclose(q.f, f3, rtol=1e-6)


def test_no_momentum_is_the_plain_loop():
    t = np.array([2.0, -1.0, 0.5])
    q = Quadratic([1.0, 0.3, 0.05], t)
    tr = run(q, 0.8, 6, 0.0)
    f = np.zeros(3)
    for k in range(6):
        f = f - 0.8 * np.array([1.0, 0.3, 0.05]) * (f - t)
        assert tr.eps[k] == pytest.approx(eps_of(f, t), rel=1e-6)


def test_restart_when_the_misfit_rises():
    """At the iterate whose misfit exceeds the previous one the update is the
    plain step alone, and the momentum after it is taken from that iterate."""
    t = np.array([1.0, 1.0])
    q = Quadratic([1.0, 0.1], t, misfits=[5.0, 4.0, 4.5, 3.0])
    run(q, 0.5, 4, 0.6)
    f0, f1, f2, f3 = q.history
    step = lambda f: f - 0.5 * np.array([1.0, 0.1]) * (f - t)  # noqa: E731
    np.testing.assert_allclose(f1, step(f0), rtol=1e-6)                 # first step: no term
    np.testing.assert_allclose(f2, step(f1) + 0.6 * (f1 - f0), rtol=1e-6)  # misfit fell: term
    np.testing.assert_allclose(f3, step(f2), rtol=1e-6)                 # misfit rose: restart
    np.testing.assert_allclose(q.f, step(f3) + 0.6 * (f3 - f2), rtol=1e-6)  # fell again: term from f2


def test_momentum_on_a_slow_mode_follows_the_recurrence():
    """One slow mode (s lam = 0.005): the error e_k of the loop solves
    e_(k+1) = (1 + beta - s lam) e_k - beta e_(k-1), e_1 = (1 - s lam) e_0, whose
    solution is A r1^k + B r2^k with r1, r2 the roots of
    r^2 - (1 + beta - s lam) r + beta; without momentum it is (1 - s lam)^k e_0.
    The misfit falls at every step here, so the momentum never restarts."""
    s_lam, beta, k = 0.005, 0.8, 60
    plain = run(Quadratic([0.01], [1.0]), 0.5, k, 0.0)
    heavy = run(Quadratic([0.01], [1.0]), 0.5, k, beta)
    assert plain.eps[-1] == pytest.approx(100 * (1 - s_lam) ** k, rel=1e-5)
    r1, r2 = np.roots([1.0, -(1 + beta - s_lam), beta]).real
    e0, e1 = -1.0, -(1 - s_lam)
    B = (e1 - r1 * e0) / (r2 - r1)
    A = e0 - B
    assert heavy.eps[-1] == pytest.approx(100 * abs(A * r1 ** k + B * r2 ** k), rel=1e-4)
    assert heavy.eps[-1] < 0.3 * plain.eps[-1]


class Stack:
    """Stand-in stacked engine: phantom p is Quadratic(lam, targets[p]) with
    one common lam, so each phantom's error and misfit rise and fall
    together and both restart rules agree.  Records every field error()
    sees."""

    def __init__(self, lam, targets):
        self.lam = float(lam)
        self.t = np.asarray(targets, np.float64)
        self.f = np.zeros_like(self.t)
        self.seen = []

    def gradient(self):
        self.d = -self.lam * (self.f - self.t)
        return 0.5 * self.lam * float(np.sum((self.f - self.t) ** 2))

    def update_uniform(self, scale, mode=None):
        self.f = self.f + scale * self.d

    def error(self):
        self.seen.append(self.f.copy())
        return 100.0 * np.linalg.norm(self.f - self.t, axis=1) / np.linalg.norm(self.t, axis=1)

    def get_field(self):
        return self.f.astype(np.float32).ravel()

    def set_field(self, f):
        self.f = np.asarray(f, np.float64).reshape(self.t.shape)


class Samples:
    def __init__(self):
        self.n, self.targets = 0, {}

    def num_samples(self):
        return self.n

    def add_sample(self, e, scale, phantom=0):
        self.n += 1

    def set_targets(self, first, count, f):
        for i in range(first, first + count):
            self.targets[i] = np.asarray(f).copy()


def test_stacked_training_trajectories_with_momentum():
    """collect_plain_stacked with momentum: each phantom's trajectory is the
    one run_plain takes on that phantom alone (restarts included), and every
    sample's target is the phantom's best iterate."""
    targets = [[1.0, 2.0, -1.0], [0.5, -0.25, 3.0]]
    st, net = Stack(0.9, targets), Samples()
    L.collect_plain_stacked(net, st, 2, 1.6, 8, momentum=0.7)
    seen = np.stack(st.seen[1:])  # the first error() is of the starting field
    for p, t in enumerate(targets):
        q = Quadratic([0.9] * 3, t)
        tr = run(q, 1.6, 8, 0.7)
        alone = np.stack(q.history[1:] + [q.f])
        np.testing.assert_allclose(seen[:, p], alone, rtol=1e-6, atol=1e-7)
        best = alone[int(np.argmin(tr.eps))]
        for k in range(8):
            np.testing.assert_allclose(net.targets[2 * k + p], best, rtol=1e-6, atol=1e-7)
    # s lam = 1.44 overshoots: the error rises somewhere, so a restart is exercised
    assert any(b > a for a, b in zip(tr.eps, tr.eps[1:]))
