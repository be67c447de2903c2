"""Learned-preconditioner workflow above the C ABI: synthetic problems,
the plain SIRT relaxation, training on plain-SIRT trajectories, and the
learned SIRT loop with its plain fallback.

Nothing here has a reference counterpart: the reference has no learned
component.  The plain iteration it accelerates is the batched uniform SIRT
update of sirt-mimo-io.cu:355, 774-796 (f += s * mask * sum_g G_g / (1 + f) /
Ng), with a fixed relaxation s per problem.

Problem family (unit-spacing preset of ``tomo_config_unit``): an n x n grid,
n/4 ring transducers at radius 0.39 n fired one at a time, 8 n time steps by
default, a Ricker source of peak frequency fp (default 0.1, 10 cells per
wavelength), a smooth random velocity perturbation (Gaussian-filtered noise,
sigma 8 cells, peak +-8 %) plus 4 sharp inclusions inside the central imaging
mask; measured traces from the FP64 solver, optional additive noise.

Relaxation: ``calibrate_scale`` finds the step s* that minimises the data
misfit along the first plain direction (doubling search, forward-only
misfits) and returns ``frac * s*``; the same relaxation drives the plain
iteration, the trajectories the network is trained on, and the learned
iteration's fallback.

Training data at scale (``train_stacked``): phantoms of one size are stacked
along the shot axis of one engine (``Engine.set_phantoms``), so one batched
gradient per iteration advances the plain-SIRT trajectories of all of them,
and every phantom's state becomes a sample (``Preconditioner.add_sample``
with its phantom index).
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field

import numpy as np

from . import (UPDATE_MEAN, Engine, Preconditioner, add_noise, config_unit, phantom_inclusions, phantom_smooth,
               transducers_ring)


def phantom(n: int, seed: int, kind: str = "smooth") -> np.ndarray:
    """Seeded phantom inside the central imaging mask of an n x n grid.

    kind "smooth": smooth random perturbation (sigma 8 cells, peak 0.08) plus
    4 inclusions (radius 4..12 cells at n = 128, scaled with n; contrast
    0.05..0.10 either sign); kind "inclusions": 2-4 inclusions only."""
    c = config_unit(n, 3)
    region = (c.mask_x0, c.mask_x1, c.mask_y0, c.mask_y1)
    fo = np.zeros(n * n, np.float32)
    if kind == "smooth":
        phantom_smooth(fo, n, n, seed, 8.0, 0.08, region)
        phantom_inclusions(fo, n, n, seed + 1, 4, 4 * n / 128, 12 * n / 128, 0.05, 0.10, region)
    elif kind == "inclusions":
        phantom_inclusions(fo, n, n, seed, 2 + seed % 3, 4 * n / 128, 12 * n / 128, 0.05, 0.10, region)
    else:
        raise ValueError(kind)
    return fo


@dataclass
class Problem:
    """One imaging problem on its own engine context."""

    n: int
    fo: np.ndarray
    steps: int = 0            # default 8 n
    shots: int = 0            # default n / 4
    checkpoint: int = -1      # default: every 200 steps (full history below 1000 steps)
    noise: float = 0.0        # additive Gaussian noise, fraction of each trace's RMS
    noise_seed: int = 0
    fp: float = 0.1           # Ricker peak frequency (0.1: 10 cells per wavelength), delay 1.5 / fp
    device: int = 0
    coverage: bool = True
    engine: Engine = field(init=False, repr=False)
    data_seconds: float = field(init=False, default=0.0)

    def __post_init__(self):
        n = self.n
        self.steps = self.steps or 8 * n
        self.shots = self.shots or n // 4
        cfg = config_unit(n, self.steps + 2)
        cfg.ricker_fp = self.fp
        cfg.ricker_t0 = 1.5 / self.fp
        if self.checkpoint < 0:
            self.checkpoint = 200 if self.steps >= 1000 else 0
        cfg.checkpoint_every = self.checkpoint
        ii, jj = transducers_ring(n, n, self.shots, 0.39 * n)
        e = Engine(cfg, device=self.device)
        e.set_transducers(ii, jj, one_based=False)
        e.set_receivers(ii, jj, one_based=False)
        e.set_groups(1)
        if self.coverage:
            e.enable_coverage(True)
        e.set_truth(self.fo)
        e.set_field(self.fo)
        t = time.perf_counter()
        meas = e.forward(fp64=True)
        self.data_seconds = time.perf_counter() - t
        if self.noise > 0:
            meas = add_noise(meas, self.noise, self.noise_seed)
        e.set_field(np.zeros(n * n, np.float32))
        e.set_measured(meas)
        self.engine = e
        self.cfg = cfg

    def reset(self):
        self.engine.set_field(np.zeros(self.n * self.n, np.float32))

    def close(self):
        self.engine.close()

    def plain_direction(self) -> np.ndarray:
        """mask * sum_g G_g / (1 + f) / Ng from the stored gradients."""
        e, c = self.engine, self.cfg
        g = np.zeros(self.n * self.n, np.float64)
        for s in range(self.shots):
            g += e.get_gradient(s)
        f = e.get_field().astype(np.float64)
        m = np.zeros((self.n, self.n))
        m[c.mask_y0:c.mask_y1, c.mask_x0:c.mask_x1] = 1.0
        return (g / self.shots / (1.0 + f) * m.ravel()).astype(np.float32)


def calibrate_scale(p: Problem, frac: float = 0.25, first_max: float = 0.01, tries: int = 12) -> float:
    """frac x the misfit-minimising step along the first plain direction
    (doubling search from the step that moves the field by `first_max`)."""
    e = p.engine
    f0 = e.get_field()
    m_best = e.gradient()
    d = p.plain_direction()
    base = first_max / max(float(np.abs(d).max()), 1e-30)
    s_best = 0.0
    for k in range(tries):
        s = base * 2.0 ** k
        e.set_field(f0 + s * d)
        m = e.misfit()
        if m < m_best:
            m_best, s_best = m, s
        else:
            break
    e.set_field(f0)
    if s_best == 0.0:
        s_best = base
    return frac * s_best


@dataclass
class Trace:
    """Per-iteration record of a run: misfit (at the field the gradient was
    taken at), relative error after the update (%), accepted learned step,
    wall seconds since the start of the run."""

    misfit: list = field(default_factory=list)
    eps: list = field(default_factory=list)
    accepted: list = field(default_factory=list)
    seconds: list = field(default_factory=list)

    def add(self, m, e, a, t):
        self.misfit.append(float(m))
        self.eps.append(float(e))
        self.accepted.append(bool(a))
        self.seconds.append(float(t))

    def time_to(self, target: float):
        """Wall seconds until eps <= target (None if never)."""
        for e, t in zip(self.eps, self.seconds):
            if e <= target:
                return t
        return None


def run_plain(p: Problem, scale: float, iters: int, weighted: bool = False, target: float | None = None,
              stop_on_rise: bool = False, momentum: float = 0.0) -> Trace:
    """Plain SIRT for up to `iters` iterations; stops early at `target`, or
    (stop_on_rise) at the first iteration whose error exceeds the previous
    one, as the reference loop does (sirt-mimo-io.cu:379-393).

    momentum > 0 adds the heavy-ball term momentum * (f_k - f_{k-1}) after
    each update (Polyak 1964), restarted (no term) at every iterate whose data
    misfit exceeds the previous iterate's (the function-value restart of
    O'Donoghue and Candes 2015).  It reads only the misfit, never the truth."""
    tr = Trace()
    e = p.engine
    t0 = time.perf_counter()
    f_prev, m_prev = None, None
    for _ in range(iters):
        f_k = e.get_field() if momentum else None
        m, eps = e.sirt_iteration(scale, weighted=weighted, mode=UPDATE_MEAN)
        if momentum:
            if f_prev is not None and m <= m_prev:
                e.set_field(e.get_field() + np.float32(momentum) * (f_k - f_prev))
                eps = e.error()
            f_prev, m_prev = f_k, m
        rose = stop_on_rise and len(tr.eps) > 0 and eps > tr.eps[-1]
        tr.add(m, eps, True, time.perf_counter() - t0)
        if (target is not None and eps <= target) or rose:
            break
    return tr


def run_learned(net: Preconditioner, p: Problem, scale: float, iters: int, alpha: float = 1.0,
                weighted: bool = True, target: float | None = None) -> Trace:
    tr = Trace()
    net.reset()
    t0 = time.perf_counter()
    for _ in range(iters):
        m, eps, acc = net.iteration(p.engine, alpha, scale, weighted=weighted)
        tr.add(m, eps, acc, time.perf_counter() - t0)
        if target is not None and eps <= target:
            break
    return tr


def collect_plain(net: Preconditioner, p: Problem, scale: float, iters: int, target: str = "best") -> float:
    """Samples of a plain SIRT trajectory of `iters` iterations: at each
    iterate the field, the plain step and the coverage (input).  Target:
    "truth" - the true field minus the iterate; "best" - the trajectory's
    iterate of lowest error (over all iterates, the last included) minus the
    iterate, i.e. where plain SIRT itself gets to; "ahead:K" - the iterate K
    iterations later (the last one near the end) minus the iterate, so that
    one learned step stands for K plain ones.  Returns the best error."""
    e = p.engine
    first = net.num_samples()
    best_eps, best_f = None, None
    ahead = int(target.split(":")[1]) if target.startswith("ahead:") else 0
    fields = []
    for _ in range(iters):
        e.gradient()
        net.add_sample(e, scale)
        e.update_uniform(scale, UPDATE_MEAN)
        eps = e.error()
        if ahead:
            fields.append(e.get_field())
        if best_eps is None or eps < best_eps:
            best_eps, best_f = eps, e.get_field()
    if target == "best":
        net.set_targets(first, iters, best_f)
    elif ahead:
        # sample i was taken at iterate i (before its update); iterate i + K
        # is fields[i + K - 1]
        for i in range(iters):
            net.set_targets(first + i, 1, fields[min(i + ahead, iters) - 1])
    elif target != "truth":
        raise ValueError(target)
    return float(best_eps)


def collect_learned(net: Preconditioner, p: Problem, scale: float, iters: int, alpha: float = 1.0,
                    weighted: bool = True) -> None:
    """Samples of the states the learned iteration itself visits (added by
    the iteration at each state its gradient is taken at)."""
    net.reset()
    net.set_sampling(True)
    try:
        for _ in range(iters):
            net.iteration(p.engine, alpha, scale, weighted=weighted)
    finally:
        net.set_sampling(False)


@dataclass
class TrainReport:
    phantoms: int = 0
    samples: int = 0
    rounds: int = 0
    losses: list = field(default_factory=list)
    seconds: dict = field(default_factory=dict)


def train(net: Preconditioner, n: int, seeds, iters: int, steps: int = 0, shots: int = 0, frac: float = 0.25,
          rounds: int = 1, train_steps: int = 600, batch: int = 8, lr: float = 1e-3, kind: str = "smooth",
          alpha: float = 1.0, target: str = "best", fp: float = 0.1, log=None) -> TrainReport:
    """Train on plain-SIRT trajectories of phantoms `seeds` (targets as in
    collect_plain), then `rounds` more rounds that add the states of the
    learned iteration on the same phantoms (target: the truth) and continue
    training.  One problem is alive at a time (its data are regenerated per
    round).  Returns losses and timings."""
    rep = TrainReport()
    t_data = t_train = 0.0
    scales = {}
    t = time.perf_counter()
    for s in seeds:
        p = Problem(n, phantom(n, int(s), kind), steps=steps, shots=shots, fp=fp)
        scales[s] = calibrate_scale(p, frac)
        p.reset()
        best = collect_plain(net, p, scales[s], iters, target=target)
        p.close()
        if log:
            log(f"plain trajectory of phantom {s}: best eps {best:.2f}, {net.num_samples()} samples")
    t_data += time.perf_counter() - t
    t = time.perf_counter()
    rep.losses += list(net.train(train_steps, batch, lr, seed=0))
    t_train += time.perf_counter() - t
    if log:
        log(f"trained on plain trajectories: loss {rep.losses[0]:.4f} -> {np.mean(rep.losses[-50:]):.4f}")
    for r in range(rounds):
        t = time.perf_counter()
        for s in seeds:
            p = Problem(n, phantom(n, int(s), kind), steps=steps, shots=shots, fp=fp)
            collect_learned(net, p, scales[s], iters, alpha=alpha)
            p.close()
            if log:
                log(f"round {r + 1}: learned trajectory of phantom {s}: {net.num_samples()} samples")
        t_data += time.perf_counter() - t
        t = time.perf_counter()
        rep.losses += list(net.train(train_steps, batch, lr, seed=r + 1))
        t_train += time.perf_counter() - t
        if log:
            log(f"round {r + 1}: {net.num_samples()} samples, loss {np.mean(rep.losses[-50:]):.4f}")
    rep.phantoms = len(scales)
    rep.samples = net.num_samples()
    rep.rounds = rounds
    rep.seconds = {"trajectories": t_data, "training": t_train}
    return rep


def stacked_problem(n: int, seeds, steps: int = 0, shots: int = 0, checkpoint: int = -1, fp: float = 0.1,
                    kind: str = "smooth", device: int = 0) -> Engine:
    """One engine holding the phantoms `seeds` (same geometry as Problem),
    stacked along the shot axis, with their FP64 data, truths and f = 0."""
    steps = steps or 8 * n
    shots = shots or n // 4
    cfg = config_unit(n, steps + 2)
    cfg.ricker_fp = fp
    cfg.ricker_t0 = 1.5 / fp
    cfg.checkpoint_every = (200 if steps >= 1000 else 0) if checkpoint < 0 else checkpoint
    ii, jj = transducers_ring(n, n, shots, 0.39 * n)
    e = Engine(cfg, device=device)
    e.set_phantoms(len(seeds))
    e.set_transducers(ii, jj, one_based=False)
    e.set_receivers(ii, jj, one_based=False)
    e.set_groups(1)
    e.enable_coverage(True)
    truth = np.concatenate([phantom(n, int(s), kind) for s in seeds])
    e.set_truth(truth)
    e.set_field(truth)
    meas = e.forward(fp64=True)
    e.set_field(np.zeros(truth.size, np.float32))
    e.set_measured(meas)
    return e


def _plain_directions(e: Engine, count: int) -> np.ndarray:
    """mask * sum_g G_g / (1 + f) / S per phantom of a stacked engine, [count][n2]."""
    c = e.cfg
    S = e.num_shots // count
    f = e.get_field().reshape(count, -1).astype(np.float64)
    m = np.zeros((c.ny, c.nx))
    m[c.mask_y0:c.mask_y1, c.mask_x0:c.mask_x1] = 1.0
    out = np.zeros_like(f)
    for p in range(count):
        g = np.zeros(f.shape[1])
        for s in range(p * S, (p + 1) * S):
            g += e.get_gradient(s)
        out[p] = g / S / (1.0 + f[p]) * m.ravel()
    return out.astype(np.float32)


def calibrate_scale_stacked(e: Engine, count: int, frac: float = 0.25, first_max: float = 0.01,
                            tries: int = 12) -> float:
    """calibrate_scale for a stack: one step for all its phantoms, the one
    that minimises their summed data misfit along their first plain
    directions."""
    f0 = e.get_field()
    m_best = e.gradient()
    d = _plain_directions(e, count).ravel()
    base = first_max / max(float(np.abs(d).max()), 1e-30)
    s_best = 0.0
    for k in range(tries):
        s = base * 2.0 ** k
        e.set_field(f0 + s * d)
        m = e.misfit()
        if m < m_best:
            m_best, s_best = m, s
        else:
            break
    e.set_field(f0)
    if s_best == 0.0:
        s_best = base
    return frac * s_best


def collect_plain_stacked(net: Preconditioner, e: Engine, count: int, scale: float, iters: int,
                          momentum: float = 0.0) -> list:
    """collect_plain (target "best") for every phantom of a stacked engine:
    at each iterate one sample per phantom, the target the phantom's own
    best iterate.  Returns the phantoms' best errors.

    momentum > 0 runs the trajectories with heavy-ball momentum (as
    run_plain), restarted per phantom at an iterate whose image error rose
    (training phantoms have a known truth; the engine reports no misfit per
    phantom)."""
    first = net.num_samples()
    best_eps = [None] * count
    best_f = [None] * count
    # the previous iterate and the errors of it and of the current one
    f_prev, eps_prev = None, None
    eps_cur = np.atleast_1d(e.error()) if momentum else None
    for _ in range(iters):
        e.gradient()
        for p in range(count):
            net.add_sample(e, scale, phantom=p)
        f_k = e.get_field().reshape(count, -1) if momentum else None
        e.update_uniform(scale, UPDATE_MEAN)
        if momentum and f_prev is not None:
            keep = (eps_cur <= eps_prev).astype(np.float32)[:, None]
            f = e.get_field().reshape(count, -1)
            e.set_field((f + np.float32(momentum) * keep * (f_k - f_prev)).ravel())
        eps = np.atleast_1d(e.error())
        if momentum:
            f_prev, eps_prev, eps_cur = f_k, eps_cur, eps
        f = e.get_field().reshape(count, -1)
        for p in range(count):
            if best_eps[p] is None or eps[p] < best_eps[p]:
                best_eps[p], best_f[p] = float(eps[p]), f[p].copy()
    for it in range(iters):
        for p in range(count):
            net.set_targets(first + it * count + p, 1, best_f[p])
    return best_eps


def train_stacked(net: Preconditioner, n: int, seeds, iters: int, stack: int = 25, steps: int = 0, shots: int = 0,
                  frac: float = 0.25, train_steps: int = 600, batch: int = 8, lr: float = 1e-3, kind: str = "smooth",
                  fp: float = 0.1, momentum: float = 0.0, log=None) -> TrainReport:
    """train() with target "best" on plain-SIRT trajectories of many
    phantoms, `stack` phantoms per engine (one relaxation per stack)."""
    rep = TrainReport()
    seeds = [int(s) for s in seeds]
    net.reserve_samples(net.num_samples() + len(seeds) * iters, n, n)
    t = time.perf_counter()
    for i in range(0, len(seeds), stack):
        chunk = seeds[i:i + stack]
        e = stacked_problem(n, chunk, steps=steps, shots=shots, fp=fp, kind=kind)
        scale = calibrate_scale_stacked(e, len(chunk), frac)
        e.set_field(np.zeros(len(chunk) * n * n, np.float32))
        best = collect_plain_stacked(net, e, len(chunk), scale, iters, momentum=momentum)
        e.close()
        if log:
            log(f"plain trajectories of phantoms {chunk[0]}..{chunk[-1]}: relaxation {scale:.4g}, best eps "
                f"{min(best):.2f}..{max(best):.2f}, {net.num_samples()} samples")
    t_data = time.perf_counter() - t
    t = time.perf_counter()
    rep.losses = list(net.train(train_steps, batch, lr, seed=0))
    rep.seconds = {"trajectories": t_data, "training": time.perf_counter() - t}
    rep.phantoms = len(seeds)
    rep.samples = net.num_samples()
    if log:
        log(f"trained on {rep.samples} samples: loss {rep.losses[0]:.4f} -> {np.mean(rep.losses[-50:]):.4f}")
    return rep
