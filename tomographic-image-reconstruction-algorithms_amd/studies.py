"""Convergence studies of the benchmark configurations (run through
``bench.py --study``).

- ``c2``: non-uniform weighted SIRT with per-pixel coverage weights against
  sequential ART on the 256^2 problem (64 transducers, 2000 steps, 3-6
  inclusions, 160^2 imaging region), with additive Gaussian noise of 0 and
  10 % of each trace's RMS; 30 iterations each at fixed relaxations (SIRT:
  1/4 of the misfit-optimal first step; ART: the same per-shot step
  divided by the shot count so one ART sweep moves the field as far as one
  SIRT step).  The Ricker peak frequency is 0.1 (10 cells per wavelength)
  unless given (``bench.py --study c2 --fp 0.05``).
- ``ttt``: time to target image error, learned vs plain SIRT, on the 512^2,
  128-source, 4000-step problem with a Ricker source of peak frequency 0.05
  (20 cells per wavelength; at 10 cells per wavelength the +-8 % smooth
  medium, about 50 wavelengths across, leaves plain SIRT from f = 0 stuck
  near 100 % error).  Relaxation 0.3 x the misfit-optimal first step.  Plain
  batched SIRT runs 50 iterations; two targets come from its error
  trajectory: the fixed one, its error after 50 iterations, and the
  reference loop's, its error where the stop rule would end it (the first
  iteration whose error exceeds the previous one, sirt-mimo-io.cu:379-393;
  the same as the fixed target when the error never rises).  A third,
  easier target, its error after 10 iterations, is reported beside them to
  show where the learned warm start helps.  Plain SIRT's time to a target is
  its wall time to the first iteration at or below it.
  The learned run takes 3 learned iterations (the network of
  ``weights/preconditioner.npy``, with its misfit-checked plain fallback)
  and continues with plain iterations until it reaches the lower target or
  50 iterations in all.  A seed whose plain run never gets below 100 % is
  flagged and left out of the medians.  Times are wall seconds of each
  method's loop.  With ``--momentum B`` the learned run's plain iterations
  carry heavy-ball momentum B (``learned.run_plain``), and plain SIRT with
  the same momentum and no network runs beside it to the same targets;
  ``--learned-alpha`` sets the learned iterations' mix of network and plain
  step.
- ``train``: the training that produced ``weights/preconditioner.npy``:
  2000 seeded 256^2 phantoms (seeds 7000-8999; 64 transducers, 2048 steps,
  checkpoint 200, source peak frequency 0.05), 20 plain-SIRT iterations
  each, every iterate a sample whose target is the trajectory's best
  iterate (lowest error).  The trajectories run 25 phantoms per engine,
  stacked along the shot axis, with one relaxation per stack (0.3 x the
  misfit-optimal first step of the stack), and the training runs in three
  phases of 667, 667 and 666 phantoms, each continuing from the previous
  phase's parameters with 8000 Adam steps of batch 8 at learning rate 1e-3
  on its own 13,340 samples (``--stack 25 --train-iters 20 --train-steps
  8000 --phantoms 667 --train-first-seed 7000``, then 7667 and 8334 with
  ``--weights`` of the previous phase).  ``weights/preconditioner_48x50.npy``
  is the earlier network: 48 phantoms (seeds 7000-7047) x 50 iterations,
  one engine each, 4000 Adam steps (``--phantoms 48``, the defaults).
- ``relax``: how fast plain SIRT itself can go on one held-out ttt phantom:
  50 iterations each of the uniform update at 1, 2 and 3 x the ttt
  relaxation, and at 1 x with per-pixel coverage (illumination) weights of
  eps 0.05 and 0.2; with ``--momentum B1,B2,...``, the uniform update at the
  ttt relaxation with each heavy-ball momentum instead.
- ``finetune``: continues training a network on plain-SIRT trajectories at
  the evaluation size, 512^2 with 128 transducers and 4000 steps (seeds from
  8000, disjoint from the held-out 9100-9107), targets as in ``train``.
"""
from __future__ import annotations

import os
import time

import numpy as np

from . import UPDATE_MEAN, Preconditioner, add_noise
from . import learned as L
from . import workloads as W

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
WEIGHTS = os.path.join(PKG_DIR, "weights", "preconditioner.npy")
# source peak frequency and relaxation of the learned-preconditioner studies
TTT_FP = 0.05
TTT_FRAC = 0.3
TRAIN_ITERS = 50


def c2_study(iters: int = 30, seed: int = 0, fp: float = 0.1, log=None) -> dict:
    c, ii, jj, fo = W.problem("c2", seed)
    c.ricker_fp = fp
    c.ricker_t0 = 1.5 / fp
    rows = []
    for noise in (0.0, 0.10):
        e = W.engine_for(c, ii, jj, fo, coverage=True)
        if noise > 0:
            e.set_field(fo)
            meas = add_noise(e.forward(fp64=True), noise, 77 + seed)
            e.set_field(np.zeros(fo.size, np.float32))
            e.set_measured(meas)
        # relaxation: misfit-optimal first step along the plain direction
        m0 = e.gradient()
        g = np.stack([e.get_gradient(s) for s in range(e.num_shots)]).astype(np.float64).sum(0) / e.num_shots
        base = 0.01 / max(float(np.abs(g).max()), 1e-30)
        best = (m0, base)
        for k in range(12):
            e.set_field((base * 2 ** k * g).astype(np.float32) * _mask(c))
            m = e.misfit()
            if m < best[0]:
                best = (m, base * 2 ** k)
            else:
                break
        scale = 0.25 * best[1]
        for method in ("weighted+coverage", "art"):
            e.set_field(np.zeros(fo.size, np.float32))
            e.set_pixel_weights(1 if method != "art" else 0)
            t0 = time.perf_counter()
            eps = []
            for _ in range(iters):
                if method == "art":
                    eps.append(e.art_iteration(scale / e.num_shots)[1])
                else:
                    eps.append(e.sirt_iteration(scale, weighted=True, mode=UPDATE_MEAN)[1])
            secs = time.perf_counter() - t0
            rows.append({"noise": noise, "method": method, "eps": [round(v, 3) for v in eps],
                         "final_eps": round(eps[-1], 3), "best_eps": round(min(eps), 3), "seconds": round(secs, 2)})
            if log:
                log(f"c2 noise {noise:.0%} {method}: eps {eps[0]:.2f} -> {eps[-1]:.2f} (best {min(eps):.2f}) "
                    f"in {secs:.1f} s")
        e.close()
    return {"study": "c2", "iterations": iters, "ricker_fp": fp, "rows": rows}


def _mask(c):
    m = np.zeros((c.ny, c.nx), np.float32)
    m[c.mask_y0:c.mask_y1, c.mask_x0:c.mask_x1] = 1.0
    return m.ravel()


def train_study(out: str, phantoms: int = 48, train_steps: int = 4000, target: str = "best", rounds: int = 0,
                iters: int = TRAIN_ITERS, stack: int = 0, first_seed: int = 7000, base: str | None = None,
                momentum: float = 0.0, log=None) -> dict:
    """Reproduce the shipped network (see the module docstring; `phantoms`
    seeded phantoms from 7000, `iters` plain iterations each); writes the
    parameters (float32, numpy .npy) to `out`.  rounds > 0 adds that many
    rounds of the learned iteration's own states (target: the truth) after
    the plain trajectories.  stack > 0 runs the trajectories `stack`
    phantoms per engine, stacked along the shot axis (target "best", one
    relaxation per stack).  base: parameters to continue from (training in
    phases over consecutive seed ranges, first_seed on).  momentum > 0 (stacked
    only) runs the trajectories with heavy-ball momentum."""
    net = Preconditioner(device=0, seed=1)
    if base:
        net.set_params(np.load(base, allow_pickle=False))
    t = time.perf_counter()
    seeds = range(first_seed, first_seed + phantoms)
    if stack > 0:
        if target != "best" or rounds:
            raise ValueError("stacked training takes target 'best' and no rounds")
        rep = L.train_stacked(net, 256, seeds, iters, stack=stack, frac=TTT_FRAC, train_steps=train_steps, batch=8,
                              lr=1e-3, fp=TTT_FP, momentum=momentum, log=log)
    else:
        if momentum:
            raise ValueError("momentum trajectories take stacked training")
        rep = L.train(net, 256, seeds, iters, frac=TTT_FRAC, rounds=rounds, train_steps=train_steps, batch=8,
                      lr=1e-3, target=target, fp=TTT_FP, log=log)
    np.save(out, net.get_params(), allow_pickle=False)
    net.close()
    return {"study": "train", "out": out, "base": base, "first_seed": first_seed, "phantoms": rep.phantoms,
            "iterations": iters, "stack": stack, "momentum": momentum,
            "samples": rep.samples,
            "loss_first": rep.losses[0], "loss_last50": float(np.mean(rep.losses[-50:])),
            "seconds": {k: round(v, 2) for k, v in rep.seconds.items()}, "total_seconds": round(time.perf_counter() - t, 2)}


def finetune_study(out: str, base: str | None = None, phantoms: int = 12, iters: int = 20,
                   train_steps: int = 2000, lr: float = 3e-4, log=None) -> dict:
    """Continue training the network of `base` (default: the shipped one) on
    `phantoms` seeded 512^2 phantoms (seeds from 8000) x `iters` plain-SIRT
    iterations; writes the parameters to `out`."""
    n = 512
    net = Preconditioner(device=0, seed=1)
    net.set_params(np.load(base or WEIGHTS, allow_pickle=False))
    t = time.perf_counter()
    rep = L.train(net, n, range(8000, 8000 + phantoms), iters, steps=4000, shots=128, frac=TTT_FRAC, rounds=0,
                  train_steps=train_steps, batch=4, lr=lr, target="best", fp=TTT_FP, log=log)
    np.save(out, net.get_params(), allow_pickle=False)
    net.close()
    return {"study": "finetune", "out": out, "base": base or WEIGHTS, "phantoms": rep.phantoms,
            "samples": rep.samples, "loss_first": rep.losses[0], "loss_last50": float(np.mean(rep.losses[-50:])),
            "seconds": {k: round(v, 2) for k, v in rep.seconds.items()},
            "total_seconds": round(time.perf_counter() - t, 2)}


def ttt_study(plain_iters: int = 50, learned_iters: int = 3, frac: float = TTT_FRAC, seed: int = 9100,
              weights: str | None = None, fp: float = TTT_FP, momentum: float = 0.0, alpha: float = 1.0,
              log=None) -> dict:
    """Plain SIRT runs `plain_iters` iterations and sets the two targets; the
    learned run takes `learned_iters` learned iterations and continues with
    plain ones until it reaches the lower target (see the module docstring).
    With `momentum`, the learned run's plain iterations carry that heavy-ball
    momentum, and plain SIRT with the same momentum (no network) runs beside
    it to the same target, so that the network's share of the gain shows."""
    n = 512
    net = Preconditioner(device=0)
    net.set_params(np.load(weights or WEIGHTS, allow_pickle=False))
    p = L.Problem(n, L.phantom(n, seed), steps=4000, shots=128, checkpoint=200, fp=fp)
    t = time.perf_counter()
    scale = L.calibrate_scale(p, frac)
    t_cal = time.perf_counter() - t
    p.reset()
    # plain SIRT needs no coverage map: it runs without one, the learned
    # iterations (whose network reads it) with
    p.engine.enable_coverage(False)
    tp = L.run_plain(p, scale, plain_iters)
    eps_p = tp.eps
    k_rise = next((k for k in range(1, len(eps_p)) if eps_p[k] > eps_p[k - 1]), None)
    k_stop = k_rise - 1 if k_rise is not None else len(eps_p) - 1
    targets = {"fixed50": eps_p[-1], "stop_rule": eps_p[k_stop], "fixed10": eps_p[min(10, len(eps_p)) - 1]}
    never_below_100 = min(eps_p) >= 100.0
    if log:
        log(f"plain: eps {eps_p[0]:.2f} -> {eps_p[-1]:.3f} in {len(eps_p)} iterations ({tp.seconds[-1]:.1f} s); "
            f"stop rule at iteration {k_stop + 1} (eps {targets['stop_rule']:.3f})"
            + ("; never below 100 %" if never_below_100 else ""))
    p.reset()
    p.engine.enable_coverage(True)
    t0 = time.perf_counter()
    goal = min(targets["fixed50"], targets["stop_rule"])
    tl = L.run_learned(net, p, scale, learned_iters, alpha=alpha, target=goal)
    eps, secs, acc = list(tl.eps), list(tl.seconds), list(tl.accepted)
    if eps[-1] > goal and len(eps) < plain_iters:
        base = time.perf_counter() - t0
        p.engine.enable_coverage(False)
        tq = L.run_plain(p, scale, plain_iters - len(eps), target=goal, momentum=momentum)
        eps += tq.eps
        secs += [base + s for s in tq.seconds]
        acc += [True] * len(tq.eps)
    tm = None
    if momentum:
        p.reset()
        p.engine.enable_coverage(False)
        tm = L.run_plain(p, scale, plain_iters, target=goal, momentum=momentum)
    p.close()
    net.close()
    res = {}
    for name, tgt in targets.items():
        k_p = next(i for i, e in enumerate(eps_p) if e <= tgt)
        k_l = next((i for i, e in enumerate(eps) if e <= tgt), None)
        t_p, t_l = tp.seconds[k_p], (secs[k_l] if k_l is not None else None)
        res[name] = {"target_eps": round(tgt, 4), "plain_iterations": k_p + 1, "plain_seconds": round(t_p, 2),
                     "learned_iterations": None if k_l is None else k_l + 1,
                     "learned_seconds": None if t_l is None else round(t_l, 2),
                     "speedup": None if (t_l is None or never_below_100) else round(t_p / t_l, 2)}
        if tm is not None:
            k_m = next((i for i, e in enumerate(tm.eps) if e <= tgt), None)
            t_m = tm.seconds[k_m] if k_m is not None else None
            res[name]["momentum_plain_iterations"] = None if k_m is None else k_m + 1
            res[name]["momentum_plain_seconds"] = None if t_m is None else round(t_m, 2)
            res[name]["momentum_plain_speedup"] = (None if (t_m is None or never_below_100)
                                                   else round(t_p / t_m, 2))
        if log:
            log(f"target {name} {tgt:.3f}: plain {k_p + 1} iterations {t_p:.1f} s, learned "
                + (f"{k_l + 1} iterations {t_l:.1f} s ({t_p / t_l:.2f}x)" if t_l is not None else "not reached")
                + ("" if tm is None else f"; plain with momentum {momentum:g}: "
                   + (f"{res[name]['momentum_plain_iterations']} iterations "
                      f"({res[name]['momentum_plain_speedup']}x)" if res[name]["momentum_plain_seconds"] is not None
                      else "not reached")))
    return {
        "study": "ttt", "grid": n, "shots": 128, "time_steps": 4000, "ricker_fp": fp, "phantom_seed": seed,
        "relaxation": scale, "relaxation_frac": frac, "calibration_seconds": round(t_cal, 2), "momentum": momentum,
        "learned_alpha": alpha,
        "plain_never_below_100": never_below_100,
        "plain_stop_iteration": k_stop + 1, "plain_error_rose": k_rise is not None,
        "plain_eps": [round(v, 3) for v in eps_p],
        "learned_eps": [round(v, 3) for v in eps], "learned_steps_accepted": acc[:learned_iters],
        "momentum_plain_eps": None if tm is None else [round(v, 3) for v in tm.eps],
        "targets": res,
        "speedup_fixed50": res["fixed50"]["speedup"], "speedup_stop_rule": res["stop_rule"]["speedup"],
        "speedup_fixed10": res["fixed10"]["speedup"],
    }


def relax_study(seed: int = 9101, iters: int = 50, fp: float = TTT_FP, momenta=None, log=None) -> dict:
    """Plain SIRT on the ttt problem of `seed` at several relaxations and with
    coverage weights (see the module docstring): the error trajectory of each
    variant and its wall seconds.  With `momenta`, the variants are instead
    the uniform update at the ttt relaxation with each heavy-ball momentum
    (``learned.run_plain``)."""
    n = 512
    p = L.Problem(n, L.phantom(n, seed), steps=4000, shots=128, checkpoint=200, fp=fp)
    scale = L.calibrate_scale(p, TTT_FRAC)
    rows = []
    variants = [("uniform x1", 0, 0.05, 1.0, 0.0), ("coverage eps 0.05", 1, 0.05, 1.0, 0.0),
                ("coverage eps 0.2", 1, 0.2, 1.0, 0.0), ("uniform x2", 0, 0.05, 2.0, 0.0),
                ("uniform x3", 0, 0.05, 3.0, 0.0)]
    if momenta:
        variants = [(f"uniform x1, momentum {b:g}", 0, 0.05, 1.0, float(b)) for b in momenta]
    for name, weights, eps_w, mult, beta in variants:
        p.reset()
        p.engine.enable_coverage(weights == 1)
        p.engine.set_pixel_weights(weights, eps_w)
        tr = L.run_plain(p, scale * mult, iters, momentum=beta)
        rows.append({"variant": name, "eps": [round(v, 3) for v in tr.eps], "seconds": round(tr.seconds[-1], 2)})
        if log:
            log(f"{name}: eps after 10 / 25 / {iters} iterations {tr.eps[9]:.2f} / {tr.eps[24]:.2f} / "
                f"{tr.eps[-1]:.2f} in {tr.seconds[-1]:.1f} s")
    p.close()
    return {"study": "relax", "grid": n, "shots": 128, "time_steps": 4000, "ricker_fp": fp, "phantom_seed": seed,
            "relaxation": scale, "relaxation_frac": TTT_FRAC, "rows": rows}
