"""Synthetic workloads of the benchmarks (unit-spacing preset of
``tomo_config_unit``: h = 1, c0 = 1, dt = 0.4, 20-cell sponge, Ricker
f_p = 0.1, point receivers at the transducers).

- ``c1``: 128^2, 32 ring transducers (radius 50), 1000 steps, 2-4 inclusions
  (radius 4-12 cells, contrast +-[5, 10] %) in the central 64^2.
- ``c2``: 256^2, 64 transducers, 2000 steps, 3-6 inclusions (contrast +-10 %)
  in the central 160^2.
- ``c3``: 512^2, 128 transducers, 4000 steps, smooth random perturbation
  (Gaussian-filtered noise, sigma 8 cells, peak 8 %) plus 4 sharp
  inclusions, forward wavefield checkpointed every 200 steps.
- ``c4``: 64 phantoms, sides drawn uniformly from {128, 192, 256, 384, 512},
  n/4 transducers, 8 n steps, inclusions as in c1; checkpoint every 128
  steps.  Either one engine per phantom (32 shots per pass, or all shots in
  one pass where the engines share wavefield workspaces), or one engine per
  side with that side's phantoms stacked along the shot axis.

Ring radius 0.39 n unless stated; transducers fire one at a time; seeds are
fixed per workload.
"""
from __future__ import annotations

import numpy as np

from . import Engine, config_unit, phantom_inclusions, phantom_smooth, transducers_ring

C4_SIDES = (128, 192, 256, 384, 512)


def c1_phantom(n: int, seed: int, region) -> np.ndarray:
    fo = np.zeros(n * n, np.float32)
    phantom_inclusions(fo, n, n, seed, 2 + seed % 3, 4.0, 12.0, 0.05, 0.10, region)
    return fo


def problem(name: str, seed: int = 0):
    """(config, ii, jj, truth) of one problem of workload c1, c2 or c3."""
    if name == "c1":
        n, steps, shots, radius = 128, 1000, 32, 50.0
    elif name == "c2":
        n, steps, shots, radius = 256, 2000, 64, 0.39 * 256
    elif name == "c3":
        n, steps, shots, radius = 512, 4000, 128, 0.39 * 512
    else:
        raise ValueError(name)
    c = config_unit(n, steps + 2)
    if name == "c2":
        c.mask_x0 = c.mask_y0 = 48
        c.mask_x1 = c.mask_y1 = 208
    if name == "c3":
        c.checkpoint_every = 200
    ii, jj = transducers_ring(n, n, shots, radius)
    region = (c.mask_x0, c.mask_x1, c.mask_y0, c.mask_y1)
    fo = np.zeros(n * n, np.float32)
    if name == "c1":
        fo = c1_phantom(n, 1000 + seed, region)
    elif name == "c2":
        phantom_inclusions(fo, n, n, 2000 + seed, 3 + seed % 4, 4.0 * n / 128, 12.0 * n / 128, 0.10, 0.10, region)
    else:
        phantom_smooth(fo, n, n, 3000 + seed, 8.0, 0.08, region)
        phantom_inclusions(fo, n, n, 3001 + seed, 4, 4 * n / 128, 12 * n / 128, 0.05, 0.10, region)
    return c, ii, jj, fo


def engine_for(c, ii, jj, fo, fp64_data: bool = True, coverage: bool = False, device: int = 0):
    """Engine with the problem's geometry, truth and measured traces, field 0."""
    e = Engine(c, device=device)
    e.set_transducers(ii, jj, one_based=False)
    e.set_receivers(ii, jj, one_based=False)
    e.set_groups(1)
    if coverage:
        e.enable_coverage(True)
    e.set_truth(fo)
    e.set_field(fo)
    meas = e.forward(fp64=fp64_data)
    e.set_field(np.zeros(fo.size, np.float32))
    e.set_measured(meas)
    return e


def c4_sides(count: int = 64, seed: int = 4000):
    rng = np.random.default_rng(seed)
    return [int(C4_SIDES[i]) for i in rng.integers(0, len(C4_SIDES), count)]


def c4_problems(count: int = 64, seed: int = 4000, device: int = 0, log=None, passes: int = 32):
    """Engines of the uneven batch (data from the FP64 solver); `passes`
    shots per pass (0: all shots in one pass, for engines that share
    wavefield workspaces)."""
    engines = []
    for i, n in enumerate(c4_sides(count, seed)):
        c = config_unit(n, 8 * n + 2)
        c.checkpoint_every = 128
        c.shots_per_pass = passes
        ii, jj = transducers_ring(n, n, n // 4, 0.39 * n)
        fo = c1_phantom(n, seed + 1 + i, (c.mask_x0, c.mask_x1, c.mask_y0, c.mask_y1))
        engines.append(engine_for(c, ii, jj, fo, device=device))
        if log:
            log(f"c4 problem {i}: {n}x{n}, {n // 4} shots, {8 * n} steps")
    return engines


def stack_pass_shots(n: int, nt: int, checkpoint: int, total: int, max_pass_bytes: float) -> int:
    """Shots per pass of a stacked engine (0: all `total` in one pass): the
    wavefield storage of a pass (two checkpoint levels per segment, the
    replayed segment and the 5-level adjoint ring, n^2 floats per shot and
    level) stays within max_pass_bytes, the passes are equal, and each is a
    whole number of 128-shot rounds of waves where the storage allows (a
    512^2 pass of 896 shots: 7 full rounds of the three-step adjoint's one
    wave per SIMD)."""
    levels = 2 * (-(-nt // checkpoint)) + checkpoint + 5
    fit = max(int(max_pass_bytes // (levels * n * n * 4)), 1)
    passes = -(-total // fit)
    if passes <= 1:
        return 0
    per = -(-total // passes)
    if -(-per // 128) * 128 <= fit:
        per = -(-per // 128) * 128
    return per


def c4_stacks(count: int = 64, seed: int = 4000, device: int = 0, log=None, max_pass_bytes: float = 190e9):
    """The uneven batch as one engine per side, that side's phantoms stacked
    along the shot axis (tomo_set_phantoms: every stencil launch advances
    every shot of every phantom of the side).  Same phantoms, geometry and
    FP64 data as c4_problems.  Shots per pass: all of them, or as many as
    keep the wavefield storage (checkpoint pairs, the replayed segment, the
    adjoint ring) within max_pass_bytes.  Returns [(side, engine, phantom
    indices)]."""
    sides = c4_sides(count, seed)
    out = []
    for n in C4_SIDES:
        idx = [i for i, v in enumerate(sides) if v == n]
        if not idx:
            continue
        c = config_unit(n, 8 * n + 2)
        c.checkpoint_every = 128
        shots = n // 4
        total = shots * len(idx)
        c.shots_per_pass = stack_pass_shots(n, c.nt, 128, total, max_pass_bytes)
        ii, jj = transducers_ring(n, n, shots, 0.39 * n)
        fos = [c1_phantom(n, seed + 1 + i, (c.mask_x0, c.mask_x1, c.mask_y0, c.mask_y1)) for i in idx]
        e = Engine(c, device=device)
        e.set_phantoms(len(idx))
        e.set_transducers(ii, jj, one_based=False)
        e.set_receivers(ii, jj, one_based=False)
        e.set_groups(1)
        truth = np.concatenate(fos)
        e.set_truth(truth)
        e.set_field(truth)
        meas = e.forward(fp64=True)
        e.set_field(np.zeros(truth.size, np.float32))
        e.set_measured(meas)
        out.append((n, e, idx))
        if log:
            log(f"c4 side {n}: {len(idx)} phantoms stacked, {total} shots, {8 * n} steps, "
                f"shots per pass {c.shots_per_pass or total}")
    return out
