"""Shared problem set-ups for the GPU parity tests, the smoke check and the
bench: the same seeded inputs go to the engine and to the oracle."""
import numpy as np

import oracle as orc
from _pkg import tomo


def ref_config(n=192, nt=401):
    """Reference preset; smaller grids scale the sensor lines and mask."""
    c = tomo.TomoConfig()
    orc.config_reference(c)
    if n != 192:
        lo = max(4, round(20 * n / 192))
        hi = n - 1 - lo
        c.nx = c.ny = n
        c.line_lo, c.line_hi = lo, hi
        c.line_first, c.line_last = lo + 1, hi - 1
        c.mask_x0 = c.mask_y0 = lo + 1
        c.mask_x1 = c.mask_y1 = hi
    c.nt = nt
    return c


def ref_geometry(c, G):
    """Square transducer ring on the sensor lines (reference layout when
    n = 192): 1-based ii, jj, and the group boxes."""
    if c.nx == 192:
        ii, jj = orc.transducers_reference()
    else:
        lo, hi = c.line_lo + 1, c.line_hi + 1          # 1-based line positions
        side = hi - lo
        p = np.arange(side)
        ii = np.concatenate([lo + 1 + p, np.full(side, hi), hi - 1 - p, np.full(side, lo)])
        jj = np.concatenate([np.full(side, hi), hi - 1 - p, np.full(side, lo), lo + 1 + p])
        ii, jj = ii.astype(np.int32), jj.astype(np.int32)
    return ii, jj, orc.groups(ii, jj, G)


def ref_phantom(c, seed=0):
    if c.nx == 192:
        return orc.phantom_reference()
    rng = np.random.default_rng(seed)
    n = c.nx
    fo = np.zeros((n, n), np.float32)
    yy, xx = np.mgrid[0:n, 0:n]
    for _ in range(3):
        cx, cy = rng.uniform(0.35 * n, 0.65 * n, 2)
        r = rng.uniform(0.05 * n, 0.12 * n)
        fo[(xx - cx) ** 2 + (yy - cy) ** 2 <= r * r] = rng.uniform(0.02, 0.06)
    return fo.ravel()


def unit_config(n=96, nt=400):
    c = tomo.TomoConfig()
    orc.config_unit(c, n, nt)
    return c


def unit_geometry(c, count=16, radius=None):
    radius = 0.27 * c.nx if radius is None else radius
    ii, jj = tomo.transducers_ring(c.nx, c.ny, count, radius)
    bbox = np.stack([ii, ii, jj, jj], axis=1).astype(np.int32).ravel()
    return ii, jj, bbox


def unit_phantom(c, seed=1, count=3):
    fo = np.zeros(c.nx * c.ny, np.float32)
    region = (c.mask_x0, c.mask_x1, c.mask_y0, c.mask_y1)
    tomo.phantom_inclusions(fo, c.nx, c.ny, seed, count, 0.04 * c.nx, 0.09 * c.nx, 0.05, 0.10, region)
    return fo


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))
