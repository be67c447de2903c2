"""ctypes wrapper of the CPU oracle (oracle/tomo_oracle.cpp).

TEST INFRASTRUCTURE ONLY.  Imported by tests/, by ``__graft_entry__.smoke()``
and by bench.py's CPU-baseline leg -- as the checker or the timed CPU
baseline, never as the product path.

Each oracle function restates a reference function (file:line in the C++
source) and is pinned as follows (tests/test_oracle.py):

=========================  ==========================================================
oracle function            pinned by
=========================  ==========================================================
orc_phantom_reference      pixel counts 599 @ 0.03 / 228 @ 0.05 (FP32 closed form),
                           committed fixture tests/golden/fo_192.txt
orc_transducers_reference  closed-form corner positions and side counts
orc_groups                 bbox of G = 1 is the transducer cell; G = 160 is a full line
orc_receivers_lines        636 receivers, anchors one cell inward
forward (orc_forward_*)    linearity in the source amplitude; mirror symmetry of a
                           symmetric source in a homogeneous field; bounded energy
                           decay after the pulse; FP32 vs FP64 agreement; the
                           data_dynamic iteration-1 KATs (whole pipeline)
adjoint (in gradient)      dot-product test <A x, y> = <x, A^T y> of the interior
                           operators (with and without sponge); fixed point f = fo
                           -> zero residual -> zero gradient; data_dynamic KATs
point receivers (record    adjoint-state dot product <traces, rr> = sum_k s(k) z(k+1)
and additive injection)    at the source cell (orc_adjoint_probe_f64), sponge on
Ricker pulse               peak 1 at t0, zeros at t0 +- 1/(sqrt(2) pi fp), zero mean,
                           spectrum peak at fp
sponge                     gamma = 0 gives the bare leapfrog bit for bit; reflected
                           energy against a wide-grid run far below the bare edge's
laplacian (in gradient)    exact 4/h^2 on u = i^2 + j^2 interior cells (via a
                           one-level probe); KATs
checkpointing              bitwise equality with the single-segment path
orc_update_uniform_*       closed form on constant gradients; KAT scale s*
orc_update_weighted_*      weights sum to 1, permutation symmetry, degenerate
                           fallback to 1/M; weighted data_dynamic KATs
orc_update_avg_weight_*    closed form for two groups (w_1 = w_2 = s ||mask (df_1 -
                           df_2)|| / 2, f += (fg_1 + fg_2) / w); permutation symmetry;
                           f enters the update whole (the reference's rule); one
                           group gives a non-finite field, as in the reference
orc_update_group_*         closed form
orc_error_*                closed form
orc_schedule               closed forms of both schedules
orc_cnn_forward            a hand-computed 1-pixel / delta-input network (closed form
                           of the zero-padded cross-correlation); per-layer shapes
orc_cnn_backward /         finite differences of orc_cnn_loss in FP64 along random
orc_cnn_loss               directions (a mirrored-tap mistake fails it); FP64 vs
                           BF16-rounding modes agree to BF16 precision
=========================  ==========================================================
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("TOMO_ORACLE_LIB") or os.path.join(HERE, "_build", "liboracle.so")

_lib = None

_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


def use_library(path: str) -> None:
    """Load the oracle from `path` (e.g. the host-native build of the timed
    CPU baseline) instead of the checker build; call before the first use."""
    global _lib, LIB_PATH
    LIB_PATH = path
    _lib = None


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (g++ + OpenMP)."""
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-C", HERE] + (["-B"] if force else []), check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        vp = ctypes.c_void_p
        L.orc_config_reference.argtypes = [vp]
        L.orc_config_unit.argtypes = [vp, ctypes.c_int, ctypes.c_int]
        L.orc_num_threads.restype = ctypes.c_int
        L.orc_set_threads.argtypes = [ctypes.c_int]
        L.orc_phantom_reference.argtypes = [_f32p, ctypes.c_int, ctypes.c_int, ctypes.c_double]
        L.orc_transducers_reference.argtypes = [_i32p, _i32p, ctypes.c_int]
        L.orc_groups.argtypes = [_i32p, _i32p, ctypes.c_int, ctypes.c_int, _i32p]
        L.orc_groups.restype = ctypes.c_int
        L.orc_receivers_lines.argtypes = [vp, _i32p, _i32p]
        L.orc_receivers_lines.restype = ctypes.c_int
        L.orc_apply_forward_op_f64.argtypes = [vp, _f64p, _f64p, _f64p]
        L.orc_apply_adjoint_op_f64.argtypes = [vp, _f64p, _f64p, _f64p]
        L.orc_laplacian_f64.argtypes = [vp, _f64p, _f64p]
        L.orc_adjoint_probe_f64.argtypes = [vp, _f64p, ctypes.c_int, _i32p, _i32p, _f64p, ctypes.c_int, _f64p]
        L.orc_schedule.argtypes = [ctypes.c_int, ctypes.c_int, _f64p]
        L.orc_schedule.restype = ctypes.c_double
        for suf, p in (("f64", _f64p), ("f32", _f32p)):
            fn = getattr(L, "orc_forward_" + suf)
            fn.argtypes = [vp, p, ctypes.c_int, _i32p, ctypes.c_int, _i32p, vp, vp]
            fn = getattr(L, "orc_gradient_" + suf)
            fn.argtypes = [vp, p, ctypes.c_int, _i32p, ctypes.c_int, _i32p, _i32p, p, p, vp, vp]
            fn = getattr(L, "orc_update_uniform_" + suf)
            fn.argtypes = [vp, p, ctypes.c_int, p, ctypes.c_double, ctypes.c_int]
            fn = getattr(L, "orc_update_weighted_" + suf)
            fn.argtypes = [vp, p, ctypes.c_int, p, ctypes.c_double, vp]
            fn = getattr(L, "orc_update_avg_weight_" + suf)
            fn.argtypes = [vp, p, ctypes.c_int, p, ctypes.c_double, vp]
            fn = getattr(L, "orc_update_group_" + suf)
            fn.argtypes = [vp, p, p, ctypes.c_double]
            fn = getattr(L, "orc_error_" + suf)
            fn.argtypes = [p, _f32p, ctypes.c_int]
            fn.restype = ctypes.c_double
            fn = getattr(L, "orc_pulse_" + suf)
            fn.argtypes = [vp, p]
        L.orc_cnn_num_params.restype = ctypes.c_int
        L.orc_cnn_forward.argtypes = [_f32p, ctypes.c_int, ctypes.c_int, ctypes.c_int, _f32p, _f32p, ctypes.c_int]
        L.orc_cnn_backward.argtypes = [_f32p, ctypes.c_int, ctypes.c_int, ctypes.c_int, _f32p, _f32p, _f64p,
                                       ctypes.c_int]
        L.orc_cnn_loss.argtypes = [_f32p, ctypes.c_int, ctypes.c_int, ctypes.c_int, _f32p, _f32p, _i32p, vp,
                                   ctypes.c_int]
        L.orc_cnn_loss.restype = ctypes.c_double
        _lib = L
    return _lib


def _dt(dtype):
    dtype = np.dtype(dtype)
    if dtype == np.float64:
        return np.float64, "f64"
    if dtype == np.float32:
        return np.float32, "f32"
    raise ValueError(dtype)


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


# ------------------------------------------------------------------ helpers

def config_reference(cfg):
    lib().orc_config_reference(ctypes.byref(cfg))
    return cfg


def config_unit(cfg, n, nt):
    lib().orc_config_unit(ctypes.byref(cfg), int(n), int(nt))
    return cfg


def set_threads(n: int) -> None:
    lib().orc_set_threads(int(n))


def num_threads() -> int:
    return int(lib().orc_num_threads())


def phantom_reference(nx=192, ny=192, h=float(np.float32(0.001))):
    fo = np.zeros(nx * ny, np.float32)
    lib().orc_phantom_reference(fo, nx, ny, h)
    return fo


def transducers_reference(ns=640):
    ii = np.zeros(ns, np.int32)
    jj = np.zeros(ns, np.int32)
    lib().orc_transducers_reference(ii, jj, ns)
    return ii, jj


def groups(ii, jj, G):
    ns = len(ii)
    bbox = np.zeros(4 * max(ns // G, 1), np.int32)
    ng = lib().orc_groups(np.ascontiguousarray(ii, np.int32), np.ascontiguousarray(jj, np.int32), ns, G, bbox)
    return bbox[: 4 * ng].copy()


def receivers_lines(cfg):
    n = 4 * (cfg.line_last - cfg.line_first + 1)
    rc = np.zeros(n, np.int32)
    ra = np.zeros(n, np.int32)
    lib().orc_receivers_lines(ctypes.byref(cfg), rc, ra)
    return rc, ra


def pulse(cfg, dtype=np.float64):
    t, suf = _dt(dtype)
    a = np.zeros(cfg.nt, t)
    getattr(lib(), "orc_pulse_" + suf)(ctypes.byref(cfg), a)
    return a


def forward(cfg, f, bbox, rcell, dtype=np.float64, history=False):
    """Traces [S][nt][nr] (and optionally the history [S][nt][ny*nx])."""
    t, suf = _dt(dtype)
    S = len(bbox) // 4
    nr = len(rcell)
    n2 = cfg.nx * cfg.ny
    tr = np.zeros((S, cfg.nt, nr), t)
    hist = np.zeros((S, cfg.nt, n2), t) if history else None
    getattr(lib(), "orc_forward_" + suf)(
        ctypes.byref(cfg), np.ascontiguousarray(f, t), S, np.ascontiguousarray(bbox, np.int32), nr,
        np.ascontiguousarray(rcell, np.int32), _ptr(tr), _ptr(hist))
    return (tr, hist) if history else tr


def gradient(cfg, f, bbox, rcell, ranchor, meas, dtype=np.float64, coverage=False):
    """Per-shot raw gradients G [S][ny*nx], misfit [S] (FP64) and optional coverage."""
    t, suf = _dt(dtype)
    S = len(bbox) // 4
    nr = len(rcell)
    n2 = cfg.nx * cfg.ny
    G = np.zeros((S, n2), t)
    mis = np.zeros(S, np.float64)
    cov = np.zeros(n2, t) if coverage else None
    getattr(lib(), "orc_gradient_" + suf)(
        ctypes.byref(cfg), np.ascontiguousarray(f, t), S, np.ascontiguousarray(bbox, np.int32), nr,
        np.ascontiguousarray(rcell, np.int32), np.ascontiguousarray(ranchor, np.int32),
        np.ascontiguousarray(meas, t), G, _ptr(mis), _ptr(cov))
    return (G, mis, cov) if coverage else (G, mis)


def update_uniform(cfg, f, G, scale, mode):
    t, suf = _dt(f.dtype)
    getattr(lib(), "orc_update_uniform_" + suf)(ctypes.byref(cfg), f, G.shape[0], np.ascontiguousarray(G, t),
                                                 float(scale), int(mode))
    return f


def update_weighted(cfg, f, G, scale):
    t, suf = _dt(f.dtype)
    w = np.zeros(G.shape[0], np.float64)
    getattr(lib(), "orc_update_weighted_" + suf)(ctypes.byref(cfg), f, G.shape[0], np.ascontiguousarray(G, t),
                                                  float(scale), _ptr(w))
    return f, w


def update_avg_weight(cfg, f, G, scale):
    """Weighted average of the intermediate fields (sirt-mimo-avg-weight.cu:
    80-203): returns f and [w_1 .. w_M, sum w]."""
    t, suf = _dt(f.dtype)
    w = np.zeros(G.shape[0] + 1, np.float64)
    getattr(lib(), "orc_update_avg_weight_" + suf)(ctypes.byref(cfg), f, G.shape[0],
                                                    np.ascontiguousarray(G, t), float(scale), _ptr(w))
    return f, w


def update_group(cfg, f, Gg, scale):
    t, suf = _dt(f.dtype)
    getattr(lib(), "orc_update_group_" + suf)(ctypes.byref(cfg), f, np.ascontiguousarray(Gg, t), float(scale))
    return f


def error(f, fo):
    t, suf = _dt(f.dtype)
    return float(getattr(lib(), "orc_error_" + suf)(f, np.ascontiguousarray(fo, np.float32), f.size))


def apply_forward_op(cfg, f, x):
    y = np.zeros(cfg.nx * cfg.ny)
    lib().orc_apply_forward_op_f64(ctypes.byref(cfg), np.ascontiguousarray(f, np.float64),
                                   np.ascontiguousarray(x, np.float64), y)
    return y


def apply_adjoint_op(cfg, f, y):
    x = np.zeros(cfg.nx * cfg.ny)
    lib().orc_apply_adjoint_op_f64(ctypes.byref(cfg), np.ascontiguousarray(f, np.float64),
                                   np.ascontiguousarray(y, np.float64), x)
    return x


def laplacian(cfg, u):
    L = np.zeros(cfg.nx * cfg.ny)
    lib().orc_laplacian_f64(ctypes.byref(cfg), np.ascontiguousarray(u, np.float64), L)
    return L


def adjoint_probe(cfg, f, rcell, ranchor, rr, cell):
    """The adjoint sweep alone, driven by residual traces rr[nt][nr]: z(k) at `cell` for every level."""
    out = np.zeros(cfg.nt)
    lib().orc_adjoint_probe_f64(ctypes.byref(cfg), np.ascontiguousarray(f, np.float64), len(rcell),
                                np.ascontiguousarray(rcell, np.int32), np.ascontiguousarray(ranchor, np.int32),
                                np.ascontiguousarray(rr, np.float64), int(cell), out)
    return out


def schedule(kind, k, params):
    return float(lib().orc_schedule(int(kind), int(k), np.ascontiguousarray(params, np.float64)))


# ------------------------------------------------------- learned preconditioner

def cnn_num_params() -> int:
    return int(lib().orc_cnn_num_params())


def cnn_forward(params, x, mode=0):
    """x[B][H][W][3] float32 -> y[B][H][W] float32 (mode 1: the GPU's BF16 rounding points)."""
    B, H, W, _ = x.shape
    y = np.zeros((B, H, W), np.float32)
    lib().orc_cnn_forward(np.ascontiguousarray(params, np.float32), B, H, W, np.ascontiguousarray(x, np.float32),
                          y, int(mode))
    return y


def cnn_backward(params, x, dy, mode=0):
    """Gradient of sum(dy * y) with respect to the parameters (float64)."""
    B, H, W, _ = x.shape
    g = np.zeros(cnn_num_params(), np.float64)
    lib().orc_cnn_backward(np.ascontiguousarray(params, np.float32), B, H, W, np.ascontiguousarray(x, np.float32),
                           np.ascontiguousarray(dy, np.float32), g, int(mode))
    return g


def cnn_loss(params, x, target, mask4, mode=0, grad=True):
    """Masked mean-square loss and (optionally) its parameter gradient."""
    B, H, W, _ = x.shape
    g = np.zeros(cnn_num_params(), np.float64) if grad else None
    m = np.ascontiguousarray(mask4, np.int32)
    loss = lib().orc_cnn_loss(np.ascontiguousarray(params, np.float32), B, H, W, np.ascontiguousarray(x, np.float32),
                              np.ascontiguousarray(target, np.float32), m, _ptr(g), int(mode))
    return (loss, g) if grad else loss
