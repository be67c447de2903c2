"""MI355X wave-tomography engine -- Python binding of ``libtomo_hip.so``.

The compute path is the HIP library built from ``csrc/`` for gfx950; this
module only marshals numpy arrays through the C ABI declared in
``include/tomo.h``.  There is no CPU fallback: if the library is missing or
no GPU is present, the engine calls raise.

The reference drives the same loop from one ``Ultrasonic_Tomography()`` host
function per executable (e.g. sirt-mimo-io.cu:97-417); :class:`Engine` exposes
its steps one by one (set field / transducers / groups, forward, residual +
adjoint + correlation, the three update rules, the error) and
:func:`Engine.sirt_iteration` / :func:`Engine.art_iteration` run whole
iterations.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG_DIR, "_build", "libtomo_hip.so")

# enum values of include/tomo.h
WAVELET_REF_GAUSS_COS, WAVELET_RICKER = 0, 1
RECV_LINES, RECV_POINTS = 0, 1
UPDATE_SUM, UPDATE_MEAN = 0, 1
SCHED_POWER, SCHED_ALTERNATING = 0, 1

_STATUS = {0: "ok", -1: "bad argument", -2: "HIP error", -3: "allocation failed",
           -4: "invalid state", -5: "I/O error", -6: "no GPU"}


class TomoConfig(ctypes.Structure):
    """Mirror of ``struct tomo_config`` (include/tomo.h)."""
    _fields_ = [
        ("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nt", ctypes.c_int32),
        ("wavelet", ctypes.c_int32),
        ("h", ctypes.c_double), ("dt", ctypes.c_double), ("c0", ctypes.c_double),
        ("wl_omega", ctypes.c_double), ("wl_tau", ctypes.c_double), ("wl_shift", ctypes.c_double),
        ("ricker_fp", ctypes.c_double), ("ricker_t0", ctypes.c_double),
        ("src_last_level", ctypes.c_int32), ("sponge_width", ctypes.c_int32),
        ("sponge_gamma", ctypes.c_double),
        ("recv_mode", ctypes.c_int32), ("record_first", ctypes.c_int32),
        ("line_lo", ctypes.c_int32), ("line_hi", ctypes.c_int32),
        ("line_first", ctypes.c_int32), ("line_last", ctypes.c_int32),
        ("recv_gain", ctypes.c_double),
        ("mask_x0", ctypes.c_int32), ("mask_x1", ctypes.c_int32),
        ("mask_y0", ctypes.c_int32), ("mask_y1", ctypes.c_int32),
        ("checkpoint_every", ctypes.c_int32), ("shots_per_pass", ctypes.c_int32),
    ]

    def copy(self) -> "TomoConfig":
        c = TomoConfig()
        ctypes.pointer(c)[0] = self
        return c

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class TomoError(RuntimeError):
    """Raised when a C-ABI call returns a negative status."""


_lib = None

_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_vp = ctypes.c_void_p
_int = ctypes.c_int

# name -> (restype, argtypes) for every function of include/tomo.h
SIGNATURES = {
    "tomo_config_reference": (None, [_vp]),
    "tomo_config_unit": (None, [_vp, _int, _int]),
    "tomo_create": (_int, [_vp, _int, ctypes.POINTER(_vp)]),
    "tomo_destroy": (None, [_vp]),
    "tomo_last_error": (ctypes.c_char_p, [_vp]),
    "tomo_strerror": (ctypes.c_char_p, [_int]),
    "tomo_abi_version": (_int, []),
    "tomo_transducers_reference": (_int, [_i32p, _i32p, _int]),
    "tomo_transducers_ring": (_int, [_int, _int, _int, ctypes.c_double, _i32p, _i32p]),
    "tomo_set_transducers": (_int, [_vp, _i32p, _i32p, _int, _int]),
    "tomo_set_groups": (_int, [_vp, _int]),
    "tomo_set_receivers": (_int, [_vp, _i32p, _i32p, _int, _int]),
    "tomo_num_shots": (_int, [_vp]),
    "tomo_num_receivers": (_int, [_vp]),
    "tomo_get_receivers": (_int, [_vp, _i32p, _i32p]),
    "tomo_set_phantoms": (_int, [_vp, _int]),
    "tomo_num_phantoms": (_int, [_vp]),
    "tomo_set_field": (_int, [_vp, _f32p]),
    "tomo_get_field": (_int, [_vp, _f32p]),
    "tomo_set_truth": (_int, [_vp, _f32p]),
    "tomo_phantom_reference": (_int, [_f32p, _int, _int, ctypes.c_double]),
    "tomo_phantom_inclusions": (_int, [_f32p, _int, _int, ctypes.c_uint64, _int, ctypes.c_double,
                                       ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                       _int, _int, _int, _int]),
    "tomo_phantom_smooth": (_int, [_f32p, _int, _int, ctypes.c_uint64, ctypes.c_double,
                                   ctypes.c_double, _int, _int, _int, _int]),
    "tomo_forward": (_int, [_vp, _int, _int, _int, _f32p]),
    "tomo_set_measured": (_int, [_vp, _f32p]),
    "tomo_add_noise": (_int, [_f32p, _int, _int, _int, ctypes.c_double, ctypes.c_uint64]),
    "tomo_traces_to_sides": (_int, [_vp, _f32p, _f32p, _f32p, _f32p, _f32p]),
    "tomo_sides_to_traces": (_int, [_vp, _f32p, _f32p, _f32p, _f32p, _f32p]),
    "tomo_gradient": (_int, [_vp, _int, _int, _vp]),
    "tomo_misfit": (_int, [_vp, ctypes.POINTER(ctypes.c_double)]),
    "tomo_get_gradient": (_int, [_vp, _int, _f32p]),
    "tomo_enable_coverage": (_int, [_vp, _int]),
    "tomo_get_coverage": (_int, [_vp, _f32p]),
    "tomo_set_pixel_weights": (_int, [_vp, _int, ctypes.c_double]),
    "tomo_line_search": (_int, [_vp, _int, ctypes.c_double, ctypes.c_double, _int,
                                ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]),
    "tomo_update_uniform": (_int, [_vp, ctypes.c_float, _int]),
    "tomo_update_weighted": (_int, [_vp, ctypes.c_float, _vp]),
    "tomo_update_uniform_each": (_int, [_vp, _f32p, _int]),
    "tomo_update_weighted_each": (_int, [_vp, _f32p, _vp]),
    # device pointers (d_f, d_grad, d_pixw, d_weights) and the stream as plain addresses
    "tomo_update_weighted_noalloc": (_int, [_vp, _vp, _int, _int, _int, _i32p, ctypes.c_double, _vp, _vp, _vp]),
    "tomo_update_avg_weight": (_int, [_vp, ctypes.c_float, _vp]),
    "tomo_update_group": (_int, [_vp, _int, ctypes.c_float]),
    "tomo_error": (_int, [_vp, _f64p]),
    "tomo_sirt_iteration": (_int, [_vp, ctypes.c_float, _int, _int, _vp, _vp]),
    "tomo_art_iteration": (_int, [_vp, ctypes.c_float, _vp, _vp]),
    "tomo_schedule": (ctypes.c_double, [_int, _int, _f64p]),
    "tomo_write_text": (_int, [ctypes.c_char_p, _f32p, _int, _int, _int, _int, _int, _int]),
    "tomo_read_text": (_int, [ctypes.c_char_p, _f32p, _int, _int, _int, _int]),
    "tomo_timing": (_int, [_vp, _f64p]),
    "tomo_timing_reset": (_int, [_vp]),
    "tomo_set_shape": (_int, [_vp, _int, _int, _int, _int, _int]),
    "tomo_get_shape": (_int, [_vp, _i32p]),
    "tomo_set_graphs": (_int, [_vp, _int]),
    "tomo_set_time_blocking": (_int, [_vp, _int]),
    "tomo_get_time_blocking": (_int, [_vp]),
    "tomo_get_adjoint_blocking": (_int, [_vp]),
    "tomo_set_adjoint_blocking": (_int, [_vp, _int]),
    "tomo_workspace_create": (_int, [_int, ctypes.POINTER(_vp)]),
    "tomo_workspace_destroy": (_int, [_vp]),
    "tomo_set_workspace": (_int, [_vp, _vp]),
    "tomo_autotune": (_int, [_vp, _int, _i32p, _f64p]),
    "tomo_timing_graphs": (_int, [_vp, _f64p]),
    # batch of problems
    "tomo_batch_create": (_int, [ctypes.POINTER(_vp), _int, ctypes.POINTER(_vp)]),
    "tomo_batch_destroy": (None, [_vp]),
    "tomo_batch_last_error": (ctypes.c_char_p, [_vp]),
    "tomo_batch_set_graph": (_int, [_vp, _int]),
    "tomo_batch_gradient": (_int, [_vp, _vp]),
    "tomo_batch_timing": (_int, [_vp, _f64p]),
    # learned preconditioner
    "tomo_cnn_num_params": (_int, []),
    "tomo_cnn_create": (_int, [_int, ctypes.POINTER(_vp)]),
    "tomo_cnn_destroy": (None, [_vp]),
    "tomo_cnn_last_error": (ctypes.c_char_p, [_vp]),
    "tomo_cnn_init": (_int, [_vp, ctypes.c_uint64]),
    "tomo_cnn_set_params": (_int, [_vp, _f32p]),
    "tomo_cnn_get_params": (_int, [_vp, _f32p]),
    "tomo_cnn_forward": (_int, [_vp, _int, _int, _int, _f32p, _f32p]),
    "tomo_cnn_backward": (_int, [_vp, _int, _int, _int, _f32p, _f32p, _f32p]),
    "tomo_cnn_loss": (_int, [_vp, _int, _int, _int, _f32p, _f32p, _i32p, ctypes.POINTER(ctypes.c_double), _vp]),
    "tomo_cnn_add_sample": (_int, [_vp, _vp, ctypes.c_float]),
    "tomo_cnn_add_sample_phantom": (_int, [_vp, _vp, _int, ctypes.c_float]),
    "tomo_cnn_reserve_samples": (_int, [_vp, _int, _int, _int]),
    "tomo_cnn_num_samples": (_int, [_vp]),
    "tomo_cnn_set_targets": (_int, [_vp, _int, _int, _f32p]),
    "tomo_cnn_clear_samples": (_int, [_vp]),
    "tomo_cnn_train": (_int, [_vp, _int, _int, ctypes.c_double, ctypes.c_uint64, _vp]),
    "tomo_learned_update": (_int, [_vp, _vp, ctypes.c_float, ctypes.c_float]),
    "tomo_learned_iteration": (_int, [_vp, _vp, ctypes.c_float, ctypes.c_float, _int, _vp, _vp, _vp]),
    "tomo_learned_reset": (_int, [_vp]),
    "tomo_cnn_set_sampling": (_int, [_vp, _int]),
}


def lib():
    """Load libtomo_hip.so (built by ``__graft_entry__.build()``)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise TomoError(f"{LIB_PATH} is not built; run __graft_entry__.build() first")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _check(status, ctx=None, net=None):
    if status < 0:
        msg = _STATUS.get(status, str(status))
        detail = None
        if ctx is not None:
            detail = lib().tomo_last_error(ctx)
        elif net is not None:
            detail = lib().tomo_cnn_last_error(net)
        if detail:
            msg += ": " + detail.decode(errors="replace")
        raise TomoError(msg)
    return status


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _ptr(a):
    return None if a is None else a.ctypes.data_as(_vp)


# --------------------------------------------------------------- presets

def config_reference() -> TomoConfig:
    c = TomoConfig()
    lib().tomo_config_reference(ctypes.byref(c))
    return c


def config_unit(n: int, nt: int) -> TomoConfig:
    c = TomoConfig()
    lib().tomo_config_unit(ctypes.byref(c), int(n), int(nt))
    return c


def transducers_reference(ns: int = 640):
    ii = np.zeros(ns, np.int32)
    jj = np.zeros(ns, np.int32)
    _check(lib().tomo_transducers_reference(ii, jj, ns))
    return ii, jj


def transducers_ring(nx: int, ny: int, count: int, radius: float):
    ii = np.zeros(count, np.int32)
    jj = np.zeros(count, np.int32)
    _check(lib().tomo_transducers_ring(nx, ny, count, float(radius), ii, jj))
    return ii, jj


def phantom_reference(nx: int = 192, ny: int = 192, h: float = float(np.float32(0.001))):
    fo = np.zeros(nx * ny, np.float32)
    _check(lib().tomo_phantom_reference(fo, nx, ny, h))
    return fo


def phantom_inclusions(fo, nx, ny, seed, count, rmin, rmax, cmin, cmax, region):
    x0, x1, y0, y1 = region
    _check(lib().tomo_phantom_inclusions(fo, nx, ny, seed, count, rmin, rmax, cmin, cmax, x0, x1, y0, y1))
    return fo


def phantom_smooth(fo, nx, ny, seed, sigma, amp, region):
    x0, x1, y0, y1 = region
    _check(lib().tomo_phantom_smooth(fo, nx, ny, seed, sigma, amp, x0, x1, y0, y1))
    return fo


def add_noise(traces, rel, seed):
    t = _f32(traces)
    S, nt, nr = t.shape
    _check(lib().tomo_add_noise(t, S, nt, nr, float(rel), int(seed)))
    return t


def schedule(kind: int, k: int, params) -> float:
    return float(lib().tomo_schedule(int(kind), int(k), np.ascontiguousarray(params, np.float64)))


def write_text(path, a, dims, precision=0, trailing_blank_lines=1):
    nx, ny, nz, nw = (list(dims) + [1, 1, 1])[:4]
    _check(lib().tomo_write_text(os.fsencode(path), _f32(a).ravel(), nx, ny, nz, nw, precision,
                                 trailing_blank_lines))


def read_text(path, dims):
    nx, ny, nz, nw = (list(dims) + [1, 1, 1])[:4]
    a = np.zeros(nx * ny * nz * nw, np.float32)
    _check(lib().tomo_read_text(os.fsencode(path), a, nx, ny, nz, nw))
    return a


# ---------------------------------------------------------------- engine

class Engine:
    """One engine context (one HIP stream on one GPU)."""

    def __init__(self, cfg: TomoConfig, device: int = -1):
        self.cfg = cfg.copy()
        self._ctx = _vp()
        _check(lib().tomo_create(ctypes.byref(self.cfg), int(device), ctypes.byref(self._ctx)))
        self.n2 = cfg.nx * cfg.ny
        self.phantoms = 1

    def close(self):
        if self._ctx:
            lib().tomo_destroy(self._ctx)
            self._ctx = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _c(self, status):
        return _check(status, self._ctx)

    # geometry
    def set_transducers(self, ii, jj, one_based=True):
        self._c(lib().tomo_set_transducers(self._ctx, _i32(ii), _i32(jj), len(ii), int(one_based)))

    def set_groups(self, group_size: int) -> int:
        return self._c(lib().tomo_set_groups(self._ctx, int(group_size)))

    def set_phantoms(self, count: int):
        """Stack `count` phantoms of this geometry along the shot axis
        (fields, truths and coverage maps become [count][ny][nx]; error()
        and sirt_iteration() return one epsilon per phantom)."""
        self._c(lib().tomo_set_phantoms(self._ctx, int(count)))
        self.phantoms = int(count)

    def set_receivers(self, ri, rj, one_based=False):
        self._c(lib().tomo_set_receivers(self._ctx, _i32(ri), _i32(rj), len(ri), int(one_based)))

    @property
    def num_shots(self) -> int:
        return self._c(lib().tomo_num_shots(self._ctx))

    @property
    def num_receivers(self) -> int:
        return self._c(lib().tomo_num_receivers(self._ctx))

    def receivers(self):
        n = self.num_receivers
        ri = np.zeros(n, np.int32)
        rj = np.zeros(n, np.int32)
        self._c(lib().tomo_get_receivers(self._ctx, ri, rj))
        return ri, rj

    # fields
    def set_field(self, f):
        self._c(lib().tomo_set_field(self._ctx, _f32(f).ravel()))

    def get_field(self):
        f = np.zeros(self.n2 * self.phantoms, np.float32)
        self._c(lib().tomo_get_field(self._ctx, f))
        return f

    def set_truth(self, fo):
        self._c(lib().tomo_set_truth(self._ctx, _f32(fo).ravel()))

    # data
    def forward(self, g0=0, ng=None, fp64=False):
        if ng is None:
            ng = self.num_shots - g0
        tr = np.zeros((ng, self.cfg.nt, self.num_receivers), np.float32)
        self._c(lib().tomo_forward(self._ctx, int(g0), int(ng), int(bool(fp64)), tr))
        return tr

    def set_measured(self, traces):
        self._c(lib().tomo_set_measured(self._ctx, _f32(traces)))

    def traces_to_sides(self, traces):
        S = traces.shape[0]
        sides = [np.zeros((S, self.cfg.nt, self.cfg.nx), np.float32) for _ in range(4)]
        self._c(lib().tomo_traces_to_sides(self._ctx, _f32(traces), *sides))
        return sides

    def sides_to_traces(self, bottom, right, top, left):
        S = bottom.shape[0]
        tr = np.zeros((S, self.cfg.nt, self.num_receivers), np.float32)
        self._c(lib().tomo_sides_to_traces(self._ctx, _f32(bottom), _f32(right), _f32(top), _f32(left), tr))
        return tr

    # gradient
    def gradient(self, g0=0, ng=None) -> float:
        if ng is None:
            ng = self.num_shots - g0
        mis = ctypes.c_double(0.0)
        self._c(lib().tomo_gradient(self._ctx, int(g0), int(ng), ctypes.cast(ctypes.byref(mis), _vp)))
        return mis.value

    def misfit(self) -> float:
        """Data misfit of all shots at the current field (forward only)."""
        m = ctypes.c_double(0.0)
        self._c(lib().tomo_misfit(self._ctx, ctypes.byref(m)))
        return m.value

    def get_gradient(self, g):
        out = np.zeros(self.n2, np.float32)
        self._c(lib().tomo_get_gradient(self._ctx, int(g), out))
        return out

    def enable_coverage(self, on=True):
        self._c(lib().tomo_enable_coverage(self._ctx, int(bool(on))))

    def get_coverage(self):
        out = np.zeros(self.n2 * self.phantoms, np.float32)
        self._c(lib().tomo_get_coverage(self._ctx, out))
        return out

    # updates
    def line_search(self, weighted=False, s_lo=0.0, s_step=100.0, count=10001):
        """Ground-truth best scale on the grid s_lo + k s_step; returns (scale, eps)."""
        sc = ctypes.c_double(0.0)
        ep = ctypes.c_double(0.0)
        self._c(lib().tomo_line_search(self._ctx, int(bool(weighted)), float(s_lo), float(s_step), int(count),
                                       ctypes.byref(sc), ctypes.byref(ep)))
        return sc.value, ep.value

    def set_pixel_weights(self, mode=1, eps=0.05):
        """Per-pixel illumination weights of the updates (1: coverage normalisation, 0: off)."""
        self._c(lib().tomo_set_pixel_weights(self._ctx, int(mode), float(eps)))

    def update_uniform(self, scale, mode=UPDATE_MEAN):
        """scale: one relaxation, or one per phantom of a stack."""
        if np.ndim(scale):
            self._c(lib().tomo_update_uniform_each(self._ctx, self._scales(scale), int(mode)))
        else:
            self._c(lib().tomo_update_uniform(self._ctx, float(scale), int(mode)))

    def update_weighted(self, scale):
        """scale: one relaxation, or one per phantom of a stack; returns the weights."""
        w = np.zeros(self.num_shots, np.float64)
        if np.ndim(scale):
            self._c(lib().tomo_update_weighted_each(self._ctx, self._scales(scale), _ptr(w)))
        else:
            self._c(lib().tomo_update_weighted(self._ctx, float(scale), _ptr(w)))
        return w

    def _scales(self, scale):
        s = np.ascontiguousarray(scale, np.float32).ravel()
        if s.size != self.phantoms:
            raise ValueError(f"{s.size} scales for {self.phantoms} phantoms")
        return s

    def update_avg_weight(self, scale):
        """Weighted average of intermediate fields (sirt-mimo-avg-weight.cu:
        80-203); returns [w_1 .. w_Ng, sum w]."""
        w = np.zeros(self.num_shots + 1, np.float64)
        self._c(lib().tomo_update_avg_weight(self._ctx, float(scale), _ptr(w)))
        return w

    def update_group(self, g, scale):
        self._c(lib().tomo_update_group(self._ctx, int(g), float(scale)))

    def error(self):
        """epsilon in %; an array of one value per phantom when phantoms are stacked."""
        e = np.zeros(self.phantoms, np.float64)
        self._c(lib().tomo_error(self._ctx, e))
        return float(e[0]) if self.phantoms == 1 else e

    def sirt_iteration(self, scale, weighted=False, mode=UPDATE_MEAN):
        mis = ctypes.c_double(0.0)
        eps = np.zeros(self.phantoms, np.float64)
        self._c(lib().tomo_sirt_iteration(self._ctx, float(scale), int(bool(weighted)), int(mode),
                                          ctypes.cast(ctypes.byref(mis), _vp), _ptr(eps)))
        return mis.value, (float(eps[0]) if self.phantoms == 1 else eps)

    def art_iteration(self, scale):
        mis = ctypes.c_double(0.0)
        eps = ctypes.c_double(0.0)
        self._c(lib().tomo_art_iteration(self._ctx, float(scale), ctypes.cast(ctypes.byref(mis), _vp),
                                         ctypes.cast(ctypes.byref(eps), _vp)))
        return mis.value, eps.value

    # instrumentation
    def timing(self) -> dict:
        out = np.zeros(6, np.float64)
        self._c(lib().tomo_timing(self._ctx, out))
        return {"fwd_ms": out[0], "fwd_launches": int(out[1]), "adj_ms": out[2],
                "adj_launches": int(out[3]), "fwd_cells": out[4], "adj_cells": out[5]}

    def timing_reset(self):
        self._c(lib().tomo_timing_reset(self._ctx))

    def timing_graphs(self) -> dict:
        out = np.zeros(3, np.float64)
        self._c(lib().tomo_timing_graphs(self._ctx, out))
        return {"ms": out[0], "replays": int(out[1]), "untimed_cells": out[2]}

    def set_time_blocking(self, steps: int):
        """Forward steps per launch in the forward passes (1, 2 or 3; default 3)."""
        self._c(lib().tomo_set_time_blocking(self._ctx, int(steps)))

    def time_blocking(self) -> int:
        return self._c(lib().tomo_get_time_blocking(self._ctx))

    def adjoint_blocking(self) -> int:
        """Adjoint steps per launch in effect (2 where the two-step adjoint applies)."""
        return self._c(lib().tomo_get_adjoint_blocking(self._ctx))

    def detach_workspace(self):
        """Back to wavefield storage of the engine's own."""
        self._c(lib().tomo_set_workspace(self._ctx, None))

    def set_adjoint_blocking(self, steps: int):
        """Most adjoint steps per launch (1, 2 or 3)."""
        self._c(lib().tomo_set_adjoint_blocking(self._ctx, int(steps)))

    def autotune(self, force=False):
        """Time the stencil shapes on this context and keep the fastest; returns (shape, (fwd_ms, adj_ms))."""
        out = np.zeros(5, np.int32)
        ms = np.zeros(2, np.float64)
        self._c(lib().tomo_autotune(self._ctx, int(bool(force)), out, ms))
        return tuple(int(v) for v in out), (float(ms[0]), float(ms[1]))

    def set_graphs(self, on=True):
        """Capture each gradient pass into a HIP graph and replay it."""
        self._c(lib().tomo_set_graphs(self._ctx, int(bool(on))))

    def set_shape(self, vec=0, fwd_rows=0, adj_rows=0, fwd_depth=0, adj_depth=0):
        self._c(lib().tomo_set_shape(self._ctx, int(vec), int(fwd_rows), int(adj_rows), int(fwd_depth),
                                     int(adj_depth)))

    def get_shape(self):
        out = np.zeros(5, np.int32)
        self._c(lib().tomo_get_shape(self._ctx, out))
        return tuple(int(v) for v in out)


# ----------------------------------------------------------- workspace

class Workspace:
    """Wavefield storage shared by engines that are never used at the same
    time (see tomo_set_workspace); destroy it after those engines."""

    def __init__(self, device: int = 0):
        self._w = _vp()
        st = lib().tomo_workspace_create(int(device), ctypes.byref(self._w))
        if st < 0:
            raise TomoError(_STATUS.get(st, str(st)))

    def attach(self, engine: "Engine"):
        engine._c(lib().tomo_set_workspace(engine._ctx, self._w))

    def close(self):
        if self._w:
            st = lib().tomo_workspace_destroy(self._w)
            if st < 0:
                detail = lib().tomo_last_error(None)
                raise TomoError(_STATUS.get(st, str(st)) + (": " + detail.decode(errors="replace") if detail else ""))
            self._w = _vp()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


# --------------------------------------------------------------- batch

class Batch:
    """Several engines (problems of any sizes) stepped together, one launch
    per time level and stencil shape.  The engines must outlive the batch."""

    def __init__(self, engines, graph: bool = False):
        self.engines = list(engines)
        arr = (_vp * len(self.engines))(*[e._ctx for e in self.engines])
        self._b = _vp()
        st = lib().tomo_batch_create(arr, len(self.engines), ctypes.byref(self._b))
        if st < 0:
            msg = _STATUS.get(st, str(st))
            detail = lib().tomo_last_error(None)
            raise TomoError(msg + (": " + detail.decode(errors="replace") if detail else ""))
        if graph:
            self._c(lib().tomo_batch_set_graph(self._b, 1))

    def _c(self, status):
        if status < 0:
            detail = lib().tomo_batch_last_error(self._b)
            raise TomoError(_STATUS.get(status, str(status)) + (": " + detail.decode(errors="replace") if detail else ""))
        return status

    def gradient(self):
        mis = np.zeros(len(self.engines), np.float64)
        self._c(lib().tomo_batch_gradient(self._b, _ptr(mis)))
        return mis

    def timing(self) -> dict:
        out = np.zeros(4, np.float64)
        self._c(lib().tomo_batch_timing(self._b, out))
        return {"ms": out[0], "cells": out[1], "launches": int(out[2]), "calls": int(out[3])}

    def close(self):
        if self._b:
            lib().tomo_batch_destroy(self._b)
            self._b = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


# ------------------------------------------------------- learned preconditioner

class Preconditioner:
    """The learned preconditioner network (one HIP stream on one GPU).

    ``x`` arrays are float32 [B][H][W][3] (field / 0.1, plain SIRT step / 0.1,
    coverage / max); outputs float32 [B][H][W]; parameters float32 in the
    layout W[o][c][ky][kx] then b[o], layer by layer (28929 values).
    """

    def __init__(self, device: int = -1, seed: int | None = None):
        self._net = _vp()
        _check(lib().tomo_cnn_create(int(device), ctypes.byref(self._net)))
        self.num_params = int(lib().tomo_cnn_num_params())
        if seed is not None:
            self.init(seed)

    def close(self):
        if self._net:
            lib().tomo_cnn_destroy(self._net)
            self._net = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _c(self, status):
        return _check(status, net=self._net)

    def init(self, seed: int):
        self._c(lib().tomo_cnn_init(self._net, int(seed)))

    def set_params(self, p):
        p = _f32(p)
        if p.size != self.num_params:
            raise TomoError(f"expected {self.num_params} parameters, got {p.size}")
        self._c(lib().tomo_cnn_set_params(self._net, p))

    def get_params(self):
        p = np.zeros(self.num_params, np.float32)
        self._c(lib().tomo_cnn_get_params(self._net, p))
        return p

    def forward(self, x):
        x = _f32(x)
        B, H, W, _ = x.shape
        y = np.zeros((B, H, W), np.float32)
        self._c(lib().tomo_cnn_forward(self._net, B, H, W, x, y))
        return y

    def backward(self, x, dy):
        x = _f32(x)
        B, H, W, _ = x.shape
        g = np.zeros(self.num_params, np.float32)
        self._c(lib().tomo_cnn_backward(self._net, B, H, W, x, _f32(dy), g))
        return g

    def loss(self, x, target, mask4, grad=True):
        x = _f32(x)
        B, H, W, _ = x.shape
        g = np.zeros(self.num_params, np.float32) if grad else None
        L = ctypes.c_double(0.0)
        self._c(lib().tomo_cnn_loss(self._net, B, H, W, x, _f32(target), _i32(mask4), ctypes.byref(L), _ptr(g)))
        return (L.value, g) if grad else L.value

    def add_sample(self, engine: "Engine", scale: float, phantom: int | None = None):
        """One training sample from the engine's state (phantom `phantom` of a
        context of stacked phantoms)."""
        if phantom is None:
            self._c(lib().tomo_cnn_add_sample(self._net, engine._ctx, float(scale)))
        else:
            self._c(lib().tomo_cnn_add_sample_phantom(self._net, engine._ctx, int(phantom), float(scale)))

    def reserve_samples(self, count: int, ny: int, nx: int):
        """Storage for `count` samples of ny x nx cells, allocated once."""
        self._c(lib().tomo_cnn_reserve_samples(self._net, int(count), int(ny), int(nx)))

    def num_samples(self) -> int:
        return int(lib().tomo_cnn_num_samples(self._net))

    def set_targets(self, first: int, count: int, field):
        """Targets (field - f_s) / 0.1 for samples [first, first + count)."""
        self._c(lib().tomo_cnn_set_targets(self._net, int(first), int(count), _f32(field)))

    def clear_samples(self):
        self._c(lib().tomo_cnn_clear_samples(self._net))

    def train(self, steps: int, batch: int, lr: float, seed: int = 0):
        losses = np.zeros(max(int(steps), 1), np.float64)
        self._c(lib().tomo_cnn_train(self._net, int(steps), int(batch), float(lr), int(seed), _ptr(losses)))
        return losses[: int(steps)]

    def update(self, engine: "Engine", alpha: float, scale: float):
        """f += mask * (alpha * 0.1 * net(input) + max(1 - alpha, 0) * plain step with `scale`)."""
        engine._c(lib().tomo_learned_update(engine._ctx, self._net, float(alpha), float(scale)))

    def iteration(self, engine: "Engine", alpha: float, scale: float, weighted: bool = True):
        """One learned iteration with the plain-SIRT fallback; returns (misfit, eps, accepted)."""
        mis = ctypes.c_double(0.0)
        eps = ctypes.c_double(0.0)
        acc = ctypes.c_int(0)
        engine._c(lib().tomo_learned_iteration(engine._ctx, self._net, float(alpha), float(scale), int(bool(weighted)),
                                               ctypes.cast(ctypes.byref(mis), _vp), ctypes.cast(ctypes.byref(eps), _vp),
                                               ctypes.cast(ctypes.byref(acc), _vp)))
        return mis.value, eps.value, bool(acc.value)

    def reset(self):
        self._c(lib().tomo_learned_reset(self._net))

    def set_sampling(self, on: bool):
        """While on, learned iterations add a training sample per visited state."""
        self._c(lib().tomo_cnn_set_sampling(self._net, int(bool(on))))
