"""tomo_update_weighted_noalloc (the weighted SIRT update on caller-owned
device arrays, scratch in a static device array) against
tomo_update_weighted of a context on the same gradients and field: the same
kernels in the same order, so the fields and weights must be equal bit for
bit.  Two calls in a row, the second on a smaller grid with fewer shots (the
static scratch keeps the first call's contents); then the per-pixel weight
path, and a size above the scratch's limit, which must be refused before
anything is launched.  The caller's device arrays are allocated here with the
HIP runtime the library itself is linked against (libamdhip64.so.7)."""
import ctypes

import numpy as np
import pytest

from _pkg import tomo

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)


def engine_with_gradients(n, shots, seed):
    c = tomo.config_unit(n, 4 * n)
    ii, jj = tomo.transducers_ring(n, n, shots, 0.39 * n)
    fo = np.zeros(n * n, np.float32)
    tomo.phantom_inclusions(fo, n, n, seed, 3, 0.04 * n, 0.09 * n, 0.05, 0.10,
                            (c.mask_x0, c.mask_x1, c.mask_y0, c.mask_y1))
    e = tomo.Engine(c, device=0)
    e.set_transducers(ii, jj, one_based=False)
    e.set_receivers(ii, jj, one_based=False)
    e.set_groups(1)
    e.set_field(fo)
    meas = e.forward(fp64=True)
    f0 = (0.3 * fo).astype(np.float32)
    e.set_field(f0)
    e.set_measured(meas)
    e.gradient()
    G = np.stack([e.get_gradient(s) for s in range(shots)])
    mask4 = np.array([c.mask_x0, c.mask_x1, c.mask_y0, c.mask_y1], np.int32)
    return e, c, f0, G, mask4


class Device:
    """Caller-side device arrays: hipMalloc / hipMemcpy / hipFree of the HIP
    runtime the engine library uses (loaded once per process)."""

    def __init__(self):
        tomo.lib()
        self.hip = ctypes.CDLL("libamdhip64.so.7")
        self.hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.hip.hipFree.argtypes = [ctypes.c_void_p]
        assert self.hip.hipSetDevice(0) == 0
        self.ptrs = []

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = ctypes.c_void_p()
        assert self.hip.hipMalloc(ctypes.byref(p), max(a.nbytes, 4)) == 0
        self.ptrs.append(p.value)
        assert self.hip.hipMemcpy(p.value, a.ctypes.data, a.nbytes, 1) == 0   # host to device
        return p.value

    def get(self, ptr, like):
        out = np.empty_like(like)
        assert self.hip.hipDeviceSynchronize() == 0
        assert self.hip.hipMemcpy(out.ctypes.data, ptr, out.nbytes, 2) == 0   # device to host
        return out

    def free(self):
        assert self.hip.hipDeviceSynchronize() == 0
        for p in self.ptrs:
            assert self.hip.hipFree(p) == 0
        self.ptrs = []


def noalloc(f0, G, nx, ny, mask4, scale, pixw=None):
    dev = Device()
    try:
        d_f = dev.put(f0)
        d_g = dev.put(G)
        w0 = np.full(G.shape[0], -1.0)
        d_w = dev.put(w0)
        d_p = dev.put(pixw) if pixw is not None else None
        rc = tomo.lib().tomo_update_weighted_noalloc(d_f, d_g, G.shape[0], nx, ny, mask4, float(scale), d_p, d_w,
                                                     None)
        assert rc == 0, (rc, tomo.lib().tomo_last_error(None))
        return dev.get(d_f, f0), dev.get(d_w, w0)
    finally:
        dev.free()


@pytest.mark.parametrize("n,shots,seed", [(128, 16, 5), (96, 8, 6)])
def test_noalloc_weighted_update_equals_context_update(n, shots, seed):
    say(f"noalloc weighted update {n}x{n}, {shots} shots")
    e, c, f0, G, mask4 = engine_with_gradients(n, shots, seed)
    # tomo_update_weighted takes a float scale: give both entry points the same value
    scale = float(np.float32(0.02 / np.abs(G.astype(np.float64).sum(0) / shots).max()))
    with e:
        w_ctx = e.update_weighted(scale)
        f_ctx = e.get_field()
    f_na, w_na = noalloc(f0, G, n, n, mask4, scale)
    np.testing.assert_array_equal(w_na, w_ctx)
    np.testing.assert_array_equal(f_na, f_ctx)
    assert (f_na != f0).any()


def test_noalloc_pixel_weights_and_size_limit():
    say("noalloc pixel weights and size limit")
    e, c, f0, G, mask4 = engine_with_gradients(64, 8, 7)
    e.close()
    scale = float(0.02 / np.abs(G.astype(np.float64).sum(0) / 8).max())
    pixw = np.random.default_rng(3).uniform(0.5, 2.0, 64 * 64).astype(np.float32)
    f_plain, w_plain = noalloc(f0, G, 64, 64, mask4, scale)
    f_pw, w_pw = noalloc(f0, G, 64, 64, mask4, scale, pixw)
    np.testing.assert_array_equal(w_pw, w_plain)
    d_plain = f_plain.astype(np.float64) - f0
    d_pw = f_pw.astype(np.float64) - f0
    np.testing.assert_allclose(d_pw, pixw * d_plain, rtol=1e-5, atol=1e-9)
    # above the static scratch: refused on the host, nothing launched
    dev = Device()
    small = dev.put(np.zeros(16, np.float32))
    rc = tomo.lib().tomo_update_weighted_noalloc(small, small, 1, 1024, 1024, np.array([0, 1024, 0, 1024], np.int32),
                                                 1.0, None, None, None)
    dev.free()
    msg = tomo.lib().tomo_last_error(None)
    msg = msg.decode() if isinstance(msg, bytes) else msg
    assert rc == -1 and "scratch" in msg, (rc, msg)  # TOMO_ERR_ARG
