"""CPU checks of the C ABI library: it loads, exports every function that
include/tomo.h declares, its host-only entry points agree with the oracle
and the reference layouts, and context creation fails cleanly without a
GPU.  No compute call is made on a device here."""
import ctypes
import os
import re

import numpy as np
import pytest

import oracle as orc
from _cases import ref_config, unit_config
from _pkg import tomo

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "tomo.h")


def header_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(tomo_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_function():
    L = tomo.lib()
    names = header_functions()
    assert len(names) > 40
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert set(names) == set(tomo.SIGNATURES), set(names) ^ set(tomo.SIGNATURES)
    assert L.tomo_abi_version() == 1


def test_config_struct_layout_matches_header():
    # offsets of the last members pin the struct layout (152 bytes)
    assert ctypes.sizeof(tomo.TomoConfig) == 152
    assert tomo.TomoConfig.shots_per_pass.offset == 148


def test_reference_preset_matches_oracle():
    a = tomo.config_reference().as_dict()
    b = orc.config_reference(tomo.TomoConfig()).as_dict()
    assert a == b
    assert a["nx"] == 192 and a["nt"] == 401 and a["mask_x0"] == 21 and a["mask_x1"] == 180
    u = tomo.config_unit(128, 1000).as_dict()
    v = orc.config_unit(tomo.TomoConfig(), 128, 1000).as_dict()
    assert u == v


def test_transducers_and_phantom_match_oracle():
    ii, jj = tomo.transducers_reference()
    oi, oj = orc.transducers_reference()
    np.testing.assert_array_equal(ii, oi)
    np.testing.assert_array_equal(jj, oj)
    fo = tomo.phantom_reference()
    np.testing.assert_array_equal(fo, orc.phantom_reference())
    assert (fo == np.float32(0.03)).sum() == 599 and (fo == np.float32(0.05)).sum() == 228


def test_ring_transducers():
    ii, jj = tomo.transducers_ring(128, 128, 32, 50.0)
    r = np.hypot(ii - 63.5, jj - 63.5)
    assert np.all(np.abs(r - 50) <= 0.75)
    assert len(set(zip(ii.tolist(), jj.tolist()))) == 32
    with pytest.raises(tomo.TomoError):
        tomo.transducers_ring(32, 32, 400, 10.0)  # more points than cells on the ring


def test_phantoms_are_seeded_and_bounded():
    n = 128
    a = np.zeros(n * n, np.float32)
    b = np.zeros(n * n, np.float32)
    region = (32, 96, 32, 96)
    tomo.phantom_inclusions(a, n, n, 1001, 4, 4, 12, 0.05, 0.10, region)
    tomo.phantom_inclusions(b, n, n, 1001, 4, 4, 12, 0.05, 0.10, region)
    np.testing.assert_array_equal(a, b)
    v = np.sqrt(1 + a) - 1                      # velocity contrast
    nz = np.abs(v[v != 0])
    assert nz.size > 0 and nz.min() >= 0.05 - 1e-6 and nz.max() <= 0.10 + 1e-6
    a2 = a.reshape(n, n)
    assert not a2[:32].any() and not a2[96:].any() and not a2[:, :32].any()
    s = np.zeros(n * n, np.float32)
    tomo.phantom_smooth(s, n, n, 7, 8.0, 0.08, region)
    c = np.sqrt(1 + s) - 1
    assert abs(np.abs(c).max() - 0.08) < 1e-6
    assert not s.reshape(n, n)[:32].any()
    for bad_sigma in (0.0, -1.0, float(n) + 1.0, 1e12):  # no filter, or one wider than the grid
        with pytest.raises(tomo.TomoError):
            tomo.phantom_smooth(s, n, n, 7, bad_sigma, 0.08, region)


def test_noise_rms_and_seed():
    rng = np.random.default_rng(0)
    tr = rng.standard_normal((3, 4000, 5)).astype(np.float32)
    a = tomo.add_noise(tr.copy(), 0.1, 42)
    b = tomo.add_noise(tr.copy(), 0.1, 42)
    np.testing.assert_array_equal(a, b)
    rms = np.sqrt((tr.astype(np.float64) ** 2).mean(axis=1))
    noise_rms = np.sqrt(((a - tr).astype(np.float64) ** 2).mean(axis=1))
    np.testing.assert_allclose(noise_rms / rms, 0.1, rtol=0.06)


def test_schedules_match_oracle():
    p = np.array([2.0, 1.0, 3.0, 1.5])
    q = np.array([5.0, 4.0, 1.0, 7.0, 6.0, 2.0])
    for k in range(6):
        assert tomo.schedule(tomo.SCHED_POWER, k, p) == orc.schedule(tomo.SCHED_POWER, k, p)
        assert tomo.schedule(tomo.SCHED_ALTERNATING, k, q) == orc.schedule(tomo.SCHED_ALTERNATING, k, q)


def test_text_io_reference_layout(tmp_path):
    a = (np.arange(2 * 3 * 4, dtype=np.float32) / 7).reshape(2, 3, 4)   # [k][j][i]
    p = tmp_path / "a.txt"
    tomo.write_text(p, a, (4, 3, 2))
    lines = open(p).read().split("\n")
    # 3 rows, blank (k block), 3 rows, blank (k block), blank (l block)
    assert lines[3] == "" and lines[7] == "" and lines[8] == ""
    assert lines[0].endswith(" ") and len(lines[0].split()) == 4
    assert lines[0].split()[1] == "0.142857149"   # max_digits10 = 9
    np.testing.assert_array_equal(tomo.read_text(p, (4, 3, 2)), a.ravel())
    # IO_Files layout: default precision, no blank lines
    q = tmp_path / "b.txt"
    tomo.write_text(q, a[0], (4, 3), precision=0, trailing_blank_lines=0)
    assert open(q).read() == "".join(" ".join(f"{v:.6g}" for v in row) + " \n" for row in a[0])
    fx = tomo.read_text(os.path.join(ROOT, "tests", "golden", "fo_192.txt"), (192, 192))
    np.testing.assert_array_equal(fx, tomo.phantom_reference())


def test_create_without_gpu_fails_cleanly():
    # ask the HIP runtime the library is linked against (not torch's, which
    # would load a second HIP runtime into the process)
    tomo.lib()
    count = ctypes.c_int(0)
    if ctypes.CDLL("libamdhip64.so.7").hipGetDeviceCount(ctypes.byref(count)) == 0 and count.value > 0:
        pytest.skip("a GPU is present")
    with pytest.raises(tomo.TomoError, match="no GPU|HIP"):
        tomo.Engine(ref_config())


def test_invalid_config_rejected():
    c = unit_config(64, 100)
    c.sponge_width = 40   # wider than half the grid
    with pytest.raises(tomo.TomoError):
        tomo.Engine(c)
