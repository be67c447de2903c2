"""The command-line drivers that mirror the reference executables: arguments,
file names, text layout and stop rules (generate-field.cu, generate-data.cu,
sirt-mimo-io.cu, sirt-mimo-norm-mem.cu, art-mimo-io.cu, sirt-mimo-seq.cu,
sirt-test-weight.cu, and the averaging / group-serial variants
sirt-mimo-avg.cu, sirt-mimo-avg-weight.cu, sirt-mimo-weight.cu, sirt-iter.cu,
sirt-iter-norm.cu)."""
import os
import subprocess

import numpy as np
import pytest

from _pkg import tomo

BIN = os.path.join(tomo.PKG_DIR, "_build")
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def run(args, cwd):
    return subprocess.run([os.path.join(BIN, args[0]), *map(str, args[1:])], cwd=cwd, capture_output=True,
                          text=True, timeout=300)


def test_generate_field_writes_reference_phantom(tmp_path):
    r = run(["generate-field"], tmp_path)
    assert r.returncode == 0, r.stderr
    fo = np.loadtxt(tmp_path / "fo.txt", dtype=np.float32).ravel()
    fx = np.loadtxt(os.path.join(ROOT, "tests", "golden", "fo_192.txt"), dtype=np.float32).ravel()
    np.testing.assert_array_equal(fo, fx)
    text = open(tmp_path / "fo.txt").read()
    assert text.endswith(" \n\n\n")


def test_usage_and_filename_errors(tmp_path):
    r = run(["sirt-mimo-io", "fo.txt"], tmp_path)
    assert r.returncode == 1 and "Usage" in r.stderr
    r = run(["sirt-mimo-io", "a.b.txt", 10, 1, 1, 1, 1, 1, 1], tmp_path)
    assert r.returncode == 1 and "only one period" in r.stderr
    # the variants' argument lists (sirt-mimo-avg.cu:113-115, sirt-mimo-avg-weight.cu:210-212,
    # sirt-mimo-weight.cu:68-70, sirt-iter.cu:410-412, sirt-iter-norm.cu:464-466)
    for exe, extra in (("sirt-mimo-avg", ""), ("sirt-mimo-avg-weight", ""),
                       ("sirt-mimo-weight", " <alpha> <beta> <omega>"),
                       ("sirt-iter", " <omega> <alpha>"), ("sirt-iter-norm", " <omega> <alpha>")):
        r = run([exe, "fo.txt", 10], tmp_path)
        assert r.returncode == 1
        assert r.stderr.startswith("Usage: ") and r.stderr.rstrip().endswith(
            "<fo_filename> <sensor group size> <target epsilon> <max iterations>" + extra), r.stderr
    r = run(["sirt-iter", "a.b.txt", 10, 1, 1, 1, 1], tmp_path)
    assert r.returncode == 1 and "only one period" in r.stderr


def table_eps(stdout):
    return [float(l.split()[1]) for l in stdout.splitlines()[2:] if l.strip()]


@pytest.mark.gpu
def test_averaging_and_group_serial_variants(tmp_path):
    """Each variant's epsilon trajectory equals the engine's update it maps
    to, run through the binding on the driver's own data files."""
    print("[case] SIRT variants G=160", flush=True)
    assert run(["generate-field"], tmp_path).returncode == 0
    assert run(["generate-data", "fo.txt", 160], tmp_path).returncode == 0
    with tomo.Engine(tomo.config_reference()) as e:
        ii, jj = tomo.transducers_reference()
        e.set_transducers(ii, jj, one_based=True)
        ng = e.set_groups(160)
        sides = [tomo.read_text(tmp_path / f"fo-data-{s}-160.txt", (192, 401, ng)).reshape(ng, 401, 192)
                 for s in ("bottom", "right", "top", "left")]
        e.set_measured(e.sides_to_traces(*sides))
        e.set_truth(tomo.read_text(tmp_path / "fo.txt", (192, 192)))
        zero = np.zeros(192 * 192, np.float32)

        def engine_eps(scales, weighted):
            e.set_field(zero)
            return [e.sirt_iteration(s, weighted=weighted, mode=tomo.UPDATE_MEAN)[1] for s in scales]

        # averaged intermediate fields at 20000: "Iter: k / epsilon = v" lines
        r = run(["sirt-mimo-avg", "fo.txt", 160, 0, 2], tmp_path)
        assert r.returncode == 0, r.stderr
        got = [float(l.split("=")[1]) for l in r.stdout.splitlines() if l.startswith("epsilon =")]
        np.testing.assert_allclose(got, engine_eps([20000.0, 20000.0], False), rtol=1e-6)
        assert os.path.exists(tmp_path / "sirt-160-fo.txt")
        # distance-weighted sum of the intermediate fields at 20000, as far as
        # the driver's stop rule lets it run
        r = run(["sirt-mimo-avg-weight", "fo.txt", 160, 0, 2], tmp_path)
        assert r.returncode == 0, r.stderr
        got = [float(l.split("=")[1]) for l in r.stdout.splitlines() if l.startswith("epsilon =")]
        e.set_field(zero)
        want = []
        for _ in got:
            e.gradient()
            e.update_avg_weight(20000.0)
            want.append(e.error())
        assert len(got) >= 1
        np.testing.assert_allclose(got, want, rtol=1e-6)
        # omega (alpha + beta / k), k = 1, 2
        r = run(["sirt-mimo-weight", "fo.txt", 160, 0, 2, 1.0, 10000.0, 2.0], tmp_path)
        assert r.returncode == 0, r.stderr
        got = [float(l.split("=")[1]) for l in r.stdout.splitlines() if l.startswith("epsilon =")]
        np.testing.assert_allclose(got, engine_eps([2 * (1 + 10000.0), 2 * (1 + 5000.0)], False), rtol=1e-6)
        assert os.path.exists(tmp_path / "sirt_weight_convergence.txt")
        # group-serial: omega (alpha + (1 - alpha) (k+1))
        r = run(["sirt-iter", "fo.txt", 160, 0, 2, 10000.0, 0.5], tmp_path)
        assert r.returncode == 0, r.stderr
        np.testing.assert_allclose(table_eps(r.stdout), engine_eps([10000.0, 15000.0], False), rtol=1e-6)
        assert os.path.exists(tmp_path / "art-160-fo.txt")
        # group-serial weighted: omega (alpha + (1 - alpha) / (k+1)^1.5)
        r = run(["sirt-iter-norm", "fo.txt", 160, 0, 2, 20000.0, 0.5], tmp_path)
        assert r.returncode == 0, r.stderr
        np.testing.assert_allclose(table_eps(r.stdout),
                                   engine_eps([20000.0, 20000.0 * (0.5 + 0.5 / 2 ** 1.5)], True), rtol=1e-6)


@pytest.mark.gpu
def test_generate_data_then_sirt_and_art(tmp_path):
    print("[case] drivers G=160", flush=True)
    assert run(["generate-field"], tmp_path).returncode == 0
    r = run(["generate-data", "fo.txt", 160], tmp_path)
    assert r.returncode == 0, r.stderr
    for side in ("bottom", "right", "top", "left"):
        a = np.loadtxt(tmp_path / f"fo-data-{side}-160.txt", dtype=np.float32)
        assert a.shape == (4 * 401, 192)
    # missing data file message (sirt-mimo-io.cu:147-150)
    r = run(["sirt-mimo-io", "fo.txt", 32, 0, 1, 1, 1, 0, 1], tmp_path)
    assert "fo-data-bottom-32.txt' file not found" in r.stderr
    # two uniform SIRT iterations at a fixed scale 22500 (omega 1, alpha 22500, beta 0)
    r = run(["sirt-mimo-io", "fo.txt", 160, 0, 2, 1, 22500, 0, 1], tmp_path)
    assert r.returncode == 0, r.stderr
    f_drv = np.loadtxt(tmp_path / "sirt-160-fo.txt", dtype=np.float32).ravel()
    lines = [l for l in r.stdout.splitlines() if l.strip()]
    assert lines[0].split() == ["iter", "epsilon", "difference"]
    eps_drv = [float(l.split()[1]) for l in lines[2:4]]
    assert os.path.exists(tmp_path / "sirt_convergence.txt")
    # the same through the binding, with the driver's data files
    with tomo.Engine(tomo.config_reference()) as e:
        ii, jj = tomo.transducers_reference()
        e.set_transducers(ii, jj, one_based=True)
        ng = e.set_groups(160)
        sides = [tomo.read_text(tmp_path / f"fo-data-{s}-160.txt", (192, 401, ng)).reshape(ng, 401, 192)
                 for s in ("bottom", "right", "top", "left")]
        e.set_measured(e.sides_to_traces(*sides))
        fo = tomo.read_text(tmp_path / "fo.txt", (192, 192))
        e.set_truth(fo)
        eps = [e.sirt_iteration(22500.0, mode=tomo.UPDATE_MEAN)[1] for _ in range(2)]
        np.testing.assert_allclose(e.get_field(), f_drv, rtol=1e-6, atol=1e-10)
    np.testing.assert_allclose(eps_drv, eps, rtol=1e-6)
    # weighted, memory-scalable passes of 3 groups; then ART
    r = run(["sirt-mimo-norm-mem", "fo.txt", 160, 3, 0, 1, 1, 22500, 0, 1], tmp_path)
    assert r.returncode == 0, r.stderr
    assert os.path.exists(tmp_path / "sirt-norm-160-fo.txt")
    r = run(["art-mimo-io", "fo.txt", 160, 0, 1, 1, 20000, 0, 1], tmp_path)
    assert r.returncode == 0, r.stderr
    f_art = np.loadtxt(tmp_path / "art-160-fo.txt", dtype=np.float32)
    assert np.isfinite(f_art).all() and np.abs(f_art).max() > 0
    # ART warm start then weighted SIRT (sirt-mimo-seq.cu): 7 iterations, the
    # first 5 ART
    r = run(["sirt-mimo-seq", "fo.txt", 160, 0, 7, 1, 20000, 0, 1], tmp_path)
    assert r.returncode == 0, r.stderr
    rows = [l.split() for l in r.stdout.splitlines()[2:] if l.strip()]
    assert len(rows) >= 1 and os.path.exists(tmp_path / "sirt-norm-160-fo.txt")
    # ground-truth line search (sirt-test-weight.cu): the first logged row of
    # sirt-dyn-160-0.txt is iteration 1, eps 83.657, scale 22500
    r = run(["sirt-test-weight", "fo.txt", 160, 0, 0, 1], tmp_path)
    assert r.returncode == 0, r.stderr
    row = [l.split() for l in r.stdout.splitlines()[2:] if l.strip()][0]
    assert row[0] == "1" and abs(float(row[1]) - 83.657) < 0.05 and abs(float(row[3]) - 22500) <= 300


def _side_traces(e, path, ng, G=160):
    sides = [tomo.read_text(path / f"fo-data-{s}-{G}.txt", (192, 401, ng)).reshape(ng, 401, 192)
             for s in ("bottom", "right", "top", "left")]
    return e.sides_to_traces(*sides)


@pytest.mark.gpu
def test_generate_data_fp64_and_noise_against_oracle(tmp_path):
    """generate-data --fp64 at G = 160 writes the oracle's FP64 traces (to the
    float rounding of the files); --noise <rel> <seed> adds to each receiver
    trace seeded Gaussian noise of standard deviation rel * RMS(trace), the
    same bits as tomo_add_noise on the clean traces (generate-data.cu:102-224;
    the reference's noisy generators are unseeded)."""
    import oracle as orc
    print("[case] generate-data --fp64 / --noise G=160", flush=True)
    clean_dir, noisy_dir, again_dir = tmp_path / "clean", tmp_path / "noisy", tmp_path / "again"
    for d in (clean_dir, noisy_dir, again_dir):
        d.mkdir()
        assert run(["generate-field"], d).returncode == 0
    r = run(["generate-data", "fo.txt", 160, "--fp64"], clean_dir)
    assert r.returncode == 0, r.stderr
    r = run(["generate-data", "fo.txt", 160, "--fp64", "--noise", 0.1, 7], noisy_dir)
    assert r.returncode == 0, r.stderr
    r = run(["generate-data", "fo.txt", 160, "--noise", 0.1, 8, "--fp64"], again_dir)
    assert r.returncode == 0, r.stderr
    r = run(["generate-data", "fo.txt", 160, "--noise", "x", 8], again_dir)
    assert r.returncode == 1 and "Usage" in r.stderr
    c = orc.config_reference(tomo.TomoConfig())
    ii, jj = orc.transducers_reference()
    rc, _ = orc.receivers_lines(c)
    bbox = orc.groups(ii, jj, 160)
    fo = orc.phantom_reference()
    want = orc.forward(c, fo, bbox, rc, np.float64)
    with tomo.Engine(tomo.config_reference()) as e:
        e.set_transducers(*tomo.transducers_reference(), one_based=True)
        ng = e.set_groups(160)
        clean = _side_traces(e, clean_dir, ng)
        noisy = _side_traces(e, noisy_dir, ng)
        other = _side_traces(e, again_dir, ng)
    rel = float(np.linalg.norm(clean - want) / np.linalg.norm(want))
    print(f"  clean FP64 files vs oracle: rel L2 {rel:.2e}", flush=True)
    assert rel < 1e-7
    np.testing.assert_array_equal(noisy, tomo.add_noise(clean.copy(), 0.1, 7))
    assert not np.array_equal(noisy, other)
    d = (noisy.astype(np.float64) - clean).transpose(0, 2, 1).reshape(-1, 401)
    rms = np.sqrt((clean.astype(np.float64).transpose(0, 2, 1).reshape(-1, 401) ** 2).mean(1))
    live = rms > 0
    ratio = d[live].std(1) / rms[live]
    print(f"  noise sd / trace RMS: mean {ratio.mean():.4f}, min {ratio.min():.3f}, max {ratio.max():.3f}", flush=True)
    assert abs(ratio.mean() - 0.1) < 0.002 and 0.07 < ratio.min() and ratio.max() < 0.13


@pytest.mark.gpu
def test_sirt_mimo_seq_trajectory_matches_oracle(tmp_path):
    """sirt-mimo-seq (sirt-mimo-seq.cu:337-596): 5 ART iterations (the 4 groups
    of G = 160 in order, f += s G_g / (1 + f) after each) then non-uniform
    weighted SIRT, s = omega (alpha + beta / (k + 1)^gamma) = 20000, 7
    iterations and the reference stop rules; the same loop on the FP64 oracle
    with the driver's own data files.  Per-iteration epsilon and the final
    field file against the oracle."""
    import oracle as orc
    print("[case] sirt-mimo-seq trajectory G=160", flush=True)
    assert run(["generate-field"], tmp_path).returncode == 0
    assert run(["generate-data", "fo.txt", 160], tmp_path).returncode == 0
    r = run(["sirt-mimo-seq", "fo.txt", 160, 0, 7, 1, 20000, 0, 1], tmp_path)
    assert r.returncode == 0, r.stderr
    eps_drv = table_eps(r.stdout)
    f_drv = np.loadtxt(tmp_path / "sirt-norm-160-fo.txt", dtype=np.float32).ravel()
    c = orc.config_reference(tomo.TomoConfig())
    ii, jj = orc.transducers_reference()
    rc, ra = orc.receivers_lines(c)
    bbox = orc.groups(ii, jj, 160)
    fo = tomo.read_text(tmp_path / "fo.txt", (192, 192))
    with tomo.Engine(tomo.config_reference()) as e:
        e.set_transducers(ii, jj, one_based=True)
        meas = _side_traces(e, tmp_path, e.set_groups(160)).astype(np.float64)
    f = np.zeros(fo.size)
    eps_orc = []
    prev = 100.0
    for it in range(7):
        if it < 5:
            for g in range(4):
                G, _ = orc.gradient(c, f, bbox[4 * g:4 * g + 4], rc, ra, meas[g:g + 1], np.float64)
                orc.update_group(c, f, G[0], 20000.0)
        else:
            G, _ = orc.gradient(c, f, bbox, rc, ra, meas, np.float64)
            orc.update_weighted(c, f, G, 20000.0)
        eps = orc.error(f, fo)
        eps_orc.append(eps)
        if eps > prev:
            break
        prev = eps
    print("  driver", [round(v, 4) for v in eps_drv], flush=True)
    print("  oracle", [round(v, 4) for v in eps_orc], flush=True)
    assert len(eps_drv) == len(eps_orc), (eps_drv, eps_orc)
    np.testing.assert_allclose(eps_drv, eps_orc, rtol=0, atol=5e-3)
    if len(eps_orc) == 7:
        rel = float(np.linalg.norm(f_drv - f) / np.linalg.norm(f))
        print(f"  final field rel L2 {rel:.2e}", flush=True)
        assert rel < 5e-3
