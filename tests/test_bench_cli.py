"""bench.py's command line on the CPU: more than one GPU is refused with a
message (the engine runs on one GPU), and the options parse."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def run(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args], capture_output=True, text=True,
                          timeout=120, cwd=ROOT)


def test_more_than_one_gpu_is_refused():
    r = run("--gpus", "2")
    assert r.returncode == 2
    assert "one GPU" in r.stdout + r.stderr
    assert not r.stdout.strip().startswith("{")


def test_options_parse():
    r = run("--help")
    assert r.returncode == 0
    for opt in ("--steps", "--warmup", "--time-blocking", "--workload", "--c4-threads", "--study", "--counter-pass",
                "--eager", "--c4-mode", "--stack", "--train-iters", "--train-first-seed", "--fp"):
        assert opt in r.stdout


def _bench_module():
    import importlib.util
    spec = importlib.util.spec_from_file_location("bench_under_test", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_roofline_figures_from_the_committed_counter_profile():
    """The bench line's traffic comes from the committed counter summary of
    the shape it runs, and its fractions follow from the launch time: 20 B
    per-unit and 12 B blocked bytes of the three-step adjoint over 128 x 512^2
    x 3 cell-updates per launch."""
    b = _bench_module()
    t_adj = b.load_traffic("adjoint3_kernel", (4, 32, 64, 1, 1), 3)
    t_fwd = b.load_traffic("forward3_kernel", (4, 32, 64, 1, 1), 3)
    cells = 128 * 512 * 512 * 3
    assert t_adj is not None and 12.0 <= t_adj / cells < 12.5
    assert t_fwd is not None and 6.0 <= t_fwd / cells < 7.0
    assert b.load_traffic("adjoint3_kernel", (4, 16, 16, 1, 1), 3) is None  # no row for another shape
    line = b.kernel_line("adjoint3_kernel", 20.0, 12.0, cells, 0.275, 1000.0, t_adj, "every launch")
    assert abs(line["achieved"] - 20.0 * cells / 0.275e-3 / 1e9) < 0.1
    assert line["frac"] == line["fractions"]["per_unit"]
    assert abs(line["fractions"]["blocked_algorithmic"] - 12.0 * cells / 0.275e-3 / 8e12) < 1e-4
    assert abs(line["fractions"]["counter_measured"] - t_adj / 0.275e-3 / 8e12) < 1e-4
    assert line["bytes_per_cell_update"]["counter_measured"] == round(t_adj / cells, 3)


def test_ttt_summary():
    """The study's summary: a seed never reaching a target counts as 0 in the
    median, a seed whose plain run never gets below 100 % is flagged and left
    out, and the momentum-only speedups are summarised beside the learned
    ones."""
    bench = _bench_module()

    def run_of(seed, never, f50, f10, m50=None, momentum=0.7):
        t = {"fixed50": {"speedup": f50, "momentum_plain_speedup": m50},
             "stop_rule": {"speedup": f50, "momentum_plain_speedup": m50},
             "fixed10": {"speedup": f10, "momentum_plain_speedup": None}}
        return {"phantom_seed": seed, "plain_never_below_100": never, "momentum": momentum, "targets": t}

    runs = [run_of(1, False, 3.5, 2.0, 2.6), run_of(2, False, None, 1.5, 2.4), run_of(3, False, 1.2, 1.0, None),
            run_of(4, True, 9.0, 9.0, 9.0)]
    out = bench.ttt_summary(runs)
    assert out["flagged_never_below_100"] == [4]
    assert out["fixed50"]["speedups"] == [3.5, None, 1.2]
    assert out["fixed50"]["reached"] == "2 of 3"
    assert out["fixed50"]["at_least_3x"] == 1
    assert out["fixed50"]["median_speedup"] == 1.2          # median of 3.5, 0, 1.2
    assert out["fixed10"]["median_speedup"] == 1.5
    mp = out["fixed50"]["momentum_plain"]
    assert mp["speedups"] == [2.6, 2.4, None] and mp["reached"] == "2 of 3" and mp["median_speedup"] == 2.4
    plain = bench.ttt_summary([run_of(1, False, 2.0, 2.0, momentum=0.0)])
    assert "momentum_plain" not in plain["fixed50"]
