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
