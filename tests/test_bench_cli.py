"""CPU checks of bench.py's command-line contract: more than one GPU is
refused with a message before any device work, and every workload and
solver choice parses."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def run(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args], capture_output=True, text=True,
                          timeout=120)


def test_more_than_one_gpu_is_refused():
    r = run("--gpus", "2")
    assert r.returncode == 2 and "one GPU" in r.stderr and r.stdout == ""


def test_options_parse():
    r = run("--help")
    assert r.returncode == 0
    for opt in ("--workload", "--solvers", "--simp-iters", "--rtol-sweep", "--err-target", "--skip-cpu"):
        assert opt in r.stdout
    for w in ("c3", "c4", "c5", "simp"):
        assert w in r.stdout
