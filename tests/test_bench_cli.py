"""bench.py's command line on the CPU: the options parse, and a request for
more than one GPU is refused with a message (the trajectory engine is
measured on a single MI355X)."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
BENCH = os.path.join(os.path.dirname(HERE), "bench.py")


def _run(*args):
    return subprocess.run([sys.executable, BENCH, *args], capture_output=True, text=True,
                          timeout=300)


def test_help_lists_the_workloads():
    r = _run("--help")
    assert r.returncode == 0
    for w in ("ellipsoid50k", "shell1m", "globule", "batch64", "c60", "ubq_he", "network"):
        assert w in r.stdout
    assert "--sur-lanes" in r.stdout


def test_more_than_one_gpu_is_refused():
    r = _run("--gpus", "2")
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout.strip() == ""
