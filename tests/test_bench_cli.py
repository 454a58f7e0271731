"""bench.py's command line on the host: one GPU only, and the help text."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _run(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args],
                          capture_output=True, text=True, timeout=120, cwd=ROOT)


def test_more_than_one_gpu_exits_with_a_message():
    r = _run("--gpus", "2")
    assert r.returncode != 0
    assert "one GPU only" in (r.stdout + r.stderr)
    assert "{" not in r.stdout  # no JSON line


def test_help_lists_workloads():
    r = _run("--help")
    assert r.returncode == 0
    for w in ("headline", "uneven", "joint", "large"):
        assert w in r.stdout
