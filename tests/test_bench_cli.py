"""bench.py's command-line contract without a GPU: more than one GPU (by flag
or by a distributed launch's WORLD_SIZE) is refused with a message and exit
status 2 before any device is touched; --help lists the workloads."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def run(args, env=None):
    e = dict(os.environ)
    e.update(env or {})
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py")] + args, capture_output=True, text=True,
                          timeout=120, env=e)


def test_more_than_one_gpu_is_refused():
    r = run(["--gpus", "2"])
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


def test_world_size_above_one_is_refused():
    r = run([], env={"WORLD_SIZE": "4", "RANK": "0", "LOCAL_RANK": "0"})
    assert r.returncode == 2
    assert "WORLD_SIZE 4" in r.stderr


def test_help_lists_workloads_and_flags():
    r = run(["--help"])
    assert r.returncode == 0
    for word in ("--workload", "large", "batch", "cycle", "--deterministic", "--steps", "--warmup"):
        assert word in r.stdout
