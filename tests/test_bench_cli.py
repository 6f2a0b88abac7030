"""bench.py's command line on the CPU: it measures one GPU only and says so
before it touches a device when asked for more."""
import os
import pathlib
import subprocess
import sys

import pytest

BENCH = pathlib.Path(__file__).resolve().parents[1] / "bench.py"


def run_bench(args, env_extra=None):
    env = dict(os.environ)
    env.update(env_extra or {})
    return subprocess.run([sys.executable, str(BENCH), *args], capture_output=True, text=True, timeout=120, env=env)


@pytest.mark.parametrize("gpus", ["2", "8"])
def test_more_than_one_gpu_is_refused_with_a_message(gpus):
    r = run_bench(["--gpus", gpus])
    assert r.returncode == 2
    assert "one GPU only" in r.stderr
    assert r.stdout == ""


def test_a_multi_rank_launch_is_refused_too():
    r = run_bench([], {"WORLD_SIZE": "2", "RANK": "0", "LOCAL_RANK": "0"})
    assert r.returncode == 2
    assert "one GPU only" in r.stderr


def test_help_lists_the_workloads():
    r = run_bench(["--help"])
    assert r.returncode == 0
    for w in ("direct", "batch", "bh", "serve"):
        assert w in r.stdout
