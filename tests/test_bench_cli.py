"""bench.py's command line on the CPU: one GPU only (--gpus above 1 and a multi-rank launch
exit with a message before any device work), and its options parse."""
import os
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]


def _bench(*args, env=None):
    return subprocess.run([sys.executable, str(ROOT / "bench.py"), *args], capture_output=True,
                          text=True, timeout=120, env=env)


def test_more_than_one_gpu_exits_with_a_message():
    r = _bench("--gpus", "2")
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


def test_multi_rank_launch_exits_with_a_message():
    env = dict(os.environ, WORLD_SIZE="2", RANK="0", LOCAL_RANK="0")
    r = _bench(env=env)
    assert r.returncode == 2
    assert "single process" in r.stderr and r.stdout == ""


def test_options_parse():
    r = _bench("--help")
    assert r.returncode == 0
    for opt in ("--steps", "--warmup", "--config", "--solver", "--no-cpu", "--batch-windows"):
        assert opt in r.stdout
    bad = _bench("--solver", "Cholesky")
    assert bad.returncode == 2 and "invalid choice" in bad.stderr
