"""bench.py refuses what it cannot measure before touching a GPU: more than one GPU (this
build measures one MI355X) and a batch beyond the members one launch advances."""

import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]


def _run(*args):
    return subprocess.run([sys.executable, str(ROOT / "bench.py"), *args], capture_output=True,
                          text=True, timeout=120, cwd=ROOT)


def test_more_than_one_gpu_exits_with_a_message():
    r = _run("--gpus", "2")
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


def test_batch_out_of_range_exits_with_a_message():
    for b in ("0", "17"):
        r = _run("--batch", b)
        assert r.returncode == 2
        assert "--batch must be 1..16" in r.stderr
