"""bench.py's command line on the CPU: the module imports, the defaults are the
one-GPU run, and more than one GPU is refused with a message (exit status 2)."""

from __future__ import annotations

import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]


def test_bench_imports_and_defaults(monkeypatch):
    sys.path.insert(0, str(ROOT))
    import bench

    monkeypatch.setattr(sys, "argv", ["bench.py"])
    a = bench.parse()
    assert (a.gpus, a.workload) == (1, "ffa")
    assert bench.EP_JERK["nseg"] == 64 and bench.EP_CNN["nseg"] == 128


def test_bench_refuses_several_gpus():
    for wl in ("ffa", "ep-taylor-jerk", "ep-circular-cnn"):
        r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2",
                            "--workload", wl], capture_output=True, text=True, timeout=120)
        assert r.returncode == 2, (wl, r.stderr)
        assert "one GPU" in r.stderr
