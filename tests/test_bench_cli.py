"""bench.py's command line without a GPU: more than one GPU is refused with a
message, before anything touches a device."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "exactly one GPU" in r.stderr


def test_help_lists_the_pipeline_options():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--help"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0
    for opt in ("--passes", "--cells-per-lane", "--draw-cells-per-lane", "--draw-ahead",
                "--shear-steps", "--no-small", "--no-transport", "--ensemble", "--serve", "--copy-calibration"):
        assert opt in r.stdout
