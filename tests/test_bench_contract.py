"""bench.py's command-line contract on the CPU: one GPU only, and the CPU
checker's sample bounds (no GPU needed)."""
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent


def test_more_than_one_gpu_is_refused_with_a_message():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120, cwd=ROOT)
    assert r.returncode != 0
    assert "--gpus 1" in r.stderr


def test_unknown_config_is_refused():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--config", "c9"],
                       capture_output=True, text=True, timeout=120, cwd=ROOT)
    assert r.returncode != 0
