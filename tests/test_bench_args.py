"""CPU checks of bench.py's command line: it runs on one GPU only, and says
so (exit status 2) before it imports torch or touches a device."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused_with_a_message():
    for n in (2, 8):
        r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", str(n),
                            "--steps", "1", "--warmup", "0"], capture_output=True, text=True,
                           timeout=120)
        assert r.returncode == 2
        assert f"one GPU only (asked for --gpus {n})" in r.stderr
        assert r.stdout == ""


def test_help_lists_the_contract_flags():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--help"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0
    for flag in ("--gpus", "--steps", "--warmup"):
        assert flag in r.stdout
