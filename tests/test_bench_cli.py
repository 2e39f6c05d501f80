"""bench.py's command line on the CPU: more than one GPU is refused with a
message and a non-zero status before anything touches a device."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_bench_refuses_several_gpus():
    p = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2", "--steps", "1"],
                       capture_output=True, text=True, timeout=60)
    assert p.returncode == 2
    assert "one GPU" in p.stderr and "--gpus 2" in p.stderr
    assert p.stdout == ""     # no JSON line


def test_bench_help_lists_the_sections():
    p = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--help"], capture_output=True, text=True,
                       timeout=60)
    assert p.returncode == 0
    for flag in ("--steps", "--warmup", "--gpus", "--no-dd", "--no-batched", "--variant"):
        assert flag in p.stdout
