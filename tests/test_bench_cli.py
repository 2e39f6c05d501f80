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


def test_bench_parses_amd_smi_power_and_clock():
    sys.path.insert(0, ROOT)
    import bench

    text = ("GPU: 0\n    POWER:\n        SOCKET_POWER: 1265 W\n        GFX_VOLTAGE: N/A\n"
            "GPU: 0\n    CLOCK:\n        GFX_0:\n            CLK: 1950 MHz\n            MIN_CLK: 500 MHz\n"
            "        GFX_1:\n            CLK: 2010 MHz\n            MAX_CLK: 2400 MHz\n")
    assert bench.parse_power_clock(text) == (1265.0, 1980.0)
    assert bench.parse_power_clock("GPU: 0\n    POWER:\n        SOCKET_POWER: N/A\n") is None
