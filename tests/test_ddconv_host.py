"""The double-double conversion of the dd product's reconstruction
(vcp_amd/csrc/vcphip_ddconv.h): its window path against its exact limb path,
bit for bit, compiled for the host (tests/cpp/test_ddconv.cpp)."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpp", "test_ddconv.cpp")


def test_window_conversion_matches_limb_path(tmp_path):
    exe = str(tmp_path / "test_ddconv")
    subprocess.run(["g++", "-O2", "-std=c++17", "-Wno-unknown-pragmas", "-o", exe, SRC], check=True)
    r = subprocess.run([exe, "2000000"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ddconv ok" in r.stdout
    cases, window = [int(x) for x in r.stdout.split()[1:4:2]]
    assert window > cases // 5   # the window path is exercised, not only the fallback
