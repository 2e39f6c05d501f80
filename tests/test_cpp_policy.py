"""The C++ policy core (include/vcp/vcphip_policy_core.hpp, used by
include/vcp/pidhip.hpp and pddhip.hpp) compiled as a C++ program against
libvcphip and the oracle (built by __graft_entry__.build())."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "test_policy_core")


def _run(*args):
    if not os.path.exists(BIN):
        subprocess.run(["make", "-s", "-C", os.path.dirname(BIN)], check=True)
    return subprocess.run([BIN, *args], capture_output=True, text=True, timeout=120)


def test_policy_core_logic():
    r = _run()
    assert r.returncode == 0, r.stderr
    assert "policy core ok" in r.stdout


@pytest.mark.gpu
def test_policy_core_on_gpu():
    r = _run("gpu")
    assert r.returncode == 0, r.stderr
    assert "with gpu" in r.stdout
