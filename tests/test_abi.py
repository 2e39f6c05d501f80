"""The C-ABI library (no GPU needed): it is built, loads, exports every
entry point include/vcphip.h declares, validates arguments before touching
a device, and derives the same Ozaki parameters as the oracle."""
import ctypes
import os
import re

import pytest

import oracle_ffi as O
import vcp_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "vcphip.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(vcphip_[a-z0-9_]+)\s*\(", text)))


def exported_symbols():
    import subprocess

    out = subprocess.run(["nm", "-D", "--defined-only", vcp_amd.LIB_PATH], capture_output=True, text=True).stdout
    return {line.split()[-1] for line in out.splitlines() if line.split() and line.split()[-1].startswith("vcphip_")}


def test_library_exists_and_loads():
    assert os.path.exists(vcp_amd.LIB_PATH)
    assert vcp_amd.lib().vcphip_version() >= 1


def test_every_declared_symbol_is_exported():
    decl = declared_symbols()
    assert len(decl) >= 25
    missing = set(decl) - exported_symbols()
    assert not missing, missing
    # and the Python binding covers exactly the declared surface
    assert set(vcp_amd.SIGNATURES) == set(decl)


def test_strerror_names_every_code():
    L = vcp_amd.lib()
    for code in range(0, -11, -1):
        assert L.vcphip_strerror(code) not in (None, b"", b"unknown error")
    assert L.vcphip_strerror(-99) == b"unknown error"


def test_null_context_is_rejected_without_device_work():
    L = vcp_amd.lib()
    assert L.vcphip_igemm(None, 0, 4, 4, 4, None, 4, None, 4, None, 4) == -2
    assert L.vcphip_gemm_rurd_dev(None, 4, 4, 4, None, 4, None, 4, None, None, 4, 0) == -2
    assert L.vcphip_ctx_synchronize(None) == -2
    assert L.vcphip_lss(None, 4, 1, None, 4, None, 4, None, 4, None, 4, None) == -2
    assert L.vcphip_lss_dev(None, 4, 1, None, 4, None, 4, None, 4, None, 4, None) == -2
    assert L.vcphip_dd_residual(None, 4, 4, 4, None, 4, None, 4, None, 4, None, 4) == -2
    assert L.vcphip_lss_batched(None, 1, None, None, None, None, None, None, None, None, None, None, None) == -2
    assert L.vcphip_lss_batched_dev(None, 1, None, None, None, None, None, None, None, None, None, None, None) == -2
    assert L.vcphip_igemm_batched(None, 1, None, None, None, None, None, None, None, None, None) == -2
    assert L.vcphip_igemm_batched_dev(None, 1, None, None, None, None, None, None, None, None, None) == -2
    assert L.vcphip_inv_mid(None, 4, None, 4, None, 4) == -2
    assert L.vcphip_dd_solve(None, 4, 1, None, 4, None, 4, None, 4, 10, None) == -2
    assert L.vcphip_dd_solve_dev(None, 4, 1, None, 4, None, 4, None, 4, 10, None) == -2
    assert L.vcphip_dd_cholesky(None, 4, None, 4, None, 4, 2) == -2
    assert L.vcphip_dd_cholesky_dev(None, 4, None, 4, None, 4, 2) == -2
    assert L.vcphip_dd_eigsym(None, 4, None, 4, None, None, 4, 4) == -2
    assert L.vcphip_dd_eigsymge_dev(None, 4, None, 4, None, 4, None, None, 4, 1) == -2
    assert L.vcphip_inv_mid_dev(None, 4, None, 4, None, 4) == -2
    assert L.vcphip_ctx_create(0, None) == -2


def test_ozaki_config_matches_oracle():
    for k in (1, 100, 1024, 8192, 65536):
        for nmod in (12, 16, 20):
            beta, mods = vcp_amd.ozaki_config(k, nmod)
            assert beta == O.ozaki_beta(k, nmod)
            assert mods == O.ozaki_moduli(nmod)
    beta, mods = vcp_amd.ozaki_config(8192)
    assert len(mods) == 16 and beta == 55


def test_ozaki_config_rejects_bad_counts():
    L = vcp_amd.lib()
    b = ctypes.c_int()
    assert L.vcphip_ozaki_config(1024, 7, ctypes.byref(b), None, 0) == -2
    assert L.vcphip_ozaki_config(1024, 21, ctypes.byref(b), None, 0) == -2


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU is present; device-error path not reachable")
def test_context_creation_without_device_reports_device_error():
    h = ctypes.c_void_p()
    assert vcp_amd.lib().vcphip_ctx_create(0, ctypes.byref(h)) == -3
