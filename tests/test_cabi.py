"""The C-ABI libraries load and export every entry point their headers declare
(no device computation: these run without a GPU)."""
import ctypes as C
import os
import re
import subprocess

import pytest

import wedge_amd as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[\w\s\*]*?\b(wg_\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def _exported(lib_path):
    out = subprocess.run(["nm", "-D", "--defined-only", lib_path], capture_output=True, text=True, check=True).stdout
    return {ln.split()[-1] for ln in out.splitlines() if ln.strip()}


def test_gpu_library_exports_every_declared_symbol():
    declared = _declared("wedge_gpu.h")
    assert len(declared) >= 30
    missing = [n for n in declared if n not in _exported(W.GPU_LIB_PATH)]
    assert not missing, missing


def test_host_library_exports_every_declared_symbol():
    declared = _declared("wedge_host.h")
    assert len(declared) >= 20
    missing = [n for n in declared if n not in _exported(W.HOST_LIB_PATH)]
    assert not missing, missing


def test_python_bindings_cover_every_declared_symbol():
    bound = {n for n, _, _ in W.GPU_SYMBOLS} | {n for n, _, _ in W.HOST_SYMBOLS}
    declared = set(_declared("wedge_gpu.h")) | set(_declared("wedge_host.h"))
    assert declared <= bound, sorted(declared - bound)


def test_libraries_load_and_error_strings():
    gpu = W._lib_gpu()
    W._lib_host()
    for code in (W.WG_OK, W.WG_E_INVAL, W.WG_E_NOMEM, W.WG_E_HIP, W.WG_E_STATE, W.WG_E_RANGE, W.WG_E_IO,
                 W.WG_E_FORMAT, W.WG_E_OVERFLOW, W.WG_E_NODEV):
        assert gpu.wg_error_string(code)
    assert gpu.wg_error_string(12345) == b"unknown error"


def test_null_arguments_are_rejected_without_a_device():
    gpu = W._lib_gpu()
    assert gpu.wg_open(None, 0) == W.WG_E_INVAL
    assert gpu.wg_sync(None) == W.WG_E_INVAL
    assert gpu.wg_load_csc(None, 1, 0, None, None, None, 0) == W.WG_E_INVAL
    assert gpu.wg_init_pr(None, 0.85, 0) == W.WG_E_STATE
    assert gpu.wg_pr_iteration(None, None) == W.WG_E_STATE
    assert gpu.wg_edge_pull_bfs(None, 0, None) == W.WG_E_STATE
    assert gpu.wg_set_option(None, 1, 1) == W.WG_E_INVAL
    gpu.wg_close(None)


def _has_gpu():
    h = C.c_void_p()
    rc = W._lib_gpu().wg_open(C.byref(h), 0)
    if rc == W.WG_OK:
        W._lib_gpu().wg_close(h)
        return True
    return False


def test_open_without_device_reports_nodev():
    if _has_gpu():
        pytest.skip("a HIP device is present")
    h = C.c_void_p()
    assert W._lib_gpu().wg_open(C.byref(h), 0) == W.WG_E_NODEV


def test_oracle_library_loads():
    import oracle

    g = W.rmat(6, 4, seed=1)
    o = oracle.Oracle.of(g, 1)
    assert o.V == 64 and o.threads == 1
