"""CPU checks of the C ABI: the libraries load, export every function their
headers declare, and fail cleanly (status code, message) without a GPU."""
import ctypes as C
import os
import re

import pytest

from retrocam_amd import _native as N

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DECL = re.compile(r"^\s*(?:const\s+)?[A-Za-z_][\w]*\s*\**\s*\b((?:rc|rcs|rct|rcj|rck)_[a-z0-9_]+)\s*\(", re.M)


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(DECL.findall(text)))


@pytest.mark.parametrize("header,path", [("retrocam_hip.h", N.HIP_LIB_PATH), ("rc_host.h", N.HOST_LIB_PATH)])
def test_library_exports_every_declared_symbol(header, path):
    names = declared(header)
    assert len(names) >= 10
    lib = C.CDLL(path)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_table_matches_header():
    assert sorted(N.HIP_SYMBOLS) == declared("retrocam_hip.h")
    assert sorted(N.HOST_SYMBOLS) == declared("rc_host.h")


def test_create_without_gpu_reports_status():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    lib = N.hip_lib()
    h = C.c_void_p()
    rc = lib.rc_create(0, C.byref(h))
    assert rc == N.RC_ERR_NO_DEVICE
    assert h.value is None
    assert b"device" in lib.rc_last_error(None)


def test_null_context_is_rejected():
    lib = N.hip_lib()
    assert lib.rc_destroy(None) == N.RC_ERR_INVALID_ARG
    assert lib.rc_reset(None) == N.RC_ERR_INVALID_ARG
    assert lib.rc_total_samples(None) == -1
    assert lib.rc_buffer_bytes(None, 0) == 0
    assert lib.rc_render_pass(None, None, None) == N.RC_ERR_INVALID_ARG
