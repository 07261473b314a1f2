"""The C-ABI libraries load and export every function their headers declare;
no compute call is made (no GPU in the CPU suite)."""
import ctypes as C
import os
import re
import subprocess

import pytest

import qubatron_amd as q
from qubatron_amd import _ffi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(qbt_[a-z0-9_]+)\s*\(", text)))


def _exported(lib):
    out = subprocess.run(["nm", "-D", "--defined-only", lib], capture_output=True, text=True, check=True).stdout
    return {ln.split()[-1] for ln in out.splitlines() if " T " in ln}


@pytest.mark.parametrize("header,lib", [("qbt_host.h", _ffi.HOST_LIB), ("qbt.h", _ffi.DEV_LIB)])
def test_exports_match_header(header, lib):
    names = _declared(header)
    assert len(names) > 10
    exported = _exported(lib)
    missing = [n for n in names if n not in exported]
    assert not missing, f"{lib} lacks {missing}"


def test_python_binding_lists_cover_headers():
    assert set(_declared("qbt_host.h")) == set(_ffi.HOST_SYMBOLS)
    assert set(_declared("qbt.h")) == set(_ffi.DEV_SYMBOLS)


def test_device_library_loads_and_reports():
    d = _ffi.dev()
    assert b"gfx950" in d.qbt_version()
    assert d.qbt_strerror(_ffi.QBT_ETREE) == b"octree references a node out of range"


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="checks the no-GPU error path")
def test_no_gpu_is_an_error_not_a_fallback():
    with pytest.raises(q.QbtError) as e:
        q.Context(0)
    assert e.value.code == _ffi.QBT_EDEVICE


def test_null_context_rejected():
    d = _ffi.dev()
    assert d.qbt_upload(None, None, _ffi.TYPE_INT, 48, 16, 0, 48, 2) == _ffi.QBT_EINVAL
    assert d.qbt_render(None, None, 1, None, 0, 0) == _ffi.QBT_EINVAL
    assert d.qbt_get_stats(None, None) == _ffi.QBT_EINVAL
    h = C.c_void_p()
    assert d.qbt_create(None, None) == _ffi.QBT_EINVAL
    del h


def test_gfx950_code_object_present():
    """libqbt.so carries a gfx950 code object (built by hipcc --offload-arch=gfx950)."""
    data = open(_ffi.DEV_LIB, "rb").read()
    assert b"gfx950" in data
