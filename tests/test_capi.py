"""C ABI checks that need no GPU: the library loads, exports every entry point that
include/omx.h declares, and reports a missing device with an instructive error (the
reference's equivalent: lib/MaxwellEigenmodes/test/runtests.jl:241-249)."""
import ctypes
import os
import re

import numpy as np
import pytest

import omx
from omx import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _header_functions():
    txt = open(os.path.join(ROOT, "include", "omx.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(omx_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    decl = _header_functions()
    assert len(decl) >= 14
    for name in decl:
        assert hasattr(L, name), name
    assert sorted(omx.exported_symbols()) == decl


def test_status_strings_and_version():
    L = _lib.lib()
    assert b"gfx950" in L.omx_version()
    for s in range(9):
        assert L.omx_status_string(s)
    assert L.omx_status_string(_lib.OMX_ERR_NO_DEVICE) == b"no usable gfx950 device"


def test_null_handle_is_invalid_argument():
    L = _lib.lib()
    assert L.omx_destroy(None) == _lib.OMX_ERR_INVALID_ARG
    assert L.omx_set_epsinv(None, None) == _lib.OMX_ERR_INVALID_ARG
    assert L.omx_create(None, 0, 4, 4, 1.0, 1.0, None) == _lib.OMX_ERR_INVALID_ARG


def test_epsinv_layout_conversion():
    g = omx.Grid(2.0, 1.0, 3, 2)
    e = np.arange(3 * 3 * 3 * 2, dtype=float).reshape(3, 3, 3, 2)
    c = omx._epsinv_c(e, g).ravel()
    for a in range(3):
        for b in range(3):
            for ix in range(3):
                for iy in range(2):
                    assert c[a + 3 * b + 9 * (ix + 3 * iy)] == e[a, b, ix, iy]


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU is present")
def test_no_device_gives_instructive_error():
    g = omx.Grid(2.0, 1.0, 8, 8)
    ei = np.zeros((3, 3, 8, 8))
    for a in range(3):
        ei[a, a] = 0.5
    with pytest.raises(omx.OmxError) as exc:
        omx.HIPSolver().handle(g, ei)
    assert exc.value.status == _lib.OMX_ERR_NO_DEVICE
    assert "device" in str(exc.value).lower()


def test_epsinv_layout_conversion_3d():
    g = omx.Grid3(2.0, 1.0, 0.5, 3, 2, 4)
    e = np.arange(3 * 3 * 3 * 2 * 4, dtype=float).reshape(3, 3, 3, 2, 4)
    c = omx._epsinv_c(e, g).ravel()
    for a in range(3):
        for b in range(3):
            for ix in range(3):
                for iy in range(2):
                    for iz in range(4):
                        assert c[a + 3 * b + 9 * (ix + 3 * (iy + 2 * iz))] == e[a, b, ix, iy, iz]
    with pytest.raises(omx.OmxError):
        omx._epsinv_c(e[..., 0], g)


def test_tensor_field_round_trip():
    e = np.random.default_rng(0).standard_normal((3, 3, 4, 3, 2))
    flat, shape = omx._tensor_field(e)
    assert flat.ravel()[1 + 3 * 2 + 9 * (3 + 4 * (1 + 3 * 1))] == e[1, 2, 3, 1, 1]
    assert np.array_equal(omx._tensor_back(flat, shape), e)


@pytest.mark.parametrize("args", [["--gpus", "2"], ["--gpus", "8"]])
def test_bench_refuses_multi_gpu(args):
    import subprocess
    import sys
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py")] + args, capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 2 and r.stderr.strip() and not r.stdout.strip()
