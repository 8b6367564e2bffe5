"""The C ABI library loads, exports exactly what include/lumagpu.h declares, and rejects
invalid descriptors before touching the GPU (so these run without one)."""
import ctypes as C
import re
from pathlib import Path

import pytest

from luma_amd import _abi

HEADERS = sorted((Path(__file__).resolve().parents[1] / "include").glob("*.h"))


def declared_symbols():
    names = set()
    for h in HEADERS:
        txt = h.read_text()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        names |= set(re.findall(r"\b(lumagpu_[a-z0-9_]+)\s*\(", txt))
    return names


def test_library_exports_every_declared_symbol():
    lib = _abi.load()
    names = declared_symbols()
    assert "lumagpu_step" in names and "lumagpu_grid_create" in names
    for n in sorted(names):
        assert hasattr(lib, n), n
    assert set(_abi.exported_symbols()) >= names
    assert lib.lumagpu_abi_version() == 1


def _desc(**kw):
    d = _abi.GridDesc()
    d.dims, d.Q, d.N, d.M, d.K = 3, 19, 8, 8, 8
    d.precision, d.omega, d.rho_in, d.rho0 = _abi.FP64, 1.0, 1.0, 1.0
    for k, v in kw.items():
        setattr(d, k, v)
    return d


@pytest.mark.parametrize("kw,code,msg", [
    (dict(omega=2.0), _abi.EPHYS, "omega too large"),          # GridObj_init_grids.cpp:353-356
    (dict(omega=-1.0), _abi.EPHYS, "omega"),
    (dict(Q=27), _abi.EUNSUPPORTED, "D2Q9 or D3Q19"),
    (dict(dims=2, Q=9, K=4), _abi.EINVAL, "K = 1 in 2D"),
    (dict(precision=7), _abi.EINVAL, "precision"),
    (dict(level=-1), _abi.EINVAL, "negative refinement level"),
])
def test_invalid_descriptor_is_rejected(kw, code, msg):
    lib = _abi.load()
    h = C.c_void_p()
    d = _desc(**kw)
    rc = lib.lumagpu_grid_create(C.byref(d), C.byref(h))
    assert rc == code
    assert msg in lib.lumagpu_last_error().decode()
    assert not h.value


def test_null_handles_are_rejected():
    lib = _abi.load()
    assert lib.lumagpu_step(None, 0) == _abi.EINVAL
    assert lib.lumagpu_grid_create(None, None) == _abi.EINVAL
    assert lib.lumagpu_grid_destroy(None) == 0
