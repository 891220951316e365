"""The C ABI: every function declared in include/fases_core.h is exported by
libfases_core.so, and argument/state errors come back as codes (no GPU needed)."""
import ctypes as C
import re

import numpy as np
import pytest

from fases_amd import _lib
from fases_amd._lib import FasesError, check, load
from tests.conftest import ROOT


def declared_functions():
    text = (ROOT / "include" / "fases_core.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(fases_[a-z0-9_]+)\s*\(", text)))


def test_every_declared_symbol_is_exported():
    lib = load()
    names = declared_functions()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # and the ctypes table covers exactly the declared API
    assert sorted(_lib.SIGNATURES) == names


def test_create_rejects_bad_params():
    lib = load()
    p = _lib.fases_params()
    check(lib.fases_params_default(C.byref(p)))
    assert p.pml_cells == 20 and p.pml_kappa_max == 5.0 and abs(p.pml_alpha_max - 20 * np.pi) < 1e-12
    ctx = C.c_void_p()
    p.h = 0.0
    rc = lib.fases_create(C.byref(p), C.byref(ctx))
    assert rc == _lib.FASES_E_ARG and not ctx.value
    assert b"h, dt, c, rho" in lib.fases_last_error(None)
    p.h, p.dt, p.precision = 0.05, 1e-4, 7
    assert lib.fases_create(C.byref(p), C.byref(ctx)) == _lib.FASES_E_ARG


def test_null_context_calls():
    lib = load()
    assert lib.fases_run(None, 10, None, None, None, None) == _lib.FASES_E_ARG
    assert lib.fases_set_geometry(None, None, None, 10, 10, 10) == _lib.FASES_E_ARG
    assert lib.fases_destroy(None) == _lib.FASES_OK


def test_boundary_map_abi():
    oid = np.array([0xFF, 1, 0, 2, 2, 0xFF, 254], np.uint8)
    from fases_amd.solver import boundary_map
    bmap, nB = boundary_map(oid)
    assert bmap.tolist() == [-1, 0, -1, 1, 2, -1, 3] and nB == 4


def test_scene_errors(tmp_path):
    from fases_amd.solver import Scene
    with pytest.raises(FasesError) as e:
        Scene.from_obj(str(tmp_path / "none.obj"), 0.05)
    assert e.value.code == _lib.FASES_E_IO
    p = tmp_path / "bad.obj"
    p.write_text("v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 9\n")
    with pytest.raises(FasesError) as e:
        Scene.from_obj(str(p), 0.05)
    assert e.value.code == _lib.FASES_E_PARSE
    with pytest.raises(FasesError):
        Scene.shoebox(10, 10, 10, 0.0)
