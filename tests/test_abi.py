"""CPU checks of the C ABI library: it loads without a GPU and exports every
function include/femeko_hip.h declares; argument validation that happens
before any device work."""
import ctypes
import os
import re

import numpy as np

from femeko_hip import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = []
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            text = open(os.path.join(ROOT, "include", fn)).read()
            names += re.findall(r"^\s*(?:int|const char \*)\s*(fk_\w+)\(", text, re.M)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    L = _lib.load()
    names = declared_functions()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    # every declared function has a ctypes signature in the binding
    assert sorted(_lib.SIGNATURES) == names


def test_version_and_status_strings():
    L = _lib.load()
    assert L.fk_version() >= 100
    assert L.fk_status_string(_lib.FK_ERR_SINGULAR) == b"singular system"
    assert L.fk_status_string(12345) == b"unknown status"


def test_null_arguments_rejected_without_gpu():
    L = _lib.load()
    out = ctypes.c_void_p()
    # null handles and out-pointers fail before any HIP call
    assert L.fk_context_destroy(None) == _lib.FK_ERR_INVALID_ARG
    assert L.fk_mesh_destroy(None) == _lib.FK_ERR_INVALID_ARG
    assert L.fk_mesh_create(None, None, 0, None, 0, 1, None, ctypes.byref(out)) == _lib.FK_ERR_INVALID_ARG
    assert L.fk_assemble(None, 0) == _lib.FK_ERR_INVALID_ARG
    assert L.fk_solve(None, 1, 1e-10, 10, 1e-5, 1, 0, None, None, None) == _lib.FK_ERR_INVALID_ARG
    assert L.fk_set_timing(None, 1) == _lib.FK_ERR_INVALID_ARG
    v = np.zeros(4)
    assert L.fk_values(None, _lib.dptr(v)) == _lib.FK_ERR_INVALID_ARG
    ms = ctypes.c_double()
    assert L.fk_copy_probe(None, 1 << 20, 1, ctypes.byref(ms)) == _lib.FK_ERR_INVALID_ARG
    assert L.fk_sell_info(None, None, None, None) == _lib.FK_ERR_INVALID_ARG
    assert L.fk_set_sell_shape(None, 256, 1) == _lib.FK_ERR_INVALID_ARG
    assert L.fk_set_spmv_format(None, 1) == _lib.FK_ERR_INVALID_ARG
    assert L.fk_batch_set_format(None, 4) == _lib.FK_ERR_INVALID_ARG
    assert L.fk_spmv_burst(None, 1, 10, ctypes.byref(ms)) == _lib.FK_ERR_INVALID_ARG
    assert L.fk_set_inner_precision(None, 1) == _lib.FK_ERR_INVALID_ARG
    assert L.fk_nonlinear_spline(None, None, None, 8, None, None, 8, None, 1e-10, 1e-4, 10, 1.0, None, None,
                                 None) == _lib.FK_ERR_INVALID_ARG
    assert L.fk_status_string(_lib.FK_ERR_DOMAIN) == b"argument outside the table range"
    assert L.fk_demag_field(None, None, None, None, 1e-8, 100, 0, None, None, None) == _lib.FK_ERR_INVALID_ARG
    assert L.fk_batch_solver_format(None, None) == _lib.FK_ERR_INVALID_ARG


def test_header_cites_reference_for_each_entry_point():
    text = open(os.path.join(ROOT, "include", "femeko_hip.h")).read()
    for ref in ("src/FEM3D.jl:40-56", "src/FEM3D.jl:59-76", "src/FEM3D.jl:79-119", "src/FEM3D.jl:274-281",
                "magnetostatics.cpp", "Fluids/fluids3D.jl:142-174", "magneticField.jl:111-130"):
        assert ref in text


def test_scalar_potential_bridge_validates_before_gpu():
    """fk_scalar_potential (context-free, like the reference's scalarPotential)
    rejects null arrays with a status and a message, before any HIP call."""
    L = _lib.load()
    u = np.zeros(4)
    s = L.fk_scalar_potential(_lib.dptr(u), None, None, None, None, 4, 1, 0, None, None, None, 0)
    assert s == _lib.FK_ERR_INVALID_ARG
    assert b"null pointer" in L.fk_scalar_potential_last_error()
