"""The C-ABI library loads and exports every entry point include/igap4h.h
declares; without a GPU, create fails with a clear error (no process exit)."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "igap4h.h")).read()
    return sorted(set(re.findall(r"\b(igap4h_[a-z_]+)\s*\(", txt)))


def test_header_matches_binding():
    from igap4standard_amd import _lib as L
    assert declared_symbols() == sorted(L.SYMBOLS)


def test_library_exports_all_symbols():
    from igap4standard_amd import _lib as L
    lib = L.lib()
    for name in declared_symbols():
        assert hasattr(lib, name), name


def test_struct_layouts_match_reference_abi():
    """Sizes/offsets of the mirrored structs on x86-64 (LP64)."""
    from igap4standard_amd import _lib as L
    assert ctypes.sizeof(L.Phys) == 4 * 4 + 8 + 4 + 4 + 6 * 8
    assert L.Phys.par_mat.offset == 16 and L.Phys.residual.offset == 32
    assert L.Nurbs.nelem.offset == 8 and L.Nurbs.nboun.offset == 40 and L.Nurbs.boun.offset == 48
    assert ctypes.sizeof(L.BC) == 32
    assert L.IGA.U_hist.offset == 24 and L.IGA.nselfIdx.offset == 32 and L.IGA.globalIdx.offset == 96


def test_create_without_gpu_is_an_error_not_an_exit():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    from igap4standard_amd import _lib as L
    from igap4standard_amd import problem as P
    from igap4standard_amd.snesiga import App
    with pytest.raises(L.IgapError, match="no HIP device"):
        App(P.reference_example(2))


def test_invalid_problem_rejected(monkeypatch):
    """Argument validation happens before any device call."""
    from igap4standard_amd import _lib as L
    lib = L.lib()
    ctx = ctypes.c_void_p()
    assert lib.igap4h_create(None, None, ctypes.byref(ctx)) == -1
    assert b"required" in lib.igap4h_last_error()
    phys = L.Phys(ndim=2, ndof=3, torder=0)
    nurbs, bc = L.Nurbs(porder=2), L.BC()
    iga = L.IGA(phys=ctypes.pointer(phys), nurbs=ctypes.pointer(nurbs), bc=ctypes.pointer(bc))
    assert lib.igap4h_create(ctypes.byref(iga), None, ctypes.byref(ctx)) == -1
    assert b"ndim" in lib.igap4h_last_error()
    assert np.int32(0) == 0
