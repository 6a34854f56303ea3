"""The structs of include/igap4h.h against the reference's own headers
(include/phys.h:11-26, include/nurbs.h:4-24, include/bc.h:4-11,
include/iga.h:8-25): sizeof and every field's offset, from two translation
units compiled separately (one with the reference headers, one with ours),
and both include orders in one translation unit (the shared include guards
let either header come first).  Needs the reference tree and gcc; skipped
elsewhere."""
import ctypes
import os
import shutil
import subprocess

import pytest

REF_INC = "/root/reference/include"
OUR_INC = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include")

FIELDS = {
    "Phys": ["ndim", "ndof", "torder", "sorder", "par_mat", "ndensity", "nfield", "residual", "tangent", "field",
             "density", "densityb", "assert_par_mat"],
    "Nurbs": ["porder", "nelem", "nbasis", "nknot", "knotVector", "nboun", "boun", "nsendPart", "sendPart",
              "nrecvPart", "recvPart", "sendBlock", "recvBlock", "universalIdx_temp", "bc_periodic", "nelem_global",
              "knotVector_global"],
    "BC": ["bc_type", "par_periodic", "par_dirichlet", "par_neumann"],
    "IGA": ["phys", "nurbs", "bc", "U_hist", "nselfIdx", "selfIdx", "nsendIdx", "sendIdx", "sendPtr", "nrecvIdx",
            "recvIdx", "recvPtr", "globalIdx"],
}


def layout_source(header, fn):
    lines = ["#include <stddef.h>", '#include "%s"' % header, "void %s(long *o)" % fn, "{", "    int k = 0;"]
    for st, fields in FIELDS.items():
        lines.append("    o[k++] = (long)sizeof(%s);" % st)
        for f in fields:
            lines.append("    o[k++] = (long)offsetof(%s, %s);" % (st, f))
    lines.append("}")
    return "\n".join(lines) + "\n"


@pytest.mark.skipif(not os.path.isdir(REF_INC) or shutil.which("gcc") is None,
                    reason="needs the reference headers and gcc")
def test_struct_layouts_equal_reference(tmp_path):
    (tmp_path / "ref.c").write_text(layout_source("iga.h", "layout_ref"))
    (tmp_path / "ours.c").write_text(layout_source("igap4h.h", "layout_ours"))
    so = tmp_path / "layout.so"
    subprocess.run(["gcc", "-std=c99", "-shared", "-fPIC", "-o", str(tmp_path / "ref.o"), "-c", str(tmp_path / "ref.c"),
                    "-I", REF_INC], check=True)
    subprocess.run(["gcc", "-std=c99", "-shared", "-fPIC", "-o", str(tmp_path / "ours.o"), "-c",
                    str(tmp_path / "ours.c"), "-I", OUR_INC], check=True)
    subprocess.run(["gcc", "-shared", "-o", str(so), str(tmp_path / "ref.o"), str(tmp_path / "ours.o")], check=True)
    lib = ctypes.CDLL(str(so))
    n = sum(1 + len(f) for f in FIELDS.values())
    a, b = (ctypes.c_long * n)(), (ctypes.c_long * n)()
    lib.layout_ref(a)
    lib.layout_ours(b)
    assert list(a) == list(b)
    assert a[0] > 0


@pytest.mark.skipif(not os.path.isdir(REF_INC) or shutil.which("gcc") is None,
                    reason="needs the reference headers and gcc")
@pytest.mark.parametrize("first", ["iga.h", "igap4h.h"])
def test_headers_coexist(tmp_path, first):
    """One translation unit with both headers, in either order, using the
    reference's structs with the library's entry points (C99, -Wall -Werror)."""
    other = "igap4h.h" if first == "iga.h" else "iga.h"
    src = ('#include "%s"\n#include "%s"\n' % (first, other) +
           "int use(IGA *app, igap4h_ctx **ctx)\n{\n"
           "    igap4h_opts o = {IGAP4H_PHYS_FROM_TORDER, 4, 0, 0};\n"
           "    return igap4h_create(app, &o, ctx) + app->phys->ndof + app->nurbs->porder;\n}\n")
    (tmp_path / "both.c").write_text(src)
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-c", "-o", str(tmp_path / "both.o"),
                    str(tmp_path / "both.c"), "-I", OUR_INC, "-I", REF_INC], check=True)
