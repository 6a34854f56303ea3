"""MI355X-native residual/tangent assembly for C1 quadratic B-spline
isogeometric gradient elasticity (IGAP4STANDARD's phys/ge and phys/getime),
behind a C-ABI library (include/igap4h.h, csrc/libigap4h.so)."""
from .problem import (GE, GETIME, PAR_MAT_REFERENCE, Problem, bc_all_clamped, bc_clamp_stretch, bc_ex1,  # noqa: F401
                      coef_from_lame, coef_from_par, reference_example, smooth_state, stretch_problem,
                      uniform_knots)
