// Host-side set-up of the assembly context: everything that does not depend
// on the iterate U and is therefore computed once at igap4h_create instead of
// inside every callback as the reference does.
#pragma once
#include <cstdint>
#include <vector>

#include "../../include/igap4h.h"

namespace igap4h {

struct Mesh {
    int ne[3];                    // elements per direction
    int nb[3];                    // basis functions per direction of the (local) window: ne + 2
    int nown[3];                  // basis functions owned (Nurbs.nbasis): nb on one rank; ne on a
                                  // rank that is not the last part of its direction
    std::vector<double> knot[3];  // ne + 5 knots each
    int nboun;
    int boun[6];
    int bc_type[36];
    double par_dirichlet[9];
    double par_periodic[9];
    double *par_neumann;          // passed through to the hook
    int64_t ndof() const { return 3LL * nb[0] * nb[1] * nb[2]; }
    int64_t nelem() const { return (int64_t)ne[0] * ne[1] * ne[2]; }
};

// Gauss-Legendre on [0,1], ascending (src/quadrature.c:381-429 values).
int gauss01(int n, double *x, double *w);

// Cox-de Boor value / k-th derivative of basis i (mesh/nurbs/evalN.c:7-62).
double evalN(const double *kv, int nk, int k, int i, int p, double xi);

// Number of AIJ nonzeros (snesiga_init.c:341-359 row counts summed).
int64_t pattern_nnz(const int nb[3]);

// Dirichlet row list in the order of snesiga_Tangent.c:214-267 (local
// window numbering): per face in boun, flag and dof, the owned face nodes.
std::vector<int32_t> dirichlet_rows(const Mesh &m);

// Final Dirichlet values per row after the INSERT sequence of
// snesiga_Residual.c:331-409 (later faces overwrite shared rows):
// mask[r] = 1 and g[r] = value for every row in the list; mask[r] = 2 for
// the other face nodes of the window on those faces (ghost nodes, whose
// Dirichlet rows the ranks owning them set).
void dirichlet_values(const Mesh &m, std::vector<unsigned char> &mask, std::vector<double> &g);

// Neumann load vector (snesiga_Residual.c:188-320); empty if hook is NULL or
// no face carries a Neumann flag.
std::vector<double> neumann_load(const Mesh &m, int nquad, igap4h_neumann_fn hook);

// Sum over the Neumann flags (face in boun, order, dof with bc_type 1) of
// the face's quadrature weights: the boundary part of the getime density
// output hh[1] (phys/getime/density.c, densityb hh[1] = 1), integrated as
// src/iga_intdensity.c:196-287 does.
double neumann_measure(const Mesh &m, int nquad);

// (A36, A37, A38, e0, e345, le^2) from par_mat (phys/ge/residual.c:45-47).
void coef_from_par(const double *par, double coef[6]);

// Coarsening of an open quadratic knot vector: kf has nef + 5 knots; kc keeps
// every other element boundary and the last one (nef odd: the last coarse
// element is one fine element), (nef + 1) / 2 + 5 knots; kc is nested in kf.
std::vector<double> coarsen_knots(const std::vector<double> &kf, int nef);

// Knot-insertion prolongation between nested quadratic spline spaces:
// P (nbf x nbc, row-major) with N^c_A = sum_a P[a][A] N^f_a, so coarse
// control values c map to fine ones P c (the effect of inserting the
// missing knots, mesh/nurbs/nurbs_knotinsert.c:7-107).  Computed exactly by
// collocation at one point inside each fine function's support.  Returns 0,
// or -1 if the spaces are not nested.
int prolongation_1d(const std::vector<double> &kf, int nef, const std::vector<double> &kc, int nec,
                    std::vector<double> &P);

} // namespace igap4h
