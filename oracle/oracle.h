/*
 * CPU oracle for residual/tangent assembly of C1 quadratic B-spline
 * isogeometric gradient elasticity.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: it is loaded by the
 * tests, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg, and by
 * nothing else.  The product (igap4standard_amd/, libigap4h.so) never links,
 * loads or calls it.
 *
 * It restates, loop for loop, the reference's assembly path:
 *   src/snesiga_Residual.c:13-417   (volume integral, Neumann, Dirichlet INSERT)
 *   src/snesiga_Tangent.c:13-284    (volume integral, naive 81x81x10x10
 *                                    contraction, MatZeroRowsColumns)
 *   src/snesiga_init.c:80-107,341-359 (globalIdx on one rank, AIJ stencil)
 *   mesh/nurbs/eval_Bspline.c:9-93, mesh/nurbs/evalN.c:7-62
 *   src/quadrature.c:381-429 (only the resulting nodes/weights, in closed form)
 *   src/iga_dirichlet.c:7-80, src/mathutil.c:6-64
 *   phys/ge/{residual,tangent,density}.c  (derived from the energy, not
 *                                          translated from the generated code)
 *   phys/getime/{residual,tangent}.c      (order-3 time-averaged series of the
 *                                          ge stress, identified numerically)
 *
 * Parity pins for every function are listed in the header of oracle.c.
 */
#ifndef IGAP4_ORACLE_H
#define IGAP4_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Same signature as the reference app hook `neumann` (include/bc.h:14). */
typedef void (*or_neumann_fn)(double *t, double *par_neumann, int face_id,
                              int bc_order, int idof, double *xq);

typedef struct {
    int nelem[3];                 /* elements per direction                   */
    const double *knot[3];        /* nelem[d]+5 knots, p = 2                  */
    int nquad;                    /* Gauss points per direction (reference 4) */
    int physics;                  /* 0 = ge (torder 0), 1 = getime (torder 2) */
    const double *par_mat;        /* 9 reference parameters (phys/ge/init.c)  */
    const double *elem_coef;      /* NULL or nelem_total x 6 coefficients     */
    int nboun;                    /* faces with boundary conditions           */
    const int *boun;              /* face ids 0..5 (nurbs_init.c:227-229)     */
    const int *bc_type;           /* 36: face*6 + order*3 + dof; 0 Dir, 1 Neu */
    const double *par_dirichlet;  /* 9: u_i = a_{iJ} X_J                      */
    const double *par_periodic;   /* 9: F0 added to the slots                 */
    double *par_neumann;          /* passed through to the hook               */
    or_neumann_fn neumann;        /* NULL => zero traction                    */
} or_problem;

/* ---- numerics --------------------------------------------------------- */
int    or_gauss(int n, double *x, double *w);               /* on [0,1]      */
void   or_eval_bspline(double N[270], const double *kV[3], const double xq[3]);
double or_evalN(const double *kV, int nk, int k, int i, int p, double xi);
void   or_matinv(double *mat, int n);
void   or_matinv_dot_vec(double *mat, double *vec, int n);

/* ---- pointwise physics ------------------------------------------------ */
/* coef = (A36, A37, A38, e0, e345, le^2); W is linear in these six.       */
void   or_coef_from_par(const double par[9], double coef[6]);
double or_ge_energy(const double u[30], const double coef[6]);
void   or_ge_residual(double r[30], const double u[30], const double coef[6]);
void   or_ge_tangent(double D[900], const double u[30], const double coef[6]);
/* u = [u(t_{n+1}) | u(t_n) | u(t_{n-1})], 90 values; tp = (rho, c, dt).    */
void   or_getime_residual(double r[30], const double u[90], const double coef[6],
                          const double tp[3]);
void   or_getime_tangent(double D[900], const double u[90], const double coef[6],
                         const double tp[3]);
/* d W / d coef at one point (W is linear in coef).                         */
void   or_ge_energy_coef_grad(double g[6], const double u[30]);

/* ---- mesh / index sets ------------------------------------------------ */
void    or_uniform_knots(int nelem, double *knot);           /* nelem+5       */
int64_t or_ndof(const or_problem *pb);
int64_t or_nnz(const or_problem *pb);
void    or_pattern(const or_problem *pb, int64_t *rowptr, int32_t *col);
int     or_dirichlet_count(const or_problem *pb);
int     or_dirichlet_rows(const or_problem *pb, int32_t *rows);
void    or_iga_dirichlet(double *out, const double *par, int face_id, int bc_order,
                         int idof, const double *kv[3], const int nk[3],
                         const int nb[3]);

/* ---- assembly --------------------------------------------------------- */
/* U_hist: (torder+1) x ndof in reference layout; R: ndof.                 */
int or_residual(const or_problem *pb, const double *U_hist, double *R);
/* vals: nnz, on the or_pattern() sparsity; Dirichlet rows/cols zeroed,
 * diagonal 1 (MatZeroRowsColumns(...,1.0,...)).                           */
int or_tangent(const or_problem *pb, const double *U_hist, const int64_t *rowptr,
               const int32_t *col, double *vals);
/* Same as or_tangent but only over elements [e_begin, e_end) in reference
 * (x slowest) order and without the Dirichlet step: the CPU baseline.     */
int or_tangent_elements(const or_problem *pb, const double *U_hist,
                        const int64_t *rowptr, const int32_t *col, double *vals,
                        int64_t e_begin, int64_t e_end);
int or_residual_elements(const or_problem *pb, const double *U_hist, double *R,
                         int64_t e_begin, int64_t e_end);

/* Element sums of the tangent over the element window [lo, hi) per
 * direction, no Dirichlet step, as a dense block: rows = the dofs of the
 * nodes whose support lies in the window (complete rows), columns = the dofs
 * of the window's nodes [lo, hi + 1]; node-major, dof fastest.            */
int or_tangent_window(const or_problem *pb, const double *U_hist, const int lo[3], const int hi[3],
                      double *out);
/* or_residual with the element loop split over nthreads threads.          */
int or_residual_mt(const or_problem *pb, const double *U_hist, double *R, int nthreads);

/* Density outputs at one point: phys/ge/density.c:6-66 (13) and
 * phys/getime/density.c:6-53 (2; tp = rho, c, dt).                       */
void or_ge_density(double hh[13], const double u[30], const double coef[6]);
void or_getime_density(double hh[2], const double u[90], const double coef[6], const double tp[3]);
/* iga_intdensity (src/iga_intdensity.c:10-293): HH[0..ndensity); returns
 * ndensity (13 ge, 2 getime).                                              */
int or_intdensity(const or_problem *pb, const double *U_hist, double *HH);

/* Energy functional (src/iga_intdensity.c:10-293, volume part, ge only). */
double or_energy(const or_problem *pb, const double *U_hist);

#ifdef __cplusplus
}
#endif
#endif
