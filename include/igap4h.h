/*
 * igap4h -- MI355X (gfx950) residual and tangent assembly for C1 quadratic
 * B-spline isogeometric gradient elasticity, as a C-ABI shared library
 * (libigap4h.so).  Plain pointers and sizes; no PETSc, no torch types.
 *
 * Drop-in boundary.  The reference assembles through two PETSc SNES
 * callbacks,
 *     PetscErrorCode snesiga_Residual(SNES, Vec U, Vec R, void *app)
 *     PetscErrorCode snesiga_Tangent(SNES, Vec U, Mat J, Mat P, void *app)
 * (include/snesiga.h:8-9; src/snesiga_Residual.c:13, src/snesiga_Tangent.c:13)
 * with `app` an IGA* (include/iga.h:8-25).  The entry points below take the
 * same IGA* at set-up and the raw arrays behind U, R and the AIJ values of P
 * per call, so a PETSc shim around them is a few lines.
 *
 * The structs Phys, Nurbs, BC and IGA are declared here with the reference's
 * layouts (include/phys.h:4-26, include/nurbs.h:4-24, include/bc.h:4-11,
 * include/iga.h:8-25) under the reference's include guards, so this header
 * can be included before or after the reference's own headers.
 *
 * One rank owning the whole patch (nbasis = nelem + 2, globalIdx the
 * identity), or one rank's block of a partitioned patch as nurbs_init and
 * snesiga_init build it for mpiexec -n P (mesh/nurbs/nurbs_init.c:9-199,
 * src/snesiga_init.c:80-339): local elements and knot window, nbasis =
 * nelem (nelem + 2 on the last part of a direction) owned basis functions,
 * the boundary faces the rank touches, selfIdx / recvIdx placing the owned
 * and the received (ghost) basis functions in the local window of
 * (nelem + 2)^3, globalIdx the PETSc rows of the window.  A block's device
 * vectors (U_hist slices, R) are in that local ghosted numbering (nselfIdx +
 * nrecvIdx entries per slice, as the callbacks build U after the exchange,
 * snesiga_Residual.c:53-74) and its CSR pattern is the window's; R and the
 * values are the rank's contributions (its elements' integrals and Neumann
 * terms on every local row, owned or ghost, as its VecSetValues /
 * MatSetValues(ADD) put them), with its own Dirichlet rows set (R = U - g;
 * rows and columns zeroed, 1 on the diagonal) and the Dirichlet rows of the
 * other ranks in the window zeroed: the ranks' contributions summed through
 * globalIdx are the assembled R and P after VecAssembly / MatAssembly and
 * MatZeroRowsColumns.  igap4h_dirichlet_rows gives the rank's row list in
 * global rows (globalIdx), as row_dirichlet.  Blocks take the device entry
 * points (and energy / intdensity: the rank's part before MPI_Allreduce);
 * the host-array entry points, the Newton-Krylov driver and the batch take
 * whole patches.  Periodic directions are not supported.
 *
 * Status codes: 0 = success, negative = error (igap4h_last_error() has the
 * message).  No entry point exits the process.
 */
#ifndef IGAP4H_H
#define IGAP4H_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- reference structs (same layout; include/phys.h:4-26) ------------- */
#if !defined(PHYS)
#define PHYS
typedef void (*residual_ptr)(double *residual_, double *u, double *par);
typedef void (*tangent_ptr)(double *tangent_, double *u, double *par);
typedef void (*field_ptr)(double *field, double *u, double *par);
typedef void (*density_ptr)(double h[], double u[], double *par);
typedef void (*densityb_ptr)(double h[], double u[], double traction, double *par, int order);
typedef void (*assert_par_mat_ptr)(double *par);
typedef struct {
    int ndim;
    int ndof;
    int torder;
    int sorder;
    double *par_mat;
    int ndensity;
    int nfield;
    residual_ptr residual;
    tangent_ptr tangent;
    field_ptr field;
    density_ptr density;
    densityb_ptr densityb;
    assert_par_mat_ptr assert_par_mat;
} Phys;
#endif

/* include/nurbs.h:4-24 */
#if !defined(NURBS)
#define NURBS
typedef struct {
    int porder;
    int *nelem;
    int *nbasis;
    int *nknot;
    double **knotVector;
    int nboun;
    int *boun;
    int nsendPart;
    int *sendPart;
    int nrecvPart;
    int *recvPart;
    int *sendBlock;
    int *recvBlock;
    int *universalIdx_temp;
    int *bc_periodic;
    int *nelem_global;
    double **knotVector_global;
} Nurbs;
#endif

/* include/bc.h:4-11 */
#if !defined(BC_)
#define BC_
typedef struct {
    int *bc_type;
    double *par_periodic;
    double *par_dirichlet;
    double *par_neumann;
} BC;
#endif

/* include/iga.h:8-25 */
#if !defined(IGA_)
#define IGA_
typedef struct {
    Phys *phys;
    Nurbs *nurbs;
    BC *bc;
    double *U_hist;
    int nselfIdx;
    int *selfIdx;
    int nsendIdx;
    int *sendIdx;
    int *sendPtr;
    int nrecvIdx;
    int *recvIdx;
    int *recvPtr;
    int *globalIdx;
} IGA;
#endif

/* ---- library types ---------------------------------------------------- */
typedef struct igap4h_ctx igap4h_ctx;

enum {
    IGAP4H_OK = 0,
    IGAP4H_EINVAL = -1,      /* bad argument or unsupported problem        */
    IGAP4H_ENOMEM = -2,      /* device or host allocation failed            */
    IGAP4H_EHIP = -3,        /* a HIP runtime call failed                   */
    IGAP4H_ENODEV = -4       /* no gfx950 device visible                    */
};

/* physics on the path (phys/ge, phys/getime).  The reference dispatches
 * through host function pointers (include/phys.h:20-25) that a GPU cannot
 * call; the device code is selected by this enum instead.                 */
enum { IGAP4H_PHYS_GE = 0, IGAP4H_PHYS_GETIME = 1, IGAP4H_PHYS_FROM_TORDER = -1 };

/* precision of the tangent contraction K_e = sum_q w B^T D B             */
enum { IGAP4H_FP64 = 0, IGAP4H_FP32 = 1 };

/* element type of the CSR values array                                    */
enum { IGAP4H_VALS_F64 = 0, IGAP4H_VALS_F32 = 1 };

/* Same signature as the reference app hook (include/bc.h:14). */
typedef void (*igap4h_neumann_fn)(double *t, double par_neumann[], int face_id, int bc_order, int idof,
                                  double xq[]);

typedef struct {
    int physics;              /* IGAP4H_PHYS_*; FROM_TORDER: torder 0 -> ge, 2 -> getime */
    int nquad;                /* Gauss points per direction, 1..4; 0 -> 4 (reference)   */
    const double *elem_coef;  /* NULL: every element uses phys->par_mat; else host array
                                 n_elem x 6 of (A36, A37, A38, e0, e345, le^2), elements
                                 in the reference order (x slowest, z fastest)           */
    igap4h_neumann_fn neumann;/* traction hook (NULL: zero traction); evaluated on the
                                 host at every face Gauss point once, at create          */
} igap4h_opts;

/* ---- life cycle ------------------------------------------------------- */
/* Replaces the per-callback set-up of snesiga_Residual.c:16-102 and
 * snesiga_init.c:80-124 (index maps, AIJ pattern): builds the CSR pattern,
 * Dirichlet row list and values, Neumann load and per-element coefficients
 * on the current HIP device.  opts may be NULL (all defaults).           */
int igap4h_create(const IGA *app, const igap4h_opts *opts, igap4h_ctx **ctx);
void igap4h_destroy(igap4h_ctx *ctx);
const char *igap4h_last_error(void);

/* ndof = 3 prod(nbasis); nnz of the AIJ pattern; ndirichlet = length of the
 * Dirichlet row list with duplicates (snesiga_Tangent.c:206-213).          */
int igap4h_sizes(const igap4h_ctx *ctx, int64_t *ndof, int64_t *nnz, int64_t *ndirichlet, int64_t *nelem,
                 int *nhist);

/* The reference re-reads app->bc on every callback; a context copies it at
 * create.  Call this after changing bc_type, par_dirichlet or par_neumann
 * (load stepping): a no-op when nothing changed, else the device is
 * synchronized (queued assembly kernels on any stream finish first) and the
 * Dirichlet rows, values and mask and the Neumann load are rebuilt.  A
 * lagged tangent of the Newton driver is discarded.                        */
int igap4h_update_bc(igap4h_ctx *ctx, const BC *bc);

/* The reference calls the app's neumann hook at every face Gauss point in
 * every residual callback (snesiga_Residual.c:188-320); a context evaluates
 * it at create (and when igap4h_update_bc changes the boundary data).  Call
 * this when the hook's values change between callbacks (a traction that
 * depends on time or on par_neumann's contents): the hook is evaluated again
 * and the Neumann load replaced (device synchronized first, as
 * igap4h_update_bc).                                                       */
int igap4h_update_neumann(igap4h_ctx *ctx);

/* Fault injection: with IGAP4H_DEBUG_ALLOC_LIMIT=<bytes> in the environment,
 * every device allocation of the library larger than that is refused with
 * IGAP4H_ENOMEM (read at each allocation).                                  */

/* ---- index sets (bit-exact with the reference) ------------------------ */
/* AIJ pattern of P (snesiga_init.c:112-124 with iga_setnz :341-359): rows in
 * globalIdx order, columns ascending.  Host outputs (ndof+1, nnz).        */
int igap4h_pattern(igap4h_ctx *ctx, int64_t *rowptr, int32_t *col);
/* Device copies owned by the context (valid until igap4h_destroy).        */
int igap4h_pattern_dev(igap4h_ctx *ctx, const int64_t **d_rowptr, const int32_t **d_col);
/* Rows passed to MatZeroRowsColumns (snesiga_Tangent.c:214-270), same order,
 * duplicates kept.  Host output of length ndirichlet.                     */
int igap4h_dirichlet_rows(igap4h_ctx *ctx, int32_t *rows);

/* ---- assembly, host arrays (PCIe-inclusive) --------------------------- */
/* snesiga_Residual (src/snesiga_Residual.c:13-417).  U: current iterate
 * (ndof); U_hist: (torder+1) x ndof, slice 0 is overwritten with U as the
 * reference does (:60), the other slices are the history; R: ndof output,
 * fully overwritten (volume + Neumann, then Dirichlet rows U - g).        */
int igap4h_residual(igap4h_ctx *ctx, const double *U, double *U_hist, double *R);
/* snesiga_Tangent (src/snesiga_Tangent.c:13-284): values of P on the
 * pattern, Dirichlet rows and columns zeroed with 1 on the diagonal.
 * prec: IGAP4H_FP64 / IGAP4H_FP32 contraction; vals_type: IGAP4H_VALS_*.  */
int igap4h_tangent(igap4h_ctx *ctx, const double *U, double *U_hist, int prec, int vals_type, void *vals);

/* ---- assembly, device arrays (inputs resident in HBM) ------------------ */
/* d_U_hist: device (torder+1) x ndof (slice 0 = current iterate); d_R:
 * device ndof; stream: hipStream_t or NULL.  Asynchronous on the stream. */
int igap4h_residual_dev(igap4h_ctx *ctx, const double *d_U_hist, double *d_R, void *stream);
int igap4h_tangent_dev(igap4h_ctx *ctx, const double *d_U_hist, int prec, int vals_type, void *d_vals,
                       void *stream);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): igap4h_residual_dev without the context's n_elem x 81
 * element-vector buffer.  Element vectors are staged in a fixed device array
 * of 32,768 element vectors (21.2 MB per device, held until the process
 * ends; its address is looked up once per device by igap4h_create); larger
 * meshes are taken in chunks of that many elements, so no mesh size is
 * refused.  d_R is zeroed and accumulated in stream order (no memset node).
 * The result equals igap4h_residual_dev bit for bit.  Allocates nothing, on
 * the host or the device; the array is shared by every context on a device,
 * so calls that overlap on one device (concurrent streams or threads) race. */
int igap4h_residual_dev_noalloc(igap4h_ctx *ctx, const double *d_U_hist, double *d_R, void *stream);

/* No counterpart in the reference (a kernel choice of this library).  The
 * element tangent has two implementations with the same results (FP64 within
 * rounding, FP32 within its 1e-5 bar): IGAP4H_TANGENT_TSTRIP builds the
 * product of the pointwise tangent with the basis slots on the vector units
 * and contracts it with the slots on MFMA; IGAP4H_TANGENT_SV factors the
 * tangent through the strain variations of the dofs and contracts those on
 * MFMA (less vector work, more MFMA work).  IGAP4H_TANGENT_AUTO (the
 * default) takes the faster one as measured on MI355X: SV for both physics
 * and precisions.                                                         */
#define IGAP4H_TANGENT_AUTO 0
#define IGAP4H_TANGENT_TSTRIP 1
#define IGAP4H_TANGENT_SV 2
int igap4h_set_tangent_form(igap4h_ctx *ctx, int form);

/* ---- material coefficients ------------------------------------------- */
/* Replace the per-element coefficients (host array n_elem x 6).  The device
 * is synchronized first; a lagged tangent of the Newton driver is discarded. */
int igap4h_set_coef(igap4h_ctx *ctx, const double *elem_coef);

/* ---- energy and density outputs -------------------------------------- */
/* iga_intdensity (src/iga_intdensity.c:10-293) on one rank: HH[k] = volume
 * integral of the physics' density output k minus its Neumann boundary term
 * (densityb), for the ndensity outputs of the physics (Phys.ndensity of
 * phys/ge/init.c:14 and phys/getime/init.c: 13 for phys/ge, 2 for
 * phys/getime; HH must hold 13 doubles; *ndensity, if not NULL, gets the
 * count).  phys/ge (phys/ge/density.c:6-84): hh[0] = W, hh[1..6] the
 * stress-like outputs 2 dW/dE00, 4 S01, 4 S02, 2 dW/dE11, 4 S12, 2 dW/dE22,
 * hh[7..12] the variant indicators; boundary: minus the traction work of
 * slice 0.  phys/getime (phys/getime/density.c:6-62): hh[0] = W((u0 + u1)/2)
 * + rho |u1 - u0|^2 / (2 dt^2), hh[1] = 0; boundary: minus the traction work
 * of (u0 + u1) / 2, and hh[1] minus the measure of the Neumann faces (one per
 * flagged face, order and dof).  Tractions come from the hook given at
 * create (zero if none), as the residual's Neumann load.  The context's
 * nquad is used (the reference hard-codes 4).  d_U_hist: device U_hist with
 * its torder + 1 slices.  Fixed-order sums: repeatable on one device.      */
int igap4h_intdensity_dev(igap4h_ctx *ctx, const double *d_U_hist, double *HH, int *ndensity, void *stream);
/* Volume part of output 0 alone (the energy Pi without boundary work).   */
int igap4h_energy_dev(igap4h_ctx *ctx, const double *d_U, double *energy, void *stream);

/* ---- host utilities (no device needed) --------------------------------- */
/* write_solution / read_solution (src/ioutil.c:89-158) on one rank: slices
 * ihist0..ihist1-1 of U_hist as raw doubles in universal order (= the local
 * order on one rank), no header; read checks the file size exactly.       */
int igap4h_write_solution(const char *fname, const double *U_hist, int64_t ndof, int ihist0, int ihist1);
int igap4h_read_solution(const char *fname, double *U_hist, int64_t ndof, int ihist0, int ihist1);
/* nurbs_knotinsert (mesh/nurbs/nurbs_knotinsert.c:7-107): refine a solution
 * nref times by midpoint knot insertion, in place; U_universal is sized for
 * the fine mesh nelem[3] and holds the coarse one (nelem / 2^nref) on entry.
 * porder 2, ndim 3, non-periodic.  Exact refinement: equal to the
 * reference's masks (:41-45) except where a coarse direction has a single
 * element, for which those masks do not reproduce the spline.             */
int igap4h_knot_insert(double *U_universal, const int nelem[3], int nref, int porder, int ndim, int ndof,
                       const int bc_periodic[3]);

/* ---- adjoint: forward/backward pair for the material coefficients ------ */
/* Forward: igap4h_residual_dev with the coefficients of igap4h_set_coef.
 * Backward (phys/ge): d_g[6e + k] = lambda^T dR/dcoef[e][k] for every element
 * e and coefficient k of (A36, A37, A38, e0, e345, le^2), at U (device, ndof,
 * slice 0 of U_hist).  Exact (W is linear in the coefficients); Dirichlet
 * rows of R do not depend on them, so lambda is ignored there.  No reference
 * counterpart: the reference has one par_mat and no derivatives.           */
int igap4h_residual_vjp_coef_dev(igap4h_ctx *ctx, const double *d_U, const double *d_lambda, double *d_g,
                                 void *stream);

/* ---- batched assembly (many independent phys/ge instances) ------------ */
/* No reference counterpart (the reference assembles one patch per rank).
 * Instance i is a context with its own device U_hist (d_U_hist[i]), residual
 * d_R[i] and CSR values d_vals[i] (nnz of vals_type); prec / vals_type as for
 * igap4h_tangent_dev.  igap4h_batch_assemble computes R and K of every
 * instance: one residual launch over all elements, one gather, one tangent
 * launch per colour over all instances, and the Dirichlet diagonals,
 * captured once into a HIP graph and replayed.  The contexts and arrays must
 * outlive the batch.  A context's igap4h_update_bc (which may reallocate its
 * Dirichlet row list and Neumann load) is picked up by the next
 * igap4h_batch_assemble / igap4h_batch_tune.                               */
typedef struct igap4h_batch igap4h_batch;
int igap4h_batch_create(igap4h_ctx *const *ctxs, int n, const double *const *d_U_hist, double *const *d_R,
                        void *const *d_vals, int prec, int vals_type, igap4h_batch **batch);
void igap4h_batch_destroy(igap4h_batch *batch);
int igap4h_batch_sizes(const igap4h_batch *batch, int64_t *nelem, int *variant);
/* Autotuner: times each tangent variant of the tuning space on this batch,
 * reps launches each, and keeps the fastest.  variant = 10 x form + Gauss
 * points per MFMA chunk, form 0 = T strip, 1 = strain variation (see
 * igap4h_set_tangent_form), 2 = strain variation with its per-point data
 * sized for 3^3 Gauss points, 4 work-groups per CU (batches whose instances
 * all have nquad <= 3): FP64 {1, 2, 3, 11, 12}, FP32 {2, 3, 4, 12, 14, 22}.
 * Forms 1 and 2 compute identical values.
 * ms[8] / variants[8] (may be NULL) get the table, variant 0 where a variant
 * does not apply and past the end.
 * A new batch starts with the variant of igap4h_tangent_dev's default form
 * (12 in both precisions), so that its results equal the single-instance
 * ones. */
#define IGAP4H_BATCH_MAX_VARIANTS 8
int igap4h_batch_tune(igap4h_batch *batch, int reps, int *best_variant, double *ms, int *variants);
int igap4h_batch_set_variant(igap4h_batch *batch, int variant);
int igap4h_batch_assemble(igap4h_batch *batch, void *stream);

/* ---- Newton-Krylov driver ------------------------------------------- */
/* The reference leaves the solve to PETSc (src/snes_init.c:8-47,
 * src/snes_solve.c:10-36: NEWTONLS + BT line search, GMRES with right
 * preconditioning, one Newton step per SNESSolve in an outer loop stopped at
 * ||F|| < 1e-10 or when |dF|/F < 1e-6).  This driver runs the same loop on
 * the GPU: residual and tangent from this library, GMRES(restart) with
 * classical Gram-Schmidt twice (in place of modified Gram-Schmidt), right-
 * preconditioned by geometric multigrid, nodal 3x3 block Jacobi or multi-
 * colour block symmetric Gauss-Seidel (GPU replacements of the ASM/ILU(0) of
 * snes_init.c:38-39), or preconditioned CG.  Every reduction is a fixed-order two-pass sum, so a
 * solve is bitwise reproducible on one device.                            */
enum { IGAP4H_PC_NONE = 0, IGAP4H_PC_BJACOBI3 = 1, IGAP4H_PC_SGS = 2, IGAP4H_PC_MG = 3 };
enum { IGAP4H_KSP_GMRES = 0, IGAP4H_KSP_CG = 1 };
enum { IGAP4H_MG_GS = 0, IGAP4H_MG_CHEBYSHEV = 1, IGAP4H_MG_CHEB4 = 2 };
enum { IGAP4H_SMOOTH_SAME = 0, IGAP4H_SMOOTH_BF16 = 1, IGAP4H_SMOOTH_F16 = 2 };

typedef struct {
    double atol;       /* stop when ||R||_2 < atol (reference 1e-10); <= 0 off  */
    double rtol;       /* stop when ||R||_2 / ||R0||_2 < rtol; <= 0 off         */
    int max_newton;    /* Newton iterations                                     */
    int tangent_prec;  /* IGAP4H_FP64 / IGAP4H_FP32 contraction                */
    int vals_type;     /* matrix values used by the Krylov solve (F64/F32)     */
    int ksp_type;      /* IGAP4H_KSP_GMRES (reference, snes_init.c:32) or _CG (the
                          alternative left commented at snes_init.c:40-43)      */
    int restart;       /* GMRES restart length (default 300).  The Krylov basis is
                          allocated on demand, 32 vectors first, doubled when a
                          cycle needs more, up to restart + 1 vectors (ndof x 8 B
                          each: 6.9 MB at 64^3, so 2.1 GB for a full 300-cycle;
                          the memory is kept by the context)                     */
    double ksp_rtol;   /* linear stop: ||b - K x|| < ksp_rtol ||b|| (PETSc 1e-5) */
    int ksp_max_it;    /* linear iterations per Newton step                    */
    int ksp_guess;     /* 1: start from the previous Newton direction
                          (KSPSetInitialGuessNonzero, snes_init.c:25); 0: zero  */
    int pc;            /* IGAP4H_PC_BJACOBI3: 3x3 nodal block Jacobi;
                          IGAP4H_PC_SGS: 27-colour nodal block symmetric
                          Gauss-Seidel; IGAP4H_PC_MG: one V-cycle of geometric
                          multigrid (knot-insertion prolongation, Galerkin
                          coarse operators P^T K P, block Gauss-Seidel
                          smoothing, dense coarsest solve); IGAP4H_PC_NONE     */
    int line_search;   /* 0 full step, 1 backtracking (quadratic then cubic)   */
    int stagnation;    /* 1: stop (as diverged) when |dF|/F < 1e-6 (snes_solve.c:30) */
    int dirichlet_lift;/* 1: set U = g on the Dirichlet rows before the first step, so
                          every Newton increment vanishes there and the zeroed columns
                          of the tangent (snesiga_Tangent.c:270) lose nothing; 0: start
                          from U as given (reference)                                */
    int mg_pre, mg_post;/* IGAP4H_PC_MG: smoothing sweeps (Gauss-Seidel) or Chebyshev
                          degree before / after the coarse correction on each level
                          (0..8 each, sum >= 1; default 5 / 5)                     */
    int mg_graph;      /* IGAP4H_PC_MG: 1 replays the V-cycle from a captured HIP graph */
    int mg_smoother;   /* IGAP4H_PC_MG: IGAP4H_MG_GS (block Gauss-Seidel, mg_pre/mg_post
                          sweeps), IGAP4H_MG_CHEBYSHEV (block-Jacobi Chebyshev of
                          degree mg_pre/mg_post over [1.1 lmax / mg_cheb_ratio,
                          1.1 lmax], lmax by power iteration) or IGAP4H_MG_CHEB4
                          (Chebyshev polynomials of the fourth kind on [0, 1.1 lmax],
                          Lottes 2023; no lower bound, mg_cheb_ratio unused)       */
    int lag_jacobian;  /* rebuild the tangent and preconditioner every lag_jacobian-th
                          Newton step (PETSc -snes_lag_jacobian; 1 = every step, the
                          reference)                                              */
    int lag_persist;   /* 1: the lag counts across solves (-snes_lag_jacobian_persists),
                          so a time stepper reuses the tangent of earlier steps    */
    int mg_smooth_vals;/* IGAP4H_PC_MG with a Chebyshev smoother: matrix of the fine-level
                          smoothing passes.  IGAP4H_SMOOTH_SAME: the Krylov matrix
                          (vals_type); IGAP4H_SMOOTH_BF16: a bfloat16 copy made when
                          the preconditioner is set up (half the bytes of FP32; the
                          Krylov matrix, residual and Newton update are unchanged);
                          IGAP4H_SMOOTH_F16: a binary16 copy of the rows scaled by a
                          power of two each (entries in (-1, 1), 2^-11 rounding).
                          The 16-bit copies need at most 617 nodes in direction 2
                          (else the solve returns IGAP4H_EINVAL)                   */
    int mg_coarse_deg; /* Chebyshev smoothers: degree before / after the coarse
                          correction on the levels below the finest, at most
                          mg_pre / mg_post (0..8; 0: mg_pre / mg_post as on the
                          finest level; default 1)                                 */
    double mg_cheb_ratio;/* IGAP4H_MG_CHEBYSHEV: the smoothing interval is
                          [1.1 lmax / mg_cheb_ratio, 1.1 lmax] (lmax of Dinv K by power
                          iteration); > 1, default 200                              */
} igap4h_newton_opts;

typedef struct {
    int converged;     /* 1 converged, 0 iteration limit, -1 stagnated, -2 linear solve failed,
                          -3 line search failed, -4 residual norm not finite
                          (SNES_DIVERGED_FNORM_NAN) */
    int newton_its;
    int ksp_its;       /* summed over Newton steps */
    double fnorm0, fnorm;
    double t_residual, t_tangent, t_solve, t_total; /* seconds (line-search residuals
                                                        count in t_residual) */
    double fnorm_hist[64];  /* ||R|| before each Newton step and after the last */
    int ksp_its_hist[64];   /* Krylov iterations of each Newton step            */
    int tangents;           /* tangent assemblies in this solve                  */
} igap4h_newton_info;

int igap4h_newton_default_opts(igap4h_newton_opts *opts);
/* d_U_hist: device (torder+1) x ndof; slice 0 is the initial guess on entry
 * and the solution on return (history slices are read only).              */
int igap4h_newton_solve_dev(igap4h_ctx *ctx, double *d_U_hist, const igap4h_newton_opts *opts,
                            igap4h_newton_info *info, void *stream);
/* y = K x on the AIJ pattern; d_vals f64 or f32 per vals_type.            */
int igap4h_spmv_dev(igap4h_ctx *ctx, const void *d_vals, int vals_type, const double *d_x, double *d_y,
                    void *stream);
/* K x = b with the options' ksp_type/restart/ksp_rtol/ksp_max_it/pc; x holds
 * the initial guess on entry.  relres = ||b - K x|| / ||b|| as the method
 * tracks it (GMRES: the least-squares estimate).                           */
int igap4h_linear_solve_dev(igap4h_ctx *ctx, const void *d_vals, int vals_type, const double *d_b, double *d_x,
                            const igap4h_newton_opts *opts, int *its, double *relres, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* IGAP4H_H */
