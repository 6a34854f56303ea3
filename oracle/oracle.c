/*
 * CPU oracle (test infrastructure only; see oracle.h).
 *
 * Parity pins, per function (tests/test_oracle_*.py):
 *   or_gauss            closed form vs reference legendre_handle (oracle/_ref),
 *                       exact integration of x^k, k <= 2n-1
 *   or_eval_bspline     reference eval_Bspline (oracle/_ref), partition of
 *                       unity, Marsden reproduction of x, x^2, FD of values
 *   or_evalN            reference evalN (oracle/_ref) incl. end-point special case
 *   or_iga_dirichlet    reference iga_dirichlet (oracle/_ref); affine
 *                       reproduction a * Greville abscissa
 *   or_ge_residual      reference ge_residual (oracle/_ref) on random u;
 *                       central FD of or_ge_energy; R(0) = 0
 *   or_ge_tangent       reference ge_tangent (oracle/_ref); FD of residual;
 *                       symmetry; zero value rows
 *   or_ge_energy        reference ge_density hh[0] (oracle/_ref); frame
 *                       indifference W(Q F) = W(F)
 *   or_getime_*         reference getime_residual/getime_tangent (oracle/_ref);
 *                       delta = 0 reduces to ge (R and D/4); FD of residual
 *   or_pattern, or_dirichlet_rows   brute-force enumeration of element
 *                       couplings; counts from src/snesiga_init.c:341-359 and
 *                       src/snesiga_Tangent.c:206-213
 *   or_residual         R = dPi/dU by FD of or_energy on interior rows; patch
 *                       test (affine U -> interior rows of R vanish for W
 *                       quadratic); Dirichlet rows U - g in closed form
 *   or_tangent          FD of or_residual (K dU ~ R(U+dU) - R(U)); symmetry;
 *                       dense brute force sum_q B^T D B on one element
 *   or_tangent_window   the rows of or_tangent on windows of a small mesh
 *                       (same element matrices, other scatter)
 *   or_residual_mt      or_residual on the same input (order of the sums
 *                       only)
 *   or_ge_density,      reference ge_density / getime_density, all outputs
 *   or_getime_density   (golden vectors from oracle/_ref, value slots set)
 *   or_intdensity       dHH[0]/dU = R off the Dirichlet rows (FD, Neumann
 *                       hook on); getime HH[1] = -(flags x face area)
 */
#include "oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define PORDER 2
#define NDOF 3
#define NDDIM 10
#define NBPE 27
#define MTOL 1.e-15 /* guard added to every knot-span denominator (eval_Bspline.c:7) */

/* ======================================================================= */
/* Gauss-Legendre on [0,1], ascending nodes (quadrature.c:381-429 produces  */
/* the same table numerically).                                            */
/* ======================================================================= */
int or_gauss(int n, double *x, double *w)
{
    switch (n) {
    case 1: x[0] = 0.5; w[0] = 1.0; return 0;
    case 2: {
        double a = 0.5 / sqrt(3.0);
        x[0] = 0.5 - a; x[1] = 0.5 + a; w[0] = w[1] = 0.5; return 0;
    }
    case 3: {
        double a = 0.5 * sqrt(0.6);
        x[0] = 0.5 - a; x[1] = 0.5; x[2] = 0.5 + a;
        w[0] = w[2] = 5.0 / 18.0; w[1] = 8.0 / 18.0; return 0;
    }
    case 4: {
        double s = 2.0 / 7.0 * sqrt(6.0 / 5.0);
        double ao = 0.5 * sqrt(3.0 / 7.0 + s), ai = 0.5 * sqrt(3.0 / 7.0 - s);
        x[0] = 0.5 - ao; x[1] = 0.5 - ai; x[2] = 0.5 + ai; x[3] = 0.5 + ao;
        w[0] = w[3] = (18.0 - sqrt(30.0)) / 72.0;
        w[1] = w[2] = (18.0 + sqrt(30.0)) / 72.0;
        return 0;
    }
    default: return -1;
    }
}

/* ======================================================================= */
/* Quadratic B-spline basis at one interior point, 27 functions x 10 slots.  */
/* Slot order (value, d/dx, d/dy, d/dz, xx, xy, xz, yy, yz, zz) and ibpe =  */
/* 9*ax + 3*ay + az as in eval_Bspline.c:77-92.  kV[d] points at the 6-knot */
/* window knot[d] + e_d; every span denominator carries +MTOL.              */
/* ======================================================================= */
static void bspline1d(const double *k, double x, double v[3], double d1[3], double d2[3])
{
    /* degree 0: only the span [k2, k3) is active */
    double n0[5] = {0, 0, 1, 0, 0};
    double n1[4], dn1[4];
    for (int i = 0; i < 4; i++) {
        n1[i] = (x - k[i]) / (k[i + 1] - k[i] + MTOL) * n0[i]
              + (k[i + 2] - x) / (k[i + 2] - k[i + 1] + MTOL) * n0[i + 1];
        dn1[i] = 1.0 / (k[i + 1] - k[i] + MTOL) * n0[i]
               - 1.0 / (k[i + 2] - k[i + 1] + MTOL) * n0[i + 1];
    }
    for (int i = 0; i < 3; i++) {
        double dl = k[i + 2] - k[i] + MTOL, dr = k[i + 3] - k[i + 1] + MTOL;
        v[i]  = (x - k[i]) / dl * n1[i] + (k[i + 3] - x) / dr * n1[i + 1];
        d1[i] = 2.0 / dl * n1[i] - 2.0 / dr * n1[i + 1];
        d2[i] = 2.0 / dl * dn1[i] - 2.0 / dr * dn1[i + 1];
    }
}

void or_eval_bspline(double N[270], const double *kV[3], const double xq[3])
{
    double v[3][3], d1[3][3], d2[3][3];
    for (int d = 0; d < 3; d++) bspline1d(kV[d], xq[d], v[d], d1[d], d2[d]);
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            for (int k = 0; k < 3; k++) {
                double *n = N + NDDIM * (9 * i + 3 * j + k);
                n[0] = v[0][i] * v[1][j] * v[2][k];
                n[1] = d1[0][i] * v[1][j] * v[2][k];
                n[2] = v[0][i] * d1[1][j] * v[2][k];
                n[3] = v[0][i] * v[1][j] * d1[2][k];
                n[4] = d2[0][i] * v[1][j] * v[2][k];
                n[5] = d1[0][i] * d1[1][j] * v[2][k];
                n[6] = d1[0][i] * v[1][j] * d1[2][k];
                n[7] = v[0][i] * d2[1][j] * v[2][k];
                n[8] = v[0][i] * d1[1][j] * d1[2][k];
                n[9] = v[0][i] * v[1][j] * d2[2][k];
            }
}

/* ======================================================================= */
/* Single basis function N_{i,p}^{(k)}(xi) by Cox-de Boor (evalN.c:7-62):   */
/* value by the recursion with 0/0 := 0 and the closed right end treated as */
/* inside the last function's support; derivative k <= p as p!/(p-k)! times */
/* a combination of degree p-k functions (The NURBS Book, eq. 2.9).         */
/* ======================================================================= */
double or_evalN(const double *kV, int nk, int k, int i, int p, double xi)
{
    if (k == 0) {
        if (i == nk - p - 2 && fabs(xi - kV[nk - 1]) < 1.e-15 &&
            fabs(kV[nk - 1] - kV[nk - 2]) < 1.e-15)
            return 1.0;
        if (p == 0) return (xi >= kV[i] && kV[i + 1] > xi) ? 1.0 : 0.0;
        int lz = (kV[i + p] == kV[i]), rz = (kV[i + p + 1] == kV[i + 1]);
        if (lz && rz) return 0.0;
        if (lz) return (kV[i + p + 1] - xi) / (kV[i + p + 1] - kV[i + 1]) * or_evalN(kV, nk, 0, i + 1, p - 1, xi);
        if (rz) return (xi - kV[i]) / (kV[i + p] - kV[i]) * or_evalN(kV, nk, 0, i, p - 1, xi);
        return (xi - kV[i]) / (kV[i + p] - kV[i]) * or_evalN(kV, nk, 0, i, p - 1, xi)
             + (kV[i + p + 1] - xi) / (kV[i + p + 1] - kV[i + 1]) * or_evalN(kV, nk, 0, i + 1, p - 1, xi);
    }
    if (k < 0 || k > p) return 0.0;
    double a[4][4];
    a[0][0] = 1.0;
    for (int r = 1; r <= k; r++) {
        double den = kV[i + p - r + 1] - kV[i];
        a[r][0] = (den == 0.0) ? 0.0 : a[r - 1][0] / den;
        for (int c = 1; c < r; c++) {
            den = kV[i + p + c - r + 1] - kV[i + c];
            a[r][c] = (den == 0.0) ? 0.0 : (a[r - 1][c] - a[r - 1][c - 1]) / den;
        }
        den = kV[i + p + 1] - kV[i + r];
        a[r][r] = (den == 0.0) ? 0.0 : -a[r - 1][r - 1] / den;
    }
    double s = 0.0;
    for (int j = 0; j <= k; j++) s += a[k][j] * or_evalN(kV, nk, 0, i + j, p - k, xi);
    double f = 1.0;
    for (int m = p; m > p - k; m--) f *= m;
    return s * f;
}

/* Closed-form inverse (mathutil.c:17-64 for n = 2, 3). */
void or_matinv(double *m, int n)
{
    double r[9];
    if (n == 2) {
        double idet = 1. / (-(m[1] * m[2]) + m[0] * m[3]);
        r[0] = idet * m[3]; r[1] = idet * (-m[1]); r[2] = idet * (-m[2]); r[3] = idet * m[0];
    } else if (n == 3) {
        double c00 = m[4] * m[8] - m[5] * m[7], c01 = m[5] * m[6] - m[3] * m[8], c02 = m[3] * m[7] - m[4] * m[6];
        double idet = 1.0 / (m[0] * c00 + m[1] * c01 + m[2] * c02);
        r[0] = idet * c00; r[1] = idet * (m[2] * m[7] - m[1] * m[8]); r[2] = idet * (m[1] * m[5] - m[2] * m[4]);
        r[3] = idet * c01; r[4] = idet * (m[0] * m[8] - m[2] * m[6]); r[5] = idet * (m[2] * m[3] - m[0] * m[5]);
        r[6] = idet * c02; r[7] = idet * (m[1] * m[6] - m[0] * m[7]); r[8] = idet * (m[0] * m[4] - m[1] * m[3]);
    } else {
        return;
    }
    memcpy(m, r, sizeof(double) * n * n);
}

void or_matinv_dot_vec(double *m, double *v, int n)
{
    double t[3] = {0, 0, 0};
    or_matinv(m, n);
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) t[i] += m[n * i + j] * v[j];
    for (int i = 0; i < n; i++) v[i] = t[i];
}

/* ======================================================================= */
/* Truncated Taylor arithmetic ("jets"): c[k] is the s^k coefficient of a   */
/* polynomial in s, products truncated at degree nd <= 3.  Degree 0 is      */
/* plain double arithmetic.                                                 */
/* ======================================================================= */
#define JN 4
typedef struct { double c[JN]; } jet;

static inline jet jcst(double x) { jet r = {{x, 0, 0, 0}}; return r; }
static inline jet jadd(jet a, jet b) { for (int k = 0; k < JN; k++) a.c[k] += b.c[k]; return a; }
static inline jet jsub(jet a, jet b) { for (int k = 0; k < JN; k++) a.c[k] -= b.c[k]; return a; }
static inline jet jsc(jet a, double s) { for (int k = 0; k < JN; k++) a.c[k] *= s; return a; }
static inline jet jmul(jet a, jet b, int nd)
{
    jet r = {{0, 0, 0, 0}};
    for (int k = 0; k <= nd; k++)
        for (int i = 0; i <= k; i++) r.c[k] += a.c[i] * b.c[k - i];
    return r;
}
static inline int jnz(jet a) { return a.c[0] != 0 || a.c[1] != 0 || a.c[2] != 0 || a.c[3] != 0; }

/* index of the symmetric second-derivative slot (I,K) within 4..9 */
static const int HIDX[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};
#define FS(a, J) (10 * (a) + 1 + (J))
#define HS(a, h) (10 * (a) + 4 + (h))

/*
 * Energy of phys/ge (phys/ge/density.c:49), with coef = (A36,A37,A38,e0,e345,le^2):
 *   W = A36 e3 (e3^2 - 3 e2^2) + A37 q + A38 q^2 + e0 e1^2 + e345 (E01^2+E02^2+E12^2)
 *       + le^2 sum_K sum_I (g_IK - mean_I g_IK)^2,
 *   F = I + grad u, E = (F^T F - I)/2, e1 = tr E/sqrt3, e2 = (E00-E11)/sqrt2,
 *   e3 = (E00+E11-2E22)/sqrt6, q = e2^2 + e3^2, g_IK = d_K E_II = sum_a F_aI u_a,IK.
 * (le^2 sum_K |dev g_K|^2 equals le^2 (|grad e2|^2 + |grad e3|^2).)
 *
 * ge_derivs evaluates W, r = dW/du_slot (30) and D = d^2W/du^2 (30x30) with
 * jets: r = sum_p dW/dE_p grad E_p + sum_IK tau_IK grad g_IK, and
 * D = J_E^T Psi'' J_E + sum_p Psi'_p hess E_p + J_g^T P J_g + sum tau hess g.
 */
typedef struct {
    jet W;
    jet r[30];
    jet D[900];
} ge_out;

static void ge_derivs(const jet u[30], const double coef[6], int nd, int want_r,
                      int want_D, ge_out *o)
{
    const double s2 = 1.0 / sqrt(2.0), s3 = 1.0 / sqrt(3.0), s6 = 1.0 / sqrt(6.0);
    const double A36 = coef[0], A37 = coef[1], A38 = coef[2], e0 = coef[3], e345 = coef[4], lsq = coef[5];
    jet F[3][3], H[3][6];
    for (int a = 0; a < 3; a++) {
        for (int J = 0; J < 3; J++) {
            F[a][J] = u[FS(a, J)];
            if (a == J) F[a][J].c[0] += 1.0;
        }
        for (int h = 0; h < 6; h++) H[a][h] = u[HS(a, h)];
    }
    /* Green-Lagrange strain, eps = (E00, E11, E22, E01, E02, E12) */
    static const int PI_[6] = {0, 1, 2, 0, 0, 1}, PJ_[6] = {0, 1, 2, 1, 2, 2};
    jet eps[6];
    for (int p = 0; p < 6; p++) {
        jet t = jcst(0.0);
        for (int a = 0; a < 3; a++) t = jadd(t, jmul(F[a][PI_[p]], F[a][PJ_[p]], nd));
        if (p < 3) t.c[0] -= 1.0;
        eps[p] = jsc(t, 0.5);
    }
    jet h1 = jsc(jadd(jadd(eps[0], eps[1]), eps[2]), s3);
    jet h2 = jsc(jsub(eps[0], eps[1]), s2);
    jet h3 = jsc(jsub(jadd(eps[0], eps[1]), jsc(eps[2], 2.0)), s6);
    jet h22 = jmul(h2, h2, nd), h33 = jmul(h3, h3, nd), q = jadd(h22, h33);
    /* normal strain gradients g_IK and their deviatoric parts */
    jet g[3][3], tau[3][3];
    for (int I = 0; I < 3; I++)
        for (int K = 0; K < 3; K++) {
            jet t = jcst(0.0);
            for (int a = 0; a < 3; a++) t = jadd(t, jmul(F[a][I], H[a][HIDX[I][K]], nd));
            g[I][K] = t;
        }
    jet gdev2 = jcst(0.0);
    for (int K = 0; K < 3; K++) {
        jet m = jsc(jadd(jadd(g[0][K], g[1][K]), g[2][K]), 1.0 / 3.0);
        for (int I = 0; I < 3; I++) {
            jet d = jsub(g[I][K], m);
            tau[I][K] = jsc(d, 2.0 * lsq);
            gdev2 = jadd(gdev2, jmul(d, d, nd));
        }
    }
    /* energy */
    {
        jet W = jmul(jsc(h3, A36), jsub(h33, jsc(h22, 3.0)), nd);
        W = jadd(W, jsc(q, A37));
        W = jadd(W, jsc(jmul(q, q, nd), A38));
        W = jadd(W, jsc(jmul(h1, h1, nd), e0));
        jet sh = jadd(jadd(jmul(eps[3], eps[3], nd), jmul(eps[4], eps[4], nd)), jmul(eps[5], eps[5], nd));
        W = jadd(W, jsc(sh, e345));
        W = jadd(W, jsc(gdev2, lsq));
        o->W = W;
    }
    if (!want_r && !want_D) return;
    /* dPsi/d(e1,e2,e3) */
    jet P1 = jsc(h1, 2.0 * e0);
    jet c2 = jadd(jadd(jsc(h3, -6.0 * A36), jcst(2.0 * A37)), jsc(q, 4.0 * A38));
    jet P2 = jmul(h2, c2, nd);
    jet P3 = jadd(jsc(jsub(h33, h22), 3.0 * A36), jmul(h3, jadd(jcst(2.0 * A37), jsc(q, 4.0 * A38)), nd));
    static const double Q[3][3] = {{0.57735026918962573, 0.57735026918962573, 0.57735026918962573},
                                   {0.70710678118654752, -0.70710678118654752, 0.0},
                                   {0.40824829046386302, 0.40824829046386302, -0.81649658092772603}};
    jet Pe[6];
    for (int p = 0; p < 3; p++) Pe[p] = jadd(jadd(jsc(P1, Q[0][p]), jsc(P2, Q[1][p])), jsc(P3, Q[2][p]));
    for (int p = 3; p < 6; p++) Pe[p] = jsc(eps[p], 2.0 * e345);
    /* grad eps_p with respect to F_aJ: de[p][a][J] */
    jet de[6][3][3];
    memset(de, 0, sizeof(de));
    for (int a = 0; a < 3; a++) {
        for (int I = 0; I < 3; I++) de[I][a][I] = F[a][I];
        de[3][a][0] = jsc(F[a][1], 0.5); de[3][a][1] = jsc(F[a][0], 0.5);
        de[4][a][0] = jsc(F[a][2], 0.5); de[4][a][2] = jsc(F[a][0], 0.5);
        de[5][a][1] = jsc(F[a][2], 0.5); de[5][a][2] = jsc(F[a][1], 0.5);
    }
    /* dense gradients of g_IK over the 30 slots */
    jet dg[3][3][30];
    memset(dg, 0, sizeof(dg));
    for (int I = 0; I < 3; I++)
        for (int K = 0; K < 3; K++)
            for (int b = 0; b < 3; b++) {
                dg[I][K][FS(b, I)] = H[b][HIDX[I][K]];
                dg[I][K][HS(b, HIDX[I][K])] = F[b][I];
            }
    if (want_r) {
        for (int s = 0; s < 30; s++) o->r[s] = jcst(0.0);
        for (int p = 0; p < 6; p++)
            for (int a = 0; a < 3; a++)
                for (int J = 0; J < 3; J++)
                    if (jnz(de[p][a][J])) o->r[FS(a, J)] = jadd(o->r[FS(a, J)], jmul(Pe[p], de[p][a][J], nd));
        for (int I = 0; I < 3; I++)
            for (int K = 0; K < 3; K++)
                for (int s = 0; s < 30; s++)
                    if (jnz(dg[I][K][s])) o->r[s] = jadd(o->r[s], jmul(tau[I][K], dg[I][K][s], nd));
    }
    if (!want_D) return;
    jet *D = o->D;
    for (int i = 0; i < 900; i++) D[i] = jcst(0.0);
    /* second derivatives of Psi in (e1,e2,e3), mapped to eps by Q^T M Q */
    jet M22 = jadd(jadd(jsc(h3, -6.0 * A36), jcst(2.0 * A37)), jsc(jadd(jsc(h22, 3.0), h33), 4.0 * A38));
    jet M23 = jmul(h2, jadd(jcst(-6.0 * A36), jsc(h3, 8.0 * A38)), nd);
    jet M33 = jadd(jadd(jsc(h3, 6.0 * A36), jcst(2.0 * A37)), jsc(jadd(h22, jsc(h33, 3.0)), 4.0 * A38));
    jet M[3][3] = {{jcst(2.0 * e0), jcst(0.0), jcst(0.0)}, {jcst(0.0), M22, M23}, {jcst(0.0), M23, M33}};
    jet Pee[6][6];
    for (int p = 0; p < 6; p++)
        for (int qq = 0; qq < 6; qq++) Pee[p][qq] = jcst(0.0);
    for (int p = 0; p < 3; p++)
        for (int qq = 0; qq < 3; qq++) {
            jet t = jcst(0.0);
            for (int k = 0; k < 3; k++)
                for (int l = 0; l < 3; l++) t = jadd(t, jsc(M[k][l], Q[k][p] * Q[l][qq]));
            Pee[p][qq] = t;
        }
    for (int p = 3; p < 6; p++) Pee[p][p] = jcst(2.0 * e345);
    /* material part J_E^T Pee J_E on the F-F block */
    jet Z[6][3][3];
    for (int p = 0; p < 6; p++)
        for (int b = 0; b < 3; b++)
            for (int L = 0; L < 3; L++) {
                jet t = jcst(0.0);
                for (int qq = 0; qq < 6; qq++)
                    if (jnz(de[qq][b][L])) t = jadd(t, jmul(Pee[p][qq], de[qq][b][L], nd));
                Z[p][b][L] = t;
            }
    for (int a = 0; a < 3; a++)
        for (int J = 0; J < 3; J++)
            for (int b = 0; b < 3; b++)
                for (int L = 0; L < 3; L++) {
                    jet t = jcst(0.0);
                    for (int p = 0; p < 6; p++)
                        if (jnz(de[p][a][J])) t = jadd(t, jmul(de[p][a][J], Z[p][b][L], nd));
                    D[30 * FS(a, J) + FS(b, L)] = jadd(D[30 * FS(a, J) + FS(b, L)], t);
                }
    /* geometric part sum_p Pe_p hess eps_p = delta_ab S_JL */
    for (int a = 0; a < 3; a++) {
        for (int I = 0; I < 3; I++) D[30 * FS(a, I) + FS(a, I)] = jadd(D[30 * FS(a, I) + FS(a, I)], Pe[I]);
        static const int OJ[3] = {0, 0, 1}, OL[3] = {1, 2, 2};
        for (int m = 0; m < 3; m++) {
            jet hs = jsc(Pe[3 + m], 0.5);
            int x = FS(a, OJ[m]), y = FS(a, OL[m]);
            D[30 * x + y] = jadd(D[30 * x + y], hs);
            D[30 * y + x] = jadd(D[30 * y + x], hs);
        }
    }
    /* gradient part: J_g^T P J_g with P = 2 le^2 (delta_II' - 1/3) per K */
    for (int K = 0; K < 3; K++)
        for (int I = 0; I < 3; I++) {
            jet Y[30];
            for (int s = 0; s < 30; s++) {
                jet t = jcst(0.0);
                for (int I2 = 0; I2 < 3; I2++) {
                    double c = 2.0 * lsq * ((I == I2 ? 1.0 : 0.0) - 1.0 / 3.0);
                    t = jadd(t, jsc(dg[I2][K][s], c));
                }
                Y[s] = t;
            }
            for (int x = 0; x < 30; x++) {
                if (!jnz(dg[I][K][x])) continue;
                for (int y = 0; y < 30; y++)
                    if (jnz(Y[y])) D[30 * x + y] = jadd(D[30 * x + y], jmul(dg[I][K][x], Y[y], nd));
            }
        }
    /* sum tau_IK hess g_IK: d^2 g_IK / dF_bI du_b,IK = 1 */
    for (int I = 0; I < 3; I++)
        for (int K = 0; K < 3; K++)
            for (int b = 0; b < 3; b++) {
                int x = FS(b, I), y = HS(b, HIDX[I][K]);
                D[30 * x + y] = jadd(D[30 * x + y], tau[I][K]);
                D[30 * y + x] = jadd(D[30 * y + x], tau[I][K]);
            }
    /* exact symmetry (upper triangle mirrored) */
    for (int x = 0; x < 30; x++)
        for (int y = x + 1; y < 30; y++) D[30 * y + x] = D[30 * x + y];
}

/* e36..e38 of phys/ge/residual.c:45-47 simplify to these closed forms. */
void or_coef_from_par(const double par[9], double coef[6])
{
    double r = par[0], c = par[1];
    coef[0] = 4.0 * (1.0 - c) / (r * r * r * (1.0 + 2.0 * c));
    coef[1] = -6.0 * c / (r * r * (1.0 + 2.0 * c));
    coef[2] = 3.0 / (r * r * r * r * (1.0 + 2.0 * c));
    coef[3] = par[2];
    coef[4] = par[3];
    coef[5] = par[4] * par[4];
}

static ge_out *ge_scratch(void)
{
    static _Thread_local ge_out *o = NULL;  /* one per thread (or_residual_mt) */
    if (!o) o = (ge_out *)malloc(sizeof(ge_out));
    return o;
}

static void to_jets(jet *uj, const double *u)
{
    for (int s = 0; s < 30; s++) uj[s] = jcst(u[s]);
}

double or_ge_energy(const double u[30], const double coef[6])
{
    jet uj[30];
    to_jets(uj, u);
    ge_out *o = ge_scratch();
    ge_derivs(uj, coef, 0, 0, 0, o);
    return o->W.c[0];
}

/* phys/ge/density.c:6-66, all 13 outputs: W; twice the derivatives of W in
 * the six strains (the off-diagonal ones appear once in W); the variant
 * indicators of (e2, e3) and the energy-based ones. */
void or_ge_density(double hh[13], const double u[30], const double coef[6])
{
    double F[3][3], E[3][3];
    for (int a = 0; a < 3; a++)
        for (int J = 0; J < 3; J++) F[a][J] = (a == J ? 1.0 : 0.0) + u[10 * a + 1 + J];
    for (int I = 0; I < 3; I++)
        for (int J = 0; J < 3; J++) {
            double t = 0.0;
            for (int a = 0; a < 3; a++) t += F[a][I] * F[a][J];
            E[I][J] = 0.5 * (t - (I == J ? 1.0 : 0.0));
        }
    const double s3 = sqrt(3.0), s2 = sqrt(2.0), s6 = sqrt(6.0);
    double e1 = (E[0][0] + E[1][1] + E[2][2]) / s3;
    double e2 = (E[0][0] - E[1][1]) / s2;
    double e3 = (E[0][0] + E[1][1] - 2.0 * E[2][2]) / s6;
    double q = e2 * e2 + e3 * e3;
    double A36 = coef[0], A37 = coef[1], A38 = coef[2], e0 = coef[3], e345 = coef[4];
    double W1 = 2.0 * e0 * e1;
    double W2 = -6.0 * A36 * e2 * e3 + 2.0 * A37 * e2 + 4.0 * A38 * q * e2;
    double W3 = 3.0 * A36 * (e3 * e3 - e2 * e2) + 2.0 * A37 * e3 + 4.0 * A38 * q * e3;
    double W = or_ge_energy(u, coef);
    hh[0] = W;
    hh[1] = 2.0 * (W1 / s3 + W2 / s2 + W3 / s6);
    hh[2] = 4.0 * e345 * E[0][1];
    hh[3] = 4.0 * e345 * E[0][2];
    hh[4] = 2.0 * (W1 / s3 - W2 / s2 + W3 / s6);
    hh[5] = 4.0 * e345 * E[1][2];
    hh[6] = 2.0 * (W1 / s3 - 2.0 * W3 / s6);
    hh[7] = (double)(e2 > 0 && e3 > -0.5 * e2);
    hh[8] = (double)(e2 < 0 && e3 > 0.5 * e2);
    hh[9] = (double)(e3 < 0.5 * e2 && e3 < -0.5 * e2);
    hh[10] = (double)(W < -0.5 && e2 > 0 && e3 > 0);
    hh[11] = (double)(W < -0.5 && e2 < 0 && e3 > 0);
    hh[12] = (double)(W < -0.5 && e3 < 0);
}

/* phys/getime/density.c:6-53: W at (u0 + u1) / 2 plus the kinetic energy
 * rho |u1 - u0|^2 / (2 dt^2) of the value slots; hh[1] = 0. */
void or_getime_density(double hh[2], const double u[90], const double coef[6], const double tp[3])
{
    double ub[30], kin = 0.0;
    for (int k = 0; k < 30; k++) ub[k] = 0.5 * (u[30 + k] + u[60 + k]);
    for (int i = 0; i < 3; i++) {
        double d = u[30 + 10 * i] - u[60 + 10 * i];
        kin += tp[0] * d * d / (2.0 * tp[2] * tp[2]);
    }
    hh[0] = or_ge_energy(ub, coef) + kin;
    hh[1] = 0.0;
}

void or_ge_residual(double r[30], const double u[30], const double coef[6])
{
    jet uj[30];
    to_jets(uj, u);
    ge_out *o = ge_scratch();
    ge_derivs(uj, coef, 0, 1, 0, o);
    for (int s = 0; s < 30; s++) r[s] = o->r[s].c[0];
}

void or_ge_tangent(double D[900], const double u[30], const double coef[6])
{
    jet uj[30];
    to_jets(uj, u);
    ge_out *o = ge_scratch();
    ge_derivs(uj, coef, 0, 0, 1, o);
    for (int i = 0; i < 900; i++) D[i] = o->D[i].c[0];
}

void or_ge_energy_coef_grad(double gr[6], const double u[30])
{
    for (int k = 0; k < 6; k++) {
        double c[6] = {0, 0, 0, 0, 0, 0};
        c[k] = 1.0;
        gr[k] = or_ge_energy(u, c);
    }
}

/*
 * phys/getime (torder 2): u = t_{n+1}, u1 = t_n, u0 = t_{n-1}.
 * Value slots (residual.c:4127,4137,4147):
 *     c (u - u0) / (2 dt) + rho (u + u0 - 2 u1) / dt^2
 * Gradient slots: the order-3 truncation of the averaged stress
 *     sum_{k=0..3} d^{k+1}W(ub)[x^k] / (k+1)!,   ub = (u0+u1)/2, x = (u-u0)/2,
 * i.e. sum_k c_k/(k+1) with c_k the Taylor coefficients of dW(ub + s x).
 * Tangent (d/du, x' = 1/2): 1/2 (H0/2 + H1/3 + H2/4) with H_m the Taylor
 * coefficients of d^2W(ub + s x); value diagonal rho/dt^2 + c/(2 dt).
 */
static void getime_jets(jet *uj, const double *u)
{
    const double *u1 = u + 30, *u0 = u + 60;
    for (int s = 0; s < 30; s++) {
        jet t = {{(u0[s] + u1[s]) / 2.0, (u[s] - u0[s]) / 2.0, 0, 0}};
        uj[s] = t;
    }
}

void or_getime_residual(double r[30], const double u[90], const double coef[6], const double tp[3])
{
    const double rho = tp[0], cd = tp[1], dt = tp[2];
    const double *u1 = u + 30, *u0 = u + 60;
    jet uj[30];
    getime_jets(uj, u);
    ge_out *o = ge_scratch();
    ge_derivs(uj, coef, 3, 1, 0, o);
    for (int s = 0; s < 30; s++) {
        if (s % 10 == 0) {
            r[s] = (cd * (u[s] - u0[s])) / (2. * dt) + (rho * (u[s] + u0[s] - 2 * u1[s])) / (dt * dt);
        } else {
            const double *c = o->r[s].c;
            r[s] = c[0] + c[1] / 2.0 + c[2] / 3.0 + c[3] / 4.0;
        }
    }
}

void or_getime_tangent(double D[900], const double u[90], const double coef[6], const double tp[3])
{
    const double rho = tp[0], cd = tp[1], dt = tp[2];
    jet uj[30];
    getime_jets(uj, u);
    ge_out *o = ge_scratch();
    ge_derivs(uj, coef, 2, 0, 1, o);
    for (int i = 0; i < 900; i++) {
        const double *c = o->D[i].c;
        D[i] = 0.5 * (c[0] / 2.0 + c[1] / 3.0 + c[2] / 4.0);
    }
    for (int a = 0; a < 3; a++) D[30 * (10 * a) + 10 * a] = rho / (dt * dt) + cd / (2. * dt);
}

/* ======================================================================= */
/* Mesh, index sets                                                         */
/* ======================================================================= */
/* Open uniform knots on [0,1] (nurbs_genuniform.c:34-38 for any nelem). */
void or_uniform_knots(int nelem, double *k)
{
    int i = 0;
    double l = 1.0 / nelem;
    while (i < PORDER + 1) k[i++] = 0.0;
    while (i < nelem + PORDER) { k[i] = (i - PORDER) * l; i++; }
    while (i < nelem + 1 + 2 * PORDER) k[i++] = 1.0;
}

static void nbasis_of(const or_problem *pb, int nb[3])
{
    for (int d = 0; d < 3; d++) nb[d] = pb->nelem[d] + PORDER;
}

int64_t or_ndof(const or_problem *pb)
{
    int nb[3];
    nbasis_of(pb, nb);
    return (int64_t)NDOF * nb[0] * nb[1] * nb[2];
}

static int stencil_w(int a, int nb) /* columns within distance p of a */
{
    int lo = a - PORDER < 0 ? 0 : a - PORDER, hi = a + PORDER > nb - 1 ? nb - 1 : a + PORDER;
    return hi - lo + 1;
}

int64_t or_nnz(const or_problem *pb)
{
    int nb[3];
    nbasis_of(pb, nb);
    int64_t s[3] = {0, 0, 0};
    for (int d = 0; d < 3; d++)
        for (int a = 0; a < nb[d]; a++) s[d] += stencil_w(a, nb[d]);
    return (int64_t)NDOF * NDOF * s[0] * s[1] * s[2];
}

/* AIJ pattern of the assembled tangent: row (node a, dof i) couples with all
 * dofs of nodes b, |a_d - b_d| <= p (snesiga_init.c:341-359 counts exactly
 * this); columns ascending in the global numbering 3*(nb1 nb2 b0 + nb2 b1 + b2)+j. */
void or_pattern(const or_problem *pb, int64_t *rowptr, int32_t *col)
{
    int nb[3];
    nbasis_of(pb, nb);
    int64_t pos = 0;
    int64_t row = 0;
    rowptr[0] = 0;
    for (int a0 = 0; a0 < nb[0]; a0++)
        for (int a1 = 0; a1 < nb[1]; a1++)
            for (int a2 = 0; a2 < nb[2]; a2++)
                for (int i = 0; i < NDOF; i++) {
                    for (int b0 = a0 - PORDER; b0 <= a0 + PORDER; b0++) {
                        if (b0 < 0 || b0 >= nb[0]) continue;
                        for (int b1 = a1 - PORDER; b1 <= a1 + PORDER; b1++) {
                            if (b1 < 0 || b1 >= nb[1]) continue;
                            for (int b2 = a2 - PORDER; b2 <= a2 + PORDER; b2++) {
                                if (b2 < 0 || b2 >= nb[2]) continue;
                                for (int j = 0; j < NDOF; j++)
                                    col[pos++] = (int32_t)(NDOF * ((nb[1] * nb[2]) * b0 + nb[2] * b1 + b2) + j);
                            }
                        }
                    }
                    rowptr[++row] = pos;
                }
}

/* Face-local axes of face id f (snesiga_Residual.c:333-364,
 * snesiga_Tangent.c:218-250): local x, y span the face, local z is normal. */
typedef struct {
    int dim[3];        /* global dimension of local x, y, z */
    int64_t stride[3]; /* basis-index stride of local x, y, z */
    int iz0, iz1;      /* normal index of the face layer and the second layer */
} face_map;

static face_map face_of(int f, const int nb[3])
{
    face_map m;
    int64_t sx = (int64_t)nb[1] * nb[2], sy = nb[2], sz = 1;
    switch (f / 2) {
    case 0: m.dim[0] = 1; m.dim[1] = 2; m.dim[2] = 0; m.stride[0] = sy; m.stride[1] = sz; m.stride[2] = sx; break;
    case 1: m.dim[0] = 0; m.dim[1] = 2; m.dim[2] = 1; m.stride[0] = sx; m.stride[1] = sz; m.stride[2] = sy; break;
    default: m.dim[0] = 0; m.dim[1] = 1; m.dim[2] = 2; m.stride[0] = sx; m.stride[1] = sy; m.stride[2] = sz; break;
    }
    int nz = nb[m.dim[2]];
    if (f % 2 == 0) { m.iz0 = 0; m.iz1 = 1; }
    else { m.iz0 = nz - 1; m.iz1 = nz - 2; }
    return m;
}

int or_dirichlet_count(const or_problem *pb)
{
    int nb[3], n = 0;
    nbasis_of(pb, nb);
    for (int ib = 0; ib < pb->nboun; ib++) {
        int f = pb->boun[ib];
        face_map m = face_of(f, nb);
        int nflag = 0;
        for (int t = 0; t < 2 * NDOF; t++) nflag += (pb->bc_type[2 * NDOF * f + t] == 0);
        n += nflag * nb[m.dim[0]] * nb[m.dim[1]];
    }
    return n;
}

/* Rows in the reference order: face (boun order) -> standard-layer dofs ->
 * high-order-layer dofs -> local x -> local y; duplicates kept. */
int or_dirichlet_rows(const or_problem *pb, int32_t *rows)
{
    int nb[3], n = 0;
    nbasis_of(pb, nb);
    for (int ib = 0; ib < pb->nboun; ib++) {
        int f = pb->boun[ib];
        face_map m = face_of(f, nb);
        for (int order = 0; order < 2; order++) {
            int iz = order == 0 ? m.iz0 : m.iz1;
            for (int i = 0; i < NDOF; i++) {
                if (pb->bc_type[2 * NDOF * f + NDOF * order + i] != 0) continue;
                for (int x = 0; x < nb[m.dim[0]]; x++)
                    for (int y = 0; y < nb[m.dim[1]]; y++) {
                        int64_t ib2 = m.stride[0] * x + m.stride[1] * y + m.stride[2] * iz;
                        rows[n++] = (int32_t)(NDOF * ib2 + i);
                    }
            }
        }
    }
    return n;
}

/* Control values of u_idof = a_{idof,J} X_J on face face_id (iga_dirichlet.c):
 * along each face direction the first p values interpolate value and slope
 * at the start, the rest follow c_k = c_{k-1} + (t_{k+p}-t_k)/p * a; the
 * normal coordinate is the face knot (far face: last knot - 1e-15).
 * bc_order 1 adds 0.0/N'_{iz1}(Z) (zero unless that derivative vanishes). */
void or_iga_dirichlet(double *out, const double *par, int face_id, int bc_order, int idof,
                      const double *kv[3], const int nk[3], const int nb[3])
{
    double Z = (face_id % 2 == 0) ? kv[2][0] : kv[2][nk[2] - 1] - 1.e-15;
    int iz1 = (face_id % 2 == 0) ? 1 : nb[2] - 2;
    double pl[3];
    switch (face_id / 2) {
    case 0: pl[0] = par[3 * idof + 1]; pl[1] = par[3 * idof + 2]; pl[2] = par[3 * idof + 0]; break;
    case 1: pl[0] = par[3 * idof + 0]; pl[1] = par[3 * idof + 2]; pl[2] = par[3 * idof + 1]; break;
    default: pl[0] = par[3 * idof + 0]; pl[1] = par[3 * idof + 1]; pl[2] = par[3 * idof + 2]; break;
    }
    double *c1d[2];
    for (int bd = 0; bd < 2; bd++) {
        c1d[bd] = (double *)malloc(sizeof(double) * nb[bd]);
        double W = kv[bd][PORDER];
        double vec[2] = {pl[bd] * W, pl[bd]};
        double mat[4];
        for (int i = 0; i < 2; i++)
            for (int j = 0; j < 2; j++) mat[2 * i + j] = or_evalN(kv[bd], nk[bd], i, j, PORDER, W);
        or_matinv_dot_vec(mat, vec, 2);
        c1d[bd][0] = vec[0];
        c1d[bd][1] = vec[1];
        for (int i = PORDER; i < nb[bd]; i++)
            c1d[bd][i] = c1d[bd][i - 1] + (kv[bd][i + PORDER] - kv[bd][i]) / ((double)PORDER) * pl[bd];
    }
    for (int i = 0; i < nb[0]; i++)
        for (int j = 0; j < nb[1]; j++) out[nb[1] * i + j] = c1d[0][i] + c1d[1][j] + pl[2] * Z;
    if (bc_order == 1) {
        double m = 0.0 / or_evalN(kv[2], nk[2], 1, iz1, PORDER, Z);
        for (int i = 0; i < nb[0] * nb[1]; i++) out[i] += m;
    }
    free(c1d[0]);
    free(c1d[1]);
}

/* ======================================================================= */
/* Volume quadrature point setup shared by residual and tangent             */
/* (snesiga_Residual.c:104-148).                                           */
/* ======================================================================= */
typedef struct {
    int torder, nhist;
    int nb[3];
    int64_t ndof;
    double cq[4], wq[4];
    double coef_glob[6], tp[3];
} asm_ctx;

static int asm_init(const or_problem *pb, asm_ctx *c)
{
    if (or_gauss(pb->nquad, c->cq, c->wq)) return -1;
    nbasis_of(pb, c->nb);
    c->ndof = or_ndof(pb);
    c->torder = pb->physics == 1 ? 2 : 0;
    c->nhist = c->torder + 1;
    or_coef_from_par(pb->par_mat, c->coef_glob);
    c->tp[0] = pb->par_mat[5]; c->tp[1] = pb->par_mat[6]; c->tp[2] = pb->par_mat[7];
    return 0;
}

static void elem_qp(const or_problem *pb, const asm_ctx *c, int ex, int ey, int ez, int qx, int qy,
                    int qz, double N[270], double xq[3], double *w)
{
    int e[3] = {ex, ey, ez}, q[3] = {qx, qy, qz};
    const double *kV[3];
    double h[3];
    for (int d = 0; d < 3; d++) {
        double X0 = pb->knot[d][e[d] + PORDER], X1 = pb->knot[d][e[d] + PORDER + 1];
        double xi = c->cq[q[d]];
        xq[d] = (-xi + 1.0) * X0 + xi * X1;
        kV[d] = pb->knot[d] + e[d];
        h[d] = X1 - X0;
    }
    or_eval_bspline(N, kV, xq);
    *w = c->wq[qx] * c->wq[qy] * c->wq[qz] * h[0] * h[1] * h[2];
}

static void interp_u(const or_problem *pb, const asm_ctx *c, const double *U_hist, int ex, int ey,
                     int ez, const double N[270], const double xq[3], double *u)
{
    int nby = c->nb[1], nbz = c->nb[2];
    for (int i = 0; i < 30 * c->nhist; i++) u[i] = 0.0;
    for (int ax = 0; ax < 3; ax++)
        for (int ay = 0; ay < 3; ay++)
            for (int az = 0; az < 3; az++) {
                int ibpe = 9 * ax + 3 * ay + az;
                int64_t ib = (int64_t)(nby * nbz) * (ex + ax) + nbz * (ey + ay) + (ez + az);
                for (int h = 0; h < c->nhist; h++)
                    for (int i = 0; i < NDOF; i++) {
                        double Uv = U_hist[c->ndof * h + NDOF * ib + i];
                        for (int s = 0; s < NDDIM; s++) u[30 * h + 10 * i + s] += N[10 * ibpe + s] * Uv;
                    }
            }
    const double *pp = pb->par_periodic;
    for (int h = 0; h < c->nhist; h++) {
        for (int i = 0; i < NDOF; i++) u[30 * h + 10 * i] += pp[3 * i] * xq[0] + pp[3 * i + 1] * xq[1] + pp[3 * i + 2] * xq[2];
        for (int i = 0; i < NDOF; i++)
            for (int d = 0; d < 3; d++) u[30 * h + 10 * i + 1 + d] += pp[3 * i + d];
    }
}

static const double *elem_coef(const or_problem *pb, const asm_ctx *c, int64_t e)
{
    return pb->elem_coef ? pb->elem_coef + 6 * e : c->coef_glob;
}

static void pointwise_r(const or_problem *pb, const asm_ctx *c, const double *coef, const double *u, double *r)
{
    if (pb->physics == 1) or_getime_residual(r, u, coef, c->tp);
    else or_ge_residual(r, u, coef);
}

static void pointwise_D(const or_problem *pb, const asm_ctx *c, const double *coef, const double *u, double *D)
{
    if (pb->physics == 1) or_getime_tangent(D, u, coef, c->tp);
    else or_ge_tangent(D, u, coef);
}

static void elem_ijk(const or_problem *pb, int64_t e, int *ex, int *ey, int *ez)
{
    int64_t nyz = (int64_t)pb->nelem[1] * pb->nelem[2];
    *ex = (int)(e / nyz);
    *ey = (int)((e % nyz) / pb->nelem[2]);
    *ez = (int)(e % pb->nelem[2]);
}

int or_residual_elements(const or_problem *pb, const double *U_hist, double *R, int64_t e_begin, int64_t e_end)
{
    asm_ctx c;
    if (asm_init(pb, &c)) return -1;
    int nq = pb->nquad, nby = c.nb[1], nbz = c.nb[2];
    double N[270], xq[3], w, u[90], r[30], val[81];
    for (int64_t e = e_begin; e < e_end; e++) {
        int ex, ey, ez;
        elem_ijk(pb, e, &ex, &ey, &ez);
        const double *coef = elem_coef(pb, &c, e);
        for (int i = 0; i < 81; i++) val[i] = 0.0;
        for (int qx = 0; qx < nq; qx++)
            for (int qy = 0; qy < nq; qy++)
                for (int qz = 0; qz < nq; qz++) {
                    elem_qp(pb, &c, ex, ey, ez, qx, qy, qz, N, xq, &w);
                    interp_u(pb, &c, U_hist, ex, ey, ez, N, xq, u);
                    pointwise_r(pb, &c, coef, u, r);
                    for (int A = 0; A < NBPE; A++)
                        for (int i = 0; i < NDOF; i++) {
                            double t = 0.0;
                            for (int s = 0; s < NDDIM; s++) t += N[10 * A + s] * r[10 * i + s];
                            val[3 * A + i] += w * t;
                        }
                }
        for (int ax = 0; ax < 3; ax++)
            for (int ay = 0; ay < 3; ay++)
                for (int az = 0; az < 3; az++) {
                    int A = 9 * ax + 3 * ay + az;
                    int64_t ib = (int64_t)(nby * nbz) * (ex + ax) + nbz * (ey + ay) + (ez + az);
                    for (int i = 0; i < NDOF; i++) R[NDOF * ib + i] += val[3 * A + i];
                }
    }
    return 0;
}

/* Neumann face integrals (snesiga_Residual.c:188-320): R -= int t N over the
 * face for standard flags, and R -= int t dN/dn over the two normal layers
 * for high-order flags; t from the app hook at each face Gauss point. */
static void neumann_add(const or_problem *pb, const asm_ctx *c, double *R)
{
    if (!pb->neumann) return;
    int nb[3], nq = pb->nquad;
    nbasis_of(pb, nb);
    int nk[3] = {pb->nelem[0] + 5, pb->nelem[1] + 5, pb->nelem[2] + 5};
    for (int ib = 0; ib < pb->nboun; ib++) {
        int f = pb->boun[ib];
        face_map m = face_of(f, nb);
        int d0 = m.dim[0], d1 = m.dim[1], d2 = m.dim[2];
        const double *k0 = pb->knot[d0], *k1 = pb->knot[d1], *k2 = pb->knot[d2];
        int iz0 = (f % 2 == 0) ? 0 : pb->nelem[d2] + PORDER - 1;
        int iz1 = (f % 2 == 0) ? 1 : pb->nelem[d2] + PORDER - 2;
        double xq[3];
        xq[d2] = (f % 2 == 0) ? k2[0] : k2[nk[d2] - 1];
        for (int order = 0; order < 2; order++)
            for (int i = 0; i < NDOF; i++) {
                if (pb->bc_type[2 * NDOF * f + NDOF * order + i] != 1) continue;
                for (int e0 = 0; e0 < pb->nelem[d0]; e0++)
                    for (int e1 = 0; e1 < pb->nelem[d1]; e1++) {
                        double X00 = k0[e0 + PORDER], X10 = k0[e0 + PORDER + 1];
                        double X01 = k1[e1 + PORDER], X11 = k1[e1 + PORDER + 1];
                        double val[18];
                        int nv = order == 0 ? 9 : 18;
                        for (int t = 0; t < nv; t++) val[t] = 0.0;
                        for (int qx = 0; qx < nq; qx++)
                            for (int qy = 0; qy < nq; qy++) {
                                double xi = c->cq[qx];
                                xq[d0] = (-xi + 1.0) * X00 + xi * X10;
                                xi = c->cq[qy];
                                xq[d1] = (-xi + 1.0) * X01 + xi * X11;
                                double Nf[18];
                                int ia = 0;
                                for (int bx = 0; bx < 3; bx++)
                                    for (int by = 0; by < 3; by++) {
                                        double nxy = or_evalN(k0, nk[d0], 0, e0 + bx, PORDER, xq[d0]) *
                                                     or_evalN(k1, nk[d1], 0, e1 + by, PORDER, xq[d1]);
                                        if (order == 0) {
                                            Nf[ia++] = nxy;
                                        } else {
                                            Nf[ia++] = nxy * or_evalN(k2, nk[d2], 1, iz0, PORDER, xq[d2]);
                                            Nf[ia++] = nxy * or_evalN(k2, nk[d2], 1, iz1, PORDER, xq[d2]);
                                        }
                                    }
                                double t;
                                pb->neumann(&t, pb->par_neumann, f, order, i, xq);
                                double w = c->wq[qx] * c->wq[qy] * (X10 - X00) * (X11 - X01);
                                for (int s = 0; s < nv; s++) val[s] -= w * Nf[s] * t;
                            }
                        int ia = 0;
                        for (int bx = 0; bx < 3; bx++)
                            for (int by = 0; by < 3; by++) {
                                int64_t base = m.stride[0] * (e0 + bx) + m.stride[1] * (e1 + by);
                                R[NDOF * (base + m.stride[2] * iz0) + i] += val[ia++];
                                if (order == 1) R[NDOF * (base + m.stride[2] * iz1) + i] += val[ia++];
                            }
                    }
            }
    }
}

/* Dirichlet rows of R: R[row] = U[row] - g, INSERT in reference order
 * (snesiga_Residual.c:331-409), later faces overwrite shared rows. */
static void dirichlet_residual(const or_problem *pb, const double *U, double *R)
{
    int nb[3];
    nbasis_of(pb, nb);
    for (int ib = 0; ib < pb->nboun; ib++) {
        int f = pb->boun[ib];
        face_map m = face_of(f, nb);
        const double *kv[3] = {pb->knot[m.dim[0]], pb->knot[m.dim[1]], pb->knot[m.dim[2]]};
        int nk[3] = {pb->nelem[m.dim[0]] + 5, pb->nelem[m.dim[1]] + 5, pb->nelem[m.dim[2]] + 5};
        int nbf[3] = {nb[m.dim[0]], nb[m.dim[1]], nb[m.dim[2]]};
        double *g = (double *)malloc(sizeof(double) * nbf[0] * nbf[1]);
        for (int order = 0; order < 2; order++) {
            int iz = order == 0 ? m.iz0 : m.iz1;
            for (int i = 0; i < NDOF; i++) {
                if (pb->bc_type[2 * NDOF * f + NDOF * order + i] != 0) continue;
                or_iga_dirichlet(g, pb->par_dirichlet, f, order, i, kv, nk, nbf);
                int ia = 0;
                for (int x = 0; x < nbf[0]; x++)
                    for (int y = 0; y < nbf[1]; y++) {
                        int64_t row = NDOF * (m.stride[0] * x + m.stride[1] * y + m.stride[2] * iz) + i;
                        R[row] = U[row] - g[ia++];
                    }
            }
        }
        free(g);
    }
}

int or_residual(const or_problem *pb, const double *U_hist, double *R)
{
    asm_ctx c;
    if (asm_init(pb, &c)) return -1;
    for (int64_t i = 0; i < c.ndof; i++) R[i] = 0.0;
    int64_t ne = (int64_t)pb->nelem[0] * pb->nelem[1] * pb->nelem[2];
    if (or_residual_elements(pb, U_hist, R, 0, ne)) return -1;
    neumann_add(pb, &c, R);
    dirichlet_residual(pb, U_hist, R);
    return 0;
}

/* AIJ-style insertion: bisection on the sorted column list of the row,
 * restarted from the previous position when the input column increases. */
static inline int64_t aij_find(const int32_t *col, int64_t lo0, int64_t hi0, int32_t c, int64_t *hint)
{
    int64_t lo = (*hint >= lo0 && col[*hint] <= c) ? *hint : lo0, hi = hi0;
    while (hi - lo > 5) {
        int64_t t = (lo + hi) / 2;
        if (col[t] > c) hi = t;
        else lo = t;
    }
    for (int64_t k = lo; k < hi; k++)
        if (col[k] == c) { *hint = k; return k; }
    return -1;
}

/* element matrix val[81 x 81] of element (ex, ey, ez), rows/columns 3 A + i
 * with A = 9 ax + 3 ay + az (snesiga_Tangent.c:104-175) */
static void elem_tangent(const or_problem *pb, const asm_ctx *c, const double *U_hist, int64_t e, int ex, int ey,
                         int ez, double *val)
{
    int nq = pb->nquad;
    double N[270], xq[3], w, u[90], D[900];
    const double *coef = elem_coef(pb, c, e);
    for (int i = 0; i < 81 * 81; i++) val[i] = 0.0;
    for (int qx = 0; qx < nq; qx++)
        for (int qy = 0; qy < nq; qy++)
            for (int qz = 0; qz < nq; qz++) {
                elem_qp(pb, c, ex, ey, ez, qx, qy, qz, N, xq, &w);
                interp_u(pb, c, U_hist, ex, ey, ez, N, xq, u);
                pointwise_D(pb, c, coef, u, D);
                /* snesiga_Tangent.c:149-175: 81 x 81 x 10 x 10 per point */
                for (int A = 0; A < NBPE; A++)
                    for (int i = 0; i < NDOF; i++) {
                        int row = 3 * A + i;
                        for (int B = 0; B < NBPE; B++)
                            for (int j = 0; j < NDOF; j++) {
                                int cl = 3 * B + j;
                                double t = 0.0;
                                for (int s = 0; s < NDDIM; s++)
                                    for (int tt = 0; tt < NDDIM; tt++)
                                        t += N[10 * A + s] * D[30 * (10 * i + s) + (10 * j + tt)] * N[10 * B + tt];
                                val[81 * row + cl] += w * t;
                            }
                    }
            }
}

int or_tangent_elements(const or_problem *pb, const double *U_hist, const int64_t *rowptr,
                        const int32_t *col, double *vals, int64_t e_begin, int64_t e_end)
{
    asm_ctx c;
    if (asm_init(pb, &c)) return -1;
    int nby = c.nb[1], nbz = c.nb[2];
    double *val = (double *)malloc(sizeof(double) * 81 * 81);
    int32_t idx[81];
    for (int64_t e = e_begin; e < e_end; e++) {
        int ex, ey, ez;
        elem_ijk(pb, e, &ex, &ey, &ez);
        elem_tangent(pb, &c, U_hist, e, ex, ey, ez, val);
        for (int ax = 0; ax < 3; ax++)
            for (int ay = 0; ay < 3; ay++)
                for (int az = 0; az < 3; az++) {
                    int A = 9 * ax + 3 * ay + az;
                    int64_t ib = (int64_t)(nby * nbz) * (ex + ax) + nbz * (ey + ay) + (ez + az);
                    for (int i = 0; i < NDOF; i++) idx[3 * A + i] = (int32_t)(NDOF * ib + i);
                }
        for (int r = 0; r < 81; r++) {
            int64_t lo = rowptr[idx[r]], hi = rowptr[idx[r] + 1], hint = lo;
            for (int cc = 0; cc < 81; cc++) {
                int64_t k = aij_find(col, lo, hi, idx[cc], &hint);
                if (k < 0) { free(val); return -2; }
                vals[k] += val[81 * r + cc];
            }
        }
    }
    free(val);
    return 0;
}

int or_tangent(const or_problem *pb, const double *U_hist, const int64_t *rowptr, const int32_t *col,
               double *vals)
{
    int64_t ndof = or_ndof(pb), nnz = or_nnz(pb);
    for (int64_t k = 0; k < nnz; k++) vals[k] = 0.0;
    int64_t ne = (int64_t)pb->nelem[0] * pb->nelem[1] * pb->nelem[2];
    int rc = or_tangent_elements(pb, U_hist, rowptr, col, vals, 0, ne);
    if (rc) return rc;
    /* MatZeroRowsColumns(P, n, rows, 1.0, NULL, NULL) (snesiga_Tangent.c:270) */
    int nd = or_dirichlet_count(pb);
    int32_t *rows = (int32_t *)malloc(sizeof(int32_t) * (nd > 0 ? nd : 1));
    or_dirichlet_rows(pb, rows);
    unsigned char *z = (unsigned char *)calloc((size_t)ndof, 1);
    for (int k = 0; k < nd; k++) z[rows[k]] = 1;
    for (int64_t r = 0; r < ndof; r++)
        for (int64_t k = rowptr[r]; k < rowptr[r + 1]; k++)
            if (z[r] || z[col[k]]) vals[k] = (z[r] && col[k] == r) ? 1.0 : 0.0;
    free(z);
    free(rows);
    return 0;
}

/* Element sums of the tangent over the elements [lo, hi) per direction,
 * without the Dirichlet step: out[r][k] for the row dofs r of the nodes a
 * with a_d in [rlo_d, rhi_d] (rlo = 0 if lo = 0 else lo + 2, rhi = nb - 1 if
 * hi = nelem else hi - 1: the nodes whose whole support lies in the window,
 * i.e. complete rows), and the column dofs k of the nodes b with b_d in
 * [lo_d, hi_d + 1]; both node-major (x slowest), dof fastest. */
int or_tangent_window(const or_problem *pb, const double *U_hist, const int lo[3], const int hi[3], double *out)
{
    asm_ctx c;
    if (asm_init(pb, &c)) return -1;
    int rlo[3], nr[3], nc[3];
    for (int d = 0; d < 3; d++) {
        if (lo[d] < 0 || hi[d] > pb->nelem[d] || hi[d] <= lo[d]) return -1;
        rlo[d] = lo[d] == 0 ? 0 : lo[d] + 2;
        int rhi = hi[d] == pb->nelem[d] ? pb->nelem[d] + 1 : hi[d] - 1;
        nr[d] = rhi - rlo[d] + 1;
        nc[d] = hi[d] - lo[d] + 2;
        if (nr[d] < 1) return -1;
    }
    const int64_t ncol = 3LL * nc[0] * nc[1] * nc[2];
    memset(out, 0, sizeof(double) * (size_t)(3LL * nr[0] * nr[1] * nr[2] * ncol));
    double *val = (double *)malloc(sizeof(double) * 81 * 81);
    for (int ex = lo[0]; ex < hi[0]; ex++)
        for (int ey = lo[1]; ey < hi[1]; ey++)
            for (int ez = lo[2]; ez < hi[2]; ez++) {
                int64_t e = ((int64_t)ex * pb->nelem[1] + ey) * pb->nelem[2] + ez;
                elem_tangent(pb, &c, U_hist, e, ex, ey, ez, val);
                for (int A = 0; A < NBPE; A++) {
                    int a0 = ex + A / 9 - rlo[0], a1 = ey + (A / 3) % 3 - rlo[1], a2 = ez + A % 3 - rlo[2];
                    if (a0 < 0 || a0 >= nr[0] || a1 < 0 || a1 >= nr[1] || a2 < 0 || a2 >= nr[2]) continue;
                    int64_t rbase = 3LL * (((int64_t)a0 * nr[1] + a1) * nr[2] + a2);
                    for (int B = 0; B < NBPE; B++) {
                        int b0 = ex + B / 9 - lo[0], b1 = ey + (B / 3) % 3 - lo[1], b2 = ez + B % 3 - lo[2];
                        int64_t cbase = 3LL * (((int64_t)b0 * nc[1] + b1) * nc[2] + b2);
                        for (int i = 0; i < NDOF; i++)
                            for (int j = 0; j < NDOF; j++)
                                out[(rbase + i) * ncol + cbase + j] += val[81 * (3 * A + i) + 3 * B + j];
                    }
                }
            }
    free(val);
    return 0;
}

/* ---- threaded residual (large-mesh checks) ----------------------------- */
#include <pthread.h>

typedef struct {
    const or_problem *pb;
    const double *U_hist;
    double *R;
    int64_t e0, e1;
    int rc;
} res_job;

static void *res_worker(void *arg)
{
    res_job *j = (res_job *)arg;
    j->rc = or_residual_elements(j->pb, j->U_hist, j->R, j->e0, j->e1);
    return NULL;
}

int or_residual_mt(const or_problem *pb, const double *U_hist, double *R, int nthreads)
{
    asm_ctx c;
    if (asm_init(pb, &c)) return -1;
    int64_t ne = (int64_t)pb->nelem[0] * pb->nelem[1] * pb->nelem[2];
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    if (nthreads > ne) nthreads = (int)ne;
    res_job *jobs = (res_job *)calloc((size_t)nthreads, sizeof(res_job));
    pthread_t *th = (pthread_t *)calloc((size_t)nthreads, sizeof(pthread_t));
    int rc = 0;
    for (int t = 0; t < nthreads; t++) {
        jobs[t].pb = pb;
        jobs[t].U_hist = U_hist;
        jobs[t].R = t == 0 ? R : (double *)calloc((size_t)c.ndof, sizeof(double));
        jobs[t].e0 = ne * t / nthreads;
        jobs[t].e1 = ne * (t + 1) / nthreads;
        if (!jobs[t].R) rc = -2;
    }
    for (int64_t i = 0; i < c.ndof; i++) R[i] = 0.0;
    for (int t = 0; t < nthreads && !rc; t++)
        if (pthread_create(&th[t], NULL, res_worker, &jobs[t])) rc = -3;
    for (int t = 0; t < nthreads; t++)
        if (th[t]) pthread_join(th[t], NULL);
    for (int t = 0; t < nthreads; t++) {
        if (jobs[t].rc) rc = jobs[t].rc;
        if (t > 0 && jobs[t].R) {
            for (int64_t i = 0; i < c.ndof; i++) R[i] += jobs[t].R[i];
            free(jobs[t].R);
        }
    }
    free(jobs);
    free(th);
    if (rc) return rc;
    neumann_add(pb, &c, R);
    dirichlet_residual(pb, U_hist, R);
    return 0;
}

double or_energy(const or_problem *pb, const double *U_hist)
{
    asm_ctx c;
    if (asm_init(pb, &c)) return NAN;
    int nq = pb->nquad;
    double N[270], xq[3], w, u[90], Pi = 0.0;
    int64_t ne = (int64_t)pb->nelem[0] * pb->nelem[1] * pb->nelem[2];
    for (int64_t e = 0; e < ne; e++) {
        int ex, ey, ez;
        elem_ijk(pb, e, &ex, &ey, &ez);
        const double *coef = elem_coef(pb, &c, e);
        for (int qx = 0; qx < nq; qx++)
            for (int qy = 0; qy < nq; qy++)
                for (int qz = 0; qz < nq; qz++) {
                    elem_qp(pb, &c, ex, ey, ez, qx, qy, qz, N, xq, &w);
                    interp_u(pb, &c, U_hist, ex, ey, ez, N, xq, u);
                    if (pb->physics == 1) {
                        /* phys/getime/density.c:6-53: W at (u0 + u1) / 2 plus kinetic energy */
                        double hh[2];
                        or_getime_density(hh, u, coef, c.tp);
                        Pi += w * hh[0];
                    } else {
                        Pi += w * or_ge_energy(u, coef);
                    }
                }
    }
    return Pi;
}

/* iga_intdensity (src/iga_intdensity.c:10-293) on one rank: the volume
 * integral of every density output, then for each face in boun and each dof
 * with a Neumann flag (standard: face values; high order: the normal
 * derivative through the two layers) the face integral of densityb
 * subtracted.  densityb: ge (phys/ge/density.c:68-84) hh[0] = u t, the rest
 * 0; getime (phys/getime/density.c:57-62) hh[0] = (u1 + u0) / 2 t, hh[1] = 1.
 * Returns ndensity (13 ge, 2 getime). */
int or_intdensity(const or_problem *pb, const double *U_hist, double *HH)
{
    asm_ctx c;
    if (asm_init(pb, &c)) return -1;
    const int nd = pb->physics == 1 ? 2 : 13;
    int nq = pb->nquad;
    double N[270], xq[3], w, u[90], hh[13];
    for (int k = 0; k < nd; k++) HH[k] = 0.0;
    int64_t ne = (int64_t)pb->nelem[0] * pb->nelem[1] * pb->nelem[2];
    for (int64_t e = 0; e < ne; e++) {
        int ex, ey, ez;
        elem_ijk(pb, e, &ex, &ey, &ez);
        const double *coef = elem_coef(pb, &c, e);
        for (int qx = 0; qx < nq; qx++)
            for (int qy = 0; qy < nq; qy++)
                for (int qz = 0; qz < nq; qz++) {
                    elem_qp(pb, &c, ex, ey, ez, qx, qy, qz, N, xq, &w);
                    interp_u(pb, &c, U_hist, ex, ey, ez, N, xq, u);
                    if (pb->physics == 1) or_getime_density(hh, u, coef, c.tp);
                    else or_ge_density(hh, u, coef);
                    for (int k = 0; k < nd; k++) HH[k] += w * hh[k];
                }
    }
    int nb[3];
    nbasis_of(pb, nb);
    int nk[3] = {pb->nelem[0] + 5, pb->nelem[1] + 5, pb->nelem[2] + 5};
    for (int ib = 0; ib < pb->nboun; ib++) {
        int f = pb->boun[ib];
        face_map m = face_of(f, nb);
        int d0 = m.dim[0], d1 = m.dim[1], d2 = m.dim[2];
        const double *k0 = pb->knot[d0], *k1 = pb->knot[d1], *k2 = pb->knot[d2];
        xq[d2] = (f % 2 == 0) ? k2[0] : k2[nk[d2] - 1];
        for (int order = 0; order < 2; order++)
            for (int i = 0; i < NDOF; i++) {
                if (pb->bc_type[2 * NDOF * f + NDOF * order + i] != 1) continue;
                for (int e0 = 0; e0 < pb->nelem[d0]; e0++)
                    for (int e1 = 0; e1 < pb->nelem[d1]; e1++) {
                        double X00 = k0[e0 + PORDER], X10 = k0[e0 + PORDER + 1];
                        double X01 = k1[e1 + PORDER], X11 = k1[e1 + PORDER + 1];
                        for (int qx = 0; qx < nq; qx++)
                            for (int qy = 0; qy < nq; qy++) {
                                double xi = c.cq[qx];
                                xq[d0] = (-xi + 1.0) * X00 + xi * X10;
                                xi = c.cq[qy];
                                xq[d1] = (-xi + 1.0) * X01 + xi * X11;
                                double ub[3] = {0.0, 0.0, 0.0};
                                for (int bx = 0; bx < 3; bx++)
                                    for (int by = 0; by < 3; by++) {
                                        double nxy = or_evalN(k0, nk[d0], 0, e0 + bx, PORDER, xq[d0]) *
                                                     or_evalN(k1, nk[d1], 0, e1 + by, PORDER, xq[d1]);
                                        int64_t base = m.stride[0] * (e0 + bx) + m.stride[1] * (e1 + by);
                                        for (int h = 0; h < c.nhist; h++) {
                                            const double *Uh = U_hist + c.ndof * h;
                                            if (order == 0) {
                                                ub[h] += nxy * Uh[NDOF * (base + m.stride[2] * m.iz0) + i];
                                            } else {
                                                ub[h] += nxy * or_evalN(k2, nk[d2], 1, m.iz0, PORDER, xq[d2]) *
                                                         Uh[NDOF * (base + m.stride[2] * m.iz0) + i];
                                                ub[h] += nxy * or_evalN(k2, nk[d2], 1, m.iz1, PORDER, xq[d2]) *
                                                         Uh[NDOF * (base + m.stride[2] * m.iz1) + i];
                                            }
                                        }
                                    }
                                double t = 0.0;
                                if (pb->neumann) pb->neumann(&t, pb->par_neumann, f, order, i, xq);
                                double wf = c.wq[qx] * c.wq[qy] * (X10 - X00) * (X11 - X01);
                                if (pb->physics == 1) {
                                    HH[0] -= wf * (ub[1] + ub[2]) / 2.0 * t;
                                    HH[1] -= wf * 1.0;
                                } else {
                                    HH[0] -= wf * ub[0] * t;
                                }
                            }
                    }
            }
    }
    return nd;
}
