// Pointwise physics of C1 quadratic B-spline gradient elasticity, device side.
//
// Slots of u (eval_Bspline.c:82-91): u[10*i + s], s = 0 value, 1..3 first
// derivatives d/dX_J, 4..9 second derivatives (00,01,02,11,12,22).
//
// Energy (phys/ge/density.c:49), coef = (A36, A37, A38, e0, e345, le^2):
//   W = A36 e3 (e3^2 - 3 e2^2) + A37 q + A38 q^2 + e0 e1^2
//       + e345 (E01^2 + E02^2 + E12^2) + le^2 sum_K |dev g_.K|^2
// with F = I + grad u, E = (F^T F - I)/2, e1 = tr E/sqrt3,
// e2 = (E00-E11)/sqrt2, e3 = (E00+E11-2E22)/sqrt6, q = e2^2 + e3^2 and the
// normal strain gradients g_IK = d_K E_II = sum_a F_aI u_a,IK.
//
// Residual r = dW/du (phys/ge/residual.c) in closed form:
//   r[F_aJ] = (F S)_aJ + sum_K tau_JK u_a,JK
//   r[H_a,h] = sum_{(I,K): {I,K} = h} tau_IK F_aI
// with S = dPsi/dE and tau_IK = 2 le^2 (g_IK - mean_I g_IK).
//
// Tangent (phys/ge/tangent.c) is never formed.  The element kernels need
// T_ij[s, B] = sum_t D[(i,s),(j,t)] N_{B,t}; for a basis function B with
// slot values n the product D_ij n is evaluated from per-point "compact"
// data (F, u_,IK, S, the 3x3 normal block of d2Psi/dE2, tau):
//   material    (F Gamma)_iJ,  Gamma = d2Psi/dE2 : sym(F_j^T grad n)
//   geometric   delta_ij (S grad n)_J
//   gradient    sum_K u_i,JK mu_JK  and  sum_{{I,K}=h} F_iI mu_IK,
//               mu = 2 le^2 dev_I(u_j,IK n_,I + F_jI n_,IK)
//   tau terms   delta_ij (sum_K tau_JK n_,JK ; sum_{{I,K}=h} tau_IK n_,I)
//
// phys/getime (torder 2) uses the same code on truncated Taylor series
// ("jets") along u(s) = (u0+u1)/2 + s (u-u0)/2: the residual's gradient
// slots are sum_k c_k/(k+1) of the series of dW (order 3), the tangent is
// (1/2)(c0/2 + c1/3 + c2/4) of the series of d2W n (order 2), and the value
// slots carry inertia and damping (phys/getime/residual.c:4127,
// phys/getime/tangent.c:4127).
#pragma once
#include <hip/hip_runtime.h>

namespace igap4h {

#define IGAP4H_HD __host__ __device__ __forceinline__

// ---------------------------------------------------------------- jets
// Truncated Taylor series in one variable, coefficients of type R (double;
// float for the FP32 tangent of phys/getime).
template <int N, class R = double>
struct Jet {
    R c[N];
};
template <class X>
struct NoDeduce {
    using t = X;
};

template <int N, class R>
IGAP4H_HD Jet<N, R> operator+(Jet<N, R> a, const Jet<N, R> &b)
{
#pragma unroll
    for (int k = 0; k < N; k++) a.c[k] += b.c[k];
    return a;
}
template <int N, class R>
IGAP4H_HD Jet<N, R> operator-(Jet<N, R> a, const Jet<N, R> &b)
{
#pragma unroll
    for (int k = 0; k < N; k++) a.c[k] -= b.c[k];
    return a;
}
template <int N, class R>
IGAP4H_HD Jet<N, R> operator*(Jet<N, R> a, typename NoDeduce<R>::t s)
{
#pragma unroll
    for (int k = 0; k < N; k++) a.c[k] *= s;
    return a;
}
template <int N, class R>
IGAP4H_HD Jet<N, R> operator*(typename NoDeduce<R>::t s, Jet<N, R> a)
{
    return a * s;
}
template <int N, class R>
IGAP4H_HD Jet<N, R> operator*(const Jet<N, R> &a, const Jet<N, R> &b)
{
    Jet<N, R> r;
#pragma unroll
    for (int k = 0; k < N; k++) {
        R t = R(0);
#pragma unroll
        for (int i = 0; i <= k; i++) t += a.c[i] * b.c[k - i];
        r.c[k] = t;
    }
    return r;
}
template <int N, class R>
IGAP4H_HD Jet<N, R> &operator+=(Jet<N, R> &a, const Jet<N, R> &b)
{
    a = a + b;
    return a;
}

template <class T>
struct Num;
template <>
struct Num<double> {
    static IGAP4H_HD double cst(double x) { return x; }
    static IGAP4H_HD double lin(double x0, double) { return x0; }
};
template <int N, class R>
struct Num<Jet<N, R>> {
    static IGAP4H_HD Jet<N, R> cst(double x)
    {
        Jet<N, R> r;
#pragma unroll
        for (int k = 0; k < N; k++) r.c[k] = R(0);
        r.c[0] = (R)x;
        return r;
    }
    static IGAP4H_HD Jet<N, R> lin(double x0, double x1)
    {
        Jet<N, R> r = cst(x0);
        if (N > 1) r.c[1] = (R)x1;
        return r;
    }
};

// (1/2)(c0/2 + c1/3 + c2/4): averaged tangent of a degree-2 series
IGAP4H_HD double avg_tangent(double x) { return x; }
IGAP4H_HD float avg_tangent(float x) { return x; }
template <int N, class R>
IGAP4H_HD R avg_tangent(const Jet<N, R> &x)
{
    R s = x.c[0] / R(2);
    if (N > 1) s += x.c[1] / R(3);
    if (N > 2) s += x.c[2] / R(4);
    return R(0.5) * s;
}
// sum_k c_k/(k+1): averaged stress of a degree-3 series
IGAP4H_HD double avg_residual(double x) { return x; }
template <int N>
IGAP4H_HD double avg_residual(const Jet<N> &x)
{
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < N; k++) s += x.c[k] / (double)(k + 1);
    return s;
}

// symmetric second-derivative slot index of (I,K) in 0..5
IGAP4H_HD int hidx(int I, int K)
{
    int a = I < K ? I : K, b = I < K ? K : I;
    return a == 0 ? b : (a == 1 ? 2 + b : 5);
}

// constants of the (e1,e2,e3) basis: Q[k][p]
#define IGAP4H_S3 0.57735026918962573
#define IGAP4H_S2 0.70710678118654752
#define IGAP4H_S6 0.40824829046386302

struct Coef {
    double A36, A37, A38, e0, e345, lsq;
};

// ---------------------------------------------------------------- compact data
// Per quadrature point, everything the tangent needs.  Layout in LDS is
// this struct verbatim (T = double or Jet<3>).
template <class T>
struct Compact {
    T F[3][3];   // F_aJ
    T H[3][6];   // u_a,h
    T S[6];      // S00 S11 S22 S01 S02 S12
    T P[6];      // d2Psi/dE_II dE_JJ: P00 P01 P02 P11 P12 P22
    T tau[3][3]; // tau_IK
};

// From the slot values (as T) compute strain quantities. Shared by the
// residual and the tangent.
template <class T>
struct Strain {
    T F[3][3], H[3][6];
    T eps[6];    // E00 E11 E22 E01 E02 E12
    T h1, h2, h3, q;
    T tau[3][3];
};

template <class T>
IGAP4H_HD void strain_of(const T *u, const Coef &c, Strain<T> &st)
{
    #pragma unroll
    for (int a = 0; a < 3; a++) {
        #pragma unroll
        for (int J = 0; J < 3; J++) {
            st.F[a][J] = u[10 * a + 1 + J];
            if (a == J) st.F[a][J] = st.F[a][J] + Num<T>::cst(1.0);
        }
        #pragma unroll
        for (int h = 0; h < 6; h++) st.H[a][h] = u[10 * a + 4 + h];
    }
    const int PI_[6] = {0, 1, 2, 0, 0, 1}, PJ_[6] = {0, 1, 2, 1, 2, 2};
    #pragma unroll
    for (int p = 0; p < 6; p++) {
        T t = st.F[0][PI_[p]] * st.F[0][PJ_[p]] + st.F[1][PI_[p]] * st.F[1][PJ_[p]] + st.F[2][PI_[p]] * st.F[2][PJ_[p]];
        if (p < 3) t = t - Num<T>::cst(1.0);
        st.eps[p] = t * 0.5;
    }
    st.h1 = (st.eps[0] + st.eps[1] + st.eps[2]) * IGAP4H_S3;
    st.h2 = (st.eps[0] - st.eps[1]) * IGAP4H_S2;
    st.h3 = (st.eps[0] + st.eps[1] - st.eps[2] * 2.0) * IGAP4H_S6;
    st.q = st.h2 * st.h2 + st.h3 * st.h3;
    #pragma unroll
    for (int K = 0; K < 3; K++) {
        T g[3];
        #pragma unroll
        for (int I = 0; I < 3; I++)
            g[I] = st.F[0][I] * st.H[0][hidx(I, K)] + st.F[1][I] * st.H[1][hidx(I, K)] + st.F[2][I] * st.H[2][hidx(I, K)];
        T m = (g[0] + g[1] + g[2]) * (1.0 / 3.0);
        #pragma unroll
        for (int I = 0; I < 3; I++) st.tau[I][K] = (g[I] - m) * (2.0 * c.lsq);
    }
}

// dPsi/dE as the symmetric S (S00 S11 S22 S01 S02 S12)
template <class T>
IGAP4H_HD void stress_of(const Strain<T> &st, const Coef &c, T S[6])
{
    T P1 = st.h1 * (2.0 * c.e0);
    T c2 = st.h3 * (-6.0 * c.A36) + Num<T>::cst(2.0 * c.A37) + st.q * (4.0 * c.A38);
    T P2 = st.h2 * c2;
    T P3 = (st.h3 * st.h3 - st.h2 * st.h2) * (3.0 * c.A36) + st.h3 * (Num<T>::cst(2.0 * c.A37) + st.q * (4.0 * c.A38));
    S[0] = P1 * IGAP4H_S3 + P2 * IGAP4H_S2 + P3 * IGAP4H_S6;
    S[1] = P1 * IGAP4H_S3 - P2 * IGAP4H_S2 + P3 * IGAP4H_S6;
    S[2] = P1 * IGAP4H_S3 - P3 * (2.0 * IGAP4H_S6);
    // Psi = e345 E01^2 + ...: dPsi/dE01 = 2 e345 E01, S01 = half of it
    S[3] = st.eps[3] * c.e345;
    S[4] = st.eps[4] * c.e345;
    S[5] = st.eps[5] * c.e345;
}

// residual in the 27 gradient slots (value slots left to the caller)
template <class T>
IGAP4H_HD void residual_grad(const T *u, const Coef &c, T r[30])
{
    Strain<T> st;
    strain_of(u, c, st);
    T S[6];
    stress_of(st, c, S);
    // full symmetric S
    T Sm[3][3] = {{S[0], S[3], S[4]}, {S[3], S[1], S[5]}, {S[4], S[5], S[2]}};
    #pragma unroll
    for (int a = 0; a < 3; a++) {
        r[10 * a] = Num<T>::cst(0.0);
        #pragma unroll
        for (int J = 0; J < 3; J++) {
            T t = st.F[a][0] * Sm[0][J] + st.F[a][1] * Sm[1][J] + st.F[a][2] * Sm[2][J];
            #pragma unroll
            for (int K = 0; K < 3; K++) t = t + st.tau[J][K] * st.H[a][hidx(J, K)];
            r[10 * a + 1 + J] = t;
        }
        // diagonal slots: {I,K} = {I,I}; mixed slots get both orders
        r[10 * a + 4] = st.tau[0][0] * st.F[a][0];
        r[10 * a + 5] = st.tau[0][1] * st.F[a][0] + st.tau[1][0] * st.F[a][1];
        r[10 * a + 6] = st.tau[0][2] * st.F[a][0] + st.tau[2][0] * st.F[a][2];
        r[10 * a + 7] = st.tau[1][1] * st.F[a][1];
        r[10 * a + 8] = st.tau[1][2] * st.F[a][1] + st.tau[2][1] * st.F[a][2];
        r[10 * a + 9] = st.tau[2][2] * st.F[a][2];
    }
}

// energy density (for tests / energy functional)
template <class T>
IGAP4H_HD T energy_of(const T *u, const Coef &c)
{
    Strain<T> st;
    strain_of(u, c, st);
    T W = st.h3 * c.A36 * (st.h3 * st.h3 - st.h2 * st.h2 * 3.0) + st.q * c.A37 + st.q * st.q * c.A38 +
          st.h1 * st.h1 * c.e0 + (st.eps[3] * st.eps[3] + st.eps[4] * st.eps[4] + st.eps[5] * st.eps[5]) * c.e345;
    T g2 = Num<T>::cst(0.0);
    for (int I = 0; I < 3; I++)
        for (int K = 0; K < 3; K++) g2 = g2 + st.tau[I][K] * st.tau[I][K];
    // tau = 2 le^2 dev g  ->  le^2 |dev g|^2 = |tau|^2 / (4 le^2)
    if (c.lsq != 0.0) W = W + g2 * (0.25 / c.lsq);
    return W;
}

// The 13 density outputs of phys/ge (phys/ge/density.c:6-66) at one point:
// hh[0] = W; hh[1..6] = 2 dW/dE00, 2 dW/dE01, 2 dW/dE02, 2 dW/dE11,
// 2 dW/dE12, 2 dW/dE22 with W written on the six independent strains (the
// off-diagonal ones once), i.e. 2 S00, 4 S01, 4 S02, 2 S11, 4 S12, 2 S22;
// hh[7..9] the variant indicators of (e2, e3) (sectors around the wells);
// hh[10..12] the same sectors where W < -1/2.
IGAP4H_HD void ge_density_outputs(const double *u, const Coef &c, double hh[13])
{
    Strain<double> st;
    strain_of<double>(u, c, st);
    double S[6];
    stress_of<double>(st, c, S);
    const double W = energy_of<double>(u, c);
    hh[0] = W;
    hh[1] = 2.0 * S[0];
    hh[2] = 4.0 * S[3];
    hh[3] = 4.0 * S[4];
    hh[4] = 2.0 * S[1];
    hh[5] = 4.0 * S[5];
    hh[6] = 2.0 * S[2];
    const double e2 = st.h2, e3 = st.h3;
    hh[7] = (e2 > 0 && e3 > -0.5 * e2) ? 1.0 : 0.0;
    hh[8] = (e2 < 0 && e3 > 0.5 * e2) ? 1.0 : 0.0;
    hh[9] = (e3 < 0.5 * e2 && e3 < -0.5 * e2) ? 1.0 : 0.0;
    hh[10] = (W < -0.5 && e2 > 0 && e3 > 0) ? 1.0 : 0.0;
    hh[11] = (W < -0.5 && e2 < 0 && e3 > 0) ? 1.0 : 0.0;
    hh[12] = (W < -0.5 && e3 < 0) ? 1.0 : 0.0;
}

// The 2 density outputs of phys/getime (phys/getime/density.c:6-53) from the
// slots of t_n (u1) and t_{n-1} (u0): hh[0] = W((u0 + u1) / 2) + rho
// |u1 - u0|^2 / (2 dt^2) (value slots), hh[1] = 0.
IGAP4H_HD void getime_density_outputs(const double *u1, const double *u0, const Coef &c, double rho, double dt,
                                             double hh[2])
{
    double ub[30];
    for (int s = 0; s < 30; s++) ub[s] = 0.5 * (u0[s] + u1[s]);
    double kin = 0.0;
    for (int i = 0; i < 3; i++) {
        const double d = u1[10 * i] - u0[10 * i];
        kin += rho * d * d / (2.0 * dt * dt);
    }
    hh[0] = energy_of<double>(ub, c) + kin;
    hh[1] = 0.0;
}

// The six energies W is linear in: W = sum_k coef_k phi_k with coef =
// (A36, A37, A38, e0, e345, le^2) and
// phi = (h3 (h3^2 - 3 h2^2), q, q^2, h1^2, E01^2 + E02^2 + E12^2, |dev g|^2).
template <class T>
IGAP4H_HD void energy_basis(const T *u, T phi[6])
{
    Coef one{0.0, 0.0, 0.0, 0.0, 0.0, 1.0};  // lsq = 1: tau = 2 dev g
    Strain<T> st;
    strain_of(u, one, st);
    phi[0] = st.h3 * (st.h3 * st.h3 - st.h2 * st.h2 * 3.0);
    phi[1] = st.q;
    phi[2] = st.q * st.q;
    phi[3] = st.h1 * st.h1;
    phi[4] = st.eps[3] * st.eps[3] + st.eps[4] * st.eps[4] + st.eps[5] * st.eps[5];
    T g2 = Num<T>::cst(0.0);
    for (int I = 0; I < 3; I++)
        for (int K = 0; K < 3; K++) g2 = g2 + st.tau[I][K] * st.tau[I][K];
    phi[5] = g2 * 0.25;
}

template <class T>
IGAP4H_HD void compact_of(const T *u, const Coef &c, Compact<T> &cp)
{
    Strain<T> st;
    strain_of(u, c, st);
    stress_of(st, c, cp.S);
    for (int a = 0; a < 3; a++) {
        for (int J = 0; J < 3; J++) cp.F[a][J] = st.F[a][J];
        for (int h = 0; h < 6; h++) cp.H[a][h] = st.H[a][h];
        for (int K = 0; K < 3; K++) cp.tau[a][K] = st.tau[a][K];
    }
    // second derivatives in (e1,e2,e3), mapped onto (E00,E11,E22): Q^T M Q
    T M22 = st.h3 * (-6.0 * c.A36) + Num<T>::cst(2.0 * c.A37) + (st.h2 * st.h2 * 3.0 + st.h3 * st.h3) * (4.0 * c.A38);
    T M23 = st.h2 * (Num<T>::cst(-6.0 * c.A36) + st.h3 * (8.0 * c.A38));
    T M33 = st.h3 * (6.0 * c.A36) + Num<T>::cst(2.0 * c.A37) + (st.h2 * st.h2 + st.h3 * st.h3 * 3.0) * (4.0 * c.A38);
    const double Q[3][3] = {{IGAP4H_S3, IGAP4H_S3, IGAP4H_S3}, {IGAP4H_S2, -IGAP4H_S2, 0.0}, {IGAP4H_S6, IGAP4H_S6, -2.0 * IGAP4H_S6}};
    const int pp[6] = {0, 0, 0, 1, 1, 2}, qq[6] = {0, 1, 2, 1, 2, 2};
    for (int m = 0; m < 6; m++) {
        int p = pp[m], r = qq[m];
        T t = Num<T>::cst(2.0 * c.e0 * Q[0][p] * Q[0][r]);
        t = t + M22 * (Q[1][p] * Q[1][r]) + M23 * (Q[1][p] * Q[2][r] + Q[2][p] * Q[1][r]) + M33 * (Q[2][p] * Q[2][r]);
        cp.P[m] = t;
    }
}

// T_ij[s, B] for i in [0, J], one basis function with slot values n[10].
// Writes out[i][s] for s = 0..9 (value slot only when mass != 0).  J is a
// compile-time constant so that out[][] stays in registers.  R is the
// arithmetic type of the product: double, or float for the FP32 tangent
// (compact data rounded once from FP64; every constant is cast to R so that
// no operation widens to double).
template <class T, int JB, int IB0 = 0, int IB1 = JB, class R = double>
IGAP4H_HD void tangent_apply(const Compact<T> &cp, const Coef &c, const R n[10], R mass, R w, R out[JB + 1][10])
{
    // rows i of [IB0, IB1] only (a caller may split one j over several lanes)
    constexpr int j = JB;
    const R half = R(0.5), third = R(1.0 / 3.0), e345 = R(c.e345), l2 = R(2.0 * c.lsq);
    // beta = sym(F_j^T grad n) in (00,11,22,01,02,12)
    T b[6];
    b[0] = cp.F[j][0] * n[1];
    b[1] = cp.F[j][1] * n[2];
    b[2] = cp.F[j][2] * n[3];
    b[3] = (cp.F[j][1] * n[1] + cp.F[j][0] * n[2]) * half;
    b[4] = (cp.F[j][2] * n[1] + cp.F[j][0] * n[3]) * half;
    b[5] = (cp.F[j][2] * n[2] + cp.F[j][1] * n[3]) * half;
    // Gamma (symmetric 3x3): normal block P b, shear e345 b
    T G[3][3];
    G[0][0] = cp.P[0] * b[0] + cp.P[1] * b[1] + cp.P[2] * b[2];
    G[1][1] = cp.P[1] * b[0] + cp.P[3] * b[1] + cp.P[4] * b[2];
    G[2][2] = cp.P[2] * b[0] + cp.P[4] * b[1] + cp.P[5] * b[2];
    G[0][1] = G[1][0] = b[3] * e345;
    G[0][2] = G[2][0] = b[4] * e345;
    G[1][2] = G[2][1] = b[5] * e345;
    // mu_IK = 2 le^2 dev_I (u_j,IK n_,I + F_jI n_,IK)
    T mu[3][3];
    for (int K = 0; K < 3; K++) {
        T nu[3];
        for (int I = 0; I < 3; I++) nu[I] = cp.H[j][hidx(I, K)] * n[1 + I] + cp.F[j][I] * n[4 + hidx(I, K)];
        T m = (nu[0] + nu[1] + nu[2]) * third;
        for (int I = 0; I < 3; I++) mu[I][K] = (nu[I] - m) * l2;
    }
    const T Sm[3][3] = {{cp.S[0], cp.S[3], cp.S[4]}, {cp.S[3], cp.S[1], cp.S[5]}, {cp.S[4], cp.S[5], cp.S[2]}};
#pragma unroll
    for (int i = IB0; i <= IB1; i++) {
        T tf[3], th[6];
        for (int J = 0; J < 3; J++) {
            T t = cp.F[i][0] * G[0][J] + cp.F[i][1] * G[1][J] + cp.F[i][2] * G[2][J];
            for (int K = 0; K < 3; K++) t = t + cp.H[i][hidx(J, K)] * mu[J][K];
            if (i == j) {
                t = t + Sm[J][0] * n[1] + Sm[J][1] * n[2] + Sm[J][2] * n[3];
                for (int K = 0; K < 3; K++) t = t + cp.tau[J][K] * n[4 + hidx(J, K)];
            }
            tf[J] = t;
        }
        th[0] = cp.F[i][0] * mu[0][0];
        th[1] = cp.F[i][0] * mu[0][1] + cp.F[i][1] * mu[1][0];
        th[2] = cp.F[i][0] * mu[0][2] + cp.F[i][2] * mu[2][0];
        th[3] = cp.F[i][1] * mu[1][1];
        th[4] = cp.F[i][1] * mu[1][2] + cp.F[i][2] * mu[2][1];
        th[5] = cp.F[i][2] * mu[2][2];
        if (i == j) {
            th[0] = th[0] + cp.tau[0][0] * n[1];
            th[1] = th[1] + cp.tau[0][1] * n[1] + cp.tau[1][0] * n[2];
            th[2] = th[2] + cp.tau[0][2] * n[1] + cp.tau[2][0] * n[3];
            th[3] = th[3] + cp.tau[1][1] * n[2];
            th[4] = th[4] + cp.tau[1][2] * n[2] + cp.tau[2][1] * n[3];
            th[5] = th[5] + cp.tau[2][2] * n[3];
        }
        out[i][0] = (i == j) ? w * mass * n[0] : R(0);
        for (int J = 0; J < 3; J++) out[i][1 + J] = w * avg_tangent(tf[J]);
        for (int h = 0; h < 6; h++) out[i][4 + h] = w * avg_tangent(th[h]);
    }
}

// ---------------------------------------------------------------- basis
// Quadratic B-spline values and first/second derivatives of the three
// functions active on the span [k2, k3) of the 6-knot window k, with the
// +1e-15 guard on each span denominator (eval_Bspline.c:69-73).
IGAP4H_HD void bspline1d(const double *k, double x, double v[3], double d1[3], double d2[3])
{
    const double tol = 1.e-15;
    double n0[5] = {0, 0, 1, 0, 0};
    double n1[4], dn1[4];
#pragma unroll
    for (int i = 0; i < 4; i++) {
        n1[i] = (x - k[i]) / (k[i + 1] - k[i] + tol) * n0[i] + (k[i + 2] - x) / (k[i + 2] - k[i + 1] + tol) * n0[i + 1];
        dn1[i] = 1.0 / (k[i + 1] - k[i] + tol) * n0[i] - 1.0 / (k[i + 2] - k[i + 1] + tol) * n0[i + 1];
    }
#pragma unroll
    for (int i = 0; i < 3; i++) {
        double dl = k[i + 2] - k[i] + tol, dr = k[i + 3] - k[i + 1] + tol;
        v[i] = (x - k[i]) / dl * n1[i] + (k[i + 3] - x) / dr * n1[i + 1];
        d1[i] = 2.0 / dl * n1[i] - 2.0 / dr * n1[i + 1];
        d2[i] = 2.0 / dl * dn1[i] - 2.0 / dr * dn1[i + 1];
    }
}

// derivative order in direction d (0..2) of slot s (0..9): slots (value, x,
// y, z, xx, xy, xz, yy, yz, zz) packed 2 bits per direction, 6 bits per
// slot, in an immediate (a per-lane table lookup would be a memory load)
IGAP4H_HD int slot_d(int s, int d)
{
    return (int)((0x814211142404040ULL >> (6 * s + 2 * d)) & 3ULL);
}

} // namespace igap4h
