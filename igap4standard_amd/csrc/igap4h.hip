// igap4h: residual and tangent assembly for C1 quadratic B-spline gradient
// elasticity on one MI355X (gfx950).  C-ABI in include/igap4h.h.
//
// Data layout in HBM (reference layout, one rank):
//   U_hist  (torder+1) x ndof doubles, dof = 3*((nb1*nb2)*a0 + nb2*a1 + a2) + i
//   R       ndof doubles
//   CSR     rowptr int64 [ndof+1], col int32 [nnz] (ascending), vals f64|f32
//   coef    n_elem x 6 doubles, element e = (ex*ne1 + ey)*ne2 + ez
//   Re      n_elem x 81 doubles (per-element residual, gathered per dof)
// Connectivity is arithmetic (structured patch); no index arrays per element.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/igap4h.h"
#include "igap4h_physics.hpp"
#include "igap4h_setup.hpp"
#include "igap4h_internal.hpp"

using namespace igap4h;

// ============================================================== errors
// a fixed buffer per thread: reporting an error allocates nothing
static thread_local char g_err[512];
int igap4h::fail(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return code;
}

extern "C" const char *igap4h_last_error(void) { return g_err; }

bool igap4h::alloc_refused(size_t bytes)
{
    const char *v = std::getenv("IGAP4H_DEBUG_ALLOC_LIMIT");
    if (!v || !*v) return false;
    char *end = nullptr;
    unsigned long long lim = std::strtoull(v, &end, 10);
    return end != v && bytes > lim;
}

// ============================================================== kernel args
struct MeshArgs {
    int ne[3];
    int nb[3];
    const double *knot[3];
    const double *phi;  // 1D basis tables, [(phioff[d] + e_d) * 4 + q][9]
    int phioff[3];
    int nq;
    double cq[4], wq[4];
    double pp[9];     // par_periodic (F0)
    long long ndof;
    double rho, cdamp, dt; // getime: par_mat[5..7]
};

template <int PK>
struct PhysT;
template <>
struct PhysT<0> { // phys/ge: torder 0, gradient slots only in the contraction
    using T = double;
    using TR = double;
    static constexpr int NS = 9, S0 = 1, NH = 1;
};
template <>
struct PhysT<1> { // phys/getime: torder 2, value slot carries the mass term
    using T = Jet<3>;
    using TR = Jet<4>;
    static constexpr int NS = 10, S0 = 0, NH = 3;
};

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));
template <class A>
struct V4;
template <>
struct V4<double> { using t = d4; };
template <>
struct V4<float> { using t = f4; };

__device__ __forceinline__ d4 mfma16(double a, double b, d4 c)
{
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f4 mfma16(float a, float b, f4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}
// Output row of accumulator register r in lane l (column is l % 16 for both):
// measured on gfx950, f64 16x16x4 interleaves rows (l/16 + 4r), f32
// 16x16x4 keeps them contiguous (4 (l/16) + r).  Operands: A row = l % 16,
// B column = l % 16, k = l / 16 in both.
__device__ __forceinline__ int acc_row(double, int l, int r) { return (l >> 4) + 4 * r; }
__device__ __forceinline__ int acc_row(float, int l, int r) { return 4 * (l >> 4) + r; }

// 1D basis values of element e_d of direction d at Gauss point q: o[a*3 + k],
// k = value, first, second derivative of function a (eval_Bspline.c:9-93)
__device__ __forceinline__ void phi_1d(const MeshArgs &m, int d, int ed, int q, double *o)
{
    const double *k = m.knot[d] + ed;
    double X0 = k[2], X1 = k[3];
    double xi = m.cq[q];
    double x = (-xi + 1.0) * X0 + xi * X1;
    double v[3], d1[3], d2[3];
    bspline1d(k, x, v, d1, d2);
    for (int a = 0; a < 3; a++) {
        o[a * 3 + 0] = v[a];
        o[a * 3 + 1] = d1[a];
        o[a * 3 + 2] = d2[a];
    }
}

// the tables of every element and direction, one thread per (d, e_d, q)
__global__ void k_phi_table(MeshArgs m, double *tab)
{
    const int n = (m.ne[0] + m.ne[1] + m.ne[2]) * 4;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int row = t / 4, q = t % 4;
    const int d = row < m.ne[0] ? 0 : (row < m.ne[0] + m.ne[1] ? 1 : 2);
    const int ed = row - m.phioff[d];
    double o[9];
    if (q < m.nq) phi_1d(m, d, ed, q, o);
    else
        for (int j = 0; j < 9; j++) o[j] = 0.0;
    for (int j = 0; j < 9; j++) tab[(size_t)t * 9 + j] = o[j];
}

// 1D basis tables of one element: phi[(d*4 + q)*9 + a*3 + k], k = value,
// first, second derivative of function a at Gauss point q in direction d,
// copied from the context's tables
__device__ __forceinline__ void fill_phi(const MeshArgs &m, const int e[3], double *sphi, int tid, int nthr)
{
    for (int idx = tid; idx < 3 * 4 * 9; idx += nthr) {
        const int d = idx / 36, r = idx % 36;
        sphi[idx] = m.phi[(size_t)(m.phioff[d] + e[d]) * 36 + r];
    }
}

// slot values of the basis function A = 9 ax + 3 ay + az at point (qx,qy,qz)
template <class R>
__device__ __forceinline__ void basis_slots(const R *sphi, int qx, int qy, int qz, int A, R n[10])
{
    const R *px = sphi + (0 * 4 + qx) * 9 + (A / 9) * 3;
    const R *py = sphi + (1 * 4 + qy) * 9 + ((A / 3) % 3) * 3;
    const R *pz = sphi + (2 * 4 + qz) * 9 + (A % 3) * 3;
    n[0] = px[0] * py[0] * pz[0];
    n[1] = px[1] * py[0] * pz[0];
    n[2] = px[0] * py[1] * pz[0];
    n[3] = px[0] * py[0] * pz[1];
    n[4] = px[2] * py[0] * pz[0];
    n[5] = px[1] * py[1] * pz[0];
    n[6] = px[1] * py[0] * pz[1];
    n[7] = px[0] * py[2] * pz[0];
    n[8] = px[0] * py[1] * pz[1];
    n[9] = px[0] * py[0] * pz[2];
}

__device__ __forceinline__ long long node_of(const MeshArgs &m, int a0, int a1, int a2)
{
    return ((long long)a0 * m.nb[1] + a1) * m.nb[2] + a2;
}

// Interpolated slots u[h][30] at one Gauss point (snesiga_Residual.c:125-144),
// F0 of par_periodic added to value and gradient slots of every slice.
template <int NH>
__device__ __forceinline__ void interp_slots(const MeshArgs &m, const double *sU, const double *sphi, const int e[3],
                                             int qx, int qy, int qz, double u[NH][30], double *w)
{
    for (int h = 0; h < NH; h++)
        for (int s = 0; s < 30; s++) u[h][s] = 0.0;
    for (int A = 0; A < 27; A++) {
        double n[10];
        basis_slots(sphi, qx, qy, qz, A, n);
        for (int h = 0; h < NH; h++)
            for (int i = 0; i < 3; i++) {
                double Uv = sU[h * 81 + 3 * A + i];
#pragma unroll
                for (int s = 0; s < 10; s++) u[h][10 * i + s] += n[s] * Uv;
            }
    }
    double xq[3], hh[3];
    int q[3] = {qx, qy, qz};
    for (int d = 0; d < 3; d++) {
        const double *k = m.knot[d] + e[d];
        double xi = m.cq[q[d]];
        xq[d] = (-xi + 1.0) * k[2] + xi * k[3];
        hh[d] = k[3] - k[2];
    }
    for (int h = 0; h < NH; h++)
        for (int i = 0; i < 3; i++) {
            u[h][10 * i] += m.pp[3 * i] * xq[0] + m.pp[3 * i + 1] * xq[1] + m.pp[3 * i + 2] * xq[2];
            for (int d = 0; d < 3; d++) u[h][10 * i + 1 + d] += m.pp[3 * i + d];
        }
    *w = m.wq[qx] * m.wq[qy] * m.wq[qz] * hh[0] * hh[1] * hh[2];
}

// u-slots of one element at all Gauss points by sum factorisation over the
// tensor-product basis, all threads of the block: z contracted first, then y,
// then x (9 + 18 + 30 products per point instead of 27 x 40).  Scratch
// (double): V1 [i][ax][ay][dz][qz] 324, V2 [i][ax][c][qy][qz] 864 with
// c = (dy,dz) in (00,10,01,20,11,02), out [30 slots][qx*16 + qy*4 + qz] 1920.
// Ends with a barrier.  Periodic F0 terms and weights are left to the caller.
constexpr int SF_SCRATCH = 324 + 864 + 1920;
template <int NTHR>
__device__ __forceinline__ void interp_slots_sf(int nq, const double *sU, const double *sphi, double *scr, int tid)
{
    double *V1 = scr, *V2 = scr + 324, *uo = scr + 324 + 864;
#pragma unroll
    for (int it = 0; it < (324 + NTHR - 1) / NTHR; it++) {
        const int idx = tid + it * NTHR;
        const int qz = idx & 3, dz = (idx >> 2) % 3, r = idx / 12; // r = (i*3 + ax)*3 + ay
        if (idx < 324 && qz < nq) {
            const int ay = r % 3, ax = (r / 3) % 3, i = r / 9;
            const double *z = sphi + (8 + qz) * 9 + dz;
            const double *u = sU + 3 * (9 * ax + 3 * ay) + i;
            V1[idx] = z[0] * u[0] + z[3] * u[3] + z[6] * u[6];
        }
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < (864 + NTHR - 1) / NTHR; it++) {
        const int idx = tid + it * NTHR;
        const int qz = idx & 3, qy = (idx >> 2) & 3, c = (idx >> 4) % 6, r = idx / 96; // r = i*3 + ax
        if (idx < 864 && qy < nq && qz < nq) {
            const int dy = (0x012010 >> (4 * c)) & 15, dz = (0x210100 >> (4 * c)) & 15;
            const double *y = sphi + (4 + qy) * 9 + dy;
            const double *v = V1 + (r * 9 + dz) * 4 + qz; // ay stride 12
            V2[idx] = y[0] * v[0] + y[3] * v[12] + y[6] * v[24];
        }
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < (1920 + NTHR - 1) / NTHR; it++) {
        const int idx = tid + it * NTHR;
        const int q = idx & 63, sl = idx >> 6, i = sl / 10, sv = sl % 10;
        const int qx = q >> 4, qy = (q >> 2) & 3, qz = q & 3;
        if (idx < 1920 && qx < nq && qy < nq && qz < nq) {
            const int dx = (int)((0x0001120010ULL >> (4 * sv)) & 15), c = (int)((0x5432102100ULL >> (4 * sv)) & 15);
            const double *x = sphi + qx * 9 + dx;
            const double *v = V2 + ((i * 3) * 6 + c) * 16 + qy * 4 + qz; // ax stride 96
            uo[idx] = x[0] * v[0] + x[3] * v[96] + x[6] * v[192];
        }
    }
    __syncthreads();
}

// As interp_slots_sf for one wave, with the last (x) stage evaluated by the
// lane of each Gauss point into its own registers: scratch is V1 and V2 only
// (1,188 doubles).  Ends without a barrier after the register stage.
constexpr int SF2_SCRATCH = 324 + 864;
template <int NTHR>
__device__ __forceinline__ void interp_slots_sf2(int nq, const double *sU, const double *sphi, double *scr, int tid,
                                                 bool active, int qx, int qy, int qz, double u[30])
{
    double *V1 = scr, *V2 = scr + 324;
#pragma unroll
    for (int it = 0; it < (324 + NTHR - 1) / NTHR; it++) {
        const int idx = tid + it * NTHR;
        const int qz2 = idx & 3, dz = (idx >> 2) % 3, r = idx / 12;
        if (idx < 324 && qz2 < nq) {
            const int ay = r % 3, ax = (r / 3) % 3, i = r / 9;
            const double *z = sphi + (8 + qz2) * 9 + dz;
            const double *uu = sU + 3 * (9 * ax + 3 * ay) + i;
            V1[idx] = z[0] * uu[0] + z[3] * uu[3] + z[6] * uu[6];
        }
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < (864 + NTHR - 1) / NTHR; it++) {
        const int idx = tid + it * NTHR;
        const int qz2 = idx & 3, qy2 = (idx >> 2) & 3, c = (idx >> 4) % 6, r = idx / 96;
        if (idx < 864 && qy2 < nq && qz2 < nq) {
            const int dy = (0x012010 >> (4 * c)) & 15, dz = (0x210100 >> (4 * c)) & 15;
            const double *y = sphi + (4 + qy2) * 9 + dy;
            const double *v = V1 + (r * 9 + dz) * 4 + qz2;
            V2[idx] = y[0] * v[0] + y[3] * v[12] + y[6] * v[24];
        }
    }
    __syncthreads();
    if (active) {
#pragma unroll
        for (int sl = 0; sl < 30; sl++) {
            const int i = sl / 10, sv = sl % 10;
            const int dx = (int)((0x0001120010ULL >> (4 * sv)) & 15), c = (int)((0x5432102100ULL >> (4 * sv)) & 15);
            const double *x = sphi + qx * 9 + dx;
            const double *v = V2 + ((i * 3) * 6 + c) * 16 + qy * 4 + qz;
            u[sl] = x[0] * v[0] + x[3] * v[96] + x[6] * v[192];
        }
    }
}

// the knot span [a_d, b_d] of an element in each direction (per-element
// constants, loaded with the element's other inputs)
struct ElemGeo {
    double a[3], b[3];
};
__device__ __forceinline__ ElemGeo elem_geo(const MeshArgs &m, const int e[3])
{
    ElemGeo g;
    for (int d = 0; d < 3; d++) {
        g.a[d] = m.knot[d][e[d] + 2];
        g.b[d] = m.knot[d][e[d] + 3];
    }
    return g;
}

// periodic F0 terms of one point (snesiga_Residual.c:125-144) and its weight
__device__ __forceinline__ void point_affine(const MeshArgs &m, const ElemGeo &g, int qx, int qy, int qz, double u[30],
                                             double *w)
{
    double xq[3], hh[3];
    int q[3] = {qx, qy, qz};
    for (int d = 0; d < 3; d++) {
        double xi = m.cq[q[d]];
        xq[d] = (-xi + 1.0) * g.a[d] + xi * g.b[d];
        hh[d] = g.b[d] - g.a[d];
    }
    for (int i = 0; i < 3; i++) {
        u[10 * i] += m.pp[3 * i] * xq[0] + m.pp[3 * i + 1] * xq[1] + m.pp[3 * i + 2] * xq[2];
        for (int d = 0; d < 3; d++) u[10 * i + 1 + d] += m.pp[3 * i + d];
    }
    *w = m.wq[qx] * m.wq[qy] * m.wq[qz] * hh[0] * hh[1] * hh[2];
}
__device__ __forceinline__ Coef load_coef(const double *coef, long long e)
{
    const double *c = coef + 6 * e;
    Coef r;
    r.A36 = c[0]; r.A37 = c[1]; r.A38 = c[2]; r.e0 = c[3]; r.e345 = c[4]; r.lsq = c[5];
    return r;
}

// ============================================================== residual
// One wave per element.  The u-slots at the Gauss points come from a sum-
// factorised interpolation (interp_slots_sf2), the pointwise stress runs with
// lane = Gauss point, and the weak-form projection
//   R_e[3A + i] = sum_q w_q sum_s N_{A,s}(q) r_{i,s}(q)   (snesiga_Residual.c:147-159)
// is sum-factorised the other way (x contracted first, then y, then z), one
// dof i at a time so that the scratch stays at 1,188 doubles.
// Writes R_e[81] to Re; k_gather sums it per dof.
template <int PK>
__device__ __forceinline__ void residual_elem_body(const MeshArgs &m, const double *__restrict__ Uh,
                                                   const double *__restrict__ coef, double *__restrict__ Re, long long e,
                                                   long long slot)
{
    constexpr int NH = PhysT<PK>::NH;
    __shared__ double sU[NH * 81];
    __shared__ double sphi[3 * 4 * 9];
    __shared__ double sScr[SF2_SCRATCH];
    const int tid = threadIdx.x;
    int ev[3];
    ev[2] = (int)(e % m.ne[2]);
    ev[1] = (int)((e / m.ne[2]) % m.ne[1]);
    ev[0] = (int)(e / ((long long)m.ne[2] * m.ne[1]));
    for (int idx = tid; idx < NH * 81; idx += 64) {
        int h = idx / 81, r = idx % 81, A = r / 3, i = r % 3;
        long long nd = node_of(m, ev[0] + A / 9, ev[1] + (A / 3) % 3, ev[2] + A % 3);
        sU[idx] = Uh[h * m.ndof + 3 * nd + i];
    }
    const Coef c = load_coef(coef, e);
    const ElemGeo geo = elem_geo(m, ev);
    fill_phi(m, ev, sphi, tid, 64);
    __syncthreads();
    const int nq = m.nq, nqp = nq * nq * nq;
    const bool active = tid < nqp;
    const int qx = tid / (nq * nq), qy = (tid / nq) % nq, qz = tid % nq, qi = qx * 16 + qy * 4 + qz;
    double u[NH][30];
    for (int h = 0; h < NH; h++) {
        interp_slots_sf2<64>(nq, sU + h * 81, sphi, sScr, tid, active, qx, qy, qz, u[h]);
        __syncthreads();
    }
    double r[30];
    for (int s = 0; s < 30; s++) r[s] = 0.0;
    double w = 0.0;
    if (active) {
        for (int h = 0; h < NH; h++) point_affine(m, geo, qx, qy, qz, u[h], &w);
        if (PK == 0) {
            residual_grad<double>(u[0], c, r);
        } else {
            using TR = typename PhysT<PK>::TR;
            TR uj[30], rj[30];
            // u = t_{n+1} (slice 0), u1 = t_n (slice 1), u0 = t_{n-1} (slice 2)
            for (int s = 0; s < 30; s++) uj[s] = Num<TR>::lin((u[NH - 1][s] + u[1 % NH][s]) / 2.0, (u[0][s] - u[NH - 1][s]) / 2.0);
            residual_grad<TR>(uj, c, rj);
            for (int s = 0; s < 30; s++) r[s] = avg_residual(rj[s]);
            for (int i = 0; i < 3; i++) {
                int s = 10 * i;
                r[s] = (m.cdamp * (u[0][s] - u[NH - 1][s])) / (2. * m.dt) +
                       (m.rho * (u[0][s] + u[NH - 1][s] - 2 * u[1 % NH][s])) / (m.dt * m.dt);
            }
        }
    }
    double *rw = sScr, *P1 = sScr + 640, *P2 = sScr + 640 + 288;
    for (int i = 0; i < 3; i++) {
        // weighted point values of dof i, [10 slots][qx*16 + qy*4 + qz]
        if (active) {
#pragma unroll
            for (int sv = 0; sv < 10; sv++) rw[sv * 64 + qi] = w * r[10 * i + sv];
        }
        __syncthreads();
        // x: P1[ax][c][qy][qz] = sum_qx sum_{s: (dy,dz)(s) = c} X(qx, ax, dx(s)) rw[s];
        //    c = (dy,dz) in (00,10,01,20,11,02); slots of c with dx = 0, 1, 2:
        //    (0,1,4) (2,5,-) (3,6,-) (7,-,-) (8,-,-) (9,-,-)
#pragma unroll
        for (int it = 0; it < 5; it++) {
            const int idx = tid + 64 * it;
            const int qz2 = idx & 3, qy2 = (idx >> 2) & 3, cc = (idx >> 4) % 6, ax = idx / 96;
            if (idx < 288 && qy2 < nq && qz2 < nq) {
                const int s0 = (0x987320 >> (4 * cc)) & 15, s1 = (0x000651 >> (4 * cc)) & 15;
                const double m1 = cc < 3 ? 1.0 : 0.0, m2 = cc == 0 ? 1.0 : 0.0;
                double t = 0.0;
#pragma unroll
                for (int qx2 = 0; qx2 < 4; qx2++) {
                    if (qx2 >= nq) break;
                    const double *x = sphi + qx2 * 9 + ax * 3;
                    const int o = qx2 * 16 + qy2 * 4 + qz2;
                    t += x[0] * rw[s0 * 64 + o] + m1 * (x[1] * rw[s1 * 64 + o]) + m2 * (x[2] * rw[4 * 64 + o]);
                }
                P1[idx] = t;
            }
        }
        __syncthreads();
        // y: P2[ax][ay][dz][qz] = sum_qy sum_{c: dz(c) = dz} Y(qy, ay, dy(c)) P1[ax][c][qy][qz];
        //    combos of dz with dy = 0, 1, 2: (0,1,3) (2,4,-) (5,-,-)
#pragma unroll
        for (int it = 0; it < 2; it++) {
            const int idx = tid + 64 * it;
            const int qz2 = idx & 3, dz = (idx >> 2) % 3, r3 = idx / 12;  // r3 = ax*3 + ay
            if (idx < 108 && qz2 < nq) {
                const int ay = r3 % 3, ax = r3 / 3;
                const int c0 = (0x520 >> (4 * dz)) & 15, c1 = (0x041 >> (4 * dz)) & 15, c2 = 3;
                const double m1 = dz < 2 ? 1.0 : 0.0, m2 = dz == 0 ? 1.0 : 0.0;
                double t = 0.0;
#pragma unroll
                for (int qy2 = 0; qy2 < 4; qy2++) {
                    if (qy2 >= nq) break;
                    const double *y = sphi + (4 + qy2) * 9 + ay * 3;
                    const double *p = P1 + ax * 96 + qy2 * 4 + qz2;
                    t += y[0] * p[c0 * 16] + m1 * (y[1] * p[c1 * 16]) + m2 * (y[2] * p[c2 * 16]);
                }
                P2[idx] = t;
            }
        }
        __syncthreads();
        // z: R_e[3A + i] = sum_qz sum_dz Z(qz, az, dz) P2[ax][ay][dz][qz]
        if (tid < 27) {
            const int A = tid, ax = A / 9, ay = (A / 3) % 3, az = A % 3;
            const double *p = P2 + (ax * 3 + ay) * 12;
            double t = 0.0;
#pragma unroll
            for (int qz2 = 0; qz2 < 4; qz2++) {
                if (qz2 >= nq) break;
                const double *z = sphi + (8 + qz2) * 9 + az * 3;
                t += z[0] * p[qz2] + z[1] * p[4 + qz2] + z[2] * p[8 + qz2];
            }
            Re[slot * 81 + 3 * A + i] = t;
        }
        __syncthreads();
    }
}

template <int PK>
__global__ __launch_bounds__(64) void k_residual_elem(MeshArgs m, const double *__restrict__ Uh,
                                                      const double *__restrict__ coef, double *__restrict__ Re)
{
    residual_elem_body<PK>(m, Uh, coef, Re, (long long)blockIdx.x, (long long)blockIdx.x);
}

// Sum of element vectors per dof in element order (the reference's ADD order),
// Neumann load, then Dirichlet rows R = U - g (INSERT, snesiga_Residual.c:382).
__device__ __forceinline__ void gather_body(const MeshArgs &m, const double *__restrict__ Re,
                                            const double *__restrict__ fneu, const unsigned char *__restrict__ dmask,
                                            const double *__restrict__ gval, const double *__restrict__ U0,
                                            double *__restrict__ R, long long dof)
{
    long long node = dof / 3;
    int i = (int)(dof % 3);
    int a2 = (int)(node % m.nb[2]);
    int a1 = (int)((node / m.nb[2]) % m.nb[1]);
    int a0 = (int)(node / ((long long)m.nb[2] * m.nb[1]));
    double s = 0.0;
    for (int ex = max(0, a0 - 2); ex <= min(a0, m.ne[0] - 1); ex++)
        for (int ey = max(0, a1 - 2); ey <= min(a1, m.ne[1] - 1); ey++)
            for (int ez = max(0, a2 - 2); ez <= min(a2, m.ne[2] - 1); ez++) {
                long long e = ((long long)ex * m.ne[1] + ey) * m.ne[2] + ez;
                int A = (a0 - ex) * 9 + (a1 - ey) * 3 + (a2 - ez);
                s += Re[e * 81 + 3 * A + i];
            }
    if (fneu) s += fneu[dof];
    if (dmask[dof] == 1) s = U0[dof] - gval[dof];
    else if (dmask[dof] == 2) s = 0.0;  // another rank's Dirichlet row: it sets the value
    R[dof] = s;
}

__global__ __launch_bounds__(256) void k_gather(MeshArgs m, const double *__restrict__ Re,
                                                const double *__restrict__ fneu, const unsigned char *__restrict__ dmask,
                                                const double *__restrict__ gval, const double *__restrict__ U0,
                                                double *__restrict__ R)
{
    long long dof = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (dof >= m.ndof) return;
    gather_body(m, Re, fneu, dmask, gval, U0, R, dof);
}

// ---------------------------------------------------------------- residual without allocation
// Element vectors of a chunk of consecutive elements [e0, e0 + n) staged in a
// fixed device array (slot = e - e0).  Elements are in the reference order (x
// slowest), so a chunk's elements all touch the x-planes of nodes
// [ex(e0), ex(e0 + n - 1) + 2] and every dof sees its contributions chunk
// after chunk in element order: adding them to the running R one at a time
// repeats k_gather's additions exactly.
constexpr long long kScratchElems = 32768;            // element vectors per chunk
static __device__ __align__(256) double scratch[kScratchElems * 81];   // 21.2 MB per device

template <int PK>
__global__ __launch_bounds__(64) void k_residual_chunk(MeshArgs m, const double *__restrict__ Uh,
                                                       const double *__restrict__ coef, double *__restrict__ Re,
                                                       long long e0)
{
    residual_elem_body<PK>(m, Uh, coef, Re, e0 + blockIdx.x, (long long)blockIdx.x);
}

// R[dof] += contributions of the chunk's elements, dof in [d0, d1)
__global__ __launch_bounds__(256) void k_gather_chunk(MeshArgs m, const double *__restrict__ Re, long long e0, long long e1,
                                                      long long d0, long long d1, double *__restrict__ R)
{
    long long dof = d0 + (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (dof >= d1) return;
    long long node = dof / 3;
    int i = (int)(dof % 3);
    int a2 = (int)(node % m.nb[2]);
    int a1 = (int)((node / m.nb[2]) % m.nb[1]);
    int a0 = (int)(node / ((long long)m.nb[2] * m.nb[1]));
    double s = R[dof];
    bool any = false;
    for (int ex = max(0, a0 - 2); ex <= min(a0, m.ne[0] - 1); ex++)
        for (int ey = max(0, a1 - 2); ey <= min(a1, m.ne[1] - 1); ey++)
            for (int ez = max(0, a2 - 2); ez <= min(a2, m.ne[2] - 1); ez++) {
                long long e = ((long long)ex * m.ne[1] + ey) * m.ne[2] + ez;
                if (e < e0 || e >= e1) continue;
                int A = (a0 - ex) * 9 + (a1 - ey) * 3 + (a2 - ez);
                s += Re[(e - e0) * 81 + 3 * A + i];
                any = true;
            }
    if (any) R[dof] = s;
}

// Neumann load, then Dirichlet rows R = U - g (as gather_body ends)
__global__ __launch_bounds__(256) void k_gather_finish(MeshArgs m, const double *__restrict__ fneu,
                                                       const unsigned char *__restrict__ dmask,
                                                       const double *__restrict__ gval, const double *__restrict__ U0,
                                                       double *__restrict__ R)
{
    long long dof = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (dof >= m.ndof) return;
    double s = R[dof];
    if (fneu) s += fneu[dof];
    if (dmask[dof] == 1) s = U0[dof] - gval[dof];
    else if (dmask[dof] == 2) s = 0.0;
    R[dof] = s;
}

// ============================================================== tangent
// One work-group of 256 threads (4 waves) per element of one colour; each
// wave owns 6 of the 24 16x16 output tiles of the i <= j dof blocks.

// Column tiles: row-block i of K holds columns (j >= i, B) padded to 32 per
// j-block: widths 96 | 64 | 32 at offsets 0 | 96 | 160 of a 192-wide strip.
__device__ __forceinline__ int ctile_block(int ct) { return ct < 6 ? 0 : (ct < 10 ? 1 : 2); }
__device__ __forceinline__ int block_off(int i) { return i == 0 ? 0 : (i == 1 ? 96 : 160); }

// Read-modify-write of the staged 81x81 element matrix into the CSR values,
// in two steps so that the loads can be issued before the element matrix is
// staged (or earlier): thread t < 243 owns column cc = t mod 81
// of the element matrix and rows t / 81 + 3 k (k = 0..26); the column's offset
// within a CSR row (x and y node steps, 3 z + dof) is decoded once, a row adds
// its base and strides.  No other element of a colour touches these values, so
// they may be read any time after the previous colour's launch.
struct ScatterCol {
    int cc, r0, cx, cy, cz;
    int kb, ks;  // staged value k of the thread's column: sKe[kb + ks k]
    bool act;
};
__device__ __forceinline__ ScatterCol scatter_col(int tid, const unsigned char *sMask)
{
    ScatterCol s;
    s.act = tid < 243;
    s.cc = tid % 81;
    s.r0 = s.act ? tid / 81 : 0;
    const int B = s.cc / 3;
    s.cx = B / 9;
    s.cy = (B / 3) % 3;
    s.cz = 3 * (B % 3) + s.cc % 3;
    s.act = s.act && !sMask[s.cc];
    s.kb = s.r0 * 81 + s.cc;  // rows 3 A + i, columns 3 B + j
    s.ks = 243;
    return s;
}
template <class VAL>
__device__ __forceinline__ void scatter_load(const ScatterCol &s, const long long *sRowBase, const int *sRowSx,
                                             const int *sRowSy, const unsigned char *sMask, const VAL *vals, VAL old[27])
{
#pragma unroll
    for (int k = 0; k < 27; k++) {
        const int r = s.r0 + 3 * k;
        if (s.act && !sMask[r]) old[k] = vals[sRowBase[r] + (s.cx * sRowSx[r] + s.cy * sRowSy[r] + s.cz)];
    }
}
template <class VAL, class ACC>
__device__ __forceinline__ void scatter_store(const ScatterCol &s, const ACC *sKe, const long long *sRowBase,
                                              const int *sRowSx, const int *sRowSy, const unsigned char *sMask, VAL *vals,
                                              const VAL old[27])
{
    // every sum first, then the stores back to back: loads and stores share one
    // counter, so a load still pending between two stores would make each store
    // wait for the ones before it
    VAL nv[27];
#pragma unroll
    for (int k = 0; k < 27; k++) nv[k] = old[k] + (VAL)sKe[s.kb + s.ks * k];
    __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0): the loads are in, none left between the stores
#pragma unroll
    for (int k = 0; k < 27; k++) {
        const int r = s.r0 + 3 * k;
        if (s.act && !sMask[r]) vals[sRowBase[r] + (s.cx * sRowSx[r] + s.cy * sRowSy[r] + s.cz)] = nv[k];
    }
}

// Loads of the element's CSR values, then stage() (which stages the element
// matrix in sKe and ends with a barrier), then the stores.
template <class ACC, class STAGE>
__device__ __forceinline__ void scatter_staged(int vals_f32, void *vals, const ACC *sKe, const long long *sRowBase,
                                               const int *sRowSx, const int *sRowSy, const unsigned char *sMask, int tid,
                                               STAGE stage)
{
    const ScatterCol sc = scatter_col(tid, sMask);
    if (vals_f32) {
        float old[27];
        scatter_load<float>(sc, sRowBase, sRowSx, sRowSy, sMask, (const float *)vals, old);
        stage();
        scatter_store<float, ACC>(sc, sKe, sRowBase, sRowSx, sRowSy, sMask, (float *)vals, old);
    } else {
        double old[27];
        scatter_load<double>(sc, sRowBase, sRowSx, sRowSy, sMask, (const double *)vals, old);
        stage();
        scatter_store<double, ACC>(sc, sKe, sRowBase, sRowSx, sRowSy, sMask, (double *)vals, old);
    }
}

// ---------------------------------------------------------------- ge, high occupancy
// phys/ge tangent, 4 waves per element with a small LDS footprint so that
// two or more elements share a CU and one element's VALU phase overlaps
// another's MFMA phase.  Per chunk of QC Gauss points:
//   VALU  waves 0-2 build T_ij (dof block j = 2 - wave, every lane the same
//         code path), wave 3 builds the A operand N^T;
//   MFMA  all 4 waves contract their 6 tiles over the chunk.
// The per-point compact data and the T strip share one LDS region that is
// reused to stage the 81x81 element matrix for a coalesced scatter.
constexpr int NCMP = 48; // doubles of Compact<double>

template <int QC>
struct GeTile {
    static constexpr int KC = QC * 9;                 // contraction depth per chunk
    static constexpr int KCP = (KC + 3) / 4 * 4;      // padded to the MFMA k of 4
    static constexpr int NTS = KCP * 192;             // T strip
};

// compact data of one Gauss point (array of structures, read in place by the
// T phase), rounded once to the tangent's arithmetic type ACC
template <class ACC>
__device__ __forceinline__ void store_compact_aos(ACC *sC, int qp, const Compact<double> &cp)
{
    const double *v = reinterpret_cast<const double *>(&cp);
#pragma unroll
    for (int f = 0; f < NCMP; f++) sC[qp * NCMP + f] = (ACC)v[f];
}

// T_ij(q, B) for dof block j = J and rows i of [I0, I1], in the arithmetic
// of ACC (FP64, or FP32 for the FP32 tangent whose MFMA operands are FP32)
template <int J, int I0, int I1, class ACC>
__device__ __forceinline__ void ge_task(const ACC *sCmp, const ACC *sphi, const int *sQ, const double *sw,
                                        ACC *sT, const Coef &c, int ql, int B, int qp, int nqp)
{
    ACC out[J + 1][10];
    if (qp < nqp) {
        ACC n[10];
        int q = sQ[qp];
        basis_slots<ACC>(sphi, q & 15, (q >> 4) & 15, (q >> 8) & 15, B, n);
        const Compact<ACC> &cp = reinterpret_cast<const Compact<ACC> *>(sCmp)[qp];  // read from LDS at use
        tangent_apply<ACC, J, I0, I1, ACC>(cp, c, n, (ACC)0, (ACC)sw[qp], out);
    } else {
#pragma unroll
        for (int i = I0; i <= I1; i++)
#pragma unroll
            for (int s = 0; s < 10; s++) out[i][s] = (ACC)0;
    }
#pragma unroll
    for (int i = I0; i <= I1; i++) {
        ACC *col = sT + block_off(i) + (J - i) * 27 + B;
#pragma unroll
        for (int s = 1; s < 10; s++) col[(ql * 9 + s - 1) * 192] = (ACC)out[i][s];
    }
}

template <class ACC, int QC>
__device__ __forceinline__ void tangent_ge_body(const MeshArgs &m, const double *__restrict__ Uh, const double *__restrict__ coef,
                                                const long long *__restrict__ rowptr, const unsigned char *__restrict__ dmask,
                                                void *__restrict__ vals, int vals_f32, const int ev[3])
{
    using G = GeTile<QC>;
    using V = typename V4<ACC>::t;
    // region: [compact SoA | T strip], reused for the element matrix
    constexpr int CMP_BYTES = NCMP * 64 * (int)sizeof(ACC);
    constexpr int OP_BYTES0 = G::NTS * (int)sizeof(ACC);
    // the operand area doubles as the interpolation scratch of the prologue
    constexpr int OP_BYTES = OP_BYTES0 > SF_SCRATCH * 8 ? OP_BYTES0 : SF_SCRATCH * 8;
    constexpr int KE_BYTES = 81 * 81 * (int)sizeof(ACC);
    constexpr int REGION = (CMP_BYTES + OP_BYTES) > KE_BYTES ? (CMP_BYTES + OP_BYTES) : KE_BYTES;
    __shared__ __attribute__((aligned(16))) unsigned char sRegion[REGION];
    __shared__ double sU[81];
    __shared__ double sphi[3 * 4 * 9];
    __shared__ ACC sphiA[3 * 4 * 9]; // the same tables in the arithmetic type
    __shared__ double sw[64];
    __shared__ int sQ[64];
    __shared__ long long sRowBase[81];
    __shared__ int sRowSx[81], sRowSy[81];
    __shared__ unsigned char sMask[81];
    ACC *sCmp = reinterpret_cast<ACC *>(sRegion);
    ACC *sT = reinterpret_cast<ACC *>(sRegion + CMP_BYTES);

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const long long e = ((long long)ev[0] * m.ne[1] + ev[1]) * m.ne[2] + ev[2];
    const int nq = m.nq, nqp = nq * nq * nq;

    if (tid < 81) {
        int A = tid / 3, i = tid % 3;
        int a0 = ev[0] + A / 9, a1 = ev[1] + (A / 3) % 3, a2 = ev[2] + A % 3;
        long long grow = 3 * node_of(m, a0, a1, a2) + i;
        sU[tid] = Uh[grow];
        int lo0 = max(0, a0 - 2), lo1 = max(0, a1 - 2), lo2 = max(0, a2 - 2);
        int w1 = min(m.nb[1] - 1, a1 + 2) - lo1 + 1, w2 = min(m.nb[2] - 1, a2 + 2) - lo2 + 1;
        sRowBase[tid] = rowptr[grow] + 3LL * (((long long)(ev[0] - lo0) * w1 + (ev[1] - lo1)) * w2 + (ev[2] - lo2));
        sRowSx[tid] = 3 * w1 * w2;
        sRowSy[tid] = 3 * w2;
        sMask[tid] = dmask[grow];
    } else if (tid >= 128 && tid < 192) {
        int qp = tid - 128;
        sQ[qp] = qp < nqp ? ((qp / (nq * nq)) | (((qp / nq) % nq) << 4) | ((qp % nq) << 8)) : 0;
    }
    // per-element constants issued with the other loads: one memory round trip before the barrier
    const Coef c = load_coef(coef, e);
    const ElemGeo geo = elem_geo(m, ev);
    fill_phi(m, ev, sphi, tid, 256);
    __syncthreads();

    double *sScr = reinterpret_cast<double *>(sRegion + CMP_BYTES);
    interp_slots_sf<256>(nq, sU, sphi, sScr, tid);
    if (tid < nqp) {
        int q = sQ[tid];
        const int qx = q & 15, qy = (q >> 4) & 15, qz = (q >> 8) & 15, qi = qx * 16 + qy * 4 + qz;
        double u[30], w;
#pragma unroll
        for (int sl = 0; sl < 30; sl++) u[sl] = sScr[324 + 864 + sl * 64 + qi];
        point_affine(m, geo, qx, qy, qz, u, &w);
        Compact<double> cp;
        compact_of<double>(u, c, cp);
        store_compact_aos<ACC>(sCmp, tid, cp);
        sw[tid] = w;
    } else if (tid >= 128 && tid < 128 + 3 * 4 * 9) {
        sphiA[tid - 128] = (ACC)sphi[tid - 128];
    }
    __syncthreads();
    // padded contraction rows of the T strip stay zero (after the scratch use)
    for (int idx = tid; idx < (G::KCP - G::KC) * 192; idx += 256) sT[G::KC * 192 + idx] = (ACC)0;

    V acc[6];
#pragma unroll
    for (int t = 0; t < 6; t++)
#pragma unroll
        for (int r = 0; r < 4; r++) acc[t][r] = (ACC)0;
    // FP32 MFMA partial sums are added into FP64 after every chunk, so the
    // FP32 rounding of the accumulation spans one chunk, not the element
    double accd[6][4];
#pragma unroll
    for (int t = 0; t < 6; t++)
#pragma unroll
        for (int r = 0; r < 4; r++) accd[t][r] = 0.0;

    // A-operand coordinates of this lane: rows A0 = lane % 16 and A1 = A0 + 16
    // (valid when < 27) of N^T, basis-table offsets per direction
    const int A0 = lane & 15, A1 = A0 + 16, krow = lane >> 4;
    const int ox0 = (A0 / 9) * 3, oy0 = ((A0 / 3) % 3) * 3, oz0 = (A0 % 3) * 3;
    const int ox1 = (A1 < 27 ? A1 / 9 : 0) * 3, oy1 = (A1 < 27 ? (A1 / 3) % 3 : 0) * 3, oz1 = (A1 < 27 ? A1 % 3 : 0) * 3;
    const ACC a1mask = A1 < 27 ? (ACC)1 : (ACC)0;

    const int nchunk = (nqp + QC - 1) / QC;
    for (int ch = 0; ch < nchunk; ch++) {
        // T_ij, lane = (ql, B); the six (i <= j) blocks over the four waves:
        // wave 0 (i = 1, 2; j = 2), wave 1 (i = 0, 1; j = 1), wave 2
        // (i = 0; j = 0), wave 3 (i = 0; j = 2)
        if (QC * 27 > 64) {
            // two lane-rounds of tasks: waves 0-1 take dof block j = 2 (rows
            // 0-2, one preamble) for the first / second round, waves 2-3 take
            // j = 1 and j = 0 (rows 0-1 and 0) for the first / second round
            const int t = ((wave & 1) ? 64 : 0) + lane;
            if (t < QC * 27) {
                const int ql = t / 27, B = t % 27, qp = ch * QC + ql;
                if (wave < 2) {
                    ge_task<2, 0, 2, ACC>(sCmp, sphiA, sQ, sw, sT, c, ql, B, qp, nqp);
                } else {
                    ge_task<1, 0, 1, ACC>(sCmp, sphiA, sQ, sw, sT, c, ql, B, qp, nqp);
                    ge_task<0, 0, 0, ACC>(sCmp, sphiA, sQ, sw, sT, c, ql, B, qp, nqp);
                }
            }
        } else {
            for (int t = lane; t < QC * 27; t += 64) {
                int ql = t / 27, B = t % 27, qp = ch * QC + ql;
                if (wave == 0) ge_task<2, 1, 2, ACC>(sCmp, sphiA, sQ, sw, sT, c, ql, B, qp, nqp);
                else if (wave == 1) ge_task<1, 0, 1, ACC>(sCmp, sphiA, sQ, sw, sT, c, ql, B, qp, nqp);
                else if (wave == 2) ge_task<0, 0, 0, ACC>(sCmp, sphiA, sQ, sw, sT, c, ql, B, qp, nqp);
                else ge_task<2, 0, 0, ACC>(sCmp, sphiA, sQ, sw, sT, c, ql, B, qp, nqp);
            }
        }
        __syncthreads();
        const int ct0 = 3 * wave;
        // all operands of the chunk first (branch-free, so that the LDS reads
        // overlap), then the MFMA stream: A fragments N_{A,s}(q) from the
        // per-element basis tables, B fragments from the T strip
        constexpr int NKS = G::KCP / 4;
        ACC fa0[NKS], fa1[NKS], fb[NKS][3];
#pragma unroll
        for (int ks = 0; ks < NKS; ks++) {
            const int k = 4 * ks + krow;
            const int ql = k / 9, sl = 1 + k % 9, qp = ch * QC + ql;
            const bool ok = k < G::KC && qp < nqp;
            const int q = sQ[ok ? qp : 0];
            const ACC *px = sphiA + (q & 15) * 9 + slot_d(sl, 0);
            const ACC *py = sphiA + (4 + ((q >> 4) & 15)) * 9 + slot_d(sl, 1);
            const ACC *pz = sphiA + (8 + ((q >> 8) & 15)) * 9 + slot_d(sl, 2);
            const ACC okm = ok ? (ACC)1 : (ACC)0;
            fa0[ks] = okm * (px[ox0] * py[oy0] * pz[oz0]);
            fa1[ks] = (okm * a1mask) * (px[ox1] * py[oy1] * pz[oz1]);
#pragma unroll
            for (int t = 0; t < 3; t++) fb[ks][t] = sT[k * 192 + (ct0 + t) * 16 + (lane & 15)];
        }
#pragma unroll
        for (int ks = 0; ks < NKS; ks++)
#pragma unroll
            for (int t = 0; t < 3; t++) {
                acc[2 * t] = mfma16(fa0[ks], fb[ks][t], acc[2 * t]);
                acc[2 * t + 1] = mfma16(fa1[ks], fb[ks][t], acc[2 * t + 1]);
            }
        if (sizeof(ACC) == 4) {
#pragma unroll
            for (int t = 0; t < 6; t++)
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    accd[t][r] += (double)acc[t][r];
                    acc[t][r] = (ACC)0;
                }
        }
        __syncthreads();
    }
    if (sizeof(ACC) == 4) {
#pragma unroll
        for (int t = 0; t < 6; t++)
#pragma unroll
            for (int r = 0; r < 4; r++) acc[t][r] = (ACC)accd[t][r];
    }

    // stage the 81x81 element matrix (mirror the i < j blocks) while the CSR values load, then scatter
    ACC *sKe = reinterpret_cast<ACC *>(sRegion);
    scatter_staged<ACC>(vals_f32, vals, sKe, sRowBase, sRowSx, sRowSy, sMask, tid, [&]() {
#pragma unroll
        for (int t = 0; t < 6; t++) {
            int ct = 3 * wave + t / 2, at = t & 1;
            int i = ctile_block(ct);
            int cl = ct * 16 + (lane & 15) - block_off(i);
            if (cl >= (3 - i) * 27) continue;
            int j = i + cl / 27, B = cl % 27;
#pragma unroll
            for (int r = 0; r < 4; r++) {
                int A = at * 16 + acc_row((ACC)0, lane, r);
                if (A >= 27) continue;
                sKe[(3 * A + i) * 81 + 3 * B + j] = acc[t][r];
                if (i != j) sKe[(3 * B + j) * 81 + 3 * A + i] = acc[t][r];
            }
        }
        __syncthreads();
    });
}

template <class ACC, int QC>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(sizeof(ACC) == 4 ? 3 : 2))) void k_tangent_ge(MeshArgs m, const double *__restrict__ Uh,
                                                    const double *__restrict__ coef,
                                                    const long long *__restrict__ rowptr,
                                                    const unsigned char *__restrict__ dmask, void *__restrict__ vals,
                                                    int vals_f32, int cx, int cy, int cz, int my, int mz)
{
    const int b = blockIdx.x;
    const int ev[3] = {cx + 3 * (b / (mz * my)), cy + 3 * ((b / mz) % my), cz + 3 * (b % mz)};
    tangent_ge_body<ACC, QC>(m, Uh, coef, rowptr, dmask, vals, vals_f32, ev);
}

// ---------------------------------------------------------------- getime
// phys/getime tangent (torder 2) on the structure of k_tangent_ge: per chunk
// of QC Gauss points the four waves build T_ij with wave-uniform dof blocks
// (lane = (point, basis function B)), then contract on MFMA.  The per-point
// compact data are jets along u(s) = ubar + s (u - u0)/2, ubar = (u0 + u1)/2
// (phys/getime/tangent.c:4127-5026: tangent = (1/2)(H0/2 + H1/3 + H2/4) of
// the series of d2W), with coefficients in the arithmetic type ACC, stored
// in LDS and read at use.  All 10 slots enter the contraction: the value
// slot carries the mass term rho/dt^2 + c/(2 dt) (phys/getime/tangent.c:4127).
// ubar and (u - u0)/2 are interpolated directly (interpolation is linear);
// the periodic F0 terms belong to ubar only.
template <class ACC>
using GtCompact = Compact<Jet<3, ACC>>;

template <int QC>
struct GtTile {
    static constexpr int NS = 10;
    static constexpr int KC = QC * NS;               // contraction depth per chunk
    static constexpr int KCP = (KC + 3) / 4 * 4;     // padded to the MFMA k of 4
    static constexpr int NTS = KCP * 192;            // T strip
};

template <int J, int I0, int I1, class ACC>
__device__ __forceinline__ void gt_task(const GtCompact<ACC> *sCmp, const ACC *sphi, const int *sQ, const double *sw,
                                        ACC *sT, const Coef &c, ACC mass, int ql, int B, int qp, int nqp)
{
    ACC out[J + 1][10];
    if (qp < nqp) {
        ACC n[10];
        int q = sQ[qp];
        basis_slots<ACC>(sphi, q & 15, (q >> 4) & 15, (q >> 8) & 15, B, n);
        tangent_apply<Jet<3, ACC>, J, I0, I1, ACC>(sCmp[qp], c, n, mass, (ACC)sw[qp], out);
    } else {
#pragma unroll
        for (int i = I0; i <= I1; i++)
#pragma unroll
            for (int s = 0; s < 10; s++) out[i][s] = (ACC)0;
    }
#pragma unroll
    for (int i = I0; i <= I1; i++) {
        ACC *col = sT + block_off(i) + (J - i) * 27 + B;
#pragma unroll
        for (int s = 0; s < 10; s++) col[(ql * 10 + s) * 192] = out[i][s];
    }
}

template <class ACC, int QC>
__device__ __forceinline__ void tangent_gt_body(const MeshArgs &m, const double *__restrict__ Uh, const double *__restrict__ coef,
                                                double mass_d, const long long *__restrict__ rowptr,
                                                const unsigned char *__restrict__ dmask, void *__restrict__ vals, int vals_f32,
                                                const int ev[3])
{
    using G = GtTile<QC>;
    using V = typename V4<ACC>::t;
    constexpr int CMP_BYTES = 64 * (int)sizeof(GtCompact<ACC>);
    constexpr int OP_BYTES0 = G::NTS * (int)sizeof(ACC);
    constexpr int OP_BYTES = OP_BYTES0 > SF_SCRATCH * 8 ? OP_BYTES0 : SF_SCRATCH * 8;
    constexpr int KE_BYTES = 81 * 81 * (int)sizeof(ACC);
    constexpr int REGION = (CMP_BYTES + OP_BYTES) > KE_BYTES ? (CMP_BYTES + OP_BYTES) : KE_BYTES;
    __shared__ __attribute__((aligned(16))) unsigned char sRegion[REGION];
    __shared__ double sU[2 * 81];  // ubar, (u - u0)/2
    __shared__ double sphi[3 * 4 * 9];
    __shared__ ACC sphiA[3 * 4 * 9];
    __shared__ double sw[64];
    __shared__ int sQ[64];
    __shared__ long long sRowBase[81];
    __shared__ int sRowSx[81], sRowSy[81];
    __shared__ unsigned char sMask[81];
    GtCompact<ACC> *sCmp = reinterpret_cast<GtCompact<ACC> *>(sRegion);
    ACC *sT = reinterpret_cast<ACC *>(sRegion + CMP_BYTES);

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const long long e = ((long long)ev[0] * m.ne[1] + ev[1]) * m.ne[2] + ev[2];
    const int nq = m.nq, nqp = nq * nq * nq;
    const ACC mass = (ACC)mass_d;

    if (tid < 81) {
        int A = tid / 3, i = tid % 3;
        int a0 = ev[0] + A / 9, a1 = ev[1] + (A / 3) % 3, a2 = ev[2] + A % 3;
        long long grow = 3 * node_of(m, a0, a1, a2) + i;
        // u = slice 0 (t_{n+1}), u1 = slice 1 (t_n), u0 = slice 2 (t_{n-1})
        const double u = Uh[grow], u1 = Uh[m.ndof + grow], u0 = Uh[2 * m.ndof + grow];
        sU[tid] = (u0 + u1) / 2.0;
        sU[81 + tid] = (u - u0) / 2.0;
        int lo0 = max(0, a0 - 2), lo1 = max(0, a1 - 2), lo2 = max(0, a2 - 2);
        int w1 = min(m.nb[1] - 1, a1 + 2) - lo1 + 1, w2 = min(m.nb[2] - 1, a2 + 2) - lo2 + 1;
        sRowBase[tid] = rowptr[grow] + 3LL * (((long long)(ev[0] - lo0) * w1 + (ev[1] - lo1)) * w2 + (ev[2] - lo2));
        sRowSx[tid] = 3 * w1 * w2;
        sRowSy[tid] = 3 * w2;
        sMask[tid] = dmask[grow];
    } else if (tid >= 128 && tid < 192) {
        int qp = tid - 128;
        sQ[qp] = qp < nqp ? ((qp / (nq * nq)) | (((qp / nq) % nq) << 4) | ((qp % nq) << 8)) : 0;
    }
    // per-element constants issued with the other loads: one memory round trip before the barrier
    const Coef c = load_coef(coef, e);
    const ElemGeo geo = elem_geo(m, ev);
    fill_phi(m, ev, sphi, tid, 256);
    __syncthreads();

    double *sScr = reinterpret_cast<double *>(sRegion + CMP_BYTES);
    double ub[30], ud[30];
    int qx = 0, qy = 0, qz = 0, qi = 0;
    if (tid < nqp) {
        int q = sQ[tid];
        qx = q & 15; qy = (q >> 4) & 15; qz = (q >> 8) & 15; qi = qx * 16 + qy * 4 + qz;
    }
    interp_slots_sf<256>(nq, sU, sphi, sScr, tid);
    if (tid < nqp) {
#pragma unroll
        for (int sl = 0; sl < 30; sl++) ub[sl] = sScr[324 + 864 + sl * 64 + qi];
    }
    __syncthreads();
    interp_slots_sf<256>(nq, sU + 81, sphi, sScr, tid);
    if (tid < nqp) {
#pragma unroll
        for (int sl = 0; sl < 30; sl++) ud[sl] = sScr[324 + 864 + sl * 64 + qi];
        double w;
        point_affine(m, geo, qx, qy, qz, ub, &w);
        Jet<3> uj[30];
#pragma unroll
        for (int sl = 0; sl < 30; sl++) uj[sl] = Num<Jet<3>>::lin(ub[sl], ud[sl]);
        Compact<Jet<3>> cp;
        compact_of<Jet<3>>(uj, c, cp);
        const double *src = reinterpret_cast<const double *>(&cp);
        ACC *dst = reinterpret_cast<ACC *>(&sCmp[tid]);
#pragma unroll
        for (int f = 0; f < 144; f++) dst[f] = (ACC)src[f];
        sw[tid] = w;
    } else if (tid >= 128 && tid < 128 + 3 * 4 * 9) {
        sphiA[tid - 128] = (ACC)sphi[tid - 128];
    }
    __syncthreads();
    for (int idx = tid; idx < (G::KCP - G::KC) * 192; idx += 256) sT[G::KC * 192 + idx] = (ACC)0;

    V acc[6];
#pragma unroll
    for (int t = 0; t < 6; t++)
#pragma unroll
        for (int r = 0; r < 4; r++) acc[t][r] = (ACC)0;
    double accd[6][4];
#pragma unroll
    for (int t = 0; t < 6; t++)
#pragma unroll
        for (int r = 0; r < 4; r++) accd[t][r] = 0.0;

    const int A0 = lane & 15, A1 = A0 + 16, krow = lane >> 4;
    const int ox0 = (A0 / 9) * 3, oy0 = ((A0 / 3) % 3) * 3, oz0 = (A0 % 3) * 3;
    const int ox1 = (A1 < 27 ? A1 / 9 : 0) * 3, oy1 = (A1 < 27 ? (A1 / 3) % 3 : 0) * 3, oz1 = (A1 < 27 ? A1 % 3 : 0) * 3;
    const ACC a1mask = A1 < 27 ? (ACC)1 : (ACC)0;

    const int nchunk = (nqp + QC - 1) / QC;
    for (int ch = 0; ch < nchunk; ch++) {
        if (QC * 27 > 64) {
            // two lane-rounds of tasks: waves 0-1 take dof block j = 2 (rows
            // 0-2, one preamble) for the first / second round, waves 2-3 take
            // j = 1 and j = 0 (rows 0-1 and 0) for the first / second round
            const int t = ((wave & 1) ? 64 : 0) + lane;
            if (t < QC * 27) {
                const int ql = t / 27, B = t % 27, qp = ch * QC + ql;
                if (wave < 2) {
                    gt_task<2, 0, 2, ACC>(sCmp, sphiA, sQ, sw, sT, c, mass, ql, B, qp, nqp);
                } else {
                    gt_task<1, 0, 1, ACC>(sCmp, sphiA, sQ, sw, sT, c, mass, ql, B, qp, nqp);
                    gt_task<0, 0, 0, ACC>(sCmp, sphiA, sQ, sw, sT, c, mass, ql, B, qp, nqp);
                }
            }
        } else {
            for (int t = lane; t < QC * 27; t += 64) {
                int ql = t / 27, B = t % 27, qp = ch * QC + ql;
                if (wave == 0) gt_task<2, 1, 2, ACC>(sCmp, sphiA, sQ, sw, sT, c, mass, ql, B, qp, nqp);
                else if (wave == 1) gt_task<1, 0, 1, ACC>(sCmp, sphiA, sQ, sw, sT, c, mass, ql, B, qp, nqp);
                else if (wave == 2) gt_task<0, 0, 0, ACC>(sCmp, sphiA, sQ, sw, sT, c, mass, ql, B, qp, nqp);
                else gt_task<2, 0, 0, ACC>(sCmp, sphiA, sQ, sw, sT, c, mass, ql, B, qp, nqp);
            }
        }
        __syncthreads();
        const int ct0 = 3 * wave;
        constexpr int NKS = G::KCP / 4;
        ACC fa0[NKS], fa1[NKS], fb[NKS][3];
#pragma unroll
        for (int ks = 0; ks < NKS; ks++) {
            const int k = 4 * ks + krow;
            const int ql = k / 10, sl = k % 10, qp = ch * QC + ql;
            const bool ok = k < G::KC && qp < nqp;
            const int q = sQ[ok ? qp : 0];
            const ACC *px = sphiA + (q & 15) * 9 + slot_d(sl, 0);
            const ACC *py = sphiA + (4 + ((q >> 4) & 15)) * 9 + slot_d(sl, 1);
            const ACC *pz = sphiA + (8 + ((q >> 8) & 15)) * 9 + slot_d(sl, 2);
            const ACC okm = ok ? (ACC)1 : (ACC)0;
            fa0[ks] = okm * (px[ox0] * py[oy0] * pz[oz0]);
            fa1[ks] = (okm * a1mask) * (px[ox1] * py[oy1] * pz[oz1]);
#pragma unroll
            for (int t = 0; t < 3; t++) fb[ks][t] = sT[k * 192 + (ct0 + t) * 16 + (lane & 15)];
        }
#pragma unroll
        for (int ks = 0; ks < NKS; ks++)
#pragma unroll
            for (int t = 0; t < 3; t++) {
                acc[2 * t] = mfma16(fa0[ks], fb[ks][t], acc[2 * t]);
                acc[2 * t + 1] = mfma16(fa1[ks], fb[ks][t], acc[2 * t + 1]);
            }
        if (sizeof(ACC) == 4) {
#pragma unroll
            for (int t = 0; t < 6; t++)
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    accd[t][r] += (double)acc[t][r];
                    acc[t][r] = (ACC)0;
                }
        }
        __syncthreads();
    }
    if (sizeof(ACC) == 4) {
#pragma unroll
        for (int t = 0; t < 6; t++)
#pragma unroll
            for (int r = 0; r < 4; r++) acc[t][r] = (ACC)accd[t][r];
    }

    ACC *sKe = reinterpret_cast<ACC *>(sRegion);
    scatter_staged<ACC>(vals_f32, vals, sKe, sRowBase, sRowSx, sRowSy, sMask, tid, [&]() {
#pragma unroll
        for (int t = 0; t < 6; t++) {
            int ct = 3 * wave + t / 2, at = t & 1;
            int i = ctile_block(ct);
            int cl = ct * 16 + (lane & 15) - block_off(i);
            if (cl >= (3 - i) * 27) continue;
            int j = i + cl / 27, B = cl % 27;
#pragma unroll
            for (int r = 0; r < 4; r++) {
                int A = at * 16 + acc_row((ACC)0, lane, r);
                if (A >= 27) continue;
                sKe[(3 * A + i) * 81 + 3 * B + j] = acc[t][r];
                if (i != j) sKe[(3 * B + j) * 81 + 3 * A + i] = acc[t][r];
            }
        }
        __syncthreads();
    });
}

// Gauss points per chunk of the getime tangent
template <class ACC>
struct GtQC { static constexpr int v = 2; };
template <>
struct GtQC<float> { static constexpr int v = 4; };
template <class ACC>
struct GtWaves { static constexpr int v = 1; };
template <>
struct GtWaves<float> { static constexpr int v = 2; };

template <class ACC, int QC>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(GtWaves<ACC>::v))) void k_tangent_gt(
    MeshArgs m, const double *__restrict__ Uh, const double *__restrict__ coef, double mass,
    const long long *__restrict__ rowptr, const unsigned char *__restrict__ dmask, void *__restrict__ vals, int vals_f32,
    int cx, int cy, int cz, int my, int mz)
{
    const int b = blockIdx.x;
    const int ev[3] = {cx + 3 * (b / (mz * my)), cy + 3 * ((b / mz) % my), cz + 3 * (b % mz)};
    tangent_gt_body<ACC, QC>(m, Uh, coef, mass, rowptr, dmask, vals, vals_f32, ev);
}

// ---------------------------------------------------------------- strain-variation tangent
// The same element tangent as tangent_apply, factored through the strain
// variation of each dof:
//   K_ij[A,B] = sum_q Y_i(A)^T (w M) Y_j(B) + delta_ij sum_q n(A)^T (w Gamma) n(B)
// phys/ge: Y_j(B) has 12 components: beta = sym(F_j^T grad n) (normal 00,
// 11, 22, shear 01, 02, 12) and the trace-free part of nu_IK = H_j,IK n_,I
// + F_jI n_,IK over I for each K, in the orthonormal basis (1,-1,0)/sqrt2,
// (1,1,-2)/sqrt6 of trace-free 3-vectors; M = diag(P, 2 e345 I3, 2 le^2 I6)
// with P = d2Psi/dE_II dE_JJ.  Gamma is the stress form on the gradient and
// second-derivative slots: grad n(A)^T S grad n(B) + sum tau_IK (n(A)_,I
// n(B)_,IK + n(B)_,I n(A)_,IK).
// phys/getime: Y = (Y(F0, H0), Y(F1, H1)), 24 components, from the jets
// F(s) = F0 + s F1, H(s) = H0 + s H1 along u(s) = ubar + s (u - u0)/2;
// with the series M(s) = C0 + s C1 + s^2 C2, the series of Y^T M Y
// truncated at s^2 and averaged as (1/2)(c0/2 + c1/3 + c2/4) is the block
// form [[M00, M01], [M01, M11]], M00 = C0/4 + C1/6 + C2/8, M01 = C0/6 + C1/8,
// M11 = C0/8; Gamma takes the averaged S and tau and the mass rho/dt^2 +
// c/(2 dt) on the value slot (phys/getime/tangent.c:4127-5026).
// MFMA: rows R = 27 i + A and columns C = 27 j + B; dofs 0..79 fill five
// 16-row tiles exactly, and the 15 tiles (rt, ct), ct >= rt, are mirrored
// into the lower triangle.  The last dof (A = 26, i = 2) is taken by the
// vector units: its column K[., 80] = sum_q Y(.)^T Z(80) is a dot product
// the lane that builds Y(.) forms with Z(80) of the point, built for every
// Gauss point in one pass after the point data.  The Gamma operands n(A) and
// w Gamma n(B) sit in columns 80 + A / 80 + B of the same strips (12 slot
// rows, the rest of those rows unused), and its 3 tiles of the 27 x 27 block
// are added to the three diagonal blocks.
constexpr int SV_G = 80;    // first Gamma column of the operand strips
constexpr int SV_RS = 112;  // k-row stride of the strips: 80 + 32; 112 = 48 mod 64 puts the 4 k-rows of a fragment on distinct banks

template <class ACC, int GT>
struct SvPoint;
template <class ACC>
struct SvPoint<ACC, 0> {
    ACC F[3][3], H[3][6];
    ACC P[6];       // w P   (P00 P01 P02 P11 P12 P22)
    ACC S[6];       // w S   (S00 S11 S22 S01 S02 S12)
    ACC tau[3][3];  // w tau
    ACC w;
    ACC zc[3];      // Z = zc Y on the shear, even and odd trace-free components: w 2 e345 s^2, w 2 le^2 s^2
};
// getime works in the basis (u, v) = (Y0 + a Y1, Y1), a = 2/3, of the jet pair:
// the averaged form's 2 x 2 factor [[1/4, 1/6], [1/6, 1/8]] on the shear and
// trace-free components becomes diag(1/4, 1/72), so Z of those components is one
// product each.  F, H hold F0 + a F1, H0 + a H1 (u = Y(F, H) by linearity).
constexpr double SV_UV = 2.0 / 3.0;
template <class ACC>
struct SvPoint<ACC, 1> {
    ACC F[3][3], H[3][6], F1[3][3], H1[3][6];
    ACC P00[6], P01[6], P11[6];  // w M_ab on the normal strains, in the (u, v) basis
    ACC zc[6];                   // Z = zc Y on the shear, even and odd trace-free components, u and v
    ACC S[6], tau[3][3];          // w x averaged series
    ACC w;
};

// the 12 strain-variation components of dof j (row F_j., H_j.) along slots
// n, each without its constant factor s_k (1/2 for the shear strains, 1/sqrt2
// and 1/sqrt6 for the trace-free basis): K = sum_k Y_k (m_k Y_k) needs s_k^2
// only, which the Z coefficients carry (SV_SH, SV_T2, SV_T6)
template <class R>
__device__ __forceinline__ void sv_y(const R F[3], const R H[6], const R n[10], R *y)
{
    y[0] = F[0] * n[1];
    y[1] = F[1] * n[2];
    y[2] = F[2] * n[3];
    y[3] = F[1] * n[1] + F[0] * n[2];
    y[4] = F[2] * n[1] + F[0] * n[3];
    y[5] = F[2] * n[2] + F[1] * n[3];
#pragma unroll
    for (int K = 0; K < 3; K++) {
        R nu[3];
#pragma unroll
        for (int I = 0; I < 3; I++) nu[I] = H[hidx(I, K)] * n[1 + I] + F[I] * n[4 + hidx(I, K)];
        y[6 + 2 * K] = nu[0] - nu[1];
        y[7 + 2 * K] = nu[0] + nu[1] - R(2) * nu[2];
    }
}
constexpr double SV_SH = 0.25, SV_T2 = 0.5, SV_T6 = 1.0 / 6.0;  // s_k^2 of the scaled components
// s_k^2 of component k >= 3
__host__ __device__ constexpr double sv_s2(int k) { return k < 6 ? SV_SH : ((k & 1) ? SV_T6 : SV_T2); }

template <class R>
__device__ __forceinline__ void sym3_apply(const R *P, const R *y, R *z)
{
    z[0] = P[0] * y[0] + P[1] * y[1] + P[2] * y[2];
    z[1] = P[1] * y[0] + P[3] * y[1] + P[4] * y[2];
    z[2] = P[2] * y[0] + P[4] * y[1] + P[5] * y[2];
}
template <class R>
__device__ __forceinline__ void sym3_apply_add(const R *P, const R *y, R *z)
{
    z[0] += P[0] * y[0] + P[1] * y[1] + P[2] * y[2];
    z[1] += P[1] * y[0] + P[3] * y[1] + P[4] * y[2];
    z[2] += P[2] * y[0] + P[4] * y[1] + P[5] * y[2];
}

// Y_j(B) and Z = w M Y_j(B) of one (point, B, j) in registers
template <class ACC>
__device__ __forceinline__ void sv_yz(const SvPoint<ACC, 0> &p, const ACC n[10], int j, ACC Y[12], ACC Z[12])
{
    sv_y<ACC>(p.F[j], p.H[j], n, Y);
    sym3_apply(p.P, Y, Z);
#pragma unroll
    for (int k = 3; k < 12; k++) Z[k] = p.zc[k < 6 ? 0 : 1 + (k & 1)] * Y[k];
}
template <class ACC>
__device__ __forceinline__ void sv_yz(const SvPoint<ACC, 1> &p, const ACC n[10], int j, ACC Y[24], ACC Z[24])
{
    sv_y<ACC>(p.F[j], p.H[j], n, Y);
    sv_y<ACC>(p.F1[j], p.H1[j], n, Y + 12);
    sym3_apply(p.P00, Y, Z);
    sym3_apply_add(p.P01, Y + 12, Z);
    sym3_apply(p.P01, Y, Z + 12);
    sym3_apply_add(p.P11, Y + 12, Z + 12);
#pragma unroll
    for (int k = 3; k < 12; k++) {
        const int c = k < 6 ? 0 : 1 + (k & 1);
        Z[k] = p.zc[2 * c] * Y[k];
        Z[12 + k] = p.zc[2 * c + 1] * Y[12 + k];
    }
}

// Y_j(B), Z_j(B) into k-rows 0..NY-1 of the strips at the column y, z point
// to; returns Y_j(B) . z80, the dof pair's entry K[(B, j), 80] at this point
template <class ACC, int GT>
__device__ __forceinline__ ACC sv_item_yz(const SvPoint<ACC, GT> &p, const ACC n[10], int j, ACC *y, ACC *z,
                                          const ACC *z80)
{
    constexpr int NY = GT ? 24 : 12;
    ACC Y[NY], Z[NY];
    sv_yz<ACC>(p, n, j, Y, Z);
    ACC d = ACC(0);
#pragma unroll
    for (int k = 0; k < NY; k++) {
        y[k * SV_RS] = Y[k];
        z[k * SV_RS] = Z[k];
        d += Y[k] * z80[k];
    }
    return d;
}

// n(B) and w Gamma n(B) on the 12 Gamma rows (ge: slots 1..9, getime: 0..9,
// the rest zero) at column B (ga, gb point there); wmass = w x mass
template <class ACC, int GT>
__device__ __forceinline__ void sv_item_gamma(const SvPoint<ACC, GT> &p, ACC wmass, const ACC n[10], ACC *ga, ACC *gb)
{
    const ACC *S = p.S;
    const ACC(&t)[3][3] = p.tau;
    ACC g[10];
    g[0] = wmass * n[0];
    g[1] = S[0] * n[1] + S[3] * n[2] + S[4] * n[3] + t[0][0] * n[4] + t[0][1] * n[5] + t[0][2] * n[6];
    g[2] = S[3] * n[1] + S[1] * n[2] + S[5] * n[3] + t[1][0] * n[5] + t[1][1] * n[7] + t[1][2] * n[8];
    g[3] = S[4] * n[1] + S[5] * n[2] + S[2] * n[3] + t[2][0] * n[6] + t[2][1] * n[8] + t[2][2] * n[9];
    g[4] = t[0][0] * n[1];
    g[5] = t[0][1] * n[1] + t[1][0] * n[2];
    g[6] = t[0][2] * n[1] + t[2][0] * n[3];
    g[7] = t[1][1] * n[2];
    g[8] = t[1][2] * n[2] + t[2][1] * n[3];
    g[9] = t[2][2] * n[3];
#pragma unroll
    for (int k = 0; k < 12; k++) {
        const int sl = GT ? k : k + 1;
        ga[k * SV_RS] = sl < 10 ? n[sl] : ACC(0);
        gb[k * SV_RS] = sl < 10 ? g[sl] : ACC(0);
    }
}

// Tile paths.  Waves 0-2 build Y and Z in the item phase, wave 3 the Gamma
// operands, and each runs its own copy of the chunk loop and staging over its
// tile path PL, with accumulators for its own slots only (no empty slots, no
// per-wave Gamma masks; fewer registers than one loop for every wave).  Slot t
// multiplies A fragment ai(t) by B fragment bi(t); the wave's fragment
// columns are ao[] / bo[] of its SvTiles (strip columns; 80, 96: the Gamma
// block).  The last NG slots are Gamma tiles, over the 12 Gamma k-rows only,
// as are the last NGA A and NGB B fragments.  A wave's largest MFMA share is
// as small as the tiles allow (the item and MFMA phases are separated by
// barriers, so a chunk lasts as long as the slowest wave's items plus the
// slowest wave's MFMAs):
//   ge:     wave 0: (0,0) (0,1) (0,2) (0,3) (0,4)
//           wave 1: (1,1) (1,2) (1,3) (1,4) (4,4)
//           wave 2: (2,2) (2,3) (2,4) (3,3) (3,4)
//           wave 3: (G0,G0) (G0,G1) (G1,G1)               15 MFMAs per point (wave 3: 9)
//   getime: wave 0: (0,0) (0,1) (0,2) (0,3) | (G0,G0)
//           wave 1: (0,4) (1,1) (1,2) (1,3) | (G0,G1)
//           wave 2: (1,4) (2,2) (2,3) (2,4) | (G1,G1)
//           wave 3: (4,4) (3,3) (3,4)                     4 x 6 + 3 = 27 MFMAs per point (wave 3: 18)
// (Fragments repeated within a wave are read again: skipping the repeats with
// wave-uniform branches split the k-step loop and cost 15 %.)
template <bool W3_, int NS_, int NA_, int NB_, int NG_, int NGA_, int NGB_>
struct SvPathDims {
    static constexpr bool W3 = W3_;  // wave 3's path
    static constexpr int NS = NS_, NA = NA_, NB = NB_, NG = NG_, NGA = NGA_, NGB = NGB_;
};
template <int GT>
struct SvPathW;  // waves 0-2
template <int GT>
struct SvPath3;  // wave 3
template <>
struct SvPathW<0> : SvPathDims<false, 5, 3, 5, 0, 0, 0> {
    __device__ static constexpr int ai(int t) { return t < 3 ? 0 : t - 2; }
    __device__ static constexpr int bi(int t) { return t; }
};
template <>
struct SvPath3<0> : SvPathDims<true, 3, 2, 2, 3, 2, 2> {
    __device__ static constexpr int ai(int t) { return t < 2 ? 0 : 1; }
    __device__ static constexpr int bi(int t) { return t == 0 ? 0 : 1; }
};
template <>
struct SvPathW<1> : SvPathDims<false, 5, 3, 5, 1, 1, 1> {
    __device__ static constexpr int ai(int t) { return t == 0 ? 0 : (t < 4 ? 1 : 2); }
    __device__ static constexpr int bi(int t) { return t; }
};
template <>
struct SvPath3<1> : SvPathDims<true, 3, 2, 3, 0, 0, 0> {
    __device__ static constexpr int ai(int t) { return t == 0 ? 0 : 1; }
    __device__ static constexpr int bi(int t) { return t; }
};
struct SvTiles {
    int ao[3], bo[5];
};
template <int GT>
__device__ __forceinline__ SvTiles sv_tiles(int wave)
{
    constexpr int G0 = SV_G, G1 = SV_G + 16;
    if (GT == 0) {
        if (wave == 0) return SvTiles{{0, 0, 0}, {0, 16, 32, 48, 64}};
        if (wave == 1) return SvTiles{{16, 16, 64}, {16, 32, 48, 64, 64}};
        if (wave == 2) return SvTiles{{32, 48, 48}, {32, 48, 64, 48, 64}};
        return SvTiles{{G0, G1, 0}, {G0, G1, 0, 0, 0}};
    }
    if (wave == 0) return SvTiles{{0, 0, G0}, {0, 16, 32, 48, G0}};
    if (wave == 1) return SvTiles{{0, 16, G0}, {64, 16, 32, 48, G1}};
    if (wave == 2) return SvTiles{{16, 32, G1}, {64, 32, 48, 64, G1}};
    return SvTiles{{64, 48, 0}, {64, 48, 64, 0, 0}};
}

// the chunk's qcnt points, the NY k-rows of each (y: A strip, z: B strip);
// Gamma fragments and slots only over the first 12 k-rows
template <class ACC, int NY, class PL>
__device__ __forceinline__ void sv_mfma_chunk(const ACC *y, const ACC *z, int qcnt, int lane, const SvTiles &T,
                                              typename V4<ACC>::t *acc)
{
    const int off = (lane >> 4) * SV_RS + (lane & 15);
    const ACC *ya = y + off, *zb = z + off;
    for (int ql = 0; ql < qcnt; ql++) {
#pragma unroll
        for (int ks = 0; ks < NY / 4; ks++) {
            const ACC *yk = ya + 4 * ks * SV_RS, *zk = zb + 4 * ks * SV_RS;
            const bool gk = ks < 3;
            ACC a[PL::NA], b[PL::NB];
#pragma unroll
            for (int i = 0; i < PL::NA; i++)
                if (gk || i < PL::NA - PL::NGA) a[i] = yk[T.ao[i]];
#pragma unroll
            for (int i = 0; i < PL::NB; i++)
                if (gk || i < PL::NB - PL::NGB) b[i] = zk[T.bo[i]];
#pragma unroll
            for (int t = 0; t < PL::NS; t++)
                if (gk || t < PL::NS - PL::NG) acc[t] = mfma16(a[PL::ai(t)], b[PL::bi(t)], acc[t]);
        }
        ya += NY * SV_RS;
        zb += NY * SV_RS;
    }
}

// the wave's material tiles into the staged 81 x 81 element matrix (both
// triangles), stored in the tiles' own order: row 27 i + A, column 27 j + B
// (the scatter reads it in that order); GAMMA: its Gamma tiles added to the
// three diagonal blocks (the three Gamma tiles touch disjoint entries, which
// no material tile writes after the barrier between the two calls)
template <class ACC, bool GAMMA, class PL>
__device__ __forceinline__ void sv_stage(ACC *sKe, int lane, const SvTiles &T, const typename V4<ACC>::t *acc)
{
#pragma unroll
    for (int t = 0; t < PL::NS; t++) {
        if ((t >= PL::NS - PL::NG) != GAMMA) continue;
        const int ar = T.ao[PL::ai(t)], bc = T.bo[PL::bi(t)];
        const int C = bc + (lane & 15);
        if (GAMMA) {
            const int B = C - SV_G;
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int A = ar + acc_row((ACC)0, lane, r) - SV_G;
                if (A < 27 && B < 27) {
#pragma unroll
                    for (int d = 0; d < 3; d++) {
                        sKe[(27 * d + A) * 81 + 27 * d + B] += acc[t][r];
                        if (ar != bc) sKe[(27 * d + B) * 81 + 27 * d + A] += acc[t][r];
                    }
                }
            }
            continue;
        }
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int R = ar + acc_row((ACC)0, lane, r);
            sKe[R * 81 + C] = acc[t][r];
            if (ar != bc) sKe[C * 81 + R] = acc[t][r];
        }
    }
}

// NQM: Gauss points per direction the point data are sized for
template <class ACC, int QC, int GT, int NQM = 4>
struct SvDims {
    static constexpr int NY = GT ? 24 : 12;
    static constexpr int YSZ = QC * NY * SV_RS;   // one A (or B) strip of a chunk, elements
    static constexpr int OPB = 2 * YSZ * (int)sizeof(ACC);
    static constexpr int PTB = NQM * NQM * NQM * (int)sizeof(SvPoint<ACC, GT>);
    // prologue use of the operand part: interpolation scratch, U, basis tables
    static constexpr int PRO = (SF_SCRATCH + 2 * 81 + 3 * 4 * 9) * 8;
    static constexpr int PART2 = OPB > PRO ? OPB : PRO;
    static constexpr int KEB = (81 * 81 * (int)sizeof(ACC) + 15) / 16 * 16;
    // per (dof block j < 3, item (point, B)) sums of the dof-80 column, FP64,
    // at the end of the region (the element matrix takes the start)
    static constexpr int S80B = 3 * 27 * QC * 8;
    static constexpr int REGION = (PTB + PART2) > KEB + S80B ? (PTB + PART2) : KEB + S80B;
    static constexpr int S80OFF = REGION - S80B;
};

template <class ACC, int QC, int GT, int NQM = 4>
__device__ __forceinline__ void tangent_sv_body(const MeshArgs &m, const double *__restrict__ Uh, const double *__restrict__ coef,
                                                double mass_d, const long long *__restrict__ rowptr,
                                                const unsigned char *__restrict__ dmask, void *__restrict__ vals, int vals_f32,
                                                const int ev[3])
{
    using D = SvDims<ACC, QC, GT, NQM>;
    using Pt = SvPoint<ACC, GT>;
    using V = typename V4<ACC>::t;
    constexpr int NY = D::NY;
    // [point data | operand strips]: the interpolation scratch of the
    // prologue sits in the operand part, the element matrix of the epilogue
    // at the start
    __shared__ __attribute__((aligned(16))) unsigned char sRegion[D::REGION];
    __shared__ ACC sphiA[3 * 4 * 9];
    __shared__ ACC sZ80[NQM * NQM * NQM][NY];  // Z of dof 80 per Gauss point
    // FP32: the slot values n(B) of the chunk's (point, B) items, built by
    // wave 3 for the next chunk after its MFMAs (it has the fewest) and read by
    // every item lane (FP64 has no room left for them at two work-groups per CU)
    constexpr bool NPRE = sizeof(ACC) == 4;
    constexpr int NW = (27 * QC + 63) / 64 * 64;  // item columns, one per lane
    __shared__ ACC sN[NPRE ? 10 : 1][NPRE ? NW : 1];
    __shared__ int sQ[64];
    __shared__ long long sRowBase[81];
    __shared__ int sRowSx[81], sRowSy[81];
    __shared__ unsigned char sMask[81];
    Pt *sPt = reinterpret_cast<Pt *>(sRegion);
    ACC *sY = reinterpret_cast<ACC *>(sRegion + D::PTB);
    ACC *sZ = sY + D::YSZ;
    // prologue only: ge U / getime (ubar, (u - u0)/2), and the basis tables, past the interpolation scratch
    double *sU = reinterpret_cast<double *>(sRegion + D::PTB) + SF_SCRATCH;
    double *sphi = sU + 2 * 81;

    const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const long long e = ((long long)ev[0] * m.ne[1] + ev[1]) * m.ne[2] + ev[2];
    const int nq = m.nq, nqp = nq * nq * nq;

    if (tid < 81) {
        int A = tid / 3, i = tid % 3;
        int a0 = ev[0] + A / 9, a1 = ev[1] + (A / 3) % 3, a2 = ev[2] + A % 3;
        long long grow = 3 * node_of(m, a0, a1, a2) + i;
        if constexpr (GT) {
            // u = slice 0 (t_{n+1}), u1 = slice 1 (t_n), u0 = slice 2 (t_{n-1})
            const double u = Uh[grow], u1 = Uh[m.ndof + grow], u0 = Uh[2 * m.ndof + grow];
            sU[tid] = (u0 + u1) / 2.0;
            sU[81 + tid] = (u - u0) / 2.0;
        } else {
            sU[tid] = Uh[grow];
        }
        int lo0 = max(0, a0 - 2), lo1 = max(0, a1 - 2), lo2 = max(0, a2 - 2);
        int w1 = min(m.nb[1] - 1, a1 + 2) - lo1 + 1, w2 = min(m.nb[2] - 1, a2 + 2) - lo2 + 1;
        sRowBase[tid] = rowptr[grow] + 3LL * (((long long)(ev[0] - lo0) * w1 + (ev[1] - lo1)) * w2 + (ev[2] - lo2));
        sRowSx[tid] = 3 * w1 * w2;
        sRowSy[tid] = 3 * w2;
        sMask[tid] = dmask[grow];
    } else if (tid >= 128 && tid < 192) {
        int qp = tid - 128;
        sQ[qp] = qp < nqp ? ((qp / (nq * nq)) | (((qp / nq) % nq) << 4) | ((qp % nq) << 8)) : 0;
    }
    // per-element constants issued with the other loads: one memory round trip before the barrier
    const Coef c = load_coef(coef, e);
    const ElemGeo geo = elem_geo(m, ev);
    fill_phi(m, ev, sphi, tid, 256);
    __syncthreads();

    // ---- per-point data
    double *sScr = reinterpret_cast<double *>(sRegion + D::PTB);
    int qx = 0, qy = 0, qz = 0, qi = 0;
    if (tid < nqp) {
        const int q = sQ[tid];
        qx = q & 15; qy = (q >> 4) & 15; qz = (q >> 8) & 15; qi = qx * 16 + qy * 4 + qz;
    }
    double ub[30];
    interp_slots_sf<256>(nq, sU, sphi, sScr, tid);
    if (tid < nqp) {
#pragma unroll
        for (int sl = 0; sl < 30; sl++) ub[sl] = sScr[324 + 864 + sl * 64 + qi];
    }
    if constexpr (GT) {
        __syncthreads();
        interp_slots_sf<256>(nq, sU + 81, sphi, sScr, tid);
    }
    if (tid < nqp) {
        double w;
        point_affine(m, geo, qx, qy, qz, ub, &w);
        Pt &P = sPt[tid];
        if constexpr (GT) {
            Jet<3> uj[30];
#pragma unroll
            for (int sl = 0; sl < 30; sl++) uj[sl] = Num<Jet<3>>::lin(ub[sl], sScr[324 + 864 + sl * 64 + qi]);
            Compact<Jet<3>> cp;
            compact_of<Jet<3>>(uj, c, cp);
            SvPoint<ACC, 1> &G = reinterpret_cast<SvPoint<ACC, 1> &>(P);
            for (int a = 0; a < 3; a++) {
                for (int J = 0; J < 3; J++) {
                    G.F[a][J] = (ACC)(cp.F[a][J].c[0] + SV_UV * cp.F[a][J].c[1]);
                    G.F1[a][J] = (ACC)cp.F[a][J].c[1];
                    G.tau[a][J] = (ACC)(w * (cp.tau[a][J].c[0] / 4.0 + cp.tau[a][J].c[1] / 6.0 + cp.tau[a][J].c[2] / 8.0));
                }
                for (int h = 0; h < 6; h++) {
                    G.H[a][h] = (ACC)(cp.H[a][h].c[0] + SV_UV * cp.H[a][h].c[1]);
                    G.H1[a][h] = (ACC)cp.H[a][h].c[1];
                }
            }
            for (int k = 0; k < 6; k++) {
                const double *pc = cp.P[k].c, *sc = cp.S[k].c;
                // the block form in the basis (u, v) = (Y0 + a Y1, Y1), a = SV_UV:
                // M00, M01 - a M00, M11 - 2 a M01 + a^2 M00
                const double m00 = pc[0] / 4.0 + pc[1] / 6.0 + pc[2] / 8.0, m01 = pc[0] / 6.0 + pc[1] / 8.0, m11 = pc[0] / 8.0;
                G.P00[k] = (ACC)(w * m00);
                G.P01[k] = (ACC)(w * (m01 - SV_UV * m00));
                G.P11[k] = (ACC)(w * (m11 - 2.0 * SV_UV * m01 + SV_UV * SV_UV * m00));
                G.S[k] = (ACC)(w * (sc[0] / 4.0 + sc[1] / 6.0 + sc[2] / 8.0));
            }
            G.w = (ACC)w;
            // shear (2 e345, 1/4), even (2 le^2, 1/2) and odd (2 le^2, 1/6) trace-free
            // components; u: x 1/4, v: x 1/72
            for (int cc = 0; cc < 3; cc++) {
                const double cw = w * (cc ? 2.0 * c.lsq : 2.0 * c.e345) * sv_s2(cc == 0 ? 3 : 5 + cc);
                G.zc[2 * cc] = (ACC)(cw / 4.0);
                G.zc[2 * cc + 1] = (ACC)(cw / 72.0);
            }
        } else {
            Compact<double> cp;
            compact_of<double>(ub, c, cp);
            SvPoint<ACC, 0> &G = reinterpret_cast<SvPoint<ACC, 0> &>(P);
            for (int a = 0; a < 3; a++) {
                for (int J = 0; J < 3; J++) {
                    G.F[a][J] = (ACC)cp.F[a][J];
                    G.tau[a][J] = (ACC)(w * cp.tau[a][J]);
                }
                for (int h = 0; h < 6; h++) G.H[a][h] = (ACC)cp.H[a][h];
            }
            for (int k = 0; k < 6; k++) {
                G.P[k] = (ACC)(w * cp.P[k]);
                G.S[k] = (ACC)(w * cp.S[k]);
            }
            G.w = (ACC)w;
            for (int cc = 0; cc < 3; cc++) G.zc[cc] = (ACC)(w * (cc ? 2.0 * c.lsq : 2.0 * c.e345) * sv_s2(cc == 0 ? 3 : 5 + cc));
        }
    } else if (tid >= 128 && tid < 128 + 3 * 4 * 9) {
        sphiA[tid - 128] = (ACC)sphi[tid - 128];
    }
    __syncthreads();
    const ACC mass = (ACC)mass_d;
    // n(B) of the items of chunk ch into sN (wave 3)
    auto fill_n = [&](int ch) {
        for (int t = lane; t < 27 * min(QC, nqp - ch * QC); t += 64) {
            const int ql = t / 27, B = t - 27 * ql, q = sQ[ch * QC + ql];
            ACC n[10];
            basis_slots<ACC>(sphiA, q & 15, (q >> 4) & 15, (q >> 8) & 15, B, n);
#pragma unroll
            for (int k = 0; k < 10; k++) sN[k][t] = n[k];
        }
    };
    // n(B) of item t of chunk ch
    auto slots = [&](int ch, int t, int B, ACC n[10]) {
        if constexpr (NPRE) {
#pragma unroll
            for (int k = 0; k < 10; k++) n[k] = sN[k][t];
        } else {
            const int q = sQ[ch * QC + t / 27];
            basis_slots<ACC>(sphiA, q & 15, (q >> 4) & 15, (q >> 8) & 15, B, n);
        }
    };
    // Z(80) of every Gauss point, one lane per point; the first chunk's n(B)
    if (tid < nqp) {
        const int q = sQ[tid];
        ACC n[10], Y[NY], Z[NY];
        basis_slots<ACC>(sphiA, q & 15, (q >> 4) & 15, (q >> 8) & 15, 26, n);
        sv_yz<ACC>(sPt[tid], n, 2, Y, Z);
#pragma unroll
        for (int k = 0; k < NY; k++) sZ80[tid][k] = Z[k];
    } else if (NPRE && wave == 3) {
        fill_n(0);
    }
    __syncthreads();

    const SvTiles T = sv_tiles<GT>(wave);
    constexpr int NSLOT = (27 * QC + 63) / 64;  // items per lane (waves 0-2)
    const int nchunk = (nqp + QC - 1) / QC;

    // FP32 values: the element's CSR values are read now and arrive during the
    // chunk loop (27 registers); FP64 values are read before the staging
    ScatterCol sc = scatter_col(tid, sMask);
    sc.kb = 27 * 81 * sc.r0 + 27 * (sc.cc % 3) + sc.cc / 3;  // row 27 i + A, column 27 j + B
    sc.ks = 81;
    float oldf[27];
    double old[27];
    if (vals_f32) scatter_load<float>(sc, sRowBase, sRowSx, sRowSy, sMask, (const float *)vals, oldf);
    ACC *sKe = reinterpret_cast<ACC *>(sRegion);
    double *s80 = reinterpret_cast<double *>(sRegion + D::S80OFF);

    // Waves 0-2 and wave 3 run their own copy of the chunk loop and staging
    // (PL: their tile path); every wave passes the same barriers.
    auto part = [&](auto pl) {
        using PL = decltype(pl);
        constexpr bool G3 = PL::W3;  // wave 3: the Gamma operands
        V acc[PL::NS];
#pragma unroll
        for (int t = 0; t < PL::NS; t++)
#pragma unroll
            for (int r = 0; r < 4; r++) acc[t][r] = (ACC)0;
        double accd[PL::NS][4];
#pragma unroll
        for (int t = 0; t < PL::NS; t++)
#pragma unroll
            for (int r = 0; r < 4; r++) accd[t][r] = 0.0;
        double acc80[NSLOT];
#pragma unroll
        for (int k = 0; k < NSLOT; k++) acc80[k] = 0.0;

        // operands of chunk ch into the strips yb, zb: wave j < 3 builds Y_j,
        // Z_j (lane = (point, B)) and the dof-80 products; wave 3 the Gamma pair
        auto items = [&](int ch, ACC *yb, ACC *zb) {
            const int qcnt = min(QC, nqp - ch * QC);
            if constexpr (!G3) {
#pragma unroll
                for (int sl = 0; sl < NSLOT; sl++) {
                    const int t = lane + 64 * sl;
                    if (t < 27 * qcnt) {
                        const int ql = t / 27, B = t - 27 * ql;
                        ACC n[10];
                        slots(ch, t, B, n);
                        // the last dof's column is the first Gamma column: its Y and Z go to
                        // the last (padding) column, whose products are discarded
                        const int col = wave == 2 && B == 26 ? SV_RS - 1 : 27 * wave + B;
                        acc80[sl] += (double)sv_item_yz<ACC, GT>(sPt[ch * QC + ql], n, wave, yb + ql * NY * SV_RS + col,
                                                                 zb + ql * NY * SV_RS + col, sZ80[ch * QC + ql]);
                    }
                }
            } else {
                for (int t = lane; t < 27 * qcnt; t += 64) {
                    const int ql = t / 27, B = t - 27 * ql, qp = ch * QC + ql;
                    ACC n[10];
                    slots(ch, t, B, n);
                    const Pt &P = sPt[qp];
                    sv_item_gamma<ACC, GT>(P, P.w * mass, n, yb + ql * NY * SV_RS + SV_G + B, zb + ql * NY * SV_RS + SV_G + B);
                }
            }
        };
        auto contract = [&](int ch, const ACC *yb, const ACC *zb) {
            sv_mfma_chunk<ACC, NY, PL>(yb, zb, min(QC, nqp - ch * QC), lane, T, acc);
            // every item lane has read sN before the barrier ahead of this phase
            if (NPRE && G3 && ch + 1 < nchunk) fill_n(ch + 1);
            // FP32: the partial sums of every 16 Gauss points go into FP64
            if (sizeof(ACC) == 4 && (((ch + 1) * QC) % 16 == 0 || ch == nchunk - 1)) {
#pragma unroll
                for (int r = 0; r < 4; r++) {
#pragma unroll
                    for (int t = 0; t < PL::NS; t++) {
                        accd[t][r] += (double)acc[t][r];
                        acc[t][r] = (ACC)0;
                    }
                }
            }
        };
        for (int ch = 0; ch < nchunk; ch++) {
            items(ch, sY, sZ);
            __syncthreads();
            contract(ch, sY, sZ);
            __syncthreads();
        }
        if (sizeof(ACC) == 4) {
#pragma unroll
            for (int r = 0; r < 4; r++) {
#pragma unroll
                for (int t = 0; t < PL::NS; t++) acc[t][r] = (ACC)accd[t][r];
            }
        }
        if (!vals_f32) scatter_load<double>(sc, sRowBase, sRowSx, sRowSy, sMask, (const double *)vals, old);

        // ---- element matrix: material tiles (both triangles) and the dof-80
        // row and column, then Gamma on the diagonal blocks
        if constexpr (!G3) {
#pragma unroll
            for (int sl = 0; sl < NSLOT; sl++) {
                const int t = lane + 64 * sl;
                if (t < 27 * QC) s80[wave * 27 * QC + t] = acc80[sl];
            }
        }
        sv_stage<ACC, false, PL>(sKe, lane, T, acc);
        __syncthreads();
        if (!G3 && tid < 81) {  // dof 27 j + B = tid
            const int j = tid / 27, B = tid - 27 * j;
            double k80 = 0.0;
            for (int ql = 0; ql < QC; ql++) k80 += s80[j * 27 * QC + ql * 27 + B];
            sKe[tid * 81 + 80] = (ACC)k80;
            sKe[80 * 81 + tid] = (ACC)k80;
        }
        __syncthreads();
        if constexpr (PL::NG > 0) sv_stage<ACC, true, PL>(sKe, lane, T, acc);
        __syncthreads();
    };
    if (wave < 3)
        part(SvPathW<GT>{});
    else
        part(SvPath3<GT>{});

    if (vals_f32)
        scatter_store<float, ACC>(sc, sKe, sRowBase, sRowSx, sRowSy, sMask, (float *)vals, oldf);
    else
        scatter_store<double, ACC>(sc, sKe, sRowBase, sRowSx, sRowSy, sMask, (double *)vals, old);
}

// two work-groups per CU (one for getime FP64: 146 KB of LDS with 2 points per chunk)
template <class ACC, int GT>
struct SvWaves { static constexpr int v = 2; };
template <>
struct SvWaves<double, 1> { static constexpr int v = 1; };

template <class ACC, int QC, int GT>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(SvWaves<ACC, GT>::v))) void k_tangent_sv(
    MeshArgs m, const double *__restrict__ Uh, const double *__restrict__ coef, double mass,
    const long long *__restrict__ rowptr, const unsigned char *__restrict__ dmask, void *__restrict__ vals, int vals_f32,
    int cx, int cy, int cz, int my, int mz)
{
    const int b = blockIdx.x;
    const int ev[3] = {cx + 3 * (b / (mz * my)), cy + 3 * ((b / mz) % my), cz + 3 * (b % mz)};
    tangent_sv_body<ACC, QC, GT>(m, Uh, coef, mass, rowptr, dmask, vals, vals_f32, ev);
}

// diagonal of Dirichlet rows = 1 (MatZeroRowsColumns(..., 1.0, ...))
__global__ void k_dirichlet_diag(MeshArgs m, const int *__restrict__ rows, int n, const long long *__restrict__ rowptr,
                                 void *vals, int vals_f32)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    long long r = rows[t];
    long long node = r / 3;
    int i = (int)(r % 3);
    int a2 = (int)(node % m.nb[2]);
    int a1 = (int)((node / m.nb[2]) % m.nb[1]);
    int a0 = (int)(node / ((long long)m.nb[2] * m.nb[1]));
    int lo0 = max(0, a0 - 2), lo1 = max(0, a1 - 2), lo2 = max(0, a2 - 2);
    int w1 = min(m.nb[1] - 1, a1 + 2) - lo1 + 1, w2 = min(m.nb[2] - 1, a2 + 2) - lo2 + 1;
    long long pos = rowptr[r] + 3LL * (((a0 - lo0) * w1 + (a1 - lo1)) * w2 + (a2 - lo2)) + i;
    if (vals_f32)
        ((float *)vals)[pos] = 1.0f;
    else
        ((double *)vals)[pos] = 1.0;
}

// zero fill of n32 32-bit words (a kernel, not a memset node, so that a
// captured graph replays it faithfully)
__global__ void k_fill_zero(unsigned int *p, long long n32)
{
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    long long stride = (long long)gridDim.x * blockDim.x;
    for (; i < n32; i += stride) p[i] = 0u;
}

// ============================================================== pattern
__device__ __forceinline__ int stencil_w(int a, int nb) { return min(nb - 1, a + 2) - max(0, a - 2) + 1; }
__device__ __forceinline__ long long stencil_prefix(int a, int nb)
{
    long long s = 0;
    for (int t = 0; t < a; t++) s += stencil_w(t, nb);
    return s;
}

// rowptr in closed form: rows ordered (a0, a1, a2, i)
__global__ void k_rowptr(MeshArgs m, long long *rowptr)
{
    long long r = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (r > m.ndof) return;
    if (r == m.ndof) {
        rowptr[r] = 9LL * stencil_prefix(m.nb[0], m.nb[0]) * stencil_prefix(m.nb[1], m.nb[1]) * stencil_prefix(m.nb[2], m.nb[2]);
        return;
    }
    long long node = r / 3;
    int i = (int)(r % 3);
    int a2 = (int)(node % m.nb[2]);
    int a1 = (int)((node / m.nb[2]) % m.nb[1]);
    int a0 = (int)(node / ((long long)m.nb[2] * m.nb[1]));
    long long T1 = stencil_prefix(m.nb[1], m.nb[1]), T2 = stencil_prefix(m.nb[2], m.nb[2]);
    long long w0 = stencil_w(a0, m.nb[0]), w1 = stencil_w(a1, m.nb[1]), w2 = stencil_w(a2, m.nb[2]);
    rowptr[r] = 9LL * (stencil_prefix(a0, m.nb[0]) * T1 * T2 + w0 * stencil_prefix(a1, m.nb[1]) * T2 +
                       w0 * w1 * stencil_prefix(a2, m.nb[2])) +
                3LL * i * w0 * w1 * w2;
}

// columns: one thread per (row, b0, b1) run of 3 * w2 ascending columns
__global__ void k_cols(MeshArgs m, const long long *__restrict__ rowptr, int *__restrict__ col)
{
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= m.ndof * 25) return;
    long long r = t / 25;
    int run = (int)(t % 25), db0 = run / 5 - 2, db1 = run % 5 - 2;
    long long node = r / 3;
    int a2 = (int)(node % m.nb[2]);
    int a1 = (int)((node / m.nb[2]) % m.nb[1]);
    int a0 = (int)(node / ((long long)m.nb[2] * m.nb[1]));
    int b0 = a0 + db0, b1 = a1 + db1;
    if (b0 < 0 || b0 >= m.nb[0] || b1 < 0 || b1 >= m.nb[1]) return;
    int lo0 = max(0, a0 - 2), lo1 = max(0, a1 - 2), lo2 = max(0, a2 - 2);
    int w1 = min(m.nb[1] - 1, a1 + 2) - lo1 + 1, w2 = min(m.nb[2] - 1, a2 + 2) - lo2 + 1;
    long long pos = rowptr[r] + 3LL * ((b0 - lo0) * w1 + (b1 - lo1)) * w2;
    for (int b2 = lo2; b2 < lo2 + w2; b2++)
        for (int j = 0; j < 3; j++) col[pos++] = (int)(3 * node_of(m, b0, b1, b2) + j);
}

// ============================================================== density outputs
// iga_intdensity (src/iga_intdensity.c:10-293), volume part: per element the
// weighted sum over its Gauss points of the physics' density outputs (13 for
// phys/ge, 2 for phys/getime), then fixed-order sums of consecutive element
// ranges (k_density_blocks) whose partials the host adds in order.  ge reads
// slice 0 of U; getime reads slices 1 (u1 = t_n) and 2 (u0 = t_{n-1}).
template <int AV>
__global__ __launch_bounds__(64) void k_density_elem(MeshArgs m, const double *__restrict__ U,
                                                     const double *__restrict__ coef, double *__restrict__ Pe)
{
    constexpr int ND = AV ? 2 : 13;
    constexpr int NS = AV ? 2 : 1;
    __shared__ double sU[NS * 81];
    __shared__ double sphi[3 * 4 * 9];
    __shared__ double sred[ND][64];
    const int tid = threadIdx.x;
    const long long e = blockIdx.x;
    int ev[3];
    ev[2] = (int)(e % m.ne[2]);
    ev[1] = (int)((e / m.ne[2]) % m.ne[1]);
    ev[0] = (int)(e / ((long long)m.ne[2] * m.ne[1]));
    for (int idx = tid; idx < NS * 81; idx += 64) {
        int h = idx / 81, r = idx % 81, A = r / 3, i = r % 3;
        sU[idx] = U[(AV ? (h + 1) * m.ndof : 0) + 3 * node_of(m, ev[0] + A / 9, ev[1] + (A / 3) % 3, ev[2] + A % 3) + i];
    }
    fill_phi(m, ev, sphi, tid, 64);
    __syncthreads();
    const int nqp = m.nq * m.nq * m.nq;
    double hh[ND];
    for (int k = 0; k < ND; k++) hh[k] = 0.0;
    if (tid < nqp) {
        int qx = tid / (m.nq * m.nq), qy = (tid / m.nq) % m.nq, qz = tid % m.nq;
        double u[NS][30], w;
        interp_slots<NS>(m, sU, sphi, ev, qx, qy, qz, u, &w);
        const Coef c = load_coef(coef, e);
        if constexpr (AV) {
            getime_density_outputs(u[0], u[NS - 1], c, m.rho, m.dt, hh);
        } else {
            ge_density_outputs(u[0], c, hh);
        }
        for (int k = 0; k < ND; k++) hh[k] *= w;
    }
    for (int k = 0; k < ND; k++) sred[k][tid] = hh[k];
    __syncthreads();
    if (tid < ND) {
        double s = 0.0;
        for (int l = 0; l < nqp; l++) s += sred[tid][l];
        Pe[e * ND + tid] = s;
    }
}

// part[b * nd + k] = sum of Pe[e * nd + k] over the elements of block b's
// range [b ne / nblk, (b + 1) ne / nblk), in element order
__global__ __launch_bounds__(64) void k_density_blocks(const double *__restrict__ Pe, long long ne, int nd,
                                                       double *__restrict__ part)
{
    const long long b = blockIdx.x, nblk = gridDim.x;
    const long long e0 = b * ne / nblk, e1 = (b + 1) * ne / nblk;
    for (int k = threadIdx.x; k < nd; k += blockDim.x) {
        double s = 0.0;
        for (long long e = e0; e < e1; e++) s += Pe[e * nd + k];
        part[b * nd + k] = s;
    }
}

// Neumann boundary work (src/iga_intdensity.c:196-287 with densityb): the
// traction integral of the face values is -fneu . U (fneu = -int t N, the
// Neumann load of the residual), U = slice 0 (ge) or (slice 1 + slice 2) / 2
// (getime, phys/getime/density.c: densityb).  Fixed-order block partials.
__global__ __launch_bounds__(256) void k_neumann_work(const double *__restrict__ fneu, const double *__restrict__ U,
                                                      long long ndof, int avg, double *__restrict__ part)
{
    __shared__ double red[256];
    double s = 0.0;
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < ndof; i += (long long)gridDim.x * 256) {
        double u = avg ? 0.5 * (U[ndof + i] + U[2 * ndof + i]) : U[i];
        s += fneu[i] * u;
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// ============================================================== adjoint (ge)
// g[e][k] = lam^T dR/dcoef[e][k] = sum_q w_q d/de phi_k(u + e lam_q) at e = 0,
// lam interpolated like u (no F0 term) and zero on the Dirichlet rows, whose
// residual U - g does not depend on the coefficients.  W is linear in the
// coefficients, so this is exact.  Basis values are recomputed per element.
__global__ __launch_bounds__(64) void k_vjp_coef(MeshArgs m, const double *__restrict__ U, const double *__restrict__ lam,
                                                 const unsigned char *__restrict__ dmask, double *__restrict__ g)
{
    __shared__ double sU[81], sL[81];
    __shared__ double sphi[3 * 4 * 9];
    __shared__ double sred[6][64];
    const int tid = threadIdx.x;
    const long long e = blockIdx.x;
    int ev[3];
    ev[2] = (int)(e % m.ne[2]);
    ev[1] = (int)((e / m.ne[2]) % m.ne[1]);
    ev[0] = (int)(e / ((long long)m.ne[2] * m.ne[1]));
    for (int idx = tid; idx < 81; idx += 64) {
        int A = idx / 3, i = idx % 3;
        long long r = 3 * node_of(m, ev[0] + A / 9, ev[1] + (A / 3) % 3, ev[2] + A % 3) + i;
        sU[idx] = U[r];
        sL[idx] = dmask[r] ? 0.0 : lam[r];
    }
    fill_phi(m, ev, sphi, tid, 64);
    __syncthreads();
    const int nqp = m.nq * m.nq * m.nq;
    double val[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    if (tid < nqp) {
        int qx = tid / (m.nq * m.nq), qy = (tid / m.nq) % m.nq, qz = tid % m.nq;
        double u[1][30], w;
        interp_slots<1>(m, sU, sphi, ev, qx, qy, qz, u, &w);
        double l[30];
        for (int s = 0; s < 30; s++) l[s] = 0.0;
        for (int A = 0; A < 27; A++) {
            double n[10];
            basis_slots(sphi, qx, qy, qz, A, n);
            for (int i = 0; i < 3; i++) {
                const double Lv = sL[3 * A + i];
#pragma unroll
                for (int s = 0; s < 10; s++) l[10 * i + s] += n[s] * Lv;
            }
        }
        Jet<2> uj[30];
        for (int s = 0; s < 30; s++) {
            uj[s].c[0] = u[0][s];
            uj[s].c[1] = l[s];
        }
        Jet<2> phi[6];
        energy_basis<Jet<2>>(uj, phi);
        for (int k = 0; k < 6; k++) val[k] = w * phi[k].c[1];
    }
    for (int k = 0; k < 6; k++) sred[k][tid] = val[k];
    __syncthreads();
    if (tid < 6) {
        double s = 0.0;
        for (int l = 0; l < nqp; l++) s += sred[tid][l];
        g[6 * e + tid] = s;
    }
}

// ============================================================== context

static MeshArgs mesh_args(const igap4h_ctx *c)
{
    MeshArgs m;
    for (int d = 0; d < 3; d++) {
        m.ne[d] = c->mesh.ne[d];
        m.nb[d] = c->mesh.nb[d];
        m.knot[d] = c->d_knot[d];
    }
    m.phi = c->d_phi;
    m.phioff[0] = 0;
    m.phioff[1] = m.ne[0];
    m.phioff[2] = m.ne[0] + m.ne[1];
    m.nq = c->nquad;
    for (int q = 0; q < 4; q++) {
        m.cq[q] = q < c->nquad ? c->cq[q] : 0.0;
        m.wq[q] = q < c->nquad ? c->wq[q] : 0.0;
    }
    for (int k = 0; k < 9; k++) m.pp[k] = c->mesh.par_periodic[k];
    m.ndof = c->mesh.ndof();
    m.rho = c->par_mat[5];
    m.cdamp = c->par_mat[6];
    m.dt = c->par_mat[7];
    return m;
}

static void ctx_free(igap4h_ctx *c)
{
    if (!c) return;
    int cur = -1;
    (void)hipGetDevice(&cur);
    (void)hipSetDevice(c->device);
    for (int d = 0; d < 3; d++) (void)hipFree(c->d_knot[d]);
    (void)hipFree(c->d_phi);
    (void)hipFree(c->d_coef);
    (void)hipFree(c->d_rowptr);
    (void)hipFree(c->d_col);
    (void)hipFree(c->d_mask);
    (void)hipFree(c->d_g);
    (void)hipFree(c->d_fneu);
    (void)hipFree(c->d_drows);
    (void)hipFree(c->d_Re);
    (void)hipFree(c->d_Uh);
    (void)hipFree(c->d_R);
    (void)hipFree(c->d_vals);
    (void)hipFree(c->d_Pe);
    (void)hipFree(c->d_part);
    solver_free(c);
    if (cur >= 0) (void)hipSetDevice(cur);
    delete c;
}

extern "C" void igap4h_destroy(igap4h_ctx *ctx) { ctx_free(ctx); }

// Device address of `scratch` on each device, looked up once (with that
// device current) by igap4h_create; the array's name is not a device address
// in host code.
constexpr int kMaxDevices = 64;
static double *g_scratch[kMaxDevices];
static std::mutex g_scratch_mu;

static int scratch_init(int dev)
{
    if (dev < 0 || dev >= kMaxDevices) return fail(IGAP4H_ENODEV, "device %d beyond the %d this library tracks", dev, kMaxDevices);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch[dev]) return 0;
    void *p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return fail(IGAP4H_EHIP, "hipGetSymbolAddress(scratch) on device %d: %s", dev, hipGetErrorString(e));
    g_scratch[dev] = (double *)p;
    return 0;
}

static int check_device(int *dev)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return fail(IGAP4H_ENODEV, "no HIP device visible");
    HIPCHK(hipGetDevice(dev));
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, *dev));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return fail(IGAP4H_ENODEV, "device %d is %s; this library is built for gfx950", *dev, prop.gcnArchName);
    return 0;
}

static int upload_coef(igap4h_ctx *c, const double *elem_coef)
{
    int64_t ne = c->mesh.nelem();
    std::vector<double> h;
    const double *src = elem_coef;
    if (!src) {
        double cf[6];
        coef_from_par(c->par_mat, cf);
        h.resize((size_t)ne * 6);
        for (int64_t e = 0; e < ne; e++)
            for (int k = 0; k < 6; k++) h[(size_t)e * 6 + k] = cf[k];
        src = h.data();
    }
    HIPCHK(hipMemcpy(c->d_coef, src, (size_t)ne * 6 * sizeof(double), hipMemcpyHostToDevice));
    return 0;
}

extern "C" int igap4h_create(const IGA *app, const igap4h_opts *opts, igap4h_ctx **out)
{
    if (!out) return fail(IGAP4H_EINVAL, "ctx output pointer is NULL");
    *out = nullptr;
    if (!app || !app->phys || !app->nurbs || !app->bc) return fail(IGAP4H_EINVAL, "IGA, Phys, Nurbs and BC are required");
    const Phys *ph = app->phys;
    const Nurbs *nu = app->nurbs;
    const BC *bc = app->bc;
    if (ph->ndim != 3 || ph->ndof != 3) return fail(IGAP4H_EINVAL, "ndim = ndof = 3 required (got %d, %d)", ph->ndim, ph->ndof);
    if (nu->porder != 2) return fail(IGAP4H_EINVAL, "porder 2 required (got %d)", nu->porder);
    if (!nu->nelem || !nu->nbasis || !nu->nknot || !nu->knotVector || !ph->par_mat || !bc->bc_type ||
        !bc->par_dirichlet || !bc->par_periodic)
        return fail(IGAP4H_EINVAL, "incomplete Nurbs/Phys/BC");
    igap4h_opts o = {IGAP4H_PHYS_FROM_TORDER, 0, nullptr, nullptr};
    if (opts) o = *opts;
    int physics = o.physics;
    if (physics == IGAP4H_PHYS_FROM_TORDER) physics = ph->torder == 2 ? IGAP4H_PHYS_GETIME : IGAP4H_PHYS_GE;
    if (physics != IGAP4H_PHYS_GE && physics != IGAP4H_PHYS_GETIME) return fail(IGAP4H_EINVAL, "unknown physics %d", physics);
    int need_torder = physics == IGAP4H_PHYS_GE ? 0 : 2;
    if (ph->torder != need_torder) return fail(IGAP4H_EINVAL, "physics %d needs torder %d (got %d)", physics, need_torder, ph->torder);
    int nquad = o.nquad ? o.nquad : 4;
    if (nquad < 1 || nquad > 4) return fail(IGAP4H_EINVAL, "nquad must be 1..4 (got %d)", nquad);

    igap4h_ctx *c = new igap4h_ctx();
    c->physics = physics;
    c->nquad = nquad;
    c->nhist = need_torder + 1;
    gauss01(nquad, c->cq, c->wq);
    for (int k = 0; k < 9; k++) c->par_mat[k] = ph->par_mat[k];
    Mesh &m = c->mesh;
    for (int d = 0; d < 3; d++) {
        m.ne[d] = nu->nelem[d];
        m.nb[d] = m.ne[d] + 2;        // the local window: every basis function of the local elements
        m.nown[d] = nu->nbasis[d];    // owned: nelem + 2 (last part of the direction) or nelem
        if (m.ne[d] < 1 || (m.nown[d] != m.ne[d] + 2 && m.nown[d] != m.ne[d]) || nu->nknot[d] != m.ne[d] + 5 ||
            !nu->knotVector[d]) {
            delete c;
            return fail(IGAP4H_EINVAL, "direction %d: need nbasis = nelem + 2 (or nelem on a rank that is not the last "
                                       "part of the direction) and nknot = nelem + 5 (no periodicity)", d);
        }
        m.knot[d].assign(nu->knotVector[d], nu->knotVector[d] + nu->nknot[d]);
    }
    // a rank's block when it owns less than its window, or (the last part in
    // every direction) its rows are not the first ones
    c->partitioned = m.nown[0] != m.nb[0] || m.nown[1] != m.nb[1] || m.nown[2] != m.nb[2];
    if (!c->partitioned && app->globalIdx) {
        const int64_t n = 3LL * m.nb[0] * m.nb[1] * m.nb[2];
        for (int64_t i = 0; i < n && !c->partitioned; i++) c->partitioned = app->globalIdx[i] != i;
    }
    if (nu->nboun < 0 || nu->nboun > 6 || (nu->nboun > 0 && !nu->boun)) {
        delete c;
        return fail(IGAP4H_EINVAL, "bad boundary face list");
    }
    m.nboun = nu->nboun;
    for (int i = 0; i < m.nboun; i++) {
        if (nu->boun[i] < 0 || nu->boun[i] > 5) {
            delete c;
            return fail(IGAP4H_EINVAL, "boundary face id %d out of range", nu->boun[i]);
        }
        m.boun[i] = nu->boun[i];
    }
    for (int i = 0; i < 36; i++) m.bc_type[i] = bc->bc_type[i];
    for (int i = 0; i < 9; i++) {
        m.par_dirichlet[i] = bc->par_dirichlet[i];
        m.par_periodic[i] = bc->par_periodic[i];
    }
    m.par_neumann = bc->par_neumann;
    const int64_t ndof = m.ndof(), ne = m.nelem();
    if (ndof > 2147483647LL) {
        delete c;
        return fail(IGAP4H_EINVAL, "ndof %lld exceeds int32 column indices", (long long)ndof);
    }
    if (!c->partitioned) {
        if (app->nselfIdx != 0 && app->nselfIdx != ndof) {
            delete c;
            return fail(IGAP4H_EINVAL, "nselfIdx %d != ndof %lld", app->nselfIdx, (long long)ndof);
        }
    } else {
        // one rank's block (src/snesiga_init.c:134-339): owned basis functions in
        // the local window through selfIdx, the received ghost layers through
        // recvIdx, globalIdx for the window
        const int64_t nself = 3LL * m.nown[0] * m.nown[1] * m.nown[2];
        if (app->nselfIdx != nself || app->nrecvIdx != ndof - nself || !app->selfIdx || (ndof > nself && !app->recvIdx) ||
            !app->globalIdx) {
            delete c;
            return fail(IGAP4H_EINVAL, "partitioned block: need nselfIdx = %lld, nrecvIdx = %lld, selfIdx, recvIdx and "
                                       "globalIdx (got %d, %d)", (long long)nself, (long long)(ndof - nself), app->nselfIdx,
                        app->nrecvIdx);
        }
        std::vector<unsigned char> seen((size_t)ndof, 0);
        int64_t ia = 0;
        for (int x = 0; x < m.nown[0]; x++)
            for (int y = 0; y < m.nown[1]; y++)
                for (int z = 0; z < m.nown[2]; z++)
                    for (int i = 0; i < 3; i++, ia++) {
                        const int64_t want = 3 * (((int64_t)x * m.nb[1] + y) * m.nb[2] + z) + i;
                        if (app->selfIdx[ia] != want) {
                            delete c;
                            return fail(IGAP4H_EINVAL, "selfIdx[%lld] = %d, expected %lld (iga_setmpicomm layout)", (long long)ia,
                                        app->selfIdx[ia], (long long)want);
                        }
                        seen[want] = 1;
                    }
        for (int64_t k = 0; k < ndof - nself; k++) {
            const int r = app->recvIdx[k];
            if (r < 0 || r >= ndof || seen[r]) {
                delete c;
                return fail(IGAP4H_EINVAL, "recvIdx[%lld] = %d is outside the window or repeats an owned/received entry",
                            (long long)k, r);
            }
            seen[r] = 1;
        }
        c->gidx.assign(app->globalIdx, app->globalIdx + ndof);
    }
    c->U_hist_app = app->U_hist;
    c->neumann = o.neumann;
    int rc = check_device(&c->device);
    if (!rc) rc = scratch_init(c->device);
    if (rc) {
        delete c;
        return rc;
    }
    c->nnz = pattern_nnz(m.nb);
    c->drows = dirichlet_rows(m);
    std::vector<unsigned char> mask;
    std::vector<double> g;
    dirichlet_values(m, mask, g);
    std::vector<double> fneu = neumann_load(m, nquad, o.neumann);
    c->neu_measure = neumann_measure(m, nquad);

#define ALLOC(p, n)                   \
    do {                              \
        int r_ = dev_alloc(&(p), (n)); \
        if (r_) {                     \
            ctx_free(c);              \
            return r_;                \
        }                             \
    } while (0)
#define CPY(dst, src, bytes)                                                          \
    do {                                                                              \
        if (hipMemcpy((dst), (src), (bytes), hipMemcpyHostToDevice) != hipSuccess) {  \
            ctx_free(c);                                                              \
            return fail(IGAP4H_EHIP, "upload failed");                                \
        }                                                                             \
    } while (0)
    for (int d = 0; d < 3; d++) {
        ALLOC(c->d_knot[d], m.knot[d].size());
        CPY(c->d_knot[d], m.knot[d].data(), m.knot[d].size() * sizeof(double));
    }
    ALLOC(c->d_coef, (size_t)ne * 6);
    rc = upload_coef(c, o.elem_coef);
    if (rc) {
        ctx_free(c);
        return rc;
    }
    ALLOC(c->d_rowptr, (size_t)ndof + 1);
    ALLOC(c->d_col, (size_t)c->nnz);
    ALLOC(c->d_mask, (size_t)ndof);
    CPY(c->d_mask, mask.data(), (size_t)ndof);
    ALLOC(c->d_g, (size_t)ndof);
    CPY(c->d_g, g.data(), (size_t)ndof * sizeof(double));
    if (!fneu.empty()) {
        ALLOC(c->d_fneu, (size_t)ndof);
        CPY(c->d_fneu, fneu.data(), (size_t)ndof * sizeof(double));
    }
    ALLOC(c->d_drows, c->drows.size());
    if (!c->drows.empty()) CPY(c->d_drows, c->drows.data(), c->drows.size() * sizeof(int32_t));
    ALLOC(c->d_Re, (size_t)ne * 81);
    // 1D basis tables of every element (the nquad points used)
    ALLOC(c->d_phi, (size_t)(m.ne[0] + m.ne[1] + m.ne[2]) * 36);
    {
        MeshArgs mt = mesh_args(c);
        const int nt = (m.ne[0] + m.ne[1] + m.ne[2]) * 4;
        hipLaunchKernelGGL(k_phi_table, dim3((unsigned)((nt + 255) / 256)), dim3(256), 0, 0, mt, c->d_phi);
    }
    // CSR pattern on the device
    MeshArgs ma = mesh_args(c);
    {
        long long nthr = ndof + 1;
        hipLaunchKernelGGL(k_rowptr, dim3((unsigned)((nthr + 255) / 256)), dim3(256), 0, 0, ma, c->d_rowptr);
        long long nt2 = ndof * 25;
        hipLaunchKernelGGL(k_cols, dim3((unsigned)((nt2 + 255) / 256)), dim3(256), 0, 0, ma, (const long long *)c->d_rowptr, c->d_col);
        if (hipDeviceSynchronize() != hipSuccess || hipGetLastError() != hipSuccess) {
            ctx_free(c);
            return fail(IGAP4H_EHIP, "pattern kernels failed");
        }
        long long last = 0;
        (void)hipMemcpy(&last, c->d_rowptr + ndof, sizeof(long long), hipMemcpyDeviceToHost);
        if (last != c->nnz) {
            ctx_free(c);
            return fail(IGAP4H_EHIP, "pattern size mismatch %lld vs %lld", last, (long long)c->nnz);
        }
    }
    *out = c;
    return 0;
#undef ALLOC
#undef CPY
}

extern "C" int igap4h_sizes(const igap4h_ctx *c, int64_t *ndof, int64_t *nnz, int64_t *ndir, int64_t *nelem, int *nhist)
{
    if (!c) return fail(IGAP4H_EINVAL, "NULL context");
    if (ndof) *ndof = c->mesh.ndof();
    if (nnz) *nnz = c->nnz;
    if (ndir) *ndir = (int64_t)c->drows.size();
    if (nelem) *nelem = c->mesh.nelem();
    if (nhist) *nhist = c->nhist;
    return 0;
}

extern "C" int igap4h_pattern_dev(igap4h_ctx *c, const int64_t **rowptr, const int32_t **col)
{
    if (!c) return fail(IGAP4H_EINVAL, "NULL context");
    if (rowptr) *rowptr = (const int64_t *)c->d_rowptr;
    if (col) *col = (const int32_t *)c->d_col;
    return 0;
}

extern "C" int igap4h_pattern(igap4h_ctx *c, int64_t *rowptr, int32_t *col)
{
    if (!c) return fail(IGAP4H_EINVAL, "NULL context");
    HIPCHK(hipSetDevice(c->device));
    if (rowptr) HIPCHK(hipMemcpy(rowptr, c->d_rowptr, ((size_t)c->mesh.ndof() + 1) * sizeof(long long), hipMemcpyDeviceToHost));
    if (col) HIPCHK(hipMemcpy(col, c->d_col, (size_t)c->nnz * sizeof(int), hipMemcpyDeviceToHost));
    return 0;
}

extern "C" int igap4h_dirichlet_rows(igap4h_ctx *c, int32_t *rows)
{
    if (!c || !rows) return fail(IGAP4H_EINVAL, "NULL argument");
    if (c->partitioned) {  // row_dirichlet holds globalIdx of the local rows (snesiga_Tangent.c:256,264)
        for (size_t k = 0; k < c->drows.size(); k++) rows[k] = c->gidx[(size_t)c->drows[k]];
    } else if (!c->drows.empty()) {
        std::memcpy(rows, c->drows.data(), c->drows.size() * sizeof(int32_t));
    }
    return 0;
}

// Re-read bc_type / par_dirichlet (the reference re-reads BC on every
// callback, snesiga_Residual.c:326-412, snesiga_Tangent.c:203-272): a no-op
// when nothing changed, else the Dirichlet rows, values and mask (and the
// Neumann load, whose faces follow bc_type) are rebuilt and uploaded.
extern "C" int igap4h_update_bc(igap4h_ctx *c, const BC *bc)
{
    if (!c || !bc || !bc->bc_type || !bc->par_dirichlet) return fail(IGAP4H_EINVAL, "NULL argument");
    Mesh &m = c->mesh;
    bool same = true;
    for (int i = 0; i < 36 && same; i++) same = m.bc_type[i] == bc->bc_type[i];
    for (int i = 0; i < 9 && same; i++) same = m.par_dirichlet[i] == bc->par_dirichlet[i];
    if (same && m.par_neumann == bc->par_neumann) return 0;
    for (int i = 0; i < 36; i++) m.bc_type[i] = bc->bc_type[i];
    for (int i = 0; i < 9; i++) m.par_dirichlet[i] = bc->par_dirichlet[i];
    m.par_neumann = bc->par_neumann;
    HIPCHK(hipSetDevice(c->device));
    // kernels still queued on any stream may read the arrays replaced below
    HIPCHK(hipDeviceSynchronize());
    solver_invalidate_jacobian(c);
    c->bc_gen++;
    const int64_t ndof = m.ndof();
    std::vector<int32_t> rows = dirichlet_rows(m);
    std::vector<unsigned char> mask;
    std::vector<double> g;
    dirichlet_values(m, mask, g);
    std::vector<double> fneu = neumann_load(m, c->nquad, c->neumann);
    c->neu_measure = neumann_measure(m, c->nquad);
    if (rows.size() != c->drows.size()) {
        (void)hipFree(c->d_drows);
        c->d_drows = nullptr;
        int r = dev_alloc(&c->d_drows, rows.size());
        if (r) return r;
    }
    c->drows = rows;
    if (!rows.empty()) HIPCHK(hipMemcpy(c->d_drows, rows.data(), rows.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->d_mask, mask.data(), (size_t)ndof, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->d_g, g.data(), (size_t)ndof * sizeof(double), hipMemcpyHostToDevice));
    if (fneu.empty()) {
        (void)hipFree(c->d_fneu);
        c->d_fneu = nullptr;
    } else {
        if (!c->d_fneu) {
            int r = dev_alloc(&c->d_fneu, (size_t)ndof);
            if (r) return r;
        }
        HIPCHK(hipMemcpy(c->d_fneu, fneu.data(), (size_t)ndof * sizeof(double), hipMemcpyHostToDevice));
    }
    return 0;
}

extern "C" int igap4h_update_neumann(igap4h_ctx *c)
{
    if (!c) return fail(IGAP4H_EINVAL, "NULL context");
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipDeviceSynchronize());
    const int64_t ndof = c->mesh.ndof();
    std::vector<double> fneu = neumann_load(c->mesh, c->nquad, c->neumann);
    if (fneu.empty()) {
        if (c->d_fneu) {
            (void)hipFree(c->d_fneu);
            c->d_fneu = nullptr;
            c->bc_gen++;
        }
        return 0;
    }
    if (!c->d_fneu) {
        int r = dev_alloc(&c->d_fneu, (size_t)ndof);
        if (r) return r;
        c->bc_gen++;
    }
    HIPCHK(hipMemcpy(c->d_fneu, fneu.data(), (size_t)ndof * sizeof(double), hipMemcpyHostToDevice));
    return 0;
}

extern "C" int igap4h_set_coef(igap4h_ctx *c, const double *elem_coef)
{
    if (!c) return fail(IGAP4H_EINVAL, "NULL context");
    HIPCHK(hipSetDevice(c->device));
    // queued kernels may still read the coefficients being replaced
    HIPCHK(hipDeviceSynchronize());
    int rc = upload_coef(c, elem_coef);
    if (rc) return rc;
    // a lagged tangent (and its preconditioner) belongs to the old coefficients
    solver_invalidate_jacobian(c);
    return 0;
}

extern "C" int igap4h_residual_dev(igap4h_ctx *c, const double *d_Uh, double *d_R, void *stream)
{
    if (!c || !d_Uh || !d_R) return fail(IGAP4H_EINVAL, "NULL argument");
    hipStream_t st = (hipStream_t)stream;
    MeshArgs m = mesh_args(c);
    long long ne = c->mesh.nelem();
    if (c->physics == IGAP4H_PHYS_GE)
        hipLaunchKernelGGL(k_residual_elem<0>, dim3((unsigned)ne), dim3(64), 0, st, m, d_Uh, (const double *)c->d_coef, c->d_Re);
    else
        hipLaunchKernelGGL(k_residual_elem<1>, dim3((unsigned)ne), dim3(64), 0, st, m, d_Uh, (const double *)c->d_coef, c->d_Re);
    HIPCHK(hipGetLastError());
    long long nd = m.ndof;
    hipLaunchKernelGGL(k_gather, dim3((unsigned)((nd + 255) / 256)), dim3(256), 0, st, m, (const double *)c->d_Re,
                       (const double *)c->d_fneu, (const unsigned char *)c->d_mask, (const double *)c->d_g, d_Uh, d_R);
    HIPCHK(hipGetLastError());
    return 0;
}

extern "C" int igap4h_residual_dev_noalloc(igap4h_ctx *c, const double *d_Uh, double *d_R, void *stream)
{
    if (!c || !d_Uh || !d_R) return fail(IGAP4H_EINVAL, "NULL argument");
    if (c->device < 0 || c->device >= kMaxDevices || !g_scratch[c->device])
        return fail(IGAP4H_EINVAL, "no scratch address for device %d (context not made by igap4h_create)", c->device);
    double *Re = g_scratch[c->device];
    hipStream_t st = (hipStream_t)stream;
    MeshArgs m = mesh_args(c);
    const long long ne = c->mesh.nelem(), nd = m.ndof;
    const long long plane = 3LL * m.nb[1] * m.nb[2];  // dofs per x-plane of nodes
    const long long eplane = (long long)m.ne[1] * m.ne[2];
    hipLaunchKernelGGL(k_fill_zero, dim3(2048), dim3(256), 0, st, (unsigned int *)d_R, 2 * nd);
    HIPCHK(hipGetLastError());
    for (long long e0 = 0; e0 < ne; e0 += kScratchElems) {
        const long long e1 = std::min(ne, e0 + kScratchElems);
        if (c->physics == IGAP4H_PHYS_GE)
            hipLaunchKernelGGL(k_residual_chunk<0>, dim3((unsigned)(e1 - e0)), dim3(64), 0, st, m, d_Uh, (const double *)c->d_coef, Re, e0);
        else
            hipLaunchKernelGGL(k_residual_chunk<1>, dim3((unsigned)(e1 - e0)), dim3(64), 0, st, m, d_Uh, (const double *)c->d_coef, Re, e0);
        HIPCHK(hipGetLastError());
        // dofs of the node planes [ex(e0), ex(e1 - 1) + 2]
        const long long d0 = (e0 / eplane) * plane, d1 = std::min(nd, ((e1 - 1) / eplane + 3) * plane);
        hipLaunchKernelGGL(k_gather_chunk, dim3((unsigned)((d1 - d0 + 255) / 256)), dim3(256), 0, st, m, (const double *)Re, e0, e1,
                           d0, d1, d_R);
        HIPCHK(hipGetLastError());
    }
    hipLaunchKernelGGL(k_gather_finish, dim3((unsigned)((nd + 255) / 256)), dim3(256), 0, st, m, (const double *)c->d_fneu,
                       (const unsigned char *)c->d_mask, (const double *)c->d_g, d_Uh, d_R);
    HIPCHK(hipGetLastError());
    return 0;
}

// Gauss points per chunk: 2 for FP64 (keeps two elements per CU), 4 for FP32.
template <class ACC>
struct GeQC { static constexpr int v = 2; };
template <>
struct GeQC<float> { static constexpr int v = 4; };

template <class ACC>
static int launch_tangent_ge(igap4h_ctx *c, const MeshArgs &m, const double *d_Uh, void *d_vals, int f32, hipStream_t st)
{
    for (int cx = 0; cx < 3; cx++)
        for (int cy = 0; cy < 3; cy++)
            for (int cz = 0; cz < 3; cz++) {
                int mx = (m.ne[0] - cx + 2) / 3, my = (m.ne[1] - cy + 2) / 3, mz = (m.ne[2] - cz + 2) / 3;
                long long nb = (long long)mx * my * mz;
                if (nb <= 0) continue;
                hipLaunchKernelGGL((k_tangent_ge<ACC, GeQC<ACC>::v>), dim3((unsigned)nb), dim3(256), 0, st, m, d_Uh,
                                   (const double *)c->d_coef, (const long long *)c->d_rowptr,
                                   (const unsigned char *)c->d_mask, d_vals, f32, cx, cy, cz, my, mz);
                HIPCHK(hipGetLastError());
            }
    return 0;
}

template <class ACC>
static int launch_tangent_gt(igap4h_ctx *c, const MeshArgs &m, const double *d_Uh, double mass, void *d_vals, int f32, hipStream_t st)
{
    for (int cx = 0; cx < 3; cx++)
        for (int cy = 0; cy < 3; cy++)
            for (int cz = 0; cz < 3; cz++) {
                int mx = (m.ne[0] - cx + 2) / 3, my = (m.ne[1] - cy + 2) / 3, mz = (m.ne[2] - cz + 2) / 3;
                long long nb = (long long)mx * my * mz;
                if (nb <= 0) continue;
                hipLaunchKernelGGL((k_tangent_gt<ACC, GtQC<ACC>::v>), dim3((unsigned)nb), dim3(256), 0, st, m, d_Uh,
                                   (const double *)c->d_coef, mass, (const long long *)c->d_rowptr,
                                   (const unsigned char *)c->d_mask, d_vals, f32, cx, cy, cz, my, mz);
                HIPCHK(hipGetLastError());
            }
    return 0;
}

// Gauss points per chunk of the strain-variation tangent, as measured (64^3
// ge FP32: 37.5 / 32.5 / 40.3 ms with 1 / 2 / 4; 48^3 getime FP32: 31.9 /
// 27.6 with 1 / 2); FP64 keeps two work-groups per CU with 1 (getime: one)
template <class ACC, int GT>
struct SvQC { static constexpr int v = 1; };
template <>
struct SvQC<float, 0> { static constexpr int v = 2; };
template <>
struct SvQC<float, 1> { static constexpr int v = 2; };
template <>
struct SvQC<double, 0> { static constexpr int v = 2; };
template <>
struct SvQC<double, 1> { static constexpr int v = 2; };


template <class ACC, int GT>
static int launch_tangent_sv(igap4h_ctx *c, const MeshArgs &m, const double *d_Uh, double mass, void *d_vals, int f32, hipStream_t st)
{
    for (int cx = 0; cx < 3; cx++)
        for (int cy = 0; cy < 3; cy++)
            for (int cz = 0; cz < 3; cz++) {
                int mx = (m.ne[0] - cx + 2) / 3, my = (m.ne[1] - cy + 2) / 3, mz = (m.ne[2] - cz + 2) / 3;
                long long nb = (long long)mx * my * mz;
                if (nb <= 0) continue;
                hipLaunchKernelGGL((k_tangent_sv<ACC, SvQC<ACC, GT>::v, GT>), dim3((unsigned)nb), dim3(256), 0, st, m, d_Uh,
                                   (const double *)c->d_coef, mass, (const long long *)c->d_rowptr,
                                   (const unsigned char *)c->d_mask, d_vals, f32, cx, cy, cz, my, mz);
                HIPCHK(hipGetLastError());
            }
    return 0;
}

extern "C" int igap4h_tangent_dev(igap4h_ctx *c, const double *d_Uh, int prec, int vals_type, void *d_vals, void *stream)
{
    if (!c || !d_Uh || !d_vals) return fail(IGAP4H_EINVAL, "NULL argument");
    if (prec != IGAP4H_FP64 && prec != IGAP4H_FP32) return fail(IGAP4H_EINVAL, "bad precision %d", prec);
    if (vals_type != IGAP4H_VALS_F64 && vals_type != IGAP4H_VALS_F32) return fail(IGAP4H_EINVAL, "bad vals type %d", vals_type);
    hipStream_t st = (hipStream_t)stream;
    MeshArgs m = mesh_args(c);
    int f32 = vals_type == IGAP4H_VALS_F32;
    long long n32 = f32 ? c->nnz : 2 * c->nnz;
    hipLaunchKernelGGL(k_fill_zero, dim3(2048), dim3(256), 0, st, (unsigned int *)d_vals, n32);
    HIPCHK(hipGetLastError());
    double mass = c->physics == IGAP4H_PHYS_GETIME ? c->par_mat[5] / (c->par_mat[7] * c->par_mat[7]) + c->par_mat[6] / (2. * c->par_mat[7]) : 0.0;
    // AUTO: the faster form, the strain variation for both physics and
    // precisions (64^3 ge, 48^3 getime, nq 4, ms per tangent, T strip / strain
    // variation): ge FP64 62.2 / 55.4, FP32 34.5 / 29.1; getime FP64 65.5 / 55.3,
    // FP32 33.1 / 23.2
    const bool gt = c->physics == IGAP4H_PHYS_GETIME, f64 = prec == IGAP4H_FP64;
    const bool sv = c->tangent_form != IGAP4H_TANGENT_TSTRIP;
    int rc;
    if (sv) {
        if (gt) rc = f64 ? launch_tangent_sv<double, 1>(c, m, d_Uh, mass, d_vals, f32, st)
                         : launch_tangent_sv<float, 1>(c, m, d_Uh, mass, d_vals, f32, st);
        else rc = f64 ? launch_tangent_sv<double, 0>(c, m, d_Uh, 0.0, d_vals, f32, st)
                      : launch_tangent_sv<float, 0>(c, m, d_Uh, 0.0, d_vals, f32, st);
    } else if (gt) {
        rc = f64 ? launch_tangent_gt<double>(c, m, d_Uh, mass, d_vals, f32, st)
                 : launch_tangent_gt<float>(c, m, d_Uh, mass, d_vals, f32, st);
    } else {
        rc = f64 ? launch_tangent_ge<double>(c, m, d_Uh, d_vals, f32, st)
                 : launch_tangent_ge<float>(c, m, d_Uh, d_vals, f32, st);
    }
    if (rc) return rc;
    int nd = (int)c->drows.size();
    if (nd > 0) {
        hipLaunchKernelGGL(k_dirichlet_diag, dim3((nd + 255) / 256), dim3(256), 0, st, m, (const int *)c->d_drows, nd,
                           (const long long *)c->d_rowptr, d_vals, f32);
        HIPCHK(hipGetLastError());
    }
    return 0;
}

extern "C" int igap4h_set_tangent_form(igap4h_ctx *c, int form)
{
    if (!c) return fail(IGAP4H_EINVAL, "NULL argument");
    if (form != IGAP4H_TANGENT_AUTO && form != IGAP4H_TANGENT_TSTRIP && form != IGAP4H_TANGENT_SV)
        return fail(IGAP4H_EINVAL, "bad tangent form %d", form);
    c->tangent_form = form;
    return 0;
}

static int stage_U(igap4h_ctx *c, const double *U, double *U_hist)
{
    int64_t ndof = c->mesh.ndof();
    if (!c->d_Uh) {
        int r = dev_alloc(&c->d_Uh, (size_t)ndof * c->nhist);
        if (r) return r;
    }
    // side effect of the reference callbacks: U_hist[0:nself] = U (:60)
    if (U_hist && U && U_hist != U) std::memcpy(U_hist, U, (size_t)ndof * sizeof(double));
    const double *src0 = U ? U : U_hist;
    if (!src0) return fail(IGAP4H_EINVAL, "need U or U_hist");
    HIPCHK(hipMemcpy(c->d_Uh, src0, (size_t)ndof * sizeof(double), hipMemcpyHostToDevice));
    if (c->nhist > 1) {
        if (!U_hist) return fail(IGAP4H_EINVAL, "getime needs U_hist history slices");
        HIPCHK(hipMemcpy(c->d_Uh + ndof, U_hist + ndof, (size_t)ndof * (c->nhist - 1) * sizeof(double), hipMemcpyHostToDevice));
    }
    return 0;
}

extern "C" int igap4h_residual(igap4h_ctx *c, const double *U, double *U_hist, double *R)
{
    if (!c || !R) return fail(IGAP4H_EINVAL, "NULL argument");
    if (c->partitioned) return fail(IGAP4H_EINVAL, "a partitioned block takes the device entry points (ghosted local vectors)");
    HIPCHK(hipSetDevice(c->device));
    int rc = stage_U(c, U, U_hist);
    if (rc) return rc;
    int64_t ndof = c->mesh.ndof();
    if (!c->d_R && (rc = dev_alloc(&c->d_R, (size_t)ndof))) return rc;
    rc = igap4h_residual_dev(c, c->d_Uh, c->d_R, nullptr);
    if (rc) return rc;
    HIPCHK(hipMemcpy(R, c->d_R, (size_t)ndof * sizeof(double), hipMemcpyDeviceToHost));
    return 0;
}

extern "C" int igap4h_tangent(igap4h_ctx *c, const double *U, double *U_hist, int prec, int vals_type, void *vals)
{
    if (!c || !vals) return fail(IGAP4H_EINVAL, "NULL argument");
    if (c->partitioned) return fail(IGAP4H_EINVAL, "a partitioned block takes the device entry points (ghosted local vectors)");
    HIPCHK(hipSetDevice(c->device));
    int rc = stage_U(c, U, U_hist);
    if (rc) return rc;
    size_t bytes = (size_t)c->nnz * (vals_type == IGAP4H_VALS_F32 ? sizeof(float) : sizeof(double));
    if (c->d_vals_bytes < bytes) {
        (void)hipFree(c->d_vals);
        c->d_vals = nullptr;
        c->d_vals_bytes = 0;
        if (dev_malloc(&c->d_vals, bytes) != hipSuccess) return fail(IGAP4H_ENOMEM, "hipMalloc of %zu bytes failed", bytes);
        c->d_vals_bytes = bytes;
    }
    rc = igap4h_tangent_dev(c, c->d_Uh, prec, vals_type, c->d_vals, nullptr);
    if (rc) return rc;
    HIPCHK(hipMemcpy(vals, c->d_vals, bytes, hipMemcpyDeviceToHost));
    return 0;
}

extern "C" int igap4h_residual_vjp_coef_dev(igap4h_ctx *c, const double *d_U, const double *d_lambda, double *d_g,
                                            void *stream)
{
    if (!c || !d_U || !d_lambda || !d_g) return fail(IGAP4H_EINVAL, "NULL argument");
    if (c->physics != IGAP4H_PHYS_GE) return fail(IGAP4H_EINVAL, "the coefficient VJP is implemented for phys/ge only");
    long long ne = c->mesh.nelem();
    MeshArgs m = mesh_args(c);
    hipLaunchKernelGGL(k_vjp_coef, dim3((unsigned)ne), dim3(64), 0, (hipStream_t)stream, m, d_U, d_lambda,
                       (const unsigned char *)c->d_mask, d_g);
    HIPCHK(hipGetLastError());
    return 0;
}

constexpr int kDensityBlocks = 1024, kWorkBlocks = 256;

// HH[0..nd) = volume integrals of the density outputs; with boundary, minus
// the Neumann boundary terms (densityb) as src/iga_intdensity.c:196-287.
static int density_integrals(igap4h_ctx *c, const double *d_U, bool boundary, double *HH, int *ndout, void *stream)
{
    HIPCHK(hipSetDevice(c->device));
    const long long ne = c->mesh.nelem();
    const int nd = c->physics == IGAP4H_PHYS_GE ? 13 : 2;
    if (c->d_Pe_n < (size_t)ne * nd) {
        (void)hipFree(c->d_Pe);
    (void)hipFree(c->d_part);
        c->d_Pe = nullptr;
        c->d_Pe_n = 0;
        int r = dev_alloc(&c->d_Pe, (size_t)ne * nd);
        if (r) return r;
        c->d_Pe_n = (size_t)ne * nd;
    }
    if (!c->d_part) {
        int r = dev_alloc(&c->d_part, (size_t)kDensityBlocks * 13 + kWorkBlocks);
        if (r) return r;
    }
    hipStream_t st = (hipStream_t)stream;
    MeshArgs m = mesh_args(c);
    if (c->physics == IGAP4H_PHYS_GE)
        hipLaunchKernelGGL(k_density_elem<0>, dim3((unsigned)ne), dim3(64), 0, st, m, d_U, (const double *)c->d_coef, c->d_Pe);
    else
        hipLaunchKernelGGL(k_density_elem<1>, dim3((unsigned)ne), dim3(64), 0, st, m, d_U, (const double *)c->d_coef, c->d_Pe);
    HIPCHK(hipGetLastError());
    const int nblk = (int)std::min<long long>(kDensityBlocks, ne);
    hipLaunchKernelGGL(k_density_blocks, dim3(nblk), dim3(64), 0, st, (const double *)c->d_Pe, ne, nd, c->d_part);
    HIPCHK(hipGetLastError());
    const bool work = boundary && c->d_fneu;
    if (work) {
        hipLaunchKernelGGL(k_neumann_work, dim3(kWorkBlocks), dim3(256), 0, st, (const double *)c->d_fneu, d_U, (long long)m.ndof,
                           c->physics == IGAP4H_PHYS_GETIME ? 1 : 0, c->d_part + (size_t)kDensityBlocks * 13);
        HIPCHK(hipGetLastError());
    }
    std::vector<double> h((size_t)kDensityBlocks * 13 + kWorkBlocks);
    HIPCHK(hipMemcpyAsync(h.data(), c->d_part, h.size() * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    for (int k = 0; k < nd; k++) {
        double s = 0.0;
        for (int b = 0; b < nblk; b++) s += h[(size_t)b * nd + k];
        HH[k] = s;
    }
    if (work) {
        double s = 0.0;
        for (int b = 0; b < kWorkBlocks; b++) s += h[(size_t)kDensityBlocks * 13 + b];
        HH[0] += s;  // minus the traction work: fneu = -int t N
    }
    if (boundary && c->physics == IGAP4H_PHYS_GETIME) HH[1] -= c->neu_measure;  // densityb hh[1] = 1
    if (ndout) *ndout = nd;
    return 0;
}

extern "C" int igap4h_energy_dev(igap4h_ctx *c, const double *d_U, double *energy, void *stream)
{
    if (!c || !d_U || !energy) return fail(IGAP4H_EINVAL, "NULL argument");
    double HH[13];
    int rc = density_integrals(c, d_U, false, HH, nullptr, stream);
    if (rc) return rc;
    *energy = HH[0];
    return 0;
}

extern "C" int igap4h_intdensity_dev(igap4h_ctx *c, const double *d_U_hist, double *HH, int *ndensity, void *stream)
{
    if (!c || !d_U_hist || !HH) return fail(IGAP4H_EINVAL, "NULL argument");
    return density_integrals(c, d_U_hist, true, HH, ndensity, stream);
}

// ============================================================== batched assembly
// Many independent phys/ge instances (own mesh sizes, coefficients, U, R and
// CSR values) assembled together: one residual launch over all elements, one
// gather launch, then per colour ONE tangent launch over that colour's
// elements of every instance (instances never share a basis function), the
// whole sequence captured once into a HIP graph and replayed per assembly.
struct BatchInst {
    MeshArgs m;
    const double *Uh;
    const double *coef;
    const long long *rowptr;
    const unsigned char *mask;
    const double *g;
    const double *fneu;
    const int *drows;
    int ndrows;
    double *Re;
    double *R;
    void *vals;
    long long nnz;
};

// largest i in [0, n) with off[i] <= x (off ascending, off[0] = 0)
__device__ __forceinline__ int batch_find(const long long *off, int n, long long x)
{
    int lo = 0, hi = n;
    while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if (off[mid] <= x) lo = mid;
        else hi = mid;
    }
    return lo;
}

__global__ __launch_bounds__(64) void k_residual_batch(const BatchInst *__restrict__ I, const long long *__restrict__ eoff, int n)
{
    const int i = batch_find(eoff, n, blockIdx.x);
    const BatchInst &b = I[i];
    const MeshArgs &m = b.m;
    const long long e = (long long)blockIdx.x - eoff[i];
    residual_elem_body<0>(m, b.Uh, b.coef, b.Re, e, e);
}

// The per-dof and per-value passes of a batch run on one 1D grid whose
// blocks are split among the instances by their own sizes (block offsets
// off[0..n], instance = batch_find): no block is spent on the largest
// instance's size for every instance.
__global__ __launch_bounds__(256) void k_gather_batch(const BatchInst *__restrict__ I, const long long *__restrict__ goff, int n)
{
    const int i = batch_find(goff, n, blockIdx.x);
    const BatchInst &b = I[i];
    const long long dof = ((long long)blockIdx.x - goff[i]) * 256 + threadIdx.x;
    if (dof < b.m.ndof) gather_body(b.m, b.Re, b.fneu, b.mask, b.g, b.Uh, b.R, dof);
}

constexpr int kFillWords = 8 * 256;  // 32-bit words zeroed per block of k_fill_zero_batch

__global__ __launch_bounds__(256) void k_fill_zero_batch(const BatchInst *__restrict__ I, const long long *__restrict__ foff,
                                                         int n, int vals_f32)
{
    const int i = batch_find(foff, n, blockIdx.x);
    const BatchInst &b = I[i];
    const long long n32 = vals_f32 ? b.nnz : 2 * b.nnz;
    unsigned int *p = (unsigned int *)b.vals;
    const long long base = ((long long)blockIdx.x - foff[i]) * kFillWords + threadIdx.x;
#pragma unroll
    for (int k = 0; k < 8; k++)
        if (base + 256 * k < n32) p[base + 256 * k] = 0u;
}

__global__ __launch_bounds__(256) void k_dirichlet_diag_batch(const BatchInst *__restrict__ I, const long long *__restrict__ doff,
                                                              int n, int vals_f32)
{
    const int i = batch_find(doff, n, blockIdx.x);
    const BatchInst &b = I[i];
    const MeshArgs &m = b.m;
    const long long t = ((long long)blockIdx.x - doff[i]) * 256 + threadIdx.x;
    if (t >= b.ndrows) return;
    long long r = b.drows[t];
    long long node = r / 3;
    int ic = (int)(r % 3);
    int a2 = (int)(node % m.nb[2]);
    int a1 = (int)((node / m.nb[2]) % m.nb[1]);
    int a0 = (int)(node / ((long long)m.nb[2] * m.nb[1]));
    int lo0 = max(0, a0 - 2), lo1 = max(0, a1 - 2), lo2 = max(0, a2 - 2);
    int w1 = min(m.nb[1] - 1, a1 + 2) - lo1 + 1, w2 = min(m.nb[2] - 1, a2 + 2) - lo2 + 1;
    long long pos = b.rowptr[r] + 3LL * (((a0 - lo0) * w1 + (a1 - lo1)) * w2 + (a2 - lo2)) + ic;
    if (vals_f32) ((float *)b.vals)[pos] = 1.0f;
    else ((double *)b.vals)[pos] = 1.0;
}

// colour (cx, cy, cz) of every instance; coff = this colour's element offsets
template <class ACC, int QC>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_tangent_ge_batch(
    const BatchInst *__restrict__ I, const long long *__restrict__ coff, int n, int vals_f32, int cx, int cy, int cz)
{
    const int i = batch_find(coff, n, blockIdx.x);
    const BatchInst &b = I[i];
    const MeshArgs &m = b.m;
    const long long t = (long long)blockIdx.x - coff[i];
    const int my = (m.ne[1] - cy + 2) / 3, mz = (m.ne[2] - cz + 2) / 3;
    const int ev[3] = {cx + 3 * (int)(t / ((long long)mz * my)), cy + 3 * (int)((t / mz) % my), cz + 3 * (int)(t % mz)};
    tangent_ge_body<ACC, QC>(m, b.Uh, b.coef, b.rowptr, b.mask, b.vals, vals_f32, ev);
}

// the same with the strain-variation form of the tangent; NQM: point data for
// up to NQM^3 Gauss points, WPE: work-groups per CU the registers are sized for
template <class ACC, int QC, int NQM, int WPE>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(WPE))) void k_tangent_sv_batch(
    const BatchInst *__restrict__ I, const long long *__restrict__ coff, int n, int vals_f32, int cx, int cy, int cz)
{
    const int i = batch_find(coff, n, blockIdx.x);
    const BatchInst &b = I[i];
    const MeshArgs &m = b.m;
    const long long t = (long long)blockIdx.x - coff[i];
    const int my = (m.ne[1] - cy + 2) / 3, mz = (m.ne[2] - cz + 2) / 3;
    const int ev[3] = {cx + 3 * (int)(t / ((long long)mz * my)), cy + 3 * (int)((t / mz) % my), cz + 3 * (int)(t % mz)};
    tangent_sv_body<ACC, QC, 0, NQM>(m, b.Uh, b.coef, 0.0, b.rowptr, b.mask, b.vals, vals_f32, ev);
}

// The tuning space, variant = 10 x form + Gauss points per MFMA chunk: form
// 0 = T strip (contraction depth 9 QC per pass), 1 = strain variation (12 QC
// per pass, 15 + 3 tiles), 2 = strain variation with the point data sized for
// 3^3 Gauss points and registers for 4 work-groups per CU (instances with
// nq <= 3 only); 0 ends a list.
constexpr int kBatchNV = 8;
static const int kBatchVariants[2][kBatchNV] = {{1, 2, 3, 11, 12, 0, 0, 0}, {2, 3, 4, 12, 14, 22, 0, 0}};  // [FP64, FP32]
static bool batch_variant_ok(int prec, int v, int nqmax)
{
    if (v / 10 == 2 && nqmax > 3) return false;
    for (int k = 0; k < kBatchNV; k++)
        if (v && kBatchVariants[prec == IGAP4H_FP32][k] == v) return true;
    return false;
}

struct igap4h_batch {
    int n = 0, prec = 0, vals_f32 = 0, device = 0, nqmax = 0;
    std::vector<BatchInst> h_inst;
    BatchInst *d_inst = nullptr;
    long long *d_off = nullptr;             // [eoff | coff (27) | goff | foff | doff], n + 1 each
    std::vector<long long> eoff, coff;      // host copies
    // block offsets per instance of the gather, fill and Dirichlet passes
    // (on the device after eoff and coff: d_off + 28 (n + 1), + 29 (n + 1), + 30 (n + 1))
    std::vector<long long> goff, foff, doff;
    int qc = 0;                             // chosen variant
    hipStream_t cap = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<igap4h_ctx *> ctx;          // the instances' contexts
    std::vector<int> bc_gen;                // their bc_gen when h_inst was filled
};

// the fields of an instance that igap4h_update_bc may change (it can
// reallocate the Dirichlet row list and the Neumann load)
static void batch_bc_fields(BatchInst &I, const igap4h_ctx *c)
{
    I.mask = c->d_mask;
    I.g = c->d_g;
    I.fneu = c->d_fneu;
    I.drows = c->d_drows;
    I.ndrows = (int)c->drows.size();
}

static void batch_drop_graph(igap4h_batch *b);

// Re-read the boundary data of instances whose context changed them since
// the descriptors were uploaded; the captured graph is dropped because the
// Dirichlet launch is sized by the largest row count.
static int batch_refresh(igap4h_batch *b)
{
    bool changed = false;
    for (int i = 0; i < b->n; i++)
        if (b->ctx[i]->bc_gen != b->bc_gen[i]) {
            batch_bc_fields(b->h_inst[i], b->ctx[i]);
            b->bc_gen[i] = b->ctx[i]->bc_gen;
            changed = true;
        }
    if (!changed) return 0;
    for (int i = 0; i < b->n; i++) b->doff[i + 1] = b->doff[i] + (b->h_inst[i].ndrows + 255) / 256;
    batch_drop_graph(b);
    HIPCHK(hipDeviceSynchronize());
    HIPCHK(hipMemcpy(b->d_inst, b->h_inst.data(), sizeof(BatchInst) * b->n, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(b->d_off + (size_t)30 * (b->n + 1), b->doff.data(), sizeof(long long) * (b->n + 1), hipMemcpyHostToDevice));
    return 0;
}

template <class ACC, int QC, int SV, int NQM = 4, int WPE = 2>
static int batch_tangent_launches(igap4h_batch *b, hipStream_t st)
{
    for (int c = 0; c < 27; c++) {
        const long long *off = b->coff.data() + (size_t)c * (b->n + 1);
        long long nb = off[b->n];
        if (nb <= 0) continue;
        const long long *d_coff = b->d_off + (b->n + 1) * (1 + c);
        if constexpr (SV)
            hipLaunchKernelGGL((k_tangent_sv_batch<ACC, QC, NQM, WPE>), dim3((unsigned)nb), dim3(256), 0, st, (const BatchInst *)b->d_inst,
                               d_coff, b->n, b->vals_f32, c / 9, (c / 3) % 3, c % 3);
        else
            hipLaunchKernelGGL((k_tangent_ge_batch<ACC, QC>), dim3((unsigned)nb), dim3(256), 0, st, (const BatchInst *)b->d_inst,
                               d_coff, b->n, b->vals_f32, c / 9, (c / 3) % 3, c % 3);
        HIPCHK(hipGetLastError());
    }
    return 0;
}

static int batch_tangent(igap4h_batch *b, int v, hipStream_t st)
{
    if (b->prec == IGAP4H_FP32) {
        if (v == 2) return batch_tangent_launches<float, 2, 0>(b, st);
        if (v == 3) return batch_tangent_launches<float, 3, 0>(b, st);
        if (v == 4) return batch_tangent_launches<float, 4, 0>(b, st);
        if (v == 12) return batch_tangent_launches<float, 2, 1>(b, st);
        if (v == 14) return batch_tangent_launches<float, 4, 1>(b, st);
        if (v == 22 && b->nqmax <= 3) return batch_tangent_launches<float, 2, 1, 3, 4>(b, st);
    } else {
        if (v == 1) return batch_tangent_launches<double, 1, 0>(b, st);
        if (v == 2) return batch_tangent_launches<double, 2, 0>(b, st);
        if (v == 3) return batch_tangent_launches<double, 3, 0>(b, st);
        if (v == 11) return batch_tangent_launches<double, 1, 1>(b, st);
        if (v == 12) return batch_tangent_launches<double, 2, 1>(b, st);
    }
    return fail(IGAP4H_EINVAL, "no batched tangent variant %d for this precision", v);
}

static int batch_sequence(igap4h_batch *b, hipStream_t st)
{
    const int n = b->n;
    const long long *off = b->d_off;
    hipLaunchKernelGGL(k_fill_zero_batch, dim3((unsigned)b->foff[n]), dim3(256), 0, st, (const BatchInst *)b->d_inst,
                       off + (size_t)29 * (n + 1), n, b->vals_f32);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(k_residual_batch, dim3((unsigned)b->eoff[n]), dim3(64), 0, st, (const BatchInst *)b->d_inst, off, n);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(k_gather_batch, dim3((unsigned)b->goff[n]), dim3(256), 0, st, (const BatchInst *)b->d_inst,
                       off + (size_t)28 * (n + 1), n);
    HIPCHK(hipGetLastError());
    int rc = batch_tangent(b, b->qc, st);
    if (rc) return rc;
    if (b->doff[n] > 0) {
        hipLaunchKernelGGL(k_dirichlet_diag_batch, dim3((unsigned)b->doff[n]), dim3(256), 0, st, (const BatchInst *)b->d_inst,
                           off + (size_t)30 * (n + 1), n, b->vals_f32);
        HIPCHK(hipGetLastError());
    }
    return 0;
}

static void batch_drop_graph(igap4h_batch *b)
{
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    b->exec = nullptr;
}

extern "C" void igap4h_batch_destroy(igap4h_batch *b)
{
    if (!b) return;
    batch_drop_graph(b);
    if (b->cap) (void)hipStreamDestroy(b->cap);
    (void)hipFree(b->d_inst);
    (void)hipFree(b->d_off);
    delete b;
}

extern "C" int igap4h_batch_create(igap4h_ctx *const *ctxs, int n, const double *const *d_U_hist, double *const *d_R,
                                   void *const *d_vals, int prec, int vals_type, igap4h_batch **out)
{
    if (!out) return fail(IGAP4H_EINVAL, "batch output pointer is NULL");
    *out = nullptr;
    if (!ctxs || !d_U_hist || !d_R || !d_vals || n < 1) return fail(IGAP4H_EINVAL, "batch needs n >= 1 contexts and arrays");
    if (prec != IGAP4H_FP64 && prec != IGAP4H_FP32) return fail(IGAP4H_EINVAL, "bad precision %d", prec);
    if (vals_type != IGAP4H_VALS_F64 && vals_type != IGAP4H_VALS_F32) return fail(IGAP4H_EINVAL, "bad vals type %d", vals_type);
    igap4h_batch *b = new igap4h_batch();
    b->n = n;
    b->prec = prec;
    b->vals_f32 = vals_type == IGAP4H_VALS_F32;
    // the single-instance default for phys/ge (IGAP4H_TANGENT_AUTO): the strain variation
    b->qc = prec == IGAP4H_FP32 ? 10 + SvQC<float, 0>::v : 10 + SvQC<double, 0>::v;
    b->eoff.assign(n + 1, 0);
    b->coff.assign((size_t)27 * (n + 1), 0);
    b->goff.assign(n + 1, 0);
    b->foff.assign(n + 1, 0);
    b->doff.assign(n + 1, 0);
    for (int i = 0; i < n; i++) {
        igap4h_ctx *c = ctxs[i];
        if (!c || !d_U_hist[i] || !d_R[i] || !d_vals[i]) {
            igap4h_batch_destroy(b);
            return fail(IGAP4H_EINVAL, "batch instance %d: NULL context or array", i);
        }
        if (c->physics != IGAP4H_PHYS_GE) {
            igap4h_batch_destroy(b);
            return fail(IGAP4H_EINVAL, "batch instance %d: batched assembly is implemented for phys/ge", i);
        }
        if (c->partitioned) {
            igap4h_batch_destroy(b);
            return fail(IGAP4H_EINVAL, "batch instance %d is a partitioned block (whole patches only)", i);
        }
        b->nqmax = std::max(b->nqmax, c->nquad);
        if (i == 0) b->device = c->device;
        else if (c->device != b->device) {
            igap4h_batch_destroy(b);
            return fail(IGAP4H_EINVAL, "batch instance %d lives on device %d, instance 0 on %d", i, c->device, b->device);
        }
        if (!c->d_Re) {
            int r = dev_alloc(&c->d_Re, (size_t)c->mesh.nelem() * 81);
            if (r) {
                igap4h_batch_destroy(b);
                return r;
            }
        }
        BatchInst I;
        I.m = mesh_args(c);
        I.Uh = d_U_hist[i];
        I.coef = c->d_coef;
        I.rowptr = c->d_rowptr;
        batch_bc_fields(I, c);
        I.Re = c->d_Re;
        I.R = d_R[i];
        I.vals = d_vals[i];
        I.nnz = c->nnz;
        b->h_inst.push_back(I);
        b->ctx.push_back(c);
        b->bc_gen.push_back(c->bc_gen);
        b->eoff[i + 1] = b->eoff[i] + c->mesh.nelem();
        for (int col = 0; col < 27; col++) {
            int cx = col / 9, cy = (col / 3) % 3, cz = col % 3;
            long long mx = (c->mesh.ne[0] - cx + 2) / 3, my = (c->mesh.ne[1] - cy + 2) / 3, mz = (c->mesh.ne[2] - cz + 2) / 3;
            long long cnt = (mx > 0 && my > 0 && mz > 0) ? mx * my * mz : 0;
            b->coff[(size_t)col * (n + 1) + i + 1] = b->coff[(size_t)col * (n + 1) + i] + cnt;
        }
        b->goff[i + 1] = b->goff[i] + (c->mesh.ndof() + 255) / 256;
        b->foff[i + 1] = b->foff[i] + ((b->vals_f32 ? c->nnz : 2 * c->nnz) + kFillWords - 1) / kFillWords;
        b->doff[i + 1] = b->doff[i] + ((long long)c->drows.size() + 255) / 256;
    }
    if (b->eoff[n] > 0x7fffffffLL || b->foff[n] > 0x7fffffffLL) {
        igap4h_batch_destroy(b);
        return fail(IGAP4H_EINVAL, "batch has %lld elements; at most 2^31 - 1", b->eoff[n]);
    }
    hipError_t e;
    if ((e = hipSetDevice(b->device)) != hipSuccess || (e = dev_malloc((void **)&b->d_inst, sizeof(BatchInst) * n)) != hipSuccess ||
        (e = dev_malloc((void **)&b->d_off, sizeof(long long) * 31 * (n + 1))) != hipSuccess ||
        (e = hipMemcpy(b->d_inst, b->h_inst.data(), sizeof(BatchInst) * n, hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(b->d_off, b->eoff.data(), sizeof(long long) * (n + 1), hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(b->d_off + (n + 1), b->coff.data(), sizeof(long long) * 27 * (n + 1), hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(b->d_off + (size_t)28 * (n + 1), b->goff.data(), sizeof(long long) * (n + 1), hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(b->d_off + (size_t)29 * (n + 1), b->foff.data(), sizeof(long long) * (n + 1), hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(b->d_off + (size_t)30 * (n + 1), b->doff.data(), sizeof(long long) * (n + 1), hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipStreamCreateWithFlags(&b->cap, hipStreamNonBlocking)) != hipSuccess) {
        igap4h_batch_destroy(b);
        return fail(IGAP4H_EHIP, "batch set-up: %s", hipGetErrorString(e));
    }
    *out = b;
    return 0;
}

extern "C" int igap4h_batch_sizes(const igap4h_batch *b, int64_t *nelem, int *qc)
{
    if (!b) return fail(IGAP4H_EINVAL, "NULL batch");
    if (nelem) *nelem = b->eoff[b->n];
    if (qc) *qc = b->qc;
    return 0;
}

extern "C" int igap4h_batch_set_variant(igap4h_batch *b, int qc)
{
    if (!b) return fail(IGAP4H_EINVAL, "NULL batch");
    if (!batch_variant_ok(b->prec, qc, b->nqmax))
        return fail(IGAP4H_EINVAL, "variant %d is not in the tuning space of this batch", qc);
    if (qc != b->qc) batch_drop_graph(b);
    b->qc = qc;
    return 0;
}

// Autotuner: time the batched tangent of every variant (reps launches each,
// HIP events on the batch's own stream), keep the fastest.  ms[k] receives
// the mean time of candidate k, qcs[k] the variant (kBatchNV entries, 0 past
// the space).
extern "C" int igap4h_batch_tune(igap4h_batch *b, int reps, int *best_qc, double *ms, int *qcs)
{
    if (!b) return fail(IGAP4H_EINVAL, "NULL batch");
    if (reps < 1) reps = 3;
    HIPCHK(hipSetDevice(b->device));
    int rc0 = batch_refresh(b);
    if (rc0) return rc0;
    const int *cand = kBatchVariants[b->prec == IGAP4H_FP32];
    if (ms)
        for (int k = 0; k < kBatchNV; k++) ms[k] = 0.0;
    if (qcs)
        for (int k = 0; k < kBatchNV; k++) qcs[k] = 0;
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    double best = 1e300;
    int bq = b->qc, rc = 0;
    for (int k = 0; k < kBatchNV && cand[k] && !rc; k++) {
        if (!batch_variant_ok(b->prec, cand[k], b->nqmax)) continue;
        if ((rc = batch_tangent(b, cand[k], b->cap))) break;  // warm-up
        if (hipEventRecord(e0, b->cap) != hipSuccess) { rc = fail(IGAP4H_EHIP, "hipEventRecord failed"); break; }
        for (int r = 0; r < reps && !rc; r++) rc = batch_tangent(b, cand[k], b->cap);
        if (rc) break;
        float t = 0.f;
        if (hipEventRecord(e1, b->cap) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
            hipEventElapsedTime(&t, e0, e1) != hipSuccess) {
            rc = fail(IGAP4H_EHIP, "batch tuning timing failed");
            break;
        }
        double mean = t / reps;
        if (ms) ms[k] = mean;
        if (qcs) qcs[k] = cand[k];
        if (mean < best) {
            best = mean;
            bq = cand[k];
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    if (bq != b->qc) batch_drop_graph(b);
    b->qc = bq;
    if (best_qc) *best_qc = bq;
    return 0;
}

// R and K of every instance: one graph replay (captured at the first call
// and after a variant change).  U_hist, R and vals are the arrays given at
// create; calls that overlap on one batch race.
extern "C" int igap4h_batch_assemble(igap4h_batch *b, void *stream)
{
    if (!b) return fail(IGAP4H_EINVAL, "NULL batch");
    HIPCHK(hipSetDevice(b->device));
    int rc0 = batch_refresh(b);
    if (rc0) return rc0;
    if (!b->exec) {
        HIPCHK(hipStreamBeginCapture(b->cap, hipStreamCaptureModeRelaxed));
        int rc = batch_sequence(b, b->cap);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(b->cap, &g);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (e != hipSuccess) return fail(IGAP4H_EHIP, "batch capture: %s", hipGetErrorString(e));
        e = hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) {
            b->exec = nullptr;
            return fail(IGAP4H_EHIP, "batch graph instantiation: %s", hipGetErrorString(e));
        }
    }
    HIPCHK(hipGraphLaunch(b->exec, (hipStream_t)stream));
    return 0;
}
