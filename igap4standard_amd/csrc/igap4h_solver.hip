// Newton-Krylov driver on one GPU: the counterpart of the PETSc SNES/KSP set-up
// of src/snes_init.c:8-47 and the outer loop of src/snes_solve.c:10-36, for the
// residual and tangent of this library.  C-ABI in include/igap4h.h.
//
// Matrix: the AIJ pattern of the assembly (rows (a0,a1,a2,i), columns the
// 3 dofs of every node b with |a_d - b_d| <= 2, ascending).  The three rows of
// a node share one column list of length L = 3 w0 w1 w2 and sit back to back
// (rowptr[3n+i] = rowptr[3n] + i L), so column indices are computed, not read:
// an SpMV streams only the values (8 or 4 bytes per nonzero).
//
// Every kernel here is one wave per node (row work) or a grid-stride vector
// pass; reductions write per-block partials that a second kernel sums in a
// fixed order, so results do not depend on scheduling.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <future>

#include "igap4h_internal.hpp"

using namespace igap4h;

namespace {

constexpr int NRB = 1024;  // blocks of a reduction pass = partials per sum
constexpr int JT = 8;      // vectors per tile of the block dot product

struct SpArgs {
    int nb[3];
    long long nnode;
    const long long *rowptr;
};

__device__ __forceinline__ double wave_sum(double v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// v of lane 15 - l (row_mirror) or 7 - l (row_half_mirror) of the lane's
// 16-lane row, or of lane l ^ 2, l ^ 1 (quad_perm): DPP moves, no LDS.
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v)
{
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}
constexpr int DPP_MIRROR = 0x140, DPP_HALF_MIRROR = 0x141, DPP_XOR2 = 0x4E, DPP_XOR1 = 0xB1;

// The three wave sums of s0, s1, s2, transposed: after the exchange over 32
// lanes the lower half carries (s0, s1), the upper half (s2, 0); after the
// one over 16 each 16-lane row carries one of s0, s1, s2, 0, which four DPP
// steps (mirror, half mirror, xor 2, xor 1: every lane of the row gathers
// all 16 once) sum within the row.  Returns the sum of s_q in every lane of
// row q = lane / 16 (row 3: 0).  2 LDS exchanges instead of 18.
__device__ __forceinline__ double wave_sum3(double s0, double s1, double s2, int lane)
{
    const bool h5 = lane & 32, h4 = lane & 16;
    const double r0 = __shfl_xor(h5 ? s0 : s2, 32, 64), r1 = __shfl_xor(h5 ? s1 : 0.0, 32, 64);
    const double A = (h5 ? s2 : s0) + r0, B = (h5 ? 0.0 : s1) + r1;
    double c = (h4 ? B : A) + __shfl_xor(h4 ? A : B, 16, 64);
    c += dpp_f64<DPP_MIRROR>(c);
    c += dpp_f64<DPP_HALF_MIRROR>(c);
    c += dpp_f64<DPP_XOR2>(c);
    c += dpp_f64<DPP_XOR1>(c);
    return c;
}

// node -> grid coordinates and the stencil box [lo, lo + w) per direction
__device__ __forceinline__ void node_box(const SpArgs &s, long long node, int a[3], int lo[3], int w[3])
{
    a[2] = (int)(node % s.nb[2]);
    a[1] = (int)((node / s.nb[2]) % s.nb[1]);
    a[0] = (int)(node / ((long long)s.nb[2] * s.nb[1]));
    for (int d = 0; d < 3; d++) {
        lo[d] = max(0, a[d] - 2);
        w[d] = min(s.nb[d] - 1, a[d] + 2) - lo[d] + 1;
    }
}

// acc[i] = sum_k K[3n+i, col_k] x[col_k] over the node's rows (one wave).
// A row has at most 375 entries, so each lane has at most 6: all loads are
// issued before the first product (one memory round trip per node).
template <class VT>
__device__ __forceinline__ void node_rows(const SpArgs &s, long long node, const VT *__restrict__ vals,
                                          const double *__restrict__ x, int lane, double acc[3])
{
    int a[3], lo[3], w[3];
    node_box(s, node, a, lo, w);
    const long long base = s.rowptr[3 * node];
    const int L = 3 * w[0] * w[1] * w[2];
    const float r2 = 1.0f / (float)w[2], r1 = 1.0f / (float)w[1];
    double xv[6];
    VT v0[6], v1[6], v2[6];
#pragma unroll
    for (int it = 0; it < 6; it++) {
        const int k = lane + 64 * it;
        xv[it] = 0.0;
        v0[it] = v1[it] = v2[it] = (VT)0;
        if (k < L) {
            int t = (k * 43691) >> 17;  // k / 3 for k < 131072
            int j = k - 3 * t;
            int t2 = (int)(((float)t + 0.5f) * r2);  // t / w2 (exact: t < 128, w2 <= 5)
            int o2 = t - t2 * w[2];
            int o0 = (int)(((float)t2 + 0.5f) * r1);
            int o1 = t2 - o0 * w[1];
            long long cn = ((long long)(lo[0] + o0) * s.nb[1] + (lo[1] + o1)) * s.nb[2] + (lo[2] + o2);
            xv[it] = x[3 * cn + j];
            v0[it] = vals[base + k];
            v1[it] = vals[base + L + k];
            v2[it] = vals[base + 2 * L + k];
        }
    }
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
#pragma unroll
    for (int it = 0; it < 6; it++) {
        s0 += (double)v0[it] * xv[it];
        s1 += (double)v1[it] * xv[it];
        s2 += (double)v2[it] * xv[it];
    }
    acc[0] = wave_sum(s0);
    acc[1] = wave_sum(s1);
    acc[2] = wave_sum(s2);
}

// Work-groups are dispatched to the 8 XCDs round robin (block b -> XCD b mod 8),
// each XCD with its own L2.  This bijection gives XCD k a contiguous range of
// block indices, so the x entries that neighbouring nodes share stay in one L2.
__device__ __forceinline__ long long xcd_remap(long long b, long long nblk)
{
    const long long q = nblk >> 3, rem = nblk & 7, x = b & 7, i = b >> 3;
    return x * q + (x < rem ? x : rem) + i;
}

// ---------------------------------------------------------------- line-staged row pass
// One work-group per segment [z0, z0 + seglen) of a z-line (a0, a1): the x
// entries of the 5 x 5 neighbouring lines over [z0 - 2, z0 + seglen + 2) are
// staged in LDS once (contiguous runs, coalesced), then the four waves take
// the segment's nodes in turn (LW waves), one node per wave, streaming the node's three
// rows and reading x from LDS.  This replaces 375 scattered L2 reads per node
// by ~5 per node and keeps the wave's memory traffic to the matrix values.
constexpr int LW = 8;  // waves (nodes in flight) per line work-group

struct LineGeom {
    int nseg, seglen;  // segments per line, nodes per segment
};

static LineGeom line_geom(const int nb[3])
{
    LineGeom g;
    g.nseg = (nb[2] + 63) / 64;
    g.seglen = (nb[2] + g.nseg - 1) / g.nseg;
    return g;
}
static size_t line_lds_bytes(const LineGeom &g) { return (size_t)25 * (g.seglen + 4) * 3 * sizeof(double); }

// sum of the stencil widths w2(z') = min(n - 1, z' + 2) - max(0, z' - 2) + 1
// over z' < z, for n >= 3 nodes in direction 2: the row offset of node z of
// a line from the line's first node is 9 w0 w1 W(z), with w0, w1 the line's
// widths in directions 0 and 1 (three rows of 3 w0 w1 w2 entries per node)
__device__ __forceinline__ long long line_W(int z, int n)
{
    const long long m = min(z, n - 2), zz = z;
    const long long lo = z >= 3 ? (zz - 3) * (zz - 2) / 2 : 0;
    return m * (m - 1) / 2 + 3 * m + (zz - m) * n - lo;
}

// stage x; returns the segment's node range and the row offset of the
// line's first node.  sx[(l0*5 + l1) * SW3 + (z - zb) * 3 + j]
__device__ __forceinline__ void line_stage(const SpArgs &s, const LineGeom g, long long blk, const double *__restrict__ x,
                                           double *sx, int &a0, int &a1, int &z0, int &z1, long long &rbase)
{
    const int seg = (int)(blk % g.nseg);
    const long long line = blk / g.nseg;
    a1 = (int)(line % s.nb[1]);
    a0 = (int)(line / s.nb[1]);
    rbase = s.rowptr[3 * line * s.nb[2]];  // in flight while x is staged
    z0 = seg * g.seglen;
    z1 = min(s.nb[2], z0 + g.seglen);
    const int zb = z0 - 2, SW3 = 3 * (g.seglen + 4);
    const int zl = max(0, z0 - 2), zh = min(s.nb[2], z1 + 2), nz3 = 3 * (zh - zl);
    for (int t = threadIdx.x; t < 25 * nz3; t += blockDim.x) {
        const int l = t / nz3, r = t - l * nz3;
        const int b0 = a0 + l / 5 - 2, b1 = a1 + l % 5 - 2;
        if (b0 < 0 || b0 >= s.nb[0] || b1 < 0 || b1 >= s.nb[1]) continue;
        sx[l * SW3 + (zl - zb) * 3 + r] = x[3 * (((long long)b0 * s.nb[1] + b1) * s.nb[2] + zl) + r];
    }
    __syncthreads();
}

// LDS index of stencil column k (k < L) of node (a0, a1, z)
__device__ __forceinline__ int line_col(const int a[3], const int lo[3], const int w[3], float r1, float r2, int zb, int SW3, int k)
{
    int t = (k * 43691) >> 17;  // k / 3 for k < 131072
    int j = k - 3 * t;
    int t2 = (int)(((float)t + 0.5f) * r2);  // t / w2 (exact: t < 128, w2 <= 5)
    int o2 = t - t2 * w[2];
    int o0 = (int)(((float)t2 + 0.5f) * r1);
    int o1 = t2 - o0 * w[1];
    return ((lo[0] + o0 - a[0] + 2) * 5 + (lo[1] + o1 - a[1] + 2)) * SW3 + (lo[2] + o2 - zb) * 3 + j;
}

__device__ __forceinline__ void line_box(const SpArgs &s, int a0, int a1, int z, int a[3], int lo[3], int w[3])
{
    a[0] = a0;
    a[1] = a1;
    a[2] = z;
    for (int d = 0; d < 3; d++) {
        lo[d] = max(0, a[d] - 2);
        w[d] = min(s.nb[d] - 1, a[d] + 2) - lo[d] + 1;
    }
}

// y = K x (RES = 0) or y = b - K x (RES = 1) on the AIJ values; with DOT,
// partial[block] = sum over the block's rows of x . y (CG's p.Kp).  The values
// are read once per pass (non-temporal loads, so they do not evict the x
// entries other blocks stage from L2).
template <class VT, int RES, int DOT>
__global__ __launch_bounds__(LW * 64) void k_spmv_line(SpArgs s, LineGeom g, const VT *__restrict__ vals, const double *__restrict__ x,
                                                   const double *__restrict__ b, double *__restrict__ y,
                                                   double *__restrict__ partial)
{
    extern __shared__ double sx[];
    __shared__ double sd[LW];
    // the wave index in an SGPR: the node loop, row addresses and stencil
    // widths are then scalar work
    const int lane = threadIdx.x & 63, wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    int a0, a1, z0, z1;
    long long rbase;
    line_stage(s, g, xcd_remap(blockIdx.x, gridDim.x), x, sx, a0, a1, z0, z1, rbase);
    const long long w01 = 9LL * (min(s.nb[0] - 1, a0 + 2) - max(0, a0 - 2) + 1) * (min(s.nb[1] - 1, a1 + 2) - max(0, a1 - 2) + 1);
    const int zb = z0 - 2, SW3 = 3 * (g.seglen + 4);
    // LDS index of the lane's columns for a node whose 5 x 5 x 5 box lies
    // inside the mesh, at z = z0 (+ 3 (z - z0) for the others)
    const bool in01 = a0 >= 2 && a0 + 2 < s.nb[0] && a1 >= 2 && a1 + 2 < s.nb[1];
    int ioff[6];
#pragma unroll
    for (int it = 0; it < 6; it++) {
        const int k = lane + 64 * it, t = k / 3, j = k - 3 * t;
        ioff[it] = ((t / 25) * 5 + (t / 5) % 5) * SW3 + (t % 5) * 3 + j;
    }
    double dsum = 0.0;
    const int w0 = min(s.nb[0] - 1, a0 + 2) - max(0, a0 - 2) + 1, w1 = min(s.nb[1] - 1, a1 + 2) - max(0, a1 - 2) + 1;
    // a node's three rows; FP32 values are issued one node ahead of their
    // contraction (two register sets, the node loop unrolled by two), so a
    // wave keeps the next node's values in flight while it reduces the current
    // one (203.6 against 206.0 us at 64^3; the same for FP64 values took
    // 375 against 371 us)
    struct Rows {
        VT v0[6], v1[6], v2[6];
    };
    auto load = [&](int z, Rows &R) {
        const int L = 3 * w0 * w1 * (min(s.nb[2] - 1, z + 2) - max(0, z - 2) + 1);
        const VT *row = vals + rbase + w01 * line_W(z, s.nb[2]);
#pragma unroll
        for (int it = 0; it < 6; it++) {
            const int k = lane + 64 * it;
            R.v0[it] = R.v1[it] = R.v2[it] = (VT)0;
            if (k < L) {
                R.v0[it] = __builtin_nontemporal_load(row + k);
                R.v1[it] = __builtin_nontemporal_load(row + L + k);
                R.v2[it] = __builtin_nontemporal_load(row + 2 * L + k);
            }
        }
    };
    auto contract = [&](int z, const Rows &R) {
        int a[3], lo[3], w[3];
        line_box(s, a0, a1, z, a, lo, w);
        const long long node = ((long long)a0 * s.nb[1] + a1) * s.nb[2] + z;
        const int L = 3 * w[0] * w[1] * w[2];
        int cx[6];
        if (in01 && z >= 2 && z + 2 < s.nb[2]) {
#pragma unroll
            for (int it = 0; it < 6; it++) cx[it] = ioff[it] + 3 * (z - z0);
        } else {
            const float r2 = 1.0f / (float)w[2], r1 = 1.0f / (float)w[1];
#pragma unroll
            for (int it = 0; it < 6; it++) cx[it] = line_col(a, lo, w, r1, r2, zb, SW3, min(lane + 64 * it, L - 1));
        }
        double s0 = 0.0, s1 = 0.0, s2 = 0.0;
#pragma unroll
        for (int it = 0; it < 6; it++) {
            const double xv = lane + 64 * it < L ? sx[cx[it]] : 0.0;
            s0 += (double)R.v0[it] * xv;
            s1 += (double)R.v1[it] * xv;
            s2 += (double)R.v2[it] * xv;
        }
        const double c = wave_sum3(s0, s1, s2, lane);
        if ((lane & 15) == 0 && lane < 48) {  // lanes 0, 16, 32: rows 3 node + 0, 1, 2
            const long long r = 3 * node + (lane >> 4);
            const double v = RES ? b[r] - c : c;
            y[r] = v;
            if (DOT) dsum += x[r] * v;
        }
    };
    if constexpr (sizeof(VT) == 4) {
        Rows A, B;
        int z = z0 + wv;
        if (z < z1) load(z, A);
        for (; z < z1; z += 2 * LW) {
            if (z + LW < z1) load(z + LW, B);
            contract(z, A);
            if (z + 2 * LW < z1) load(z + 2 * LW, A);
            if (z + LW < z1) contract(z + LW, B);
        }
    } else {
        for (int z = z0 + wv; z < z1; z += LW) {
            Rows A;
            load(z, A);
            contract(z, A);
        }
    }
    if (DOT) {
        dsum = wave_sum(dsum);
        if (lane == 0) sd[wv] = dsum;
        __syncthreads();
        if (threadIdx.x == 0) {
            double t = 0.0;
            for (int i = 0; i < LW; i++) t += sd[i];
            partial[blockIdx.x] = t;
        }
    }
}

// ---------------------------------------------------------------- 16-bit smoothing copies
// The copy is interleaved for one node per lane: nodes are taken in packs of
// 64 consecutive indices, and word u (8 values of 16 bits) of node 64 p + l
// sits at v4[(p * NU16 + u) * 64 + l], so a wave's load of word u is 1 KiB
// contiguous.  The words come in 5 phases of NUP, one per neighbour plane
// d0 = -2..2; in phase d0, value v = 9 o + 3 i + j (v < 225) is
// K[(n, i), (n + d, j)] with d = (d0, o / 5 - 2, o % 5 - 2) over the full
// stencil; entries outside the mesh and the phase's last 7 values are zero.
// bfloat16 = the upper half of the FP32 bit pattern, rounded to nearest
// even.  binary16 (F16 = 1): row r is divided by rs[r] = 2^e with
// max_k |K_rk| < 2^e first, exactly, so every entry lies in (-1, 1).
constexpr int NUP = 29;          // words per phase: ceil(225 / 8)
constexpr int NU16 = 5 * NUP;    // words per node
constexpr int CH_NT = 256;       // k_cheb16 threads (nodes) per block

// x slab of one phase of k_cheb16, in doubles: 3 components x (CH_NT + 4 nb2 + 4) nodes
inline size_t cheb16_lds_bytes(int nb2) { return (size_t)3 * (CH_NT + 4 * nb2 + 4) * sizeof(double); }

__device__ __forceinline__ unsigned short f32_to_bf16(float f)
{
    unsigned u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}

__device__ __forceinline__ double wave_max(double v)
{
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}

// One wave per node, PK_NODES = 8 consecutive nodes per block: each word u
// of the 8 nodes is one 128-byte line of the copy.  The waves read their
// node's three AIJ rows coalesced and place every value in the block's
// lines in LDS (zero-filled first); the lines are then written out whole.
// F16: rs[3 n + r] = 2^e > max_k |K_(3n+r),k| (1 for a zero row) from a
// wave reduction first.
constexpr int PK_NODES = 8;
template <class VT, int F16>
__global__ __launch_bounds__(64 * PK_NODES) void k_pack_16(SpArgs s, const VT *__restrict__ vals, double *__restrict__ rs,
                                                           uint4 *__restrict__ v4)
{
    __shared__ uint4 sl[NU16 * PK_NODES];  // [u][node in block]
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const long long n0 = (long long)blockIdx.x * PK_NODES, node = n0 + wv;
    for (int k = threadIdx.x; k < NU16 * PK_NODES; k += blockDim.x) sl[k] = make_uint4(0u, 0u, 0u, 0u);
    __syncthreads();
    unsigned short *po = (unsigned short *)sl + wv * 8;  // half e of word u of this node: po[u * 64 + e]
    for (int r = 0; r < 3 && node < s.nnode; r++) {
        int a[3], lo[3], w[3];
        node_box(s, node, a, lo, w);
        const int L = 3 * w[0] * w[1] * w[2];
        const long long base = s.rowptr[3 * node];
        const VT *row = vals + base + (long long)r * L;
        double inv = 1.0;
        if (F16) {
            double m = 0.0;
            for (int k = lane; k < L; k += 64) m = fmax(m, fabs((double)row[k]));
            m = wave_max(m);
            int e = 0;
            (void)frexp(m, &e);  // m < 2^e
            const double sc = m > 0.0 ? ldexp(1.0, e) : 1.0;
            inv = 1.0 / sc;  // exact (power of two)
            if (lane == 0) rs[3 * node + r] = sc;
        }
        for (int k = lane; k < L; k += 64) {
            const int t = k / 3, j = k - 3 * t, t2 = t / w[2], o2 = t - t2 * w[2], o0 = t2 / w[1], o1 = t2 - o0 * w[1];
            const int ph = lo[0] + o0 - a[0] + 2, ol = (lo[1] + o1 - a[1] + 2) * 5 + (lo[2] + o2 - a[2] + 2);
            const int v = 9 * ol + 3 * r + j;
            unsigned short h;
            if (F16) {
                _Float16 f = (_Float16)((double)row[k] * inv);
                h = __builtin_bit_cast(unsigned short, f);
            } else {
                h = f32_to_bf16((float)row[k]);
            }
            po[(ph * NUP + (v >> 3)) * 64 + (v & 7)] = h;
        }
    }
    __syncthreads();
    // node n0 + q's word u -> v4[(p * NU16 + u) * 64 + l], l = (n0 + q) mod 64 (n0 is a multiple of 8)
    const long long pb = (n0 >> 6) * NU16 * 64 + (n0 & 63);
    for (int k = threadIdx.x; k < NU16 * PK_NODES; k += blockDim.x) {
        const int u = k / PK_NODES, q = k - u * PK_NODES;
        v4[pb + (long long)u * 64 + q] = sl[k];
    }
}

template <int F16>
__device__ __forceinline__ float half_at(const uint4 &w, int e)
{
    const unsigned u = (e >> 1) == 0 ? w.x : (e >> 1) == 1 ? w.y : (e >> 1) == 2 ? w.z : w.w;
    if (F16) return (float)__builtin_bit_cast(_Float16, (unsigned short)((e & 1) ? u >> 16 : u & 0xFFFFu));
    return __uint_as_float((e & 1) ? u & 0xFFFF0000u : u << 16);
}

// acc[i] += the node's 3 x 3 blocks at in-plane offsets O0 .. O0 + NO - 1 of
// one phase times x from the slab; w points at the group's first word (a
// stream read once: non-temporal), sx at the lane's node in the slab.
template <int NO, int O0, int F16>
__device__ __forceinline__ void sm16_group(const uint4 *__restrict__ w, const double *sx, int Z, int SL, double acc[3])
{
    typedef unsigned u4v __attribute__((ext_vector_type(4)));
    constexpr int NW = (9 * NO + 7) / 8;
    uint4 wd[NW];
#pragma unroll
    for (int u = 0; u < NW; u++) {
        const u4v t4 = __builtin_nontemporal_load((const u4v *)(w + u * 64));
        wd[u] = make_uint4(t4.x, t4.y, t4.z, t4.w);
    }
    double xv[NO][3];
#pragma unroll
    for (int k = 0; k < NO; k++) {
        const int o = O0 + k;
        const double *xs = sx + (o / 5) * Z + (o % 5);
#pragma unroll
        for (int j = 0; j < 3; j++) xv[k][j] = xs[j * SL];
    }
#pragma unroll
    for (int v = 0; v < 9 * NO; v++) acc[(v % 9) / 3] += (double)half_at<F16>(wd[v >> 3], v & 7) * xv[v / 9][v % 3];
}

// Fine-level Chebyshev pass on the 16-bit copy, one lane per node (rows 3n..3n+2):
//   MODE 0: d = c2 Dinv K~ x                                  (power iteration)
//   MODE 1: r = b - K~ x, d = c2 Dinv r                       (start from x)
//   MODE 2: r -= K~ din, x += din, and if newd: dout = c1 din + c2 Dinv r
// For each phase d0 the block stages the x entries of nodes
// [n0 + d0 P - 2 Z - 2, n0 + CH_NT + d0 P + 2 Z + 2) (P = nb1 nb2, Z = nb2, n0 the
// block's first node) in LDS, one array per component; nodes outside the
// mesh stage zeros.  Blocks take XCD-contiguous node ranges.
template <int MODE, int F16>
__global__ __launch_bounds__(CH_NT) void k_cheb16(SpArgs s, const uint4 *__restrict__ v4, const double *__restrict__ rs,
                                                  const double *__restrict__ Dinv, const double *__restrict__ xin,
                                                  const double *__restrict__ b, double *__restrict__ r, double *__restrict__ x,
                                                  double *__restrict__ dout, double c1, double c2, int newd)
{
    extern __shared__ double sx[];
    const long long n0 = xcd_remap(blockIdx.x, gridDim.x) * CH_NT;
    const long long t = n0 + threadIdx.x;
    const int l = threadIdx.x & 63;
    const long long p = t >> 6;
    const int Z = s.nb[2], P = s.nb[1] * s.nb[2], SL = CH_NT + 4 * Z + 4;
    const long long nx = 3 * s.nnode;
    const uint4 *w = v4 + p * NU16 * 64 + l;
    double acc[3] = {0.0, 0.0, 0.0};
#pragma unroll 1
    for (int ph = 0; ph < 5; ph++) {
        __syncthreads();
        const long long g0 = 3 * (n0 + (long long)(ph - 2) * P - 2 * Z - 2);
        for (int k = threadIdx.x; k < 3 * SL; k += CH_NT) {
            const long long gi = g0 + k;
            const int m = k / 3;
            sx[(k - 3 * m) * SL + m] = gi >= 0 && gi < nx ? xin[gi] : 0.0;
        }
        __syncthreads();
        const uint4 *wp = w + ph * NUP * 64;
        const double *xs = sx + threadIdx.x;
        sm16_group<8, 0, F16>(wp, xs, Z, SL, acc);
        sm16_group<8, 8, F16>(wp + 9 * 64, xs, Z, SL, acc);
        sm16_group<8, 16, F16>(wp + 18 * 64, xs, Z, SL, acc);
        sm16_group<1, 24, F16>(wp + 27 * 64, xs, Z, SL, acc);
    }
    if (t >= s.nnode) return;
    if (F16)
        for (int i = 0; i < 3; i++) acc[i] *= rs[3 * t + i];
    double q[3];
    for (int i = 0; i < 3; i++) {
        if (MODE == 0) {
            q[i] = acc[i];
            continue;
        }
        if (MODE == 1) {
            q[i] = b[3 * t + i] - acc[i];
        } else {
            q[i] = r[3 * t + i] - acc[i];
            x[3 * t + i] += xin[3 * t + i];
        }
        r[3 * t + i] = q[i];
    }
    if (MODE != 2 || newd) {
        const double *B = Dinv + 9 * t;
        for (int i = 0; i < 3; i++) {
            const double dq = B[3 * i] * q[0] + B[3 * i + 1] * q[1] + B[3 * i + 2] * q[2];
            dout[3 * t + i] = MODE == 2 ? c1 * xin[3 * t + i] + c2 * dq : c2 * dq;
        }
    }
}

__device__ __forceinline__ void inv3(const double Ain[9], double B[9])
{
    // a dof whose row is empty (coarse functions supported only on Dirichlet
    // dofs) gets a unit diagonal
    double A[9];
    for (int k = 0; k < 9; k++) A[k] = Ain[k];
    for (int i = 0; i < 3; i++)
        if (A[3 * i] == 0.0 && A[3 * i + 1] == 0.0 && A[3 * i + 2] == 0.0) A[4 * i] = 1.0;
    double c00 = A[4] * A[8] - A[5] * A[7], c01 = A[5] * A[6] - A[3] * A[8], c02 = A[3] * A[7] - A[4] * A[6];
    double det = A[0] * c00 + A[1] * c01 + A[2] * c02;
    if (det == 0.0 || !isfinite(det)) {
        for (int k = 0; k < 9; k++) B[k] = (k % 4 == 0) ? 1.0 : 0.0;
        return;
    }
    double id = 1.0 / det;
    B[0] = c00 * id;
    B[1] = (A[2] * A[7] - A[1] * A[8]) * id;
    B[2] = (A[1] * A[5] - A[2] * A[4]) * id;
    B[3] = c01 * id;
    B[4] = (A[0] * A[8] - A[2] * A[6]) * id;
    B[5] = (A[2] * A[3] - A[0] * A[5]) * id;
    B[6] = c02 * id;
    B[7] = (A[1] * A[6] - A[0] * A[7]) * id;
    B[8] = (A[0] * A[4] - A[1] * A[3]) * id;
}

// nodal 3x3 diagonal blocks of K and their inverses (9 doubles per node, row-major)
template <class VT>
__global__ void k_block_diag(SpArgs s, const VT *__restrict__ vals, double *__restrict__ Dblk, double *__restrict__ Dinv)
{
    long long node = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (node >= s.nnode) return;
    int a[3], lo[3], w[3];
    node_box(s, node, a, lo, w);
    long long L = 3LL * w[0] * w[1] * w[2];
    long long off = 3LL * (((a[0] - lo[0]) * w[1] + (a[1] - lo[1])) * w[2] + (a[2] - lo[2]));
    long long base = s.rowptr[3 * node];
    double A[9], B[9];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) A[3 * i + j] = (double)vals[base + i * L + off + j];
    inv3(A, B);
    for (int k = 0; k < 9; k++) {
        Dblk[9 * node + k] = A[k];
        Dinv[9 * node + k] = B[k];
    }
}

// z = Dinv r per node
__global__ void k_bjacobi(long long nnode, const double *__restrict__ Dinv, const double *__restrict__ r,
                          double *__restrict__ z)
{
    long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= nnode) return;
    const double *B = Dinv + 9 * n;
    double r0 = r[3 * n], r1 = r[3 * n + 1], r2 = r[3 * n + 2];
    z[3 * n] = B[0] * r0 + B[1] * r1 + B[2] * r2;
    z[3 * n + 1] = B[3] * r0 + B[4] * r1 + B[5] * r2;
    z[3 * n + 2] = B[6] * r0 + B[7] * r1 + B[8] * r2;
}

// One colour of a block Gauss-Seidel sweep.  v holds the sweep's output so
// far and is zero on nodes not yet visited, including this colour's, so the
// full row product K v is the off-diagonal sum over visited colours.
//   forward  (BWD = 0): v_n = Dinv_n (r_n - (K v)_n)
//   backward (BWD = 1): v_n = Dinv_n (D_n y_n - (K v)_n)
// Nodes of one colour (a_d = c_d mod 3) share no matrix entry.
template <class VT, int BWD>
__global__ __launch_bounds__(256) void k_sgs_colour(SpArgs s, int c0, int c1, int c2, int m1, int m2, long long ncol,
                                                    const VT *__restrict__ vals, const double *__restrict__ Dblk,
                                                    const double *__restrict__ Dinv, const double *__restrict__ rhs,
                                                    double *v)
{
    const int lane = threadIdx.x & 63;
    const long long t = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (t >= ncol) return;
    long long a2 = c2 + 3 * (t % m2), a1 = c1 + 3 * ((t / m2) % m1), a0 = c0 + 3 * (t / ((long long)m1 * m2));
    long long node = (a0 * s.nb[1] + a1) * s.nb[2] + a2;
    double acc[3];
    node_rows<VT>(s, node, vals, v, lane, acc);
    if (lane == 0) {
        double q[3];
        const double *D = Dblk + 9 * node;
        for (int i = 0; i < 3; i++) {
            double rh = BWD ? D[3 * i] * rhs[3 * node] + D[3 * i + 1] * rhs[3 * node + 1] + D[3 * i + 2] * rhs[3 * node + 2]
                            : rhs[3 * node + i];
            q[i] = rh - acc[i];
        }
        const double *B = Dinv + 9 * node;
        for (int i = 0; i < 3; i++) v[3 * node + i] = B[3 * i] * q[0] + B[3 * i + 1] * q[1] + B[3 * i + 2] * q[2];
    }
}

__global__ void k_zero(double *x, long long n)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        x[i] = 0.0;
}

// out = x + a y   (a from the host)
__global__ void k_waxpy(long long n, double a, const double *__restrict__ x, const double *__restrict__ y,
                        double *__restrict__ out)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        out[i] = x[i] + a * y[i];
}

// x *= 1 / sc[0]  (device scalar)
__global__ void k_scale_inv(long long n, const double *__restrict__ sc, double *__restrict__ x)
{
    double f = 1.0 / sc[0];
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        x[i] *= f;
}

// deterministic block sum of a per-thread value into partial[blockIdx.x]
__device__ __forceinline__ void block_partial(double v, double *sred, double *partial_slot)
{
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    v = wave_sum(v);
    if (lane == 0) sred[wv] = v;
    __syncthreads();
    if (threadIdx.x == 0) *partial_slot = ((sred[0] + sred[1]) + sred[2]) + sred[3];
    __syncthreads();
}

// partial dots of up to two pairs: partial[p * NRB + b]
__global__ __launch_bounds__(256) void k_dot2(long long n, const double *__restrict__ x0, const double *__restrict__ y0,
                                              const double *__restrict__ x1, const double *__restrict__ y1,
                                              double *__restrict__ partial)
{
    __shared__ double sred[4];
    double a = 0.0, b = 0.0;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        a += x0[i] * y0[i];
        if (x1) b += x1[i] * y1[i];
    }
    block_partial(a, sred, &partial[blockIdx.x]);
    if (x1) block_partial(b, sred, &partial[NRB + blockIdx.x]);
}

// out[j] = sum_b partial[j * nparts + b]; SQRT: out[j] = sqrt(...)
template <int SQRT>
__global__ __launch_bounds__(256) void k_finish(const double *__restrict__ partial, int nparts, double *__restrict__ out)
{
    __shared__ double sred[4];
    const double *p = partial + (long long)blockIdx.x * nparts;
    double v = 0.0;
    for (int i = threadIdx.x; i < nparts; i += blockDim.x) v += p[i];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    v = wave_sum(v);
    if (lane == 0) sred[wv] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = ((sred[0] + sred[1]) + sred[2]) + sred[3];
        out[blockIdx.x] = SQRT ? sqrt(s) : s;
    }
}

// partial[j * NRB + b] = sum over block b's range of V_j . w, j < k
// (V row-major k x ld).  Tiles of JT vectors share one read of w.
__global__ __launch_bounds__(256) void k_mdot(const double *__restrict__ V, long long ld, int k, const double *__restrict__ w,
                                              long long n, double *__restrict__ partial)
{
    __shared__ double sred[4][JT];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int j0 = 0; j0 < k; j0 += JT) {
        double acc[JT];
#pragma unroll
        for (int t = 0; t < JT; t++) acc[t] = 0.0;
        // four grid-stride steps per pass, all their loads issued together
        const long long stride = (long long)gridDim.x * blockDim.x;
        for (long long i0 = (long long)blockIdx.x * blockDim.x + threadIdx.x; i0 < n; i0 += 4 * stride) {
            double wi[4], vt[4][JT];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const long long i = i0 + u * stride;
                wi[u] = i < n ? w[i] : 0.0;
#pragma unroll
                for (int t = 0; t < JT; t++) vt[u][t] = (i < n && j0 + t < k) ? V[(long long)(j0 + t) * ld + i] : 0.0;
            }
#pragma unroll
            for (int u = 0; u < 4; u++)
#pragma unroll
                for (int t = 0; t < JT; t++) acc[t] += vt[u][t] * wi[u];
        }
#pragma unroll
        for (int t = 0; t < JT; t++) {
            double v = wave_sum(acc[t]);
            if (lane == 0) sred[wv][t] = v;
        }
        __syncthreads();
        if (threadIdx.x < JT && j0 + (int)threadIdx.x < k) {
            int t = threadIdx.x;
            partial[(long long)(j0 + t) * NRB + blockIdx.x] = ((sred[0][t] + sred[1][t]) + sred[2][t]) + sred[3][t];
        }
        __syncthreads();
    }
}

// w += alpha * sum_{j<k} V_j h[j]   (h in device memory)
__global__ __launch_bounds__(256) void k_mgemv(const double *__restrict__ V, long long ld, int k, const double *__restrict__ h,
                                               double alpha, double *__restrict__ w, long long n)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        double s = 0.0;
        // eight basis vectors per step, their loads issued together
        for (int j0 = 0; j0 < k; j0 += 8) {
            double v[8];
#pragma unroll
            for (int t = 0; t < 8; t++) v[t] = j0 + t < k ? V[(long long)(j0 + t) * ld + i] : 0.0;
#pragma unroll
            for (int t = 0; t < 8; t++)
                if (j0 + t < k) s += v[t] * h[j0 + t];
        }
        w[i] += alpha * s;
    }
}

// CG update with device scalars: alpha = sc[irz] / sc[ipq];
// x += alpha p, r -= alpha q; with BJ: z = Dinv r and partials of r.z, r.r
// (launched with NRB blocks, like k_dot2, so that every partial is written)
template <int BJ>
__global__ __launch_bounds__(256) void k_cg_update(long long nnode, const double *__restrict__ sc, int irz, int ipq,
                                                   const double *__restrict__ p, const double *__restrict__ q,
                                                   double *__restrict__ x, double *__restrict__ r,
                                                   const double *__restrict__ Dinv, double *__restrict__ z,
                                                   double *__restrict__ partial)
{
    __shared__ double sred[4];
    const double alpha = sc[irz] / sc[ipq];
    double rz = 0.0, rr = 0.0;
    for (long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x; n < nnode; n += (long long)gridDim.x * blockDim.x) {
        double rv[3];
        for (int i = 0; i < 3; i++) {
            long long g = 3 * n + i;
            x[g] += alpha * p[g];
            rv[i] = r[g] - alpha * q[g];
            r[g] = rv[i];
            rr += rv[i] * rv[i];
        }
        if (BJ) {
            const double *B = Dinv + 9 * n;
            for (int i = 0; i < 3; i++) {
                double zi = B[3 * i] * rv[0] + B[3 * i + 1] * rv[1] + B[3 * i + 2] * rv[2];
                z[3 * n + i] = zi;
                rz += rv[i] * zi;
            }
        }
    }
    block_partial(rz, sred, &partial[blockIdx.x]);
    block_partial(rr, sred, &partial[NRB + blockIdx.x]);
}

// p = z + beta p, beta = sc[inew] / sc[iold]
__global__ void k_cg_dir(long long n, const double *__restrict__ sc, int inew, int iold, const double *__restrict__ z,
                         double *__restrict__ p)
{
    const double beta = sc[inew] / sc[iold];
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        p[i] = z[i] + beta * p[i];
}

// U = g on the Dirichlet rows (mask)
__global__ void k_lift(long long n, const unsigned char *__restrict__ mask, const double *__restrict__ g, double *__restrict__ U)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        if (mask[i]) U[i] = g[i];
}

inline unsigned grid_for(long long n, int per_block) { return (unsigned)((n + per_block - 1) / per_block); }
inline unsigned vgrid(long long n) { return (unsigned)std::min<long long>(NRB, (n + 255) / 256); }


// ====================================================================== multigrid
// Stencil storage of a level matrix ("SM"): for node n, row i, stencil offset
// o = ((d0+2)*5 + (d1+2))*5 + (d2+2), column dof j: sm[n*1125 + i*375 + o*3 + j]
// = K[(n,i),(n+d,j)], zero where n+d is outside the grid.  The AIJ pattern
// of every level is this stencil cut at the boundary, so the Galerkin
// operator P^T K P of a level keeps it.
constexpr int MGW = 4;  // nonzeros per row/column of a 1D transfer
// value type of the coarse levels' stencil storage: the hierarchy only
// preconditions, so FP32 values (products and sums in FP64) halve its traffic
using SMT = float;

struct Xfer {                 // one direction: fine nf <-> coarse nc
    const int *c2f_lo;        // [nc]: first fine index of coarse A's support
    const double *c2f_w;      // [nc * MGW]: P[c2f_lo + k][A]
    const int *f2c_lo;        // [nf]
    const double *f2c_w;      // [nf * MGW]: P[a][f2c_lo + k]
};
struct Xfer3 {
    Xfer d[3];
    int nbf[3], nbc[3];
};

__device__ __forceinline__ void decode_node(const int nb[3], long long node, int a[3])
{
    a[2] = (int)(node % nb[2]);
    a[1] = (int)((node / nb[2]) % nb[1]);
    a[0] = (int)(node / ((long long)nb[2] * nb[1]));
}

// acc[i] = sum over the stencil of K[(n,i),(n+d,j)] x[n+d, j] (one wave;
// loads issued together as in node_rows)
__device__ __forceinline__ void node_rows_sm(const SpArgs &s, long long node, const SMT *__restrict__ sm,
                                             const double *__restrict__ x, int lane, double acc[3])
{
    int a[3];
    decode_node(s.nb, node, a);
    const SMT *v = sm + node * 1125;
    double xv[6], v0[6], v1[6], v2[6];
#pragma unroll
    for (int it = 0; it < 6; it++) {
        const int k = lane + 64 * it;
        xv[it] = v0[it] = v1[it] = v2[it] = 0.0;
        int o = (k * 43691) >> 17;     // k / 3
        int j = k - 3 * o;
        int q0 = (o * 1311) >> 15;     // o / 25 (o < 125)
        int r0 = o - 25 * q0;
        int q1 = (r0 * 13108) >> 16;   // r0 / 5
        int b0 = a[0] + q0 - 2, b1 = a[1] + q1 - 2, b2 = a[2] + (r0 - 5 * q1) - 2;
        if (k < 375 && b0 >= 0 && b0 < s.nb[0] && b1 >= 0 && b1 < s.nb[1] && b2 >= 0 && b2 < s.nb[2]) {
            xv[it] = x[3 * (((long long)b0 * s.nb[1] + b1) * s.nb[2] + b2) + j];
            v0[it] = v[k];
            v1[it] = v[375 + k];
            v2[it] = v[750 + k];
        }
    }
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
#pragma unroll
    for (int it = 0; it < 6; it++) {
        s0 += v0[it] * xv[it];
        s1 += v1[it] * xv[it];
        s2 += v2[it] * xv[it];
    }
    acc[0] = wave_sum(s0);
    acc[1] = wave_sum(s1);
    acc[2] = wave_sum(s2);
}

template <int SM, class VT>
__device__ __forceinline__ void level_rows(const SpArgs &s, long long node, const VT *__restrict__ vals,
                                           const double *__restrict__ x, int lane, double acc[3])
{
    if (SM) node_rows_sm(s, node, (const SMT *)vals, x, lane, acc);
    else node_rows<VT>(s, node, vals, x, lane, acc);
}

// y = b - K x on a level in stencil storage
__global__ __launch_bounds__(256) void k_resid_sm(SpArgs s, const SMT *__restrict__ sm, const double *__restrict__ x,
                                                  const double *__restrict__ b, double *__restrict__ y)
{
    const int lane = threadIdx.x & 63;
    const long long node = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (node >= s.nnode) return;
    double acc[3];
    node_rows_sm(s, node, sm, x, lane, acc);
    if (lane == 0)
        for (int i = 0; i < 3; i++) y[3 * node + i] = b[3 * node + i] - acc[i];
}

// one colour of block Gauss-Seidel: x_n += Dinv_n (b_n - (K x)_n)
template <int SM, class VT>
__global__ __launch_bounds__(256) void k_gs_colour(SpArgs s, int c0, int c1, int c2, int m1, int m2, long long ncol,
                                                   const VT *__restrict__ vals, const double *__restrict__ Dinv,
                                                   const double *__restrict__ b, double *x)
{
    const int lane = threadIdx.x & 63;
    const long long t = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (t >= ncol) return;
    long long a2 = c2 + 3 * (t % m2), a1 = c1 + 3 * ((t / m2) % m1), a0 = c0 + 3 * (t / ((long long)m1 * m2));
    long long node = (a0 * s.nb[1] + a1) * s.nb[2] + a2;
    double acc[3];
    level_rows<SM, VT>(s, node, vals, x, lane, acc);
    if (lane == 0) {
        double q[3];
        for (int i = 0; i < 3; i++) q[i] = b[3 * node + i] - acc[i];
        const double *B = Dinv + 9 * node;
        for (int i = 0; i < 3; i++) x[3 * node + i] += B[3 * i] * q[0] + B[3 * i + 1] * q[1] + B[3 * i + 2] * q[2];
    }
}

// A whole block Gauss-Seidel sweep (27 colours in order) by one work-group:
// for the small coarse levels, where a launch per colour costs more than
// the colour's work.  The work-group barrier orders the colours.
template <int BWD>
__global__ __launch_bounds__(1024) void k_gs_sweep_wg(SpArgs s, const SMT *__restrict__ sm, const double *__restrict__ Dinv,
                                                      const double *__restrict__ b, double *x)
{
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
    for (int cc = 0; cc < 27; cc++) {
        const int col = BWD ? 26 - cc : cc;
        const int c0 = col / 9, c1 = (col / 3) % 3, c2 = col % 3;
        const int m0 = (s.nb[0] - c0 + 2) / 3, m1 = (s.nb[1] - c1 + 2) / 3, m2 = (s.nb[2] - c2 + 2) / 3;
        const long long ncol = (long long)m0 * m1 * m2;
        for (long long t = wv; t < ncol; t += nw) {
            long long a2 = c2 + 3 * (t % m2), a1 = c1 + 3 * ((t / m2) % m1), a0 = c0 + 3 * (t / ((long long)m1 * m2));
            long long node = (a0 * s.nb[1] + a1) * s.nb[2] + a2;
            double acc[3];
            node_rows_sm(s, node, sm, x, lane, acc);
            if (lane == 0) {
                double q[3];
                for (int i = 0; i < 3; i++) q[i] = b[3 * node + i] - acc[i];
                const double *B = Dinv + 9 * node;
                for (int i = 0; i < 3; i++) x[3 * node + i] += B[3 * i] * q[0] + B[3 * i + 1] * q[1] + B[3 * i + 2] * q[2];
            }
        }
        __threadfence_block();
        __syncthreads();
    }
}

// ---- Chebyshev smoothing preconditioned by the nodal block diagonal
// (Saad, "Iterative Methods for Sparse Linear Systems", 2nd ed., Alg. 12.1):
// on [a, b] with theta = (a+b)/2, delta = (b-a)/2, sigma = theta/delta:
//   r0 = b - K x0, d0 = Dinv r0 / theta;
//   x_{k+1} = x_k + d_k, r_{k+1} = r_k - K d_k,
//   d_{k+1} = rho_{k+1} rho_k d_k + (2 rho_{k+1} / delta) Dinv r_{k+1},
//   rho_0 = 1/sigma, rho_{k+1} = 1 / (2 sigma - rho_k).
// Each kernel reads neighbours only of a vector it does not write.

// r = b - K x, d = Dinv r / theta (one wave per node)
template <int SM, class VT>
__global__ __launch_bounds__(256) void k_cheb_init(SpArgs s, const VT *__restrict__ vals, const double *__restrict__ Dinv,
                                                   const double *__restrict__ b, const double *__restrict__ x,
                                                   double *__restrict__ r, double *__restrict__ d, double inv_theta)
{
    const int lane = threadIdx.x & 63;
    const long long node = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (node >= s.nnode) return;
    double acc[3];
    level_rows<SM, VT>(s, node, vals, x, lane, acc);
    if (lane == 0) {
        double q[3];
        for (int i = 0; i < 3; i++) {
            q[i] = b[3 * node + i] - acc[i];
            r[3 * node + i] = q[i];
        }
        const double *B = Dinv + 9 * node;
        for (int i = 0; i < 3; i++) d[3 * node + i] = inv_theta * (B[3 * i] * q[0] + B[3 * i + 1] * q[1] + B[3 * i + 2] * q[2]);
    }
}

// start from x = 0: r = b, d = Dinv b / theta, x = 0 (thread per node)
__global__ void k_cheb_zero(long long nnode, const double *__restrict__ Dinv, const double *__restrict__ b, double *__restrict__ r,
                            double *__restrict__ d, double *__restrict__ x, double inv_theta)
{
    long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= nnode) return;
    double q[3] = {b[3 * n], b[3 * n + 1], b[3 * n + 2]};
    const double *B = Dinv + 9 * n;
    for (int i = 0; i < 3; i++) {
        r[3 * n + i] = q[i];
        x[3 * n + i] = 0.0;
        d[3 * n + i] = inv_theta * (B[3 * i] * q[0] + B[3 * i + 1] * q[1] + B[3 * i + 2] * q[2]);
    }
}

// x += din, r -= K din, and if newd: dout = c1 din + c2 Dinv r
template <int SM, class VT>
__global__ __launch_bounds__(256) void k_cheb_step(SpArgs s, const VT *__restrict__ vals, const double *__restrict__ Dinv,
                                                   const double *__restrict__ din, double *__restrict__ dout,
                                                   double *__restrict__ r, double *__restrict__ x, double c1, double c2,
                                                   int newd)
{
    const int lane = threadIdx.x & 63;
    const long long node = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (node >= s.nnode) return;
    double acc[3];
    level_rows<SM, VT>(s, node, vals, din, lane, acc);
    if (lane == 0) {
        double q[3];
        for (int i = 0; i < 3; i++) {
            q[i] = r[3 * node + i] - acc[i];
            r[3 * node + i] = q[i];
            x[3 * node + i] += din[3 * node + i];
        }
        if (newd) {
            const double *B = Dinv + 9 * node;
            for (int i = 0; i < 3; i++)
                dout[3 * node + i] = c1 * din[3 * node + i] + c2 * (B[3 * i] * q[0] + B[3 * i + 1] * q[1] + B[3 * i + 2] * q[2]);
        }
    }
}

// fine level with the Krylov matrix: the row pass is k_spmv_line, these apply
// the per-node updates (the 16-bit copy's k_cheb16 fuses them)
// d = Dinv r / theta
__global__ void k_cheb_dfromr(long long nnode, const double *__restrict__ Dinv, const double *__restrict__ r,
                              double *__restrict__ d, double inv_theta)
{
    long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= nnode) return;
    const double *B = Dinv + 9 * n;
    double q0 = r[3 * n], q1 = r[3 * n + 1], q2 = r[3 * n + 2];
    for (int i = 0; i < 3; i++) d[3 * n + i] = inv_theta * (B[3 * i] * q0 + B[3 * i + 1] * q1 + B[3 * i + 2] * q2);
}
// r -= q (q = K din), x += din, and if newd: dout = c1 din + c2 Dinv r
__global__ void k_cheb_update(long long nnode, const double *__restrict__ Dinv, const double *__restrict__ q,
                              const double *__restrict__ din, double *__restrict__ dout, double *__restrict__ r,
                              double *__restrict__ x, double c1, double c2, int newd)
{
    long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= nnode) return;
    double rv[3];
    for (int i = 0; i < 3; i++) {
        rv[i] = r[3 * n + i] - q[3 * n + i];
        r[3 * n + i] = rv[i];
        x[3 * n + i] += din[3 * n + i];
    }
    if (newd) {
        const double *B = Dinv + 9 * n;
        for (int i = 0; i < 3; i++)
            dout[3 * n + i] = c1 * din[3 * n + i] + c2 * (B[3 * i] * rv[0] + B[3 * i + 1] * rv[1] + B[3 * i + 2] * rv[2]);
    }
}

// out = Dinv in per node (thread per node)
__global__ void k_apply_dinv(long long nnode, const double *__restrict__ Dinv, const double *__restrict__ in, double *__restrict__ out)
{
    const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= nnode) return;
    const double *B = Dinv + 9 * n;
    const double q0 = in[3 * n], q1 = in[3 * n + 1], q2 = in[3 * n + 2];
    for (int i = 0; i < 3; i++) out[3 * n + i] = B[3 * i] * q0 + B[3 * i + 1] * q1 + B[3 * i + 2] * q2;
}

// y = Dinv K x (power iteration for the largest eigenvalue of Dinv K)
template <int SM, class VT>
__global__ __launch_bounds__(256) void k_dinv_k(SpArgs s, const VT *__restrict__ vals, const double *__restrict__ Dinv,
                                                const double *__restrict__ x, double *__restrict__ y)
{
    const int lane = threadIdx.x & 63;
    const long long node = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (node >= s.nnode) return;
    double acc[3];
    level_rows<SM, VT>(s, node, vals, x, lane, acc);
    if (lane == 0) {
        const double *B = Dinv + 9 * node;
        for (int i = 0; i < 3; i++) y[3 * node + i] = B[3 * i] * acc[0] + B[3 * i + 1] * acc[1] + B[3 * i + 2] * acc[2];
    }
}

// deterministic pseudo-random start vector in [-1, 1]
__global__ void k_hash_vec(long long n, double *v)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        unsigned long long z = (unsigned long long)i * 0x9E3779B97F4A7C15ULL + 0x5EEDULL;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        z ^= z >> 31;
        v[i] = 2.0 * ((double)(z >> 11) * (1.0 / 9007199254740992.0)) - 1.0;
    }
}

// diagonal 3x3 blocks of a level in stencil storage, and their inverses
__global__ void k_block_diag_sm(long long nnode, const SMT *__restrict__ sm, double *__restrict__ Dblk, double *__restrict__ Dinv)
{
    long long node = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (node >= nnode) return;
    double A[9], B[9];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) A[3 * i + j] = sm[node * 1125 + i * 375 + 62 * 3 + j];
    inv3(A, B);
    for (int k = 0; k < 9; k++) {
        Dblk[9 * node + k] = A[k];
        Dinv[9 * node + k] = B[k];
    }
}

// One direction d of the Galerkin product: out = (R_d (x) I) in (R_d (x) I)^T
// with R_d = P_d^T, the other directions unchanged.  in: stencil storage on
// grid nbi (IN_CSR: the assembled AIJ values of the fine level, with dmask
// bit i of node n set when dof 3n + i is a Dirichlet dof); out: stencil
// storage on nbo (nbo[d] coarse).  One thread per (node, offset).
template <int IN_CSR, class VT>
__global__ __launch_bounds__(256) void k_galerkin(int d, int3 nbi, int3 nbo, const long long *__restrict__ rowptr,
                                                  const VT *__restrict__ in, Xfer x, SMT *__restrict__ out,
                                                  const unsigned char *__restrict__ dmask)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int nbi_[3] = {nbi.x, nbi.y, nbi.z}, nbo_[3] = {nbo.x, nbo.y, nbo.z};
    const long long nout = (long long)nbo.x * nbo.y * nbo.z;
    if (t >= nout * 125) return;
    const long long N = t / 125;
    const int o = (int)(t - N * 125);
    int A[3], D[3] = {o / 25 - 2, (o / 5) % 5 - 2, o % 5 - 2};
    decode_node(nbo_, N, A);
    double acc[9];
#pragma unroll
    for (int k = 0; k < 9; k++) acc[k] = 0.0;
    bool ok = true;
    for (int k = 0; k < 3; k++) ok = ok && A[k] + D[k] >= 0 && A[k] + D[k] < nbo_[k];
    if (ok) {
        const int Ad = A[d], Bd = A[d] + D[d];
        const int alo = x.c2f_lo[Ad], blo = x.c2f_lo[Bd];
        for (int ka = 0; ka < MGW; ka++) {
            const double wa = x.c2f_w[Ad * MGW + ka];
            if (wa == 0.0) continue;
            for (int kb = 0; kb < MGW; kb++) {
                const double wb = x.c2f_w[Bd * MGW + kb];
                const int off = (blo + kb) - (alo + ka);
                if (wb == 0.0 || off < -2 || off > 2) continue;
                int Ai[3] = {A[0], A[1], A[2]}, Di[3] = {D[0], D[1], D[2]};
                Ai[d] = alo + ka;
                Di[d] = off;
                const long long nin = ((long long)Ai[0] * nbi_[1] + Ai[1]) * nbi_[2] + Ai[2];
                const double w = wa * wb;
                if (IN_CSR) {
                    int lo[3], wd[3];
                    for (int k = 0; k < 3; k++) {
                        lo[k] = max(0, Ai[k] - 2);
                        wd[k] = min(nbi_[k] - 1, Ai[k] + 2) - lo[k] + 1;
                    }
                    const long long L = 3LL * wd[0] * wd[1] * wd[2];
                    const long long pos = rowptr[3 * nin] +
                        3LL * (((Ai[0] + Di[0] - lo[0]) * wd[1] + (Ai[1] + Di[1] - lo[1])) * wd[2] + (Ai[2] + Di[2] - lo[2]));
                    // Dirichlet rows and columns (identity in K) are left out of the
                    // coarse operators: the coarse problem sees them as removed dofs
                    const long long ncl = ((long long)(Ai[0] + Di[0]) * nbi_[1] + (Ai[1] + Di[1])) * nbi_[2] + (Ai[2] + Di[2]);
                    const int mi = dmask ? dmask[nin] : 0, mj = dmask ? dmask[ncl] : 0;
#pragma unroll
                    for (int i = 0; i < 3; i++)
#pragma unroll
                        for (int j = 0; j < 3; j++)
                            if (!(((mi >> i) | (mj >> j)) & 1)) acc[3 * i + j] += w * (double)in[pos + i * L + j];
                } else {
                    const int oi = ((Di[0] + 2) * 5 + (Di[1] + 2)) * 5 + (Di[2] + 2);
                    const VT *v = in + nin * 1125 + oi * 3;
#pragma unroll
                    for (int i = 0; i < 3; i++)
#pragma unroll
                        for (int j = 0; j < 3; j++) acc[3 * i + j] += w * (double)v[i * 375 + j];
                }
            }
        }
    }
    SMT *po = out + N * 1125 + o * 3;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) po[i * 375 + j] = (SMT)acc[3 * i + j];
}

__global__ void k_node_mask(long long nnode, const unsigned char *__restrict__ dmask, unsigned char *__restrict__ nm)
{
    const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (n < nnode) nm[n] = (unsigned char)((dmask[3 * n] ? 1 : 0) | (dmask[3 * n + 1] ? 2 : 0) | (dmask[3 * n + 2] ? 4 : 0));
}

// bc = P^T rf (restriction), one thread per coarse node
// (dmask: the fine residual's Dirichlet entries are left out)
// bc = P^T rf (restriction), one wave per coarse node: lane = one of the
// MGW^3 = 64 fine nodes of its support, then a wave sum (dmask: the fine
// Dirichlet dofs contribute nothing)
static_assert(MGW * MGW * MGW == 64, "one lane per fine node of a coarse support");
// With Dc: also the coarse level's Chebyshev start from zero (rc = bc, xc = 0,
// dc = inv_theta Dc bc) that k_cheb_zero would make.
__global__ __launch_bounds__(256) void k_restrict(Xfer3 x, const double *__restrict__ rf, double *__restrict__ bc,
                                                  const unsigned char *__restrict__ dmask, const double *__restrict__ Dc,
                                                  double *__restrict__ rc, double *__restrict__ xc, double *__restrict__ dc,
                                                  double inv_theta)
{
    const long long nc = (long long)x.nbc[0] * x.nbc[1] * x.nbc[2];
    const long long C = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (C >= nc) return;
    const int lane = threadIdx.x & 63, k0 = lane >> 4, k1 = (lane >> 2) & 3, k2 = lane & 3;
    int A[3];
    decode_node(x.nbc, C, A);
    const double w = (x.d[0].c2f_w[A[0] * MGW + k0] * x.d[1].c2f_w[A[1] * MGW + k1]) * x.d[2].c2f_w[A[2] * MGW + k2];
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    if (w != 0.0) {
        const long long f = (((long long)(x.d[0].c2f_lo[A[0]] + k0)) * x.nbf[1] + (x.d[1].c2f_lo[A[1]] + k1)) * x.nbf[2] +
                            (x.d[2].c2f_lo[A[2]] + k2);
        s0 = w * rf[3 * f];
        s1 = w * rf[3 * f + 1];
        s2 = w * rf[3 * f + 2];
        if (dmask) {
            if (dmask[3 * f]) s0 = 0.0;
            if (dmask[3 * f + 1]) s1 = 0.0;
            if (dmask[3 * f + 2]) s2 = 0.0;
        }
    }
    s0 = wave_sum(s0);
    s1 = wave_sum(s1);
    s2 = wave_sum(s2);
    if (lane == 0) {
        const double q[3] = {s0, s1, s2};
        for (int i = 0; i < 3; i++) bc[3 * C + i] = q[i];
        if (Dc) {
            const double *B = Dc + 9 * C;
            for (int i = 0; i < 3; i++) {
                rc[3 * C + i] = q[i];
                xc[3 * C + i] = 0.0;
                dc[3 * C + i] = inv_theta * (B[3 * i] * q[0] + B[3 * i + 1] * q[1] + B[3 * i + 2] * q[2]);
            }
        }
    }
}

// xf += P (xc + dc) (prolongation; dc, if given, is the coarse smoother's last
// update, not yet added to xc), one thread per fine node (dmask: no
// correction on the fine Dirichlet dofs)
__global__ void k_prolong_add(Xfer3 x, const double *__restrict__ xc, const double *__restrict__ dc, double *__restrict__ xf,
                              const unsigned char *__restrict__ dmask)
{
    const long long nf = (long long)x.nbf[0] * x.nbf[1] * x.nbf[2];
    const long long F = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (F >= nf) return;
    int a[3];
    decode_node(x.nbf, F, a);
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    const int l0 = x.d[0].f2c_lo[a[0]], l1 = x.d[1].f2c_lo[a[1]], l2 = x.d[2].f2c_lo[a[2]];
    for (int k0 = 0; k0 < MGW; k0++) {
        double w0 = x.d[0].f2c_w[a[0] * MGW + k0];
        if (w0 == 0.0) continue;
        for (int k1 = 0; k1 < MGW; k1++) {
            double w1 = w0 * x.d[1].f2c_w[a[1] * MGW + k1];
            if (w1 == 0.0) continue;
            for (int k2 = 0; k2 < MGW; k2++) {
                double w = w1 * x.d[2].f2c_w[a[2] * MGW + k2];
                if (w == 0.0) continue;
                long long C = (((long long)(l0 + k0)) * x.nbc[1] + (l1 + k1)) * x.nbc[2] + (l2 + k2);
                double v[3] = {xc[3 * C], xc[3 * C + 1], xc[3 * C + 2]};
                if (dc)
                    for (int i = 0; i < 3; i++) v[i] += dc[3 * C + i];
                s0 += w * v[0];
                s1 += w * v[1];
                s2 += w * v[2];
            }
        }
    }
    if (dmask) {
        if (!dmask[3 * F]) xf[3 * F] += s0;
        if (!dmask[3 * F + 1]) xf[3 * F + 1] += s1;
        if (!dmask[3 * F + 2]) xf[3 * F + 2] += s2;
    } else {
        xf[3 * F] += s0;
        xf[3 * F + 1] += s1;
        xf[3 * F + 2] += s2;
    }
}

// x = Ainv b (dense n x n), one wave per row
__global__ __launch_bounds__(256) void k_dense_mv(int n, const double *__restrict__ Ainv, const double *__restrict__ b,
                                                  double *__restrict__ x)
{
    const int lane = threadIdx.x & 63;
    const int r = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (r >= n) return;
    double s = 0.0;
    for (int c = lane; c < n; c += 64) s += Ainv[(long long)r * n + c] * b[c];
    s = wave_sum(s);
    if (lane == 0) x[r] = s;
}

}  // namespace

// ====================================================================== work space
namespace {
struct MGLevel {
    int ne[3], nb[3];
    long long nnode = 0, ndof = 0;
    SMT *sm = nullptr;                    // stencil storage (levels >= 1)
    double *Dblk = nullptr, *Dinv = nullptr;
    double *x = nullptr, *b = nullptr, *r = nullptr;
    SMT *stage[2] = {nullptr, nullptr};   // Galerkin intermediates towards the next level
    int *xfer_i = nullptr;                // transfer tables to the next level (device)
    double *xfer_w = nullptr;
    Xfer3 T;                              // views into xfer_i / xfer_w
    double *Ainv = nullptr;               // coarsest level: dense inverse
    bool dense = false;
    double *d[2] = {nullptr, nullptr};    // Chebyshev directions
    double lmax = 0.0;                    // largest eigenvalue estimate of Dinv K
    double *ev = nullptr;                 // its eigenvector estimate, the next set-up's start
    bool ev_ok = false;                   // ev holds one from this solve
};
}  // namespace

struct igap4h::Solver {
    std::vector<MGLevel> mg;  // mg[0] is the fine level (its matrix is the assembled AIJ values)
    // the V-cycle captured as a graph (in: mg_in, out: z); rebuilt when the key changes
    hipStream_t cap = nullptr;
    hipGraphExec_t mg_exec = nullptr;
    const void *key_vals = nullptr;
    int key_f32 = -1, key_pre = -1, key_post = -1, key_smoother = -1;
    int key_svt = -1, key_cdeg = -1;
    double key_cratio = -1.0;
    const void *pc_vals = nullptr;  // matrix the preconditioner was last set up on
    // options the preconditioner was last set up with (pc, smoother, smoothing
    // values, coarse degree, pre, post; Chebyshev ratio)
    int pc_key[6] = {-1, -1, -1, -1, -1, -1};
    double pc_key_ratio = -1.0;
    uint4 *sm16 = nullptr;          // 16-bit copy of the fine matrix for smoothing (node-interleaved)
    double *rs16 = nullptr;         // row scales of the binary16 copy
    unsigned char *nmask = nullptr;
    bool jac_valid = false;  // s->vals holds a tangent (of jac_key) for lagged Newton steps
    long long jac_key = -1;
    long long jac_age = 0;
    double *mg_in = nullptr;
    long long ndof = 0, nnode = 0;
    int restart = 0;
    size_t vals_bytes = 0;
    void *vals = nullptr;  // tangent values used by the Newton solve
    double *Dblk = nullptr, *Dinv = nullptr;
    double *V = nullptr;   // Krylov basis: vcap x ndof, grown on demand up to restart + 1
    int vcap = 0;
    double *r = nullptr, *z = nullptr, *p = nullptr, *q = nullptr, *w = nullptr, *t = nullptr;
    double *Y = nullptr, *U0 = nullptr, *R = nullptr;
    double *partial = nullptr;  // (restart + 2) x NRB
    double *pspmv = nullptr;    // one partial per SpMV block
    double *hd = nullptr;       // GMRES: two Gram-Schmidt columns and y, 3 (restart + 1)
    double *hh = nullptr;       // pinned host mirror of hd
    double *sc = nullptr;       // device scalars
    double *hp = nullptr;       // pinned host mirror of sc
};

void igap4h::solver_invalidate_jacobian(igap4h_ctx *c)
{
    if (c && c->solver) c->solver->jac_valid = false;
}

void igap4h::solver_free(igap4h_ctx *c)
{
    Solver *s = c ? c->solver : nullptr;
    if (!s) return;
    (void)hipFree(s->vals);
    (void)hipFree(s->sm16);
    (void)hipFree(s->nmask);
    (void)hipFree(s->rs16);
    for (double *p : {s->Dblk, s->Dinv, s->V, s->r, s->z, s->p, s->q, s->w, s->t, s->Y, s->U0, s->R, s->partial,
                      s->pspmv, s->hd, s->sc})
        (void)hipFree(p);
    if (s->hp) (void)hipHostFree(s->hp);
    if (s->hh) (void)hipHostFree(s->hh);
    if (s->mg_exec) (void)hipGraphExecDestroy(s->mg_exec);
    if (s->cap) (void)hipStreamDestroy(s->cap);
    (void)hipFree(s->mg_in);
    for (MGLevel &L : s->mg) {
        for (double *p : {L.Dblk, L.Dinv, L.x, L.b, L.r, L.xfer_w, L.Ainv, L.d[0], L.d[1], L.ev}) (void)hipFree(p);
        for (SMT *p : {L.sm, L.stage[0], L.stage[1]}) (void)hipFree(p);
        (void)hipFree(L.xfer_i);
    }
    delete s;
    c->solver = nullptr;
}

static int solver_ws(igap4h_ctx *c, int restart, size_t vals_bytes)
{
    if (!c->solver) {
        c->solver = new Solver();
        Solver *s = c->solver;
        s->ndof = c->mesh.ndof();
        s->nnode = s->ndof / 3;
        size_t n = (size_t)s->ndof;
        int rc;
        if ((rc = dev_alloc(&s->Dblk, 9 * (size_t)s->nnode)) || (rc = dev_alloc(&s->Dinv, 9 * (size_t)s->nnode)) ||
            (rc = dev_alloc(&s->r, n)) || (rc = dev_alloc(&s->z, n)) || (rc = dev_alloc(&s->p, n)) ||
            (rc = dev_alloc(&s->q, n)) || (rc = dev_alloc(&s->w, n)) || (rc = dev_alloc(&s->t, n)) ||
            (rc = dev_alloc(&s->Y, n)) || (rc = dev_alloc(&s->U0, n)) || (rc = dev_alloc(&s->R, n)) ||
            (rc = dev_alloc(&s->sc, 64)) || (rc = dev_alloc(&s->pspmv, (size_t)s->nnode + 1))) {
            solver_free(c);  // no half-built work space survives a failure
            return rc;
        }
        if (hipHostMalloc((void **)&s->hp, 64 * sizeof(double)) != hipSuccess) {
            s->hp = nullptr;
            solver_free(c);
            return fail(IGAP4H_ENOMEM, "hipHostMalloc of 512 bytes failed");
        }
    }
    Solver *s = c->solver;
    if (restart > s->restart) {
        (void)hipFree(s->V);
        (void)hipFree(s->partial);
        (void)hipFree(s->hd);
        if (s->hh) (void)hipHostFree(s->hh);
        s->V = s->partial = s->hd = s->hh = nullptr;
        s->restart = 0;
        s->vcap = 0;
        int rc;
        // the basis starts small (most solves need few vectors) and grows in
        // basis_reserve
        const int v0 = std::min(restart + 1, 32);
        if ((rc = dev_alloc(&s->V, (size_t)v0 * s->ndof))) return rc;
        s->vcap = v0;
        if ((rc = dev_alloc(&s->partial, (size_t)(restart + 2) * NRB))) return rc;
        if ((rc = dev_alloc(&s->hd, 3 * (size_t)(restart + 1)))) return rc;
        if (hipHostMalloc((void **)&s->hh, 3 * (size_t)(restart + 1) * sizeof(double)) != hipSuccess) {
            s->hh = nullptr;
            return fail(IGAP4H_ENOMEM, "hipHostMalloc failed");
        }
        s->restart = restart;
    }
    if (vals_bytes > s->vals_bytes) {
        (void)hipFree(s->vals);
        s->vals = nullptr;
        s->vals_bytes = 0;
        if (dev_malloc(&s->vals, vals_bytes) != hipSuccess) return fail(IGAP4H_ENOMEM, "hipMalloc of %zu bytes failed", vals_bytes);
        s->vals_bytes = vals_bytes;
    }
    return 0;
}

// make room for nvec Krylov vectors (at most restart + 1): doubles the
// allocation and moves the vectors already built, in stream order
static int basis_reserve(Solver *s, int nvec, hipStream_t st)
{
    if (nvec <= s->vcap) return 0;
    const int cap = std::min(s->restart + 1, std::max(nvec, 2 * s->vcap));
    double *V = nullptr;
    int rc = dev_alloc(&V, (size_t)cap * s->ndof);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(V, s->V, (size_t)s->vcap * s->ndof * sizeof(double), hipMemcpyDeviceToDevice, st));
    HIPCHK(hipStreamSynchronize(st));
    (void)hipFree(s->V);
    s->V = V;
    s->vcap = cap;
    return 0;
}

static SpArgs sp_args(const igap4h_ctx *c)
{
    SpArgs a;
    for (int d = 0; d < 3; d++) a.nb[d] = c->mesh.nb[d];
    a.nnode = c->mesh.ndof() / 3;
    a.rowptr = c->d_rowptr;
    return a;
}

// ====================================================================== multigrid set-up
// Levels by coarsening the knot vectors (every other element boundary kept,
// and the last one: n elements become (n + 1) / 2; a direction with one
// element stays) while the level has more than MG_DENSE dofs; the coarsest
// level is solved with a dense inverse when it has at most MG_DENSE dofs,
// else by symmetric Gauss-Seidel sweeps.
// The inverse is formed on the host at every set-up (Gauss-Jordan, 2 n^3):
// 200 stops the coarsening at 2^3 elements (192 dofs, ~3 ms); with 400 a 48^3
// hierarchy ended at 3^3 elements (375 dofs) and the inverse took 24 ms per
// Newton step, a quarter of a getime time step, for the same iteration
// count.  On small meshes such a level can be singular (its Galerkin
// operator sees the fine level's Dirichlet rows masked out, and a coarse
// space of mostly boundary functions has more of them than the fine
// interior has dofs): the set-up then drops it once and inverts the level
// above, as long as that has at most MG_DENSE_MAX dofs.
constexpr long long MG_DENSE = 200;
constexpr long long MG_DENSE_MAX = 400;

static int mg_build(igap4h_ctx *c, Solver *s)
{
    if (!s->mg.empty()) return 0;
    std::vector<double> kf[3];
    MGLevel L0;
    for (int d = 0; d < 3; d++) {
        L0.ne[d] = c->mesh.ne[d];
        L0.nb[d] = c->mesh.nb[d];
        kf[d] = c->mesh.knot[d];
    }
    L0.nnode = (long long)L0.nb[0] * L0.nb[1] * L0.nb[2];
    L0.ndof = 3 * L0.nnode;
    s->mg.push_back(L0);
    int rc;
    if ((rc = dev_alloc(&s->mg[0].r, (size_t)L0.ndof)) || (rc = dev_alloc(&s->mg[0].d[0], (size_t)L0.ndof)) ||
        (rc = dev_alloc(&s->mg[0].d[1], (size_t)L0.ndof)) || (rc = dev_alloc(&s->mg[0].x, (size_t)L0.ndof)) ||
        (rc = dev_alloc(&s->mg[0].ev, (size_t)L0.ndof)))
        return rc;  // mg[0].x: fine-level product buffer (the level's solution is the caller's z)
    while (s->mg.size() < 12) {
        MGLevel &F = s->mg.back();
        bool any = false;
        for (int d = 0; d < 3; d++) any = any || F.ne[d] >= 2;
        // the fine level (its Dirichlet rows masked in the Galerkin product)
        // is coarsened above MG_DENSE_MAX dofs, the coarse levels above MG_DENSE
        if (F.ndof <= (s->mg.size() == 1 ? MG_DENSE_MAX : MG_DENSE) || !any) break;
        MGLevel C;
        std::vector<double> kc[3], P[3];
        for (int d = 0; d < 3; d++) {
            C.ne[d] = F.ne[d] >= 2 ? (F.ne[d] + 1) / 2 : F.ne[d];
            C.nb[d] = C.ne[d] + 2;
            kc[d] = F.ne[d] >= 2 ? coarsen_knots(kf[d], F.ne[d]) : kf[d];
            if (prolongation_1d(kf[d], F.ne[d], kc[d], C.ne[d], P[d]))
                return fail(IGAP4H_EINVAL, "multigrid: knot vectors of direction %d are not nested", d);
        }
        C.nnode = (long long)C.nb[0] * C.nb[1] * C.nb[2];
        C.ndof = 3 * C.nnode;
        // transfer tables: [c2f_lo d0 d1 d2 | f2c_lo d0 d1 d2], weights alike
        size_t ni = 0;
        for (int d = 0; d < 3; d++) ni += (size_t)C.nb[d] + F.nb[d];
        std::vector<int> hi(ni, 0);
        std::vector<double> hw(ni * MGW, 0.0);
        size_t oi = 0;
        size_t off_c[3], off_f[3];
        for (int d = 0; d < 3; d++) {
            const int nbf = F.nb[d], nbc = C.nb[d];
            off_c[d] = oi;
            for (int A = 0; A < nbc; A++, oi++) {
                int lo = -1, hi_ = -1;
                for (int a = 0; a < nbf; a++)
                    if (P[d][(size_t)a * nbc + A] != 0.0) { if (lo < 0) lo = a; hi_ = a; }
                if (lo < 0 || hi_ - lo >= MGW) return fail(IGAP4H_EINVAL, "multigrid: transfer wider than %d", MGW);
                hi[oi] = lo;
                for (int k = 0; lo + k <= hi_; k++) hw[oi * MGW + k] = P[d][(size_t)(lo + k) * nbc + A];
            }
        }
        for (int d = 0; d < 3; d++) {
            const int nbf = F.nb[d], nbc = C.nb[d];
            off_f[d] = oi;
            for (int a = 0; a < nbf; a++, oi++) {
                int lo = -1, hi_ = -1;
                for (int A = 0; A < nbc; A++)
                    if (P[d][(size_t)a * nbc + A] != 0.0) { if (lo < 0) lo = A; hi_ = A; }
                if (lo < 0 || hi_ - lo >= MGW) return fail(IGAP4H_EINVAL, "multigrid: transfer wider than %d", MGW);
                hi[oi] = lo;
                for (int k = 0; lo + k <= hi_; k++) hw[oi * MGW + k] = P[d][(size_t)a * nbc + lo + k];
            }
        }
        if ((rc = dev_alloc(&F.xfer_i, ni)) || (rc = dev_alloc(&F.xfer_w, ni * MGW))) return rc;
        HIPCHK(hipMemcpy(F.xfer_i, hi.data(), ni * sizeof(int), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(F.xfer_w, hw.data(), ni * MGW * sizeof(double), hipMemcpyHostToDevice));
        for (int d = 0; d < 3; d++) {
            F.T.d[d].c2f_lo = F.xfer_i + off_c[d];
            F.T.d[d].c2f_w = F.xfer_w + off_c[d] * MGW;
            F.T.d[d].f2c_lo = F.xfer_i + off_f[d];
            F.T.d[d].f2c_w = F.xfer_w + off_f[d] * MGW;
            F.T.nbf[d] = F.nb[d];
            F.T.nbc[d] = C.nb[d];
        }
        const size_t st0 = (size_t)C.nb[0] * F.nb[1] * F.nb[2] * 1125, st1 = (size_t)C.nb[0] * C.nb[1] * F.nb[2] * 1125;
        if ((rc = dev_alloc(&F.stage[0], st0)) || (rc = dev_alloc(&F.stage[1], st1))) return rc;
        if ((rc = dev_alloc(&C.sm, (size_t)C.nnode * 1125)) || (rc = dev_alloc(&C.Dblk, 9 * (size_t)C.nnode)) ||
            (rc = dev_alloc(&C.Dinv, 9 * (size_t)C.nnode)) || (rc = dev_alloc(&C.x, (size_t)C.ndof)) ||
            (rc = dev_alloc(&C.b, (size_t)C.ndof)) || (rc = dev_alloc(&C.r, (size_t)C.ndof)) ||
            (rc = dev_alloc(&C.d[0], (size_t)C.ndof)) || (rc = dev_alloc(&C.d[1], (size_t)C.ndof)) ||
            (rc = dev_alloc(&C.ev, (size_t)C.ndof)))
            return rc;
        for (int d = 0; d < 3; d++) kf[d] = kc[d];
        s->mg.push_back(C);
    }
    MGLevel &B = s->mg.back();
    if (s->mg.size() > 1 && B.ndof <= MG_DENSE) {
        B.dense = true;
        if ((rc = dev_alloc(&B.Ainv, (size_t)B.ndof * B.ndof))) return rc;
    }
    return 0;
}

static SpArgs level_args(const MGLevel &L)
{
    SpArgs a;
    for (int d = 0; d < 3; d++) a.nb[d] = L.nb[d];
    a.nnode = L.nnode;
    a.rowptr = nullptr;
    return a;
}

// Gauss-Jordan inverse with partial pivoting (row-major n x n, in place into inv)
static bool dense_inverse(std::vector<double> &A, int n, std::vector<double> &inv)
{
    inv.assign((size_t)n * n, 0.0);
    for (int i = 0; i < n; i++) inv[(size_t)i * n + i] = 1.0;
    for (int k = 0; k < n; k++) {
        int piv = k;
        for (int r = k + 1; r < n; r++)
            if (std::fabs(A[(size_t)r * n + k]) > std::fabs(A[(size_t)piv * n + k])) piv = r;
        double pv = A[(size_t)piv * n + k];
        if (!(std::fabs(pv) > 0.0) || !std::isfinite(pv)) return false;
        if (piv != k)
            for (int c2 = 0; c2 < n; c2++) {
                std::swap(A[(size_t)k * n + c2], A[(size_t)piv * n + c2]);
                std::swap(inv[(size_t)k * n + c2], inv[(size_t)piv * n + c2]);
            }
        double ip = 1.0 / A[(size_t)k * n + k];
        for (int c2 = 0; c2 < n; c2++) {
            A[(size_t)k * n + c2] *= ip;
            inv[(size_t)k * n + c2] *= ip;
        }
        // columns c < k of A are unit vectors by now (zero in row k): only c >= k change
        for (int r = 0; r < n; r++) {
            if (r == k) continue;
            double f = A[(size_t)r * n + k];
            if (f == 0.0) continue;
            for (int c2 = k; c2 < n; c2++) A[(size_t)r * n + c2] -= f * A[(size_t)k * n + c2];
            for (int c2 = 0; c2 < n; c2++) inv[(size_t)r * n + c2] -= f * inv[(size_t)k * n + c2];
        }
    }
    return true;
}

// ====================================================================== operators
namespace {

constexpr long long MG_WG_NODES = 512;  // levels up to this size smooth in one work-group
constexpr int CHEB_POWER_ITS = 12;      // power iterations for lmax from a pseudo-random start
constexpr int CHEB_POWER_WARM = 4;      // from the previous Newton step's eigenvector

// launch a one-wave-per-node level kernel KERNEL<SM, VT> on level l of Op *this
#define LEVEL_LAUNCH(l, KERNEL, ...)                                                                                   \
    do {                                                                                                               \
        const MGLevel &L_ = s->mg[l];                                                                                  \
        SpArgs a_ = (l) ? level_args(L_) : sp;                                                                         \
        dim3 g_(grid_for(a_.nnode, 4));                                                                                \
        if (l) hipLaunchKernelGGL((KERNEL<1, SMT>), g_, dim3(256), 0, st, a_, (const SMT *)L_.sm, __VA_ARGS__);        \
        else if (f32) hipLaunchKernelGGL((KERNEL<0, float>), g_, dim3(256), 0, st, a_, (const float *)vals, __VA_ARGS__); \
        else hipLaunchKernelGGL((KERNEL<0, double>), g_, dim3(256), 0, st, a_, (const double *)vals, __VA_ARGS__);    \
        HIPCHK(hipGetLastError());                                                                                     \
    } while (0)

struct Op {
    igap4h_ctx *c;
    Solver *s;
    SpArgs sp;
    const void *vals;
    int f32;
    int pc;
    hipStream_t st;
    int nu_pre = 1, nu_post = 1;  // Gauss-Seidel sweeps before / after the coarse correction
    int mg_graph = 1;             // replay the V-cycle from a captured graph
    int smoother = 0;             // 0 block Gauss-Seidel, 1 Chebyshev (degree nu_pre / nu_post), 2 fourth-kind Chebyshev
    int svt = IGAP4H_SMOOTH_SAME; // matrix of the fine-level Chebyshev passes
    int cdeg = 0;                 // Chebyshev degree cap below the finest level (0: none)
    double cratio = 200.0;        // Chebyshev interval [lmax' / cratio, lmax'], lmax' = 1.1 lmax

    // smoothing degree (Chebyshev) or sweeps (Gauss-Seidel) nu on level l
    bool chebyshev() const { return smoother == IGAP4H_MG_CHEBYSHEV || smoother == IGAP4H_MG_CHEB4; }
    int deg(size_t l, int nu) const { return l && chebyshev() && cdeg > 0 ? std::min(nu, cdeg) : nu; }

    // fine-level pass on the 16-bit copy (k_cheb16): MODE 0 is d = c2 Dinv K~ x,
    // MODE 1 starts Chebyshev from x (r = b - K~ x, d = c2 Dinv r), MODE 2 is one step on din
    template <int MODE>
    int cheb16(const double *xin, const double *b, double *r, double *x, double *dout, double c1, double c2, int newd) const
    {
        const unsigned g = (unsigned)((sp.nnode + CH_NT - 1) / CH_NT);
        const size_t sh = cheb16_lds_bytes(sp.nb[2]);
        if (svt == IGAP4H_SMOOTH_F16)
            hipLaunchKernelGGL((k_cheb16<MODE, 1>), dim3(g), dim3(CH_NT), sh, st, sp, (const uint4 *)s->sm16, (const double *)s->rs16,
                               (const double *)s->Dinv, xin, b, r, x, dout, c1, c2, newd);
        else
            hipLaunchKernelGGL((k_cheb16<MODE, 0>), dim3(g), dim3(CH_NT), sh, st, sp, (const uint4 *)s->sm16, (const double *)s->rs16,
                               (const double *)s->Dinv, xin, b, r, x, dout, c1, c2, newd);
        HIPCHK(hipGetLastError());
        return 0;
    }
    // the 16-bit copy of the current matrix; packs of 64 nodes, their number a
    // multiple of CH_NT / 64 so that k_cheb16's blocks stay inside the copy
    int pack_16() const
    {
        const long long npk = (sp.nnode + CH_NT - 1) / CH_NT * (CH_NT / 64);
        if (!s->sm16) {
            if (cheb16_lds_bytes(sp.nb[2]) > 65536)
                return fail(IGAP4H_EINVAL, "16-bit smoothing copy: %d nodes in direction 2 (limit 617)", sp.nb[2]);
            const size_t bytes = (size_t)npk * NU16 * 64 * sizeof(uint4);
            if (dev_malloc((void **)&s->sm16, bytes) != hipSuccess) {
                s->sm16 = nullptr;
                return fail(IGAP4H_ENOMEM, "hipMalloc of %zu bytes failed", bytes);
            }
            int rc;
            if ((rc = dev_alloc(&s->rs16, 3 * (size_t)sp.nnode))) return rc;
        }
        // every block writes its 8 nodes' words whole (nodes past nnode as
        // zeros); the packs past the last block are zeroed once
        const long long nblk = grid_for(sp.nnode, PK_NODES), done = nblk * PK_NODES / 64 * NU16 * 64 * 2;
        const long long nw = npk * NU16 * 64 * 2;  // the copy in 8-byte words
        if (done < nw)
            hipLaunchKernelGGL(k_zero, dim3(vgrid(nw - done)), dim3(256), 0, st, (double *)s->sm16 + done, nw - done);
        const dim3 gp((unsigned)nblk), bp(64 * PK_NODES);
        uint4 *o16 = s->sm16;
        if (svt == IGAP4H_SMOOTH_F16) {
            if (f32) hipLaunchKernelGGL((k_pack_16<float, 1>), gp, bp, 0, st, sp, (const float *)vals, s->rs16, o16);
            else hipLaunchKernelGGL((k_pack_16<double, 1>), gp, bp, 0, st, sp, (const double *)vals, s->rs16, o16);
        } else {
            if (f32) hipLaunchKernelGGL((k_pack_16<float, 0>), gp, bp, 0, st, sp, (const float *)vals, s->rs16, o16);
            else hipLaunchKernelGGL((k_pack_16<double, 0>), gp, bp, 0, st, sp, (const double *)vals, s->rs16, o16);
        }
        HIPCHK(hipGetLastError());
        return 0;
    }

    // line-staged row pass: one work-group per z-line segment
    int spmv(const double *x, double *y, const double *b = nullptr) const
    {
        const LineGeom lg = line_geom(sp.nb);
        const unsigned g = (unsigned)((long long)sp.nb[0] * sp.nb[1] * lg.nseg);
        const size_t sh = line_lds_bytes(lg);
        if (f32) {
            if (b) hipLaunchKernelGGL((k_spmv_line<float, 1, 0>), dim3(g), dim3(LW * 64), sh, st, sp, lg, (const float *)vals, x, b, y, nullptr);
            else hipLaunchKernelGGL((k_spmv_line<float, 0, 0>), dim3(g), dim3(LW * 64), sh, st, sp, lg, (const float *)vals, x, b, y, nullptr);
        } else {
            if (b) hipLaunchKernelGGL((k_spmv_line<double, 1, 0>), dim3(g), dim3(LW * 64), sh, st, sp, lg, (const double *)vals, x, b, y, nullptr);
            else hipLaunchKernelGGL((k_spmv_line<double, 0, 0>), dim3(g), dim3(LW * 64), sh, st, sp, lg, (const double *)vals, x, b, y, nullptr);
        }
        HIPCHK(hipGetLastError());
        return 0;
    }
    // y = K x and partial[block] of x.y; returns the number of partials
    int spmv_dot(const double *x, double *y, double *partial, int *nparts) const
    {
        const LineGeom lg = line_geom(sp.nb);
        const unsigned g = (unsigned)((long long)sp.nb[0] * sp.nb[1] * lg.nseg);
        const size_t sh = line_lds_bytes(lg);
        if (f32) hipLaunchKernelGGL((k_spmv_line<float, 0, 1>), dim3(g), dim3(LW * 64), sh, st, sp, lg, (const float *)vals, x, nullptr, y, partial);
        else hipLaunchKernelGGL((k_spmv_line<double, 0, 1>), dim3(g), dim3(LW * 64), sh, st, sp, lg, (const double *)vals, x, nullptr, y, partial);
        HIPCHK(hipGetLastError());
        *nparts = (int)g;
        return 0;
    }
    int gs_sweep(size_t l, bool bwd, const double *b, double *x) const
    {
        const MGLevel &L = s->mg[l];
        SpArgs a = l ? level_args(L) : sp;
        if (l && L.nnode <= MG_WG_NODES) {
            if (bwd) hipLaunchKernelGGL(k_gs_sweep_wg<1>, dim3(1), dim3(1024), 0, st, a, (const SMT *)L.sm, (const double *)L.Dinv, b, x);
            else hipLaunchKernelGGL(k_gs_sweep_wg<0>, dim3(1), dim3(1024), 0, st, a, (const SMT *)L.sm, (const double *)L.Dinv, b, x);
            HIPCHK(hipGetLastError());
            return 0;
        }
        const double *Dinv = l ? L.Dinv : s->Dinv;
        for (int cc = 0; cc < 27; cc++) {
            int col = bwd ? 26 - cc : cc;
            int c0 = col / 9, c1 = (col / 3) % 3, c2 = col % 3;
            int m0 = (a.nb[0] - c0 + 2) / 3, m1 = (a.nb[1] - c1 + 2) / 3, m2 = (a.nb[2] - c2 + 2) / 3;
            long long ncol = (long long)m0 * m1 * m2;
            if (ncol <= 0) continue;
            dim3 g(grid_for(ncol, 4));
            if (l) hipLaunchKernelGGL((k_gs_colour<1, SMT>), g, dim3(256), 0, st, a, c0, c1, c2, m1, m2, ncol, (const SMT *)L.sm, Dinv, b, x);
            else if (f32) hipLaunchKernelGGL((k_gs_colour<0, float>), g, dim3(256), 0, st, a, c0, c1, c2, m1, m2, ncol, (const float *)vals, Dinv, b, x);
            else hipLaunchKernelGGL((k_gs_colour<0, double>), g, dim3(256), 0, st, a, c0, c1, c2, m1, m2, ncol, (const double *)vals, Dinv, b, x);
            HIPCHK(hipGetLastError());
        }
        return 0;
    }
    // the coarsest level's operator on the host (stencil storage)
    int dense_fetch(const MGLevel &B, std::vector<SMT> &h) const
    {
        h.resize((size_t)B.nnode * 1125);
        HIPCHK(hipMemcpyAsync(h.data(), B.sm, h.size() * sizeof(SMT), hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        return 0;
    }
    // its dense form and inverse (host only); false when it is singular
    static bool dense_invert(const MGLevel &B, const std::vector<SMT> &h, std::vector<double> &inv)
    {
        std::vector<double> A((size_t)B.ndof * B.ndof, 0.0);
        for (long long n = 0; n < B.nnode; n++) {
            int a[3] = {(int)(n / ((long long)B.nb[2] * B.nb[1])), (int)((n / B.nb[2]) % B.nb[1]), (int)(n % B.nb[2])};
            for (int o = 0; o < 125; o++) {
                int b0 = a[0] + o / 25 - 2, b1 = a[1] + (o / 5) % 5 - 2, b2 = a[2] + o % 5 - 2;
                if (b0 < 0 || b0 >= B.nb[0] || b1 < 0 || b1 >= B.nb[1] || b2 < 0 || b2 >= B.nb[2]) continue;
                long long m = ((long long)b0 * B.nb[1] + b1) * B.nb[2] + b2;
                for (int i = 0; i < 3; i++)
                    for (int j = 0; j < 3; j++)
                        A[(size_t)(3 * n + i) * B.ndof + 3 * m + j] = h[(size_t)n * 1125 + i * 375 + o * 3 + j];
            }
        }
        for (long long r = 0; r < B.ndof; r++) {  // dofs removed with the Dirichlet rows
            bool empty = true;
            for (long long q = 0; q < B.ndof && empty; q++) empty = A[(size_t)r * B.ndof + q] == 0.0;
            if (empty) A[(size_t)r * B.ndof + r] = 1.0;
        }
        return dense_inverse(A, (int)B.ndof, inv);
    }
    int dense_upload(MGLevel &B, const std::vector<double> &inv) const
    {
        HIPCHK(hipMemcpyAsync(B.Ainv, inv.data(), inv.size() * sizeof(double), hipMemcpyHostToDevice, st));
        HIPCHK(hipStreamSynchronize(st));  // inv is freed by the caller
        return 0;
    }
    // dense inverse of level B, in order: 0, 1 when the operator is singular,
    // or an error code
    int dense_setup(MGLevel &B) const
    {
        std::vector<SMT> h;
        std::vector<double> inv;
        int rc;
        if ((rc = dense_fetch(B, h))) return rc;
        if (!dense_invert(B, h, inv)) return 1;
        return dense_upload(B, inv);
    }
    static void free_level(MGLevel &L)
    {
        for (double *p : {L.Dblk, L.Dinv, L.x, L.b, L.r, L.xfer_w, L.Ainv, L.d[0], L.d[1], L.ev}) (void)hipFree(p);
        for (SMT *p : {L.sm, L.stage[0], L.stage[1]}) (void)hipFree(p);
        (void)hipFree(L.xfer_i);
    }
    int mg_setup() const
    {
        int rc = 0;
        for (size_t l = 0; l + 1 < s->mg.size(); l++) {
            const MGLevel &F = s->mg[l];
            const MGLevel &C = s->mg[l + 1];
            int3 g0 = make_int3(F.nb[0], F.nb[1], F.nb[2]);
            int3 g1 = make_int3(C.nb[0], F.nb[1], F.nb[2]);
            int3 g2 = make_int3(C.nb[0], C.nb[1], F.nb[2]);
            int3 g3 = make_int3(C.nb[0], C.nb[1], C.nb[2]);
            auto nthr = [](int3 g) { return grid_for((long long)g.x * g.y * g.z * 125, 256); };
            if (l == 0) {
                if (!s->nmask && dev_malloc((void **)&s->nmask, (size_t)sp.nnode) != hipSuccess) {
                    s->nmask = nullptr;
                    return fail(IGAP4H_ENOMEM, "hipMalloc of %lld bytes failed", sp.nnode);
                }
                hipLaunchKernelGGL(k_node_mask, dim3(grid_for(sp.nnode, 256)), dim3(256), 0, st, sp.nnode, (const unsigned char *)c->d_mask, s->nmask);
                const unsigned char *dm = s->nmask;
                if (f32) hipLaunchKernelGGL((k_galerkin<1, float>), dim3(nthr(g1)), dim3(256), 0, st, 0, g0, g1, sp.rowptr, (const float *)vals, F.T.d[0], F.stage[0], dm);
                else hipLaunchKernelGGL((k_galerkin<1, double>), dim3(nthr(g1)), dim3(256), 0, st, 0, g0, g1, sp.rowptr, (const double *)vals, F.T.d[0], F.stage[0], dm);
            } else {
                hipLaunchKernelGGL((k_galerkin<0, SMT>), dim3(nthr(g1)), dim3(256), 0, st, 0, g0, g1, (const long long *)nullptr, (const SMT *)F.sm, F.T.d[0], F.stage[0], (const unsigned char *)nullptr);
            }
            HIPCHK(hipGetLastError());
            hipLaunchKernelGGL((k_galerkin<0, SMT>), dim3(nthr(g2)), dim3(256), 0, st, 1, g1, g2, (const long long *)nullptr, (const SMT *)F.stage[0], F.T.d[1], F.stage[1], (const unsigned char *)nullptr);
            hipLaunchKernelGGL((k_galerkin<0, SMT>), dim3(nthr(g3)), dim3(256), 0, st, 2, g2, g3, (const long long *)nullptr, (const SMT *)F.stage[1], F.T.d[2], C.sm, (const unsigned char *)nullptr);
            HIPCHK(hipGetLastError());
            hipLaunchKernelGGL(k_block_diag_sm, dim3(grid_for(C.nnode, 256)), dim3(256), 0, st, C.nnode, (const SMT *)C.sm, C.Dblk, C.Dinv);
            HIPCHK(hipGetLastError());
        }
        // the coarsest level's inverse is formed on a host thread while the
        // device runs the eigenvalue estimates
        MGLevel &B0 = s->mg.back();
        std::vector<SMT> hB;
        std::vector<double> hinv;
        std::future<bool> job;
        if (B0.dense) {
            if ((rc = dense_fetch(B0, hB))) return rc;
            try {
                job = std::async(std::launch::async, [&B0, &hB, &hinv] { return dense_invert(B0, hB, hinv); });
            } catch (...) {  // no thread: inverted below, in order
            }
        }
        if (chebyshev())
            for (size_t l = 0; l + 1 < s->mg.size(); l++)
                if ((rc = estimate_lmax(l))) return rc;  // the job's future waits for it
        int dr = 0;
        if (B0.dense) {
            const bool ok = job.valid() ? job.get() : dense_invert(B0, hB, hinv);
            dr = ok ? dense_upload(B0, hinv) : 1;
        }
        if (dr == 1 && s->mg.size() > 2 && s->mg[s->mg.size() - 2].ndof <= MG_DENSE_MAX) {
            free_level(s->mg.back());
            s->mg.pop_back();
            MGLevel &N = s->mg.back();
            if ((rc = dev_alloc(&N.Ainv, (size_t)N.ndof * N.ndof))) return rc;
            N.dense = true;
            dr = dense_setup(N);
        }
        if (dr == 1) return fail(IGAP4H_EINVAL, "multigrid: coarsest operator is singular");
        if (dr) return dr;
        (void)rc;
        return 0;
    }
    int norm_n(const double *x, long long n, double *out_dev) const
    {
        hipLaunchKernelGGL(k_dot2, dim3(NRB), dim3(256), 0, st, n, x, x, (const double *)nullptr, (const double *)nullptr, s->partial);
        HIPCHK(hipGetLastError());
        return finish(s->partial, 1, NRB, out_dev, true);
    }
    // largest eigenvalue of Dinv K on level l by power iteration (on the
    // fine level with a 16-bit smoothing copy K~: of Dinv K~)
    int estimate_lmax(size_t l) const
    {
        MGLevel &L = s->mg[l];
        const double *Dinv = l ? L.Dinv : s->Dinv;
        double *v = L.d[0], *w = L.d[1];
        const long long n = L.ndof;
        int rc;
        if (L.ev_ok) {  // a later Newton step of the same solve: start from the last estimate
            HIPCHK(hipMemcpyAsync(v, L.ev, n * sizeof(double), hipMemcpyDeviceToDevice, st));
        } else {
            hipLaunchKernelGGL(k_hash_vec, dim3(vgrid(n)), dim3(256), 0, st, n, v);
            HIPCHK(hipGetLastError());
        }
        if ((rc = norm_n(v, n, s->sc + 12))) return rc;
        hipLaunchKernelGGL(k_scale_inv, dim3(vgrid(n)), dim3(256), 0, st, n, (const double *)(s->sc + 12), v);
        const int its = L.ev_ok ? CHEB_POWER_WARM : CHEB_POWER_ITS;
        for (int it = 0; it < its; it++) {
            if (l == 0 && svt != IGAP4H_SMOOTH_SAME) {  // the smoother's operator: Dinv times the 16-bit copy
                if ((rc = cheb16<0>(v, nullptr, nullptr, nullptr, w, 0.0, 1.0, 1))) return rc;
            } else if (l == 0) {  // the line pass, then Dinv per node
                if ((rc = spmv(v, L.x))) return rc;
                hipLaunchKernelGGL(k_apply_dinv, dim3(grid_for(L.nnode, 256)), dim3(256), 0, st, L.nnode, Dinv, (const double *)L.x, w);
                HIPCHK(hipGetLastError());
            } else {
                LEVEL_LAUNCH(l, k_dinv_k, Dinv, (const double *)v, w);
            }
            if ((rc = norm_n(w, n, s->sc + 12))) return rc;
            hipLaunchKernelGGL(k_scale_inv, dim3(vgrid(n)), dim3(256), 0, st, n, (const double *)(s->sc + 12), w);
            HIPCHK(hipGetLastError());
            std::swap(v, w);
        }
        HIPCHK(hipMemcpyAsync(L.ev, v, n * sizeof(double), hipMemcpyDeviceToDevice, st));
        L.ev_ok = true;
        if ((rc = read_scalars(12, 1))) return rc;
        L.lmax = s->hp[12];
        if (!(L.lmax > 0.0) || !std::isfinite(L.lmax)) return fail(IGAP4H_EINVAL, "multigrid: eigenvalue estimate %g on level %zu", L.lmax, l);
        return 0;
    }
    // Chebyshev smoothing of degree m on level l for K x = b (from x = 0 if
    // from_zero; started: that start was made by the restriction); with
    // want_res, L.r = b - K x on return.  With pending, the last update
    // x += d is left to the caller: *pending = d.  Both kinds run
    //   d_0 = s0 Dinv r_0;  x += d_{k-1}, r -= K d_{k-1}, d_k = c1_k d_{k-1} + c2_k Dinv r_k
    // with the coefficients of cheb_s0 / cheb_c.
    int cheb(size_t l, const double *b, double *x, int m, bool from_zero, bool want_res, bool started = false,
             const double **pending = nullptr) const
    {
        const MGLevel &L = s->mg[l];
        const double *Dinv = l ? L.Dinv : s->Dinv;
        const double s0 = cheb_s0(l);
        double *r = L.r, *d0 = L.d[0], *d1 = L.d[1];
        const dim3 gn(grid_for(L.nnode, 256));
        int rc;
        if (from_zero) {
            if (!started) {
                hipLaunchKernelGGL(k_cheb_zero, gn, dim3(256), 0, st, L.nnode, Dinv, b, r, d0, x, s0);
                HIPCHK(hipGetLastError());
            }
        } else if (l == 0 && svt != IGAP4H_SMOOTH_SAME) {
            if ((rc = cheb16<1>(x, b, r, nullptr, d0, 0.0, s0, 1))) return rc;
        } else if (l == 0) {  // fine level: the row pass is the SpMV kernel
            if ((rc = spmv(x, r, b))) return rc;
            hipLaunchKernelGGL(k_cheb_dfromr, gn, dim3(256), 0, st, L.nnode, Dinv, (const double *)r, d0, s0);
            HIPCHK(hipGetLastError());
        } else {
            LEVEL_LAUNCH(l, k_cheb_init, Dinv, b, (const double *)x, r, d0, s0);
        }
        double rho = 0.0;
        for (int k = 1; k <= m; k++) {
            const bool last = k == m;
            if (last && !want_res && pending) {
                *pending = d0;
                break;
            }
            if (last && !want_res) {
                hipLaunchKernelGGL(k_waxpy, dim3(vgrid(L.ndof)), dim3(256), 0, st, L.ndof, 1.0, (const double *)x, (const double *)d0, x);
                HIPCHK(hipGetLastError());
                break;
            }
            double c1, c2;
            cheb_c(l, k, &rho, &c1, &c2);
            if (l == 0 && svt != IGAP4H_SMOOTH_SAME) {
                if ((rc = cheb16<2>(d0, nullptr, r, x, d1, c1, c2, last ? 0 : 1))) return rc;
            } else if (l == 0) {
                if ((rc = spmv(d0, L.x))) return rc;
                hipLaunchKernelGGL(k_cheb_update, gn, dim3(256), 0, st, L.nnode, Dinv, (const double *)L.x, (const double *)d0, d1, r, x,
                                   c1, c2, last ? 0 : 1);
                HIPCHK(hipGetLastError());
            } else {
                LEVEL_LAUNCH(l, k_cheb_step, Dinv, (const double *)d0, d1, r, x, c1, c2, last ? 0 : 1);
            }
            std::swap(d0, d1);
        }
        return 0;
    }
    // the first direction's factor s0 on level l: 1 / theta, theta = (bmax +
    // amin) / 2 over the interval [amin, bmax] = [1.1 lmax / ratio, 1.1 lmax]
    // (first kind); 4 / (3 bmax) (fourth kind)
    double cheb_s0(size_t l) const
    {
        const double bmax = 1.1 * s->mg[l].lmax, amin = bmax / cratio;
        return smoother == IGAP4H_MG_CHEB4 ? 4.0 / (3.0 * bmax) : 2.0 / (bmax + amin);
    }
    // step k's (c1, c2); rho carries the first kind's recurrence (0 before step 1).
    // First kind, delta = (bmax - amin) / 2, sigma = theta / delta:
    //   rho_0 = 1 / sigma, rho_k = 1 / (2 sigma - rho_{k-1}), c1 = rho_k rho_{k-1}, c2 = 2 rho_k / delta.
    // Fourth kind (Lottes, "Optimal polynomial smoothers for multigrid
    // V-cycles", Numer. Linear Algebra Appl. 2023, unit weights):
    //   c1 = (2k - 1) / (2k + 3), c2 = (8k + 4) / ((2k + 3) bmax).
    void cheb_c(size_t l, int k, double *rho, double *c1, double *c2) const
    {
        const double bmax = 1.1 * s->mg[l].lmax, amin = bmax / cratio;
        if (smoother == IGAP4H_MG_CHEB4) {
            *c1 = (2.0 * k - 1.0) / (2.0 * k + 3.0);
            *c2 = (8.0 * k + 4.0) / ((2.0 * k + 3.0) * bmax);
            return;
        }
        const double theta = 0.5 * (bmax + amin), delta = 0.5 * (bmax - amin), sigma = theta / delta;
        if (k == 1) *rho = 1.0 / sigma;
        const double rn = 1.0 / (2.0 * sigma - *rho);
        *c1 = rn * *rho;
        *c2 = 2.0 * rn / delta;
        *rho = rn;
    }
    // x = V-cycle(b) on level l (x overwritten); started and pending as in cheb
    int vcycle(size_t l, const double *b, double *x, bool started = false, const double **pending = nullptr) const
    {
        const MGLevel &L = s->mg[l];
        const long long n = L.ndof;
        int rc;
        if (l + 1 == s->mg.size()) {
            if (L.dense) {
                hipLaunchKernelGGL(k_dense_mv, dim3(grid_for(n, 4)), dim3(256), 0, st, (int)n, (const double *)L.Ainv, b, x);
                HIPCHK(hipGetLastError());
                return 0;
            }
            hipLaunchKernelGGL(k_zero, dim3(vgrid(n)), dim3(256), 0, st, x, n);
            HIPCHK(hipGetLastError());
            for (int it = 0; it < 4; it++)
                if ((rc = gs_sweep(l, false, b, x)) || (rc = gs_sweep(l, true, b, x))) return rc;
            return 0;
        }
        const MGLevel &C = s->mg[l + 1];
        if (chebyshev()) {
            const double *res = b;
            const int pre = deg(l, nu_pre), post = deg(l, nu_post);
            if (pre > 0) {
                if ((rc = cheb(l, b, x, pre, true, true, started))) return rc;
                res = L.r;
            } else {
                hipLaunchKernelGGL(k_zero, dim3(vgrid(n)), dim3(256), 0, st, x, n);
                HIPCHK(hipGetLastError());
            }
            // the restriction also makes the coarse level's Chebyshev start
            // (unless that level is the coarsest), and the prolongation adds
            // the coarse smoother's last update
            const bool cstart = deg(l + 1, nu_pre) > 0 && l + 2 < s->mg.size();
            const double cinv = cstart ? cheb_s0(l + 1) : 0.0;
            hipLaunchKernelGGL(k_restrict, dim3(grid_for(C.nnode, 4)), dim3(256), 0, st, L.T, res, C.b,
                               l ? (const unsigned char *)nullptr : (const unsigned char *)c->d_mask,
                               cstart ? (const double *)C.Dinv : nullptr, C.r, C.x, C.d[0], cinv);
            HIPCHK(hipGetLastError());
            const double *cpend = nullptr;
            if ((rc = vcycle(l + 1, C.b, C.x, cstart, &cpend))) return rc;
            hipLaunchKernelGGL(k_prolong_add, dim3(grid_for(L.nnode, 256)), dim3(256), 0, st, L.T, (const double *)C.x, cpend, x,
                               l ? (const unsigned char *)nullptr : (const unsigned char *)c->d_mask);
            HIPCHK(hipGetLastError());
            if (post > 0) return cheb(l, b, x, post, false, false, false, pending);
            return 0;
        }
        hipLaunchKernelGGL(k_zero, dim3(vgrid(n)), dim3(256), 0, st, x, n);
        HIPCHK(hipGetLastError());
        for (int k = 0; k < nu_pre; k++)
            if ((rc = gs_sweep(l, false, b, x))) return rc;
        const double *res = b;  // with no pre-smoothing x = 0 and the residual is b
        if (nu_pre > 0) {
            if (l) {
                hipLaunchKernelGGL(k_resid_sm, dim3(grid_for(L.nnode, 4)), dim3(256), 0, st, level_args(L), (const SMT *)L.sm, (const double *)x, b, L.r);
                HIPCHK(hipGetLastError());
            } else if ((rc = spmv(x, L.r, b))) {
                return rc;
            }
            res = L.r;
        }
        hipLaunchKernelGGL(k_restrict, dim3(grid_for(C.nnode, 4)), dim3(256), 0, st, L.T, res, C.b,
                           l ? (const unsigned char *)nullptr : (const unsigned char *)c->d_mask, (const double *)nullptr,
                           (double *)nullptr, (double *)nullptr, (double *)nullptr, 0.0);
        HIPCHK(hipGetLastError());
        if ((rc = vcycle(l + 1, C.b, C.x))) return rc;
        hipLaunchKernelGGL(k_prolong_add, dim3(grid_for(L.nnode, 256)), dim3(256), 0, st, L.T, (const double *)C.x, (const double *)nullptr, x,
                           l ? (const unsigned char *)nullptr : (const unsigned char *)c->d_mask);
        HIPCHK(hipGetLastError());
        for (int k = 0; k < nu_post; k++)
            if ((rc = gs_sweep(l, true, b, x))) return rc;
        return 0;
    }
    // z = V-cycle(r), replayed from a captured graph (z must be s->z)
    int mg_apply(const double *r, double *z) const
    {
        if (z != s->z) return vcycle(0, r, z);
        const long long n = 3 * sp.nnode;
        if (!s->mg_exec || s->key_vals != vals || s->key_f32 != f32 || s->key_pre != nu_pre || s->key_post != nu_post ||
            s->key_smoother != smoother || s->key_svt != svt || s->key_cdeg != cdeg || s->key_cratio != cratio) {
            if (s->mg_exec) {
                HIPCHK(hipGraphExecDestroy(s->mg_exec));
                s->mg_exec = nullptr;
            }
            int rc;
            if (!s->mg_in && (rc = dev_alloc(&s->mg_in, (size_t)n))) return rc;
            if (!s->cap) HIPCHK(hipStreamCreateWithFlags(&s->cap, hipStreamNonBlocking));
            Op cop = *this;
            cop.st = s->cap;
            HIPCHK(hipStreamBeginCapture(s->cap, hipStreamCaptureModeRelaxed));
            rc = cop.vcycle(0, s->mg_in, s->z);
            hipGraph_t g = nullptr;
            hipError_t e = hipStreamEndCapture(s->cap, &g);
            if (rc) {
                if (g) (void)hipGraphDestroy(g);
                return rc;
            }
            if (e != hipSuccess) return fail(IGAP4H_EHIP, "V-cycle capture: %s", hipGetErrorString(e));
            e = hipGraphInstantiate(&s->mg_exec, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (e != hipSuccess) {
                s->mg_exec = nullptr;
                return fail(IGAP4H_EHIP, "V-cycle graph instantiation: %s", hipGetErrorString(e));
            }
            s->key_vals = vals;
            s->key_f32 = f32;
            s->key_pre = nu_pre;
            s->key_post = nu_post;
            s->key_smoother = smoother;
            s->key_svt = svt;
            s->key_cdeg = cdeg;
            s->key_cratio = cratio;
        }
        HIPCHK(hipMemcpyAsync(s->mg_in, r, n * sizeof(double), hipMemcpyDeviceToDevice, st));
        HIPCHK(hipGraphLaunch(s->mg_exec, st));
        return 0;
    }
    bool pc_key_matches() const
    {
        const int k[6] = {pc, smoother, svt, cdeg, nu_pre, nu_post};
        for (int i = 0; i < 6; i++)
            if (s->pc_key[i] != k[i]) return false;
        return s->pc_key_ratio == cratio;
    }
    int setup_pc() const
    {
        s->pc_vals = vals;
        const int k[6] = {pc, smoother, svt, cdeg, nu_pre, nu_post};
        for (int i = 0; i < 6; i++) s->pc_key[i] = k[i];
        s->pc_key_ratio = cratio;
        if (pc == IGAP4H_PC_NONE) return 0;
        unsigned g = grid_for(sp.nnode, 256);
        if (f32) hipLaunchKernelGGL((k_block_diag<float>), dim3(g), dim3(256), 0, st, sp, (const float *)vals, s->Dblk, s->Dinv);
        else hipLaunchKernelGGL((k_block_diag<double>), dim3(g), dim3(256), 0, st, sp, (const double *)vals, s->Dblk, s->Dinv);
        HIPCHK(hipGetLastError());
        if (pc == IGAP4H_PC_MG) {
            int rc = mg_build(c, s);
            if (rc) return rc;
            if (chebyshev() && svt != IGAP4H_SMOOTH_SAME && (rc = pack_16())) return rc;
            // the V-cycle graph holds this set-up's Chebyshev coefficients (from
            // lmax) as kernel arguments: capture it again
            s->key_vals = nullptr;
            return mg_setup();
        }
        return 0;
    }
    template <class VT, int BWD>
    int sgs_pass(const double *rhs, double *v) const
    {
        for (int cc = 0; cc < 27; cc++) {
            int col = BWD ? 26 - cc : cc;
            int c0 = col / 9, c1 = (col / 3) % 3, c2 = col % 3;
            int m0 = (sp.nb[0] - c0 + 2) / 3, m1 = (sp.nb[1] - c1 + 2) / 3, m2 = (sp.nb[2] - c2 + 2) / 3;
            long long ncol = (long long)m0 * m1 * m2;
            if (ncol <= 0) continue;
            hipLaunchKernelGGL((k_sgs_colour<VT, BWD>), dim3(grid_for(ncol, 4)), dim3(256), 0, st, sp, c0, c1, c2, m1, m2, ncol,
                               (const VT *)vals, (const double *)s->Dblk, (const double *)s->Dinv, rhs, v);
            HIPCHK(hipGetLastError());
        }
        return 0;
    }
    // z = M^{-1} r (z != r)
    int apply_pc(const double *r, double *z) const
    {
        long long n = 3 * sp.nnode;
        if (pc == IGAP4H_PC_NONE) {
            HIPCHK(hipMemcpyAsync(z, r, n * sizeof(double), hipMemcpyDeviceToDevice, st));
            return 0;
        }
        if (pc == IGAP4H_PC_MG) return mg_graph ? mg_apply(r, z) : vcycle(0, r, z);
        if (pc == IGAP4H_PC_BJACOBI3) {
            hipLaunchKernelGGL(k_bjacobi, dim3(grid_for(sp.nnode, 256)), dim3(256), 0, st, sp.nnode, (const double *)s->Dinv, r, z);
            HIPCHK(hipGetLastError());
            return 0;
        }
        // symmetric Gauss-Seidel: forward into s->t, backward into z
        double *y = s->t;
        hipLaunchKernelGGL(k_zero, dim3(vgrid(n)), dim3(256), 0, st, y, n);
        hipLaunchKernelGGL(k_zero, dim3(vgrid(n)), dim3(256), 0, st, z, n);
        HIPCHK(hipGetLastError());
        int rc = f32 ? sgs_pass<float, 0>(r, y) : sgs_pass<double, 0>(r, y);
        if (rc) return rc;
        return f32 ? sgs_pass<float, 1>(y, z) : sgs_pass<double, 1>(y, z);
    }
    // out[0..np) = sums of np consecutive partial rows of nparts each
    int finish(const double *partial, int np, int nparts, double *out, bool sq = false) const
    {
        if (sq) hipLaunchKernelGGL(k_finish<1>, dim3(np), dim3(256), 0, st, partial, nparts, out);
        else hipLaunchKernelGGL(k_finish<0>, dim3(np), dim3(256), 0, st, partial, nparts, out);
        HIPCHK(hipGetLastError());
        return 0;
    }
    int norm(const double *x, double *out_dev) const
    {
        long long n = 3 * sp.nnode;
        hipLaunchKernelGGL(k_dot2, dim3(NRB), dim3(256), 0, st, n, x, x, (const double *)nullptr, (const double *)nullptr, s->partial);
        HIPCHK(hipGetLastError());
        return finish(s->partial, 1, NRB, out_dev, true);
    }
    int read_scalars(int i0, int cnt) const
    {
        HIPCHK(hipMemcpyAsync(s->hp + i0, s->sc + i0, cnt * sizeof(double), hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        return 0;
    }
};

// right-preconditioned GMRES(m), classical Gram-Schmidt with one
// re-orthogonalisation; stop at ||b - K x|| <= rtol ||b|| (the least-squares
// estimate inside a cycle, the true residual at each restart)
int gmres(const Op &A, const double *b, double *x, int m, double rtol, int maxit, int *its_out, double *relres)
{
    Solver *s = A.s;
    const long long n = s->ndof;
    hipStream_t st = A.st;
    double *sc = s->sc, *hp = s->hp;
    double *h1 = s->hd, *h2 = s->hd + (m + 1), *ydev = s->hd + 2 * (m + 1);  // device
    double *hh = s->hh;                                                       // pinned host
    std::vector<double> H((size_t)(m + 1) * m), cs(m), sn(m), g(m + 1), y(m);
    int rc;
    if ((rc = A.norm(b, sc + 0))) return rc;
    if ((rc = A.read_scalars(0, 1))) return rc;
    const double bnorm = hp[0];
    *its_out = 0;
    *relres = 0.0;
    if (bnorm == 0.0) {
        hipLaunchKernelGGL(k_zero, dim3(vgrid(n)), dim3(256), 0, st, x, n);
        HIPCHK(hipGetLastError());
        return 0;
    }
    const double tol = rtol * bnorm;
    int its = 0;
    double res = 0.0;
    while (true) {
        double *V0 = s->V;
        if ((rc = A.spmv(x, V0, b))) return rc;  // r0 = b - K x
        if ((rc = A.norm(V0, sc + 1))) return rc;
        if ((rc = A.read_scalars(1, 1))) return rc;
        const double beta = hp[1];
        res = beta;
        if (beta <= tol || its >= maxit) break;
        hipLaunchKernelGGL(k_scale_inv, dim3(vgrid(n)), dim3(256), 0, st, n, (const double *)(sc + 1), V0);
        HIPCHK(hipGetLastError());
        std::fill(H.begin(), H.end(), 0.0);
        std::fill(g.begin(), g.end(), 0.0);
        g[0] = beta;
        int j = 0;
        for (; j < m;) {
            if ((rc = basis_reserve(s, j + 2, st))) return rc;
            double *Vj = s->V + (long long)j * n, *W = s->V + (long long)(j + 1) * n;
            if ((rc = A.apply_pc(Vj, s->z))) return rc;
            if ((rc = A.spmv(s->z, W))) return rc;
            for (int pass = 0; pass < 2; pass++) {  // h = V^T w; w -= V h
                double *hc = pass ? h2 : h1;
                hipLaunchKernelGGL(k_mdot, dim3(NRB), dim3(256), 0, st, (const double *)s->V, n, j + 1, (const double *)W, n, s->partial);
                HIPCHK(hipGetLastError());
                if ((rc = A.finish(s->partial, j + 1, NRB, hc))) return rc;
                hipLaunchKernelGGL(k_mgemv, dim3(vgrid(n)), dim3(256), 0, st, (const double *)s->V, n, j + 1, (const double *)hc, -1.0, W, n);
                HIPCHK(hipGetLastError());
            }
            if ((rc = A.norm(W, sc + 2))) return rc;
            HIPCHK(hipMemcpyAsync(hh, s->hd, 2 * (m + 1) * sizeof(double), hipMemcpyDeviceToHost, st));
            if ((rc = A.read_scalars(2, 1))) return rc;  // also waits for hh
            const double hn = hp[2];
            double *h = &H[(size_t)j * (m + 1)];
            for (int i = 0; i <= j; i++) h[i] = hh[i] + hh[m + 1 + i];
            h[j + 1] = hn;
            if (hn > 0.0) {
                hipLaunchKernelGGL(k_scale_inv, dim3(vgrid(n)), dim3(256), 0, st, n, (const double *)(sc + 2), W);
                HIPCHK(hipGetLastError());
            }
            for (int i = 0; i < j; i++) {  // previous Givens rotations
                double t0 = cs[i] * h[i] + sn[i] * h[i + 1];
                h[i + 1] = -sn[i] * h[i] + cs[i] * h[i + 1];
                h[i] = t0;
            }
            double den = std::hypot(h[j], h[j + 1]);
            cs[j] = den == 0.0 ? 1.0 : h[j] / den;
            sn[j] = den == 0.0 ? 0.0 : h[j + 1] / den;
            h[j] = den;
            h[j + 1] = 0.0;
            g[j + 1] = -sn[j] * g[j];
            g[j] = cs[j] * g[j];
            its++;
            j++;
            res = std::fabs(g[j]);
            if (res <= tol || its >= maxit || hn == 0.0) break;
        }
        // y = R^{-1} g; x += M^{-1} V y
        for (int i = j - 1; i >= 0; i--) {
            double v = g[i];
            for (int k = i + 1; k < j; k++) v -= H[(size_t)k * (m + 1) + i] * y[k];
            double d = H[(size_t)i * (m + 1) + i];
            y[i] = d == 0.0 ? 0.0 : v / d;
        }
        for (int i = 0; i < j; i++) hh[2 * (m + 1) + i] = y[i];
        HIPCHK(hipMemcpyAsync(ydev, hh + 2 * (m + 1), j * sizeof(double), hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(k_zero, dim3(vgrid(n)), dim3(256), 0, st, s->w, n);
        hipLaunchKernelGGL(k_mgemv, dim3(vgrid(n)), dim3(256), 0, st, (const double *)s->V, n, j, (const double *)ydev, 1.0, s->w, n);
        HIPCHK(hipGetLastError());
        if ((rc = A.apply_pc(s->w, s->z))) return rc;
        hipLaunchKernelGGL(k_waxpy, dim3(vgrid(n)), dim3(256), 0, st, n, 1.0, (const double *)x, (const double *)s->z, x);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(st));  // hh is reused by the next cycle
        if (res <= tol || its >= maxit) break;
    }
    *its_out = its;
    *relres = res / bnorm;
    return 0;
}

// preconditioned conjugate gradients; device scalars sc[0..1] (r.z, two
// slots alternating), sc[2] (p.Kp), sc[3] (r.r)
int pcg(const Op &A, const double *b, double *x, double rtol, int maxit, int *its_out, double *relres)
{
    Solver *s = A.s;
    const long long n = s->ndof, nn = s->nnode;
    hipStream_t st = A.st;
    double *sc = s->sc, *hp = s->hp;
    int rc;
    if ((rc = A.norm(b, sc + 4))) return rc;
    if ((rc = A.read_scalars(4, 1))) return rc;
    const double bnorm = hp[4];
    *its_out = 0;
    *relres = 0.0;
    if (bnorm == 0.0) {
        hipLaunchKernelGGL(k_zero, dim3(vgrid(n)), dim3(256), 0, st, x, n);
        HIPCHK(hipGetLastError());
        return 0;
    }
    const double tol = rtol * bnorm;
    if ((rc = A.spmv(x, s->r, b))) return rc;  // r = b - K x
    if ((rc = A.apply_pc(s->r, s->z))) return rc;
    hipLaunchKernelGGL(k_dot2, dim3(NRB), dim3(256), 0, st, n, (const double *)s->r, (const double *)s->z,
                       (const double *)s->r, (const double *)s->r, s->partial);
    HIPCHK(hipGetLastError());
    if ((rc = A.finish(s->partial, 1, NRB, sc + 0))) return rc;
    if ((rc = A.finish(s->partial + NRB, 1, NRB, sc + 3))) return rc;
    HIPCHK(hipMemcpyAsync(s->p, s->z, n * sizeof(double), hipMemcpyDeviceToDevice, st));
    if ((rc = A.read_scalars(3, 1))) return rc;
    double res = std::sqrt(hp[3]);
    int its = 0, cur = 0;
    const bool bj = A.pc == IGAP4H_PC_BJACOBI3;
    while (res > tol && its < maxit) {
        int np;
        if ((rc = A.spmv_dot(s->p, s->q, s->pspmv, &np))) return rc;
        // p.Kp from the spmv's per-block partials (np of them)
        if ((rc = A.finish(s->pspmv, 1, np, sc + 2))) return rc;
        if (bj) {
            hipLaunchKernelGGL(k_cg_update<1>, dim3(NRB), dim3(256), 0, st, nn, (const double *)sc, cur, 2,
                               (const double *)s->p, (const double *)s->q, x, s->r, (const double *)s->Dinv, s->z, s->partial);
            HIPCHK(hipGetLastError());
        } else {
            hipLaunchKernelGGL(k_cg_update<0>, dim3(NRB), dim3(256), 0, st, nn, (const double *)sc, cur, 2,
                               (const double *)s->p, (const double *)s->q, x, s->r, (const double *)s->Dinv, s->z, s->partial);
            HIPCHK(hipGetLastError());
            if ((rc = A.apply_pc(s->r, s->z))) return rc;
            hipLaunchKernelGGL(k_dot2, dim3(NRB), dim3(256), 0, st, n, (const double *)s->r, (const double *)s->z,
                               (const double *)nullptr, (const double *)nullptr, s->partial);
            HIPCHK(hipGetLastError());
        }
        const int nxt = cur ^ 1;
        if ((rc = A.finish(s->partial, 1, NRB, sc + nxt))) return rc;
        if ((rc = A.finish(s->partial + NRB, 1, NRB, sc + 3))) return rc;
        hipLaunchKernelGGL(k_cg_dir, dim3(vgrid(n)), dim3(256), 0, st, n, (const double *)sc, nxt, cur, (const double *)s->z, s->p);
        HIPCHK(hipGetLastError());
        its++;
        cur = nxt;
        if ((rc = A.read_scalars(2, 2))) return rc;
        res = std::sqrt(hp[3]);
        if (!(hp[2] > 0.0)) {  // K (or M) not positive definite along p
            *its_out = its;
            *relres = res / bnorm;
            return fail(IGAP4H_EINVAL, "CG breakdown: p.Kp = %g at iteration %d (matrix not SPD; use GMRES)", hp[2], its);
        }
    }
    *its_out = its;
    *relres = res / bnorm;
    return 0;
}

int check_opts(const igap4h_newton_opts *o)
{
    if (!o) return fail(IGAP4H_EINVAL, "NULL options");
    if (o->ksp_type != IGAP4H_KSP_GMRES && o->ksp_type != IGAP4H_KSP_CG) return fail(IGAP4H_EINVAL, "bad ksp_type %d", o->ksp_type);
    if (o->pc < IGAP4H_PC_NONE || o->pc > IGAP4H_PC_MG) return fail(IGAP4H_EINVAL, "bad pc %d", o->pc);
    if (o->ksp_type == IGAP4H_KSP_GMRES && (o->restart < 1 || o->restart > 4096))
        return fail(IGAP4H_EINVAL, "GMRES restart %d outside 1..4096", o->restart);
    if (o->ksp_max_it < 1) return fail(IGAP4H_EINVAL, "ksp_max_it must be >= 1");
    if (o->pc == IGAP4H_PC_MG && (o->mg_pre < 0 || o->mg_pre > 8 || o->mg_post < 0 || o->mg_post > 8 || o->mg_pre + o->mg_post < 1))
        return fail(IGAP4H_EINVAL, "multigrid sweeps pre %d / post %d outside 0..8 (sum >= 1)", o->mg_pre, o->mg_post);
    if (o->mg_smoother != IGAP4H_MG_GS && o->mg_smoother != IGAP4H_MG_CHEBYSHEV && o->mg_smoother != IGAP4H_MG_CHEB4)
        return fail(IGAP4H_EINVAL, "bad mg_smoother %d", o->mg_smoother);
    if (o->mg_smooth_vals != IGAP4H_SMOOTH_SAME && o->mg_smooth_vals != IGAP4H_SMOOTH_BF16 && o->mg_smooth_vals != IGAP4H_SMOOTH_F16)
        return fail(IGAP4H_EINVAL, "bad mg_smooth_vals %d", o->mg_smooth_vals);
    if (o->mg_coarse_deg < 0 || o->mg_coarse_deg > 8) return fail(IGAP4H_EINVAL, "mg_coarse_deg %d outside 0..8", o->mg_coarse_deg);
    if (o->pc == IGAP4H_PC_MG && o->mg_smoother == IGAP4H_MG_CHEBYSHEV && !(o->mg_cheb_ratio > 1.0 && o->mg_cheb_ratio <= 1e6))
        return fail(IGAP4H_EINVAL, "mg_cheb_ratio %g outside (1, 1e6]", o->mg_cheb_ratio);
    if (o->vals_type != IGAP4H_VALS_F64 && o->vals_type != IGAP4H_VALS_F32) return fail(IGAP4H_EINVAL, "bad vals_type %d", o->vals_type);
    if (o->tangent_prec != IGAP4H_FP64 && o->tangent_prec != IGAP4H_FP32) return fail(IGAP4H_EINVAL, "bad tangent_prec %d", o->tangent_prec);
    return 0;
}

int linear_solve(const Op &A, const double *b, double *x, const igap4h_newton_opts *o, int *its, double *relres)
{
    if (o->ksp_type == IGAP4H_KSP_CG) return pcg(A, b, x, o->ksp_rtol, o->ksp_max_it, its, relres);
    return gmres(A, b, x, o->restart, o->ksp_rtol, o->ksp_max_it, its, relres);
}

using clk = std::chrono::steady_clock;
inline double secs(clk::time_point a, clk::time_point b) { return std::chrono::duration<double>(b - a).count(); }

}  // namespace

// ====================================================================== C-ABI
extern "C" int igap4h_newton_default_opts(igap4h_newton_opts *o)
{
    if (!o) return fail(IGAP4H_EINVAL, "NULL options");
    std::memset(o, 0, sizeof(*o));
    o->atol = 1e-10;  // snes_init.c:18
    o->rtol = 0.0;
    o->max_newton = 50;
    o->tangent_prec = IGAP4H_FP64;
    o->vals_type = IGAP4H_VALS_F64;
    o->ksp_type = IGAP4H_KSP_GMRES;  // snes_init.c:32
    o->restart = 300;                 // the reference's 10000 keeps every Krylov vector; 64^3
                                      // linear solves take 20-50 iterations, but the first Newton
                                      // step's (indefinite) system on finer meshes needs hundreds:
                                      // 96^3 1,852 iterations with restart 50, 525 with 300.
                                      // The basis grows on demand up to restart + 1 vectors
    o->ksp_rtol = 1e-5;               // PETSc default (snes_init.c:26 keeps it)
    o->ksp_max_it = 20000;
    o->ksp_guess = 1;                 // snes_init.c:25
    o->pc = IGAP4H_PC_MG;
    o->line_search = 1;               // snes_init.c:21-22
    o->stagnation = 1;                // snes_solve.c:30
    o->mg_pre = 5;                    // Chebyshev degree 5 before and after the coarse correction
    o->mg_post = 5;                   // on the finest level (64^3: fastest of degrees 2..6)
    o->mg_graph = 1;
    o->mg_smoother = IGAP4H_MG_CHEBYSHEV;
    o->lag_jacobian = 1;
    o->lag_persist = 0;
    o->mg_smooth_vals = IGAP4H_SMOOTH_SAME;
    o->mg_coarse_deg = 1;             // 64^3: 2 % more GMRES iterations than degree 3, 3-8 % less time
    o->mg_cheb_ratio = 200.0;         // with degree 5: 64^3 fastest of 20..400
    return 0;
}

extern "C" int igap4h_spmv_dev(igap4h_ctx *c, const void *d_vals, int vals_type, const double *d_x, double *d_y, void *stream)
{
    if (!c || !d_vals || !d_x || !d_y) return fail(IGAP4H_EINVAL, "NULL argument");
    if (vals_type != IGAP4H_VALS_F64 && vals_type != IGAP4H_VALS_F32) return fail(IGAP4H_EINVAL, "bad vals type %d", vals_type);
    Op A{c, nullptr, sp_args(c), d_vals, vals_type == IGAP4H_VALS_F32, IGAP4H_PC_NONE, (hipStream_t)stream};
    return A.spmv(d_x, d_y);
}

extern "C" int igap4h_linear_solve_dev(igap4h_ctx *c, const void *d_vals, int vals_type, const double *d_b, double *d_x,
                                       const igap4h_newton_opts *o, int *its, double *relres, void *stream)
{
    if (c && c->partitioned)
        return fail(IGAP4H_EINVAL, "the Newton-Krylov driver solves whole patches; a partitioned block only assembles");
    if (!c || !d_vals || !d_b || !d_x || !its || !relres) return fail(IGAP4H_EINVAL, "NULL argument");
    int rc = check_opts(o);
    if (rc) return rc;
    if (vals_type != IGAP4H_VALS_F64 && vals_type != IGAP4H_VALS_F32) return fail(IGAP4H_EINVAL, "bad vals type %d", vals_type);
    HIPCHK(hipSetDevice(c->device));
    if ((rc = solver_ws(c, o->ksp_type == IGAP4H_KSP_GMRES ? o->restart : 1, 0))) return rc;
    Op A{c, c->solver, sp_args(c), d_vals, vals_type == IGAP4H_VALS_F32, o->pc, (hipStream_t)stream};
    for (MGLevel &L : c->solver->mg) L.ev_ok = false;  // warm starts only within one solve
    A.nu_pre = o->mg_pre;
    A.nu_post = o->mg_post;
    A.mg_graph = o->mg_graph;
    A.smoother = o->mg_smoother;
    A.svt = o->mg_smooth_vals;
    A.cdeg = o->mg_coarse_deg;
    A.cratio = o->mg_cheb_ratio;
    if ((rc = A.setup_pc())) return rc;
    return linear_solve(A, d_b, d_x, o, its, relres);
}

extern "C" int igap4h_newton_solve_dev(igap4h_ctx *c, double *d_Uh, const igap4h_newton_opts *o, igap4h_newton_info *info,
                                       void *stream)
{
    if (c && c->partitioned)
        return fail(IGAP4H_EINVAL, "the Newton-Krylov driver solves whole patches; a partitioned block only assembles");
    if (!c || !d_Uh || !info) return fail(IGAP4H_EINVAL, "NULL argument");
    int rc = check_opts(o);
    if (rc) return rc;
    HIPCHK(hipSetDevice(c->device));
    std::memset(info, 0, sizeof(*info));
    const int f32 = o->vals_type == IGAP4H_VALS_F32;
    if ((rc = solver_ws(c, o->ksp_type == IGAP4H_KSP_GMRES ? o->restart : 1, (size_t)c->nnz * (f32 ? 4 : 8)))) return rc;
    Solver *s = c->solver;
    hipStream_t st = (hipStream_t)stream;
    const long long n = s->ndof;
    Op A{c, s, sp_args(c), s->vals, f32, o->pc, st};
    for (MGLevel &L : s->mg) L.ev_ok = false;  // warm starts only within one solve (repeatable solves)
    A.nu_pre = o->mg_pre;
    A.nu_post = o->mg_post;
    A.mg_graph = o->mg_graph;
    A.smoother = o->mg_smoother;
    A.svt = o->mg_smooth_vals;
    A.cdeg = o->mg_coarse_deg;
    A.cratio = o->mg_cheb_ratio;
    double *U = d_Uh;  // slice 0: current iterate

    auto t_all = clk::now();
    if (o->dirichlet_lift && !c->drows.empty()) {
        hipLaunchKernelGGL(k_lift, dim3(vgrid(n)), dim3(256), 0, st, n, (const unsigned char *)c->d_mask, (const double *)c->d_g, U);
        HIPCHK(hipGetLastError());
    }
    auto residual = [&](double *fn) -> int {
        auto t0 = clk::now();
        int r = igap4h_residual_dev(c, d_Uh, s->R, st);
        if (r) return r;
        if ((r = A.norm(s->R, s->sc + 5))) return r;
        if ((r = A.read_scalars(5, 1))) return r;
        *fn = s->hp[5];
        info->t_residual += secs(t0, clk::now());
        return 0;
    };
    double fnorm;
    if ((rc = residual(&fnorm))) return rc;
    info->fnorm0 = fnorm;
    info->fnorm_hist[0] = fnorm;
    hipLaunchKernelGGL(k_zero, dim3(vgrid(n)), dim3(256), 0, st, s->Y, n);
    HIPCHK(hipGetLastError());
    double fold = 999.0;  // snes_solve.c:23
    info->converged = 0;
    for (int it = 0; it < o->max_newton; it++) {
        if (!std::isfinite(fnorm)) {  // SNES_DIVERGED_FNORM_NAN
            info->converged = -4;
            break;
        }
        if ((o->atol > 0.0 && fnorm < o->atol) || (o->rtol > 0.0 && fnorm < o->rtol * info->fnorm0)) {
            info->converged = 1;
            break;
        }
        auto t0 = clk::now();
        // the tangent (and the preconditioner built on it) is rebuilt every
        // lag_jacobian-th Newton step, counted across solves with lag_persist
        const int lag = o->lag_jacobian > 1 ? o->lag_jacobian : 1;
        const long long key = ((long long)o->tangent_prec << 1) | o->vals_type;
        const bool stale = !s->jac_valid || s->jac_key != key || (!o->lag_persist && it == 0);
        if (stale) s->jac_age = 0;
        const bool rebuild = stale || s->jac_age % lag == 0;
        if (rebuild) {
            if ((rc = igap4h_tangent_dev(c, d_Uh, o->tangent_prec, o->vals_type, s->vals, st))) return rc;
            HIPCHK(hipStreamSynchronize(st));
            s->jac_valid = true;
            s->jac_key = key;
            info->tangents++;
        }
        s->jac_age++;
        auto t1 = clk::now();
        info->t_tangent += secs(t0, t1);
        if (!o->ksp_guess) {
            hipLaunchKernelGGL(k_zero, dim3(vgrid(n)), dim3(256), 0, st, s->Y, n);
            HIPCHK(hipGetLastError());
        }
        // a reused tangent keeps its preconditioner only if that was built with
        // this solve's preconditioner options
        if ((rebuild || s->pc_vals != s->vals || !A.pc_key_matches()) && (rc = A.setup_pc())) return rc;
        int kits = 0;
        double relres = 0.0;
        rc = linear_solve(A, s->R, s->Y, o, &kits, &relres);
        HIPCHK(hipStreamSynchronize(st));
        info->t_solve += secs(t1, clk::now());
        info->ksp_its += kits;
        if (it < 64) info->ksp_its_hist[it] = kits;
        if (rc) {
            info->converged = -2;
            info->newton_its = it;
            info->fnorm = fnorm;
            info->t_total = secs(t_all, clk::now());
            return rc;
        }
        // U_new = U - lambda Y with backtracking on f = |F|^2 / 2
        HIPCHK(hipMemcpyAsync(s->U0, U, n * sizeof(double), hipMemcpyDeviceToDevice, st));
        auto trial = [&](double lam, double *gn) -> int {
            hipLaunchKernelGGL(k_waxpy, dim3(vgrid(n)), dim3(256), 0, st, n, -lam, (const double *)s->U0, (const double *)s->Y, U);
            HIPCHK(hipGetLastError());
            return residual(gn);
        };
        double gnorm;
        if (!o->line_search) {
            if ((rc = trial(1.0, &gnorm))) return rc;
        } else {
            // slope of f along -Y: -(F, K Y)   (Dennis & Schnabel A6.3.1)
            if ((rc = A.spmv(s->Y, s->w))) return rc;
            hipLaunchKernelGGL(k_dot2, dim3(NRB), dim3(256), 0, st, n, (const double *)s->R, (const double *)s->w,
                               (const double *)nullptr, (const double *)nullptr, s->partial);
            HIPCHK(hipGetLastError());
            if ((rc = A.finish(s->partial, 1, NRB, s->sc + 6))) return rc;
            if ((rc = A.read_scalars(6, 1))) return rc;
            double slope = s->hp[6];
            if (slope > 0.0) slope = -slope;
            if (slope == 0.0) slope = -1.0;
            const double alpha = 1e-4, f0 = 0.5 * fnorm * fnorm;
            double lam = 1.0;
            if ((rc = trial(lam, &gnorm))) return rc;
            bool ok = std::isfinite(gnorm) && 0.5 * gnorm * gnorm <= f0 + lam * alpha * slope;
            double lprev = lam, gprev = gnorm;
            if (!ok) {  // quadratic model
                double lt = -slope / (gnorm * gnorm - fnorm * fnorm - 2.0 * lam * slope);
                if (!std::isfinite(gnorm) || !std::isfinite(lt)) lt = 0.1 * lam;
                lprev = lam;
                gprev = gnorm;
                lam = lt > 0.5 * lam ? 0.5 * lam : (lt <= 0.1 * lam ? 0.1 * lam : lt);
                if ((rc = trial(lam, &gnorm))) return rc;
                ok = std::isfinite(gnorm) && 0.5 * gnorm * gnorm < f0 + lam * alpha * slope;
            }
            for (int cnt = 0; !ok && cnt < 40; cnt++) {  // cubic model
                if (lam < 1e-12) break;
                double t1c = 0.5 * (gnorm * gnorm - fnorm * fnorm) - lam * slope;
                double t2c = 0.5 * (gprev * gprev - fnorm * fnorm) - lprev * slope;
                double a = (t1c / (lam * lam) - t2c / (lprev * lprev)) / (lam - lprev);
                double bb = (-lprev * t1c / (lam * lam) + lam * t2c / (lprev * lprev)) / (lam - lprev);
                double d = std::max(bb * bb - 3.0 * a * slope, 0.0);
                double lt = a == 0.0 ? -slope / (2.0 * bb) : (-bb + std::sqrt(d)) / (3.0 * a);
                if (!std::isfinite(lt) || !std::isfinite(gnorm)) lt = 0.1 * lam;
                lprev = lam;
                gprev = gnorm;
                lam = lt > 0.5 * lam ? 0.5 * lam : (lt <= 0.1 * lam ? 0.1 * lam : lt);
                if ((rc = trial(lam, &gnorm))) return rc;
                ok = std::isfinite(gnorm) && 0.5 * gnorm * gnorm < f0 + lam * alpha * slope;
            }
            if (!ok) {  // keep the last accepted iterate
                HIPCHK(hipMemcpyAsync(U, s->U0, n * sizeof(double), hipMemcpyDeviceToDevice, st));
                double f;
                if ((rc = residual(&f))) return rc;
                info->converged = -3;
                info->newton_its = it + 1;
                fnorm = f;
                break;
            }
        }
        fnorm = gnorm;
        info->newton_its = it + 1;
        if (it + 1 < 64) info->fnorm_hist[it + 1] = fnorm;
        if (o->stagnation && std::fabs(fold - fnorm) / fold < 1e-6 &&
            !((o->atol > 0.0 && fnorm < o->atol) || (o->rtol > 0.0 && fnorm < o->rtol * info->fnorm0))) {
            info->converged = -1;
            break;
        }
        fold = fnorm;
        if ((o->atol > 0.0 && fnorm < o->atol) || (o->rtol > 0.0 && fnorm < o->rtol * info->fnorm0)) {
            info->converged = 1;
            break;
        }
    }
    HIPCHK(hipStreamSynchronize(st));
    info->fnorm = fnorm;
    info->t_total = secs(t_all, clk::now());
    return 0;
}
