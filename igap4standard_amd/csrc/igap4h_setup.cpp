// Host-side set-up of the assembly context (see igap4h_setup.hpp).
#include "igap4h_setup.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace igap4h {

static const int P = 2; // quadratic

int gauss01(int n, double *x, double *w)
{
    switch (n) {
    case 1: x[0] = 0.5; w[0] = 1.0; return 0;
    case 2: {
        double a = 0.5 / std::sqrt(3.0);
        x[0] = 0.5 - a; x[1] = 0.5 + a; w[0] = w[1] = 0.5; return 0;
    }
    case 3: {
        double a = 0.5 * std::sqrt(0.6);
        x[0] = 0.5 - a; x[1] = 0.5; x[2] = 0.5 + a;
        w[0] = w[2] = 5.0 / 18.0; w[1] = 8.0 / 18.0; return 0;
    }
    case 4: {
        double s = 2.0 / 7.0 * std::sqrt(6.0 / 5.0);
        double ao = 0.5 * std::sqrt(3.0 / 7.0 + s), ai = 0.5 * std::sqrt(3.0 / 7.0 - s);
        x[0] = 0.5 - ao; x[1] = 0.5 - ai; x[2] = 0.5 + ai; x[3] = 0.5 + ao;
        w[0] = w[3] = (18.0 - std::sqrt(30.0)) / 72.0;
        w[1] = w[2] = (18.0 + std::sqrt(30.0)) / 72.0;
        return 0;
    }
    default: return -1;
    }
}

double evalN(const double *kv, int nk, int k, int i, int p, double xi)
{
    if (k == 0) {
        // closed right end belongs to the last function
        if (i == nk - p - 2 && std::fabs(xi - kv[nk - 1]) < 1.e-15 && std::fabs(kv[nk - 1] - kv[nk - 2]) < 1.e-15)
            return 1.0;
        if (p == 0) return (xi >= kv[i] && kv[i + 1] > xi) ? 1.0 : 0.0;
        bool lz = kv[i + p] == kv[i], rz = kv[i + p + 1] == kv[i + 1];
        if (lz && rz) return 0.0;
        double left = lz ? 0.0 : (xi - kv[i]) / (kv[i + p] - kv[i]) * evalN(kv, nk, 0, i, p - 1, xi);
        double right = rz ? 0.0 : (kv[i + p + 1] - xi) / (kv[i + p + 1] - kv[i + 1]) * evalN(kv, nk, 0, i + 1, p - 1, xi);
        if (lz) return right;
        if (rz) return left;
        return left + right;
    }
    if (k < 0 || k > p) return 0.0;
    // N^(k)_{i,p} = p!/(p-k)! sum_j a_{k,j} N_{i+j,p-k} (The NURBS Book eq. 2.9)
    double a[4][4];
    a[0][0] = 1.0;
    for (int r = 1; r <= k; r++) {
        double den = kv[i + p - r + 1] - kv[i];
        a[r][0] = den == 0.0 ? 0.0 : a[r - 1][0] / den;
        for (int c = 1; c < r; c++) {
            den = kv[i + p + c - r + 1] - kv[i + c];
            a[r][c] = den == 0.0 ? 0.0 : (a[r - 1][c] - a[r - 1][c - 1]) / den;
        }
        den = kv[i + p + 1] - kv[i + r];
        a[r][r] = den == 0.0 ? 0.0 : -a[r - 1][r - 1] / den;
    }
    double s = 0.0;
    for (int j = 0; j <= k; j++) s += a[k][j] * evalN(kv, nk, 0, i + j, p - k, xi);
    double f = 1.0;
    for (int m = p; m > p - k; m--) f *= m;
    return s * f;
}

int64_t pattern_nnz(const int nb[3])
{
    int64_t s[3];
    for (int d = 0; d < 3; d++) {
        s[d] = 0;
        for (int a = 0; a < nb[d]; a++) {
            int lo = a - P < 0 ? 0 : a - P, hi = a + P > nb[d] - 1 ? nb[d] - 1 : a + P;
            s[d] += hi - lo + 1;
        }
    }
    return 9LL * s[0] * s[1] * s[2];
}

namespace {
struct FaceMap {
    int dim[3];
    int64_t stride[3];
    int iz0, iz1;
};

// strides from the window's basis counts nb; the far face's layers from the
// owned counts nown (snesiga_Residual.c:352-361 uses Nurbs.nbasis)
FaceMap face_map(int f, const int nb[3], const int nown[3])
{
    FaceMap m;
    int64_t sx = (int64_t)nb[1] * nb[2], sy = nb[2], sz = 1;
    if (f / 2 == 0) { m.dim[0] = 1; m.dim[1] = 2; m.dim[2] = 0; m.stride[0] = sy; m.stride[1] = sz; m.stride[2] = sx; }
    else if (f / 2 == 1) { m.dim[0] = 0; m.dim[1] = 2; m.dim[2] = 1; m.stride[0] = sx; m.stride[1] = sz; m.stride[2] = sy; }
    else { m.dim[0] = 0; m.dim[1] = 1; m.dim[2] = 2; m.stride[0] = sx; m.stride[1] = sy; m.stride[2] = sz; }
    int nz = nown[m.dim[2]];
    m.iz0 = (f % 2 == 0) ? 0 : nz - 1;
    m.iz1 = (f % 2 == 0) ? 1 : nz - 2;
    return m;
}

// Control values of the affine field u_idof = a_{idof,J} X_J on face f
// (src/iga_dirichlet.c:7-80), local layout [x][y].
void face_dirichlet(const Mesh &m, int f, int order, int idof, const FaceMap &fm, std::vector<double> &out)
{
    const double *kv[3];
    int nk[3], nbf[3];
    for (int t = 0; t < 3; t++) {
        kv[t] = m.knot[fm.dim[t]].data();
        nk[t] = (int)m.knot[fm.dim[t]].size();
        nbf[t] = m.nown[fm.dim[t]];   // iga_dirichlet gets Nurbs.nbasis (owned counts)
    }
    const double *par = m.par_dirichlet;
    double Z = (f % 2 == 0) ? kv[2][0] : kv[2][nk[2] - 1] - 1.e-15;
    int iz1 = (f % 2 == 0) ? 1 : nbf[2] - 2;
    double pl[3];
    if (f / 2 == 0) { pl[0] = par[3 * idof + 1]; pl[1] = par[3 * idof + 2]; pl[2] = par[3 * idof + 0]; }
    else if (f / 2 == 1) { pl[0] = par[3 * idof + 0]; pl[1] = par[3 * idof + 2]; pl[2] = par[3 * idof + 1]; }
    else { pl[0] = par[3 * idof + 0]; pl[1] = par[3 * idof + 1]; pl[2] = par[3 * idof + 2]; }
    std::vector<double> c1d[2];
    for (int bd = 0; bd < 2; bd++) {
        c1d[bd].assign(nbf[bd], 0.0);
        double W = kv[bd][P];
        // value and slope at the start: 2x2 solve with the closed-form inverse
        double mt[4];
        for (int i = 0; i < 2; i++)
            for (int j = 0; j < 2; j++) mt[2 * i + j] = evalN(kv[bd], nk[bd], i, j, P, W);
        double vec[2] = {pl[bd] * W, pl[bd]};
        double idet = 1. / (-(mt[1] * mt[2]) + mt[0] * mt[3]);
        double inv[4] = {idet * mt[3], idet * (-mt[1]), idet * (-mt[2]), idet * mt[0]};
        double v0 = 0.0, v1 = 0.0;
        v0 += inv[0] * vec[0]; v0 += inv[1] * vec[1];
        v1 += inv[2] * vec[0]; v1 += inv[3] * vec[1];
        c1d[bd][0] = v0;
        c1d[bd][1] = v1;
        for (int i = P; i < nbf[bd]; i++) c1d[bd][i] = c1d[bd][i - 1] + (kv[bd][i + P] - kv[bd][i]) / ((double)P) * pl[bd];
    }
    out.assign((size_t)nbf[0] * nbf[1], 0.0);
    for (int i = 0; i < nbf[0]; i++)
        for (int j = 0; j < nbf[1]; j++) out[(size_t)nbf[1] * i + j] = c1d[0][i] + c1d[1][j] + pl[2] * Z;
    if (order == 1) {
        double mm = 0.0 / evalN(kv[2], nk[2], 1, iz1, P, Z);
        for (auto &v : out) v += mm;
    }
}
} // namespace

std::vector<int32_t> dirichlet_rows(const Mesh &m)
{
    std::vector<int32_t> rows;
    for (int ib = 0; ib < m.nboun; ib++) {
        int f = m.boun[ib];
        FaceMap fm = face_map(f, m.nb, m.nown);
        for (int order = 0; order < 2; order++) {
            int iz = order == 0 ? fm.iz0 : fm.iz1;
            for (int i = 0; i < 3; i++) {
                if (m.bc_type[6 * f + 3 * order + i] != 0) continue;
                for (int x = 0; x < m.nown[fm.dim[0]]; x++)
                    for (int y = 0; y < m.nown[fm.dim[1]]; y++)
                        rows.push_back((int32_t)(3 * (fm.stride[0] * x + fm.stride[1] * y + fm.stride[2] * iz) + i));
            }
        }
    }
    return rows;
}

void dirichlet_values(const Mesh &m, std::vector<unsigned char> &mask, std::vector<double> &g)
{
    mask.assign((size_t)m.ndof(), 0);
    g.assign((size_t)m.ndof(), 0.0);
    std::vector<double> vals;
    for (int ib = 0; ib < m.nboun; ib++) {
        int f = m.boun[ib];
        FaceMap fm = face_map(f, m.nb, m.nown);
        for (int order = 0; order < 2; order++) {
            int iz = order == 0 ? fm.iz0 : fm.iz1;
            for (int i = 0; i < 3; i++) {
                if (m.bc_type[6 * f + 3 * order + i] != 0) continue;
                face_dirichlet(m, f, order, i, fm, vals);
                size_t ia = 0;
                for (int x = 0; x < m.nown[fm.dim[0]]; x++)
                    for (int y = 0; y < m.nown[fm.dim[1]]; y++) {
                        int64_t r = 3 * (fm.stride[0] * x + fm.stride[1] * y + fm.stride[2] * iz) + i;
                        mask[r] = 1;
                        g[r] = vals[ia++];
                    }
            }
        }
    }
    // ghost nodes of the window on the same faces: Dirichlet rows of the ranks
    // owning them (the same face and flags; they set those rows)
    for (int ib = 0; ib < m.nboun; ib++) {
        int f = m.boun[ib];
        FaceMap fm = face_map(f, m.nb, m.nown);
        for (int order = 0; order < 2; order++) {
            int iz = order == 0 ? fm.iz0 : fm.iz1;
            for (int i = 0; i < 3; i++) {
                if (m.bc_type[6 * f + 3 * order + i] != 0) continue;
                for (int x = 0; x < m.nb[fm.dim[0]]; x++)
                    for (int y = 0; y < m.nb[fm.dim[1]]; y++) {
                        if (x < m.nown[fm.dim[0]] && y < m.nown[fm.dim[1]]) continue;
                        int64_t r = 3 * (fm.stride[0] * x + fm.stride[1] * y + fm.stride[2] * iz) + i;
                        if (!mask[r]) mask[r] = 2;
                    }
            }
        }
    }
}

std::vector<double> neumann_load(const Mesh &m, int nquad, igap4h_neumann_fn hook)
{
    std::vector<double> R;
    if (!hook) return R;
    bool any = false;
    for (int ib = 0; ib < m.nboun; ib++)
        for (int t = 0; t < 6; t++) any |= m.bc_type[6 * m.boun[ib] + t] == 1;
    if (!any) return R;
    R.assign((size_t)m.ndof(), 0.0);
    double cq[4], wq[4];
    gauss01(nquad, cq, wq);
    for (int ib = 0; ib < m.nboun; ib++) {
        int f = m.boun[ib];
        FaceMap fm = face_map(f, m.nb, m.nown);
        int d0 = fm.dim[0], d1 = fm.dim[1], d2 = fm.dim[2];
        const double *k0 = m.knot[d0].data(), *k1 = m.knot[d1].data(), *k2 = m.knot[d2].data();
        int nk0 = (int)m.knot[d0].size(), nk1 = (int)m.knot[d1].size(), nk2 = (int)m.knot[d2].size();
        int iz0 = (f % 2 == 0) ? 0 : m.ne[d2] + P - 1;
        int iz1 = (f % 2 == 0) ? 1 : m.ne[d2] + P - 2;
        double xq[3];
        xq[d2] = (f % 2 == 0) ? k2[0] : k2[nk2 - 1];
        for (int order = 0; order < 2; order++)
            for (int i = 0; i < 3; i++) {
                if (m.bc_type[6 * f + 3 * order + i] != 1) continue;
                for (int e0 = 0; e0 < m.ne[d0]; e0++)
                    for (int e1 = 0; e1 < m.ne[d1]; e1++) {
                        double X00 = k0[e0 + P], X10 = k0[e0 + P + 1], X01 = k1[e1 + P], X11 = k1[e1 + P + 1];
                        double val[18] = {0};
                        int nv = order == 0 ? 9 : 18;
                        for (int qx = 0; qx < nquad; qx++)
                            for (int qy = 0; qy < nquad; qy++) {
                                xq[d0] = (-cq[qx] + 1.0) * X00 + cq[qx] * X10;
                                xq[d1] = (-cq[qy] + 1.0) * X01 + cq[qy] * X11;
                                double Nf[18];
                                int ia = 0;
                                for (int bx = 0; bx < 3; bx++)
                                    for (int by = 0; by < 3; by++) {
                                        double nxy = evalN(k0, nk0, 0, e0 + bx, P, xq[d0]) * evalN(k1, nk1, 0, e1 + by, P, xq[d1]);
                                        if (order == 0) {
                                            Nf[ia++] = nxy;
                                        } else {
                                            Nf[ia++] = nxy * evalN(k2, nk2, 1, iz0, P, xq[d2]);
                                            Nf[ia++] = nxy * evalN(k2, nk2, 1, iz1, P, xq[d2]);
                                        }
                                    }
                                double t = 0.0;
                                hook(&t, m.par_neumann, f, order, i, xq);
                                double w = wq[qx] * wq[qy] * (X10 - X00) * (X11 - X01);
                                for (int s = 0; s < nv; s++) val[s] -= w * Nf[s] * t;
                            }
                        int ia = 0;
                        for (int bx = 0; bx < 3; bx++)
                            for (int by = 0; by < 3; by++) {
                                int64_t base = fm.stride[0] * (e0 + bx) + fm.stride[1] * (e1 + by);
                                R[3 * (base + fm.stride[2] * iz0) + i] += val[ia++];
                                if (order == 1) R[3 * (base + fm.stride[2] * iz1) + i] += val[ia++];
                            }
                    }
            }
    }
    return R;
}

double neumann_measure(const Mesh &m, int nquad)
{
    double cq[4], wq[4], total = 0.0;
    gauss01(nquad, cq, wq);
    for (int ib = 0; ib < m.nboun; ib++) {
        int f = m.boun[ib];
        FaceMap fm = face_map(f, m.nb, m.nown);
        int d0 = fm.dim[0], d1 = fm.dim[1];
        const double *k0 = m.knot[d0].data(), *k1 = m.knot[d1].data();
        for (int order = 0; order < 2; order++)
            for (int i = 0; i < 3; i++) {
                if (m.bc_type[6 * f + 3 * order + i] != 1) continue;
                for (int e0 = 0; e0 < m.ne[d0]; e0++)
                    for (int e1 = 0; e1 < m.ne[d1]; e1++)
                        for (int qx = 0; qx < nquad; qx++)
                            for (int qy = 0; qy < nquad; qy++)
                                total += wq[qx] * wq[qy] * (k0[e0 + P + 1] - k0[e0 + P]) * (k1[e1 + P + 1] - k1[e1 + P]);
            }
    }
    return total;
}

void coef_from_par(const double *par, double coef[6])
{
    double r = par[0], c = par[1];
    coef[0] = 4.0 * (1.0 - c) / (r * r * r * (1.0 + 2.0 * c));
    coef[1] = -6.0 * c / (r * r * (1.0 + 2.0 * c));
    coef[2] = 3.0 / (r * r * r * r * (1.0 + 2.0 * c));
    coef[3] = par[2];
    coef[4] = par[3];
    coef[5] = par[4] * par[4];
}

std::vector<double> coarsen_knots(const std::vector<double> &kf, int nef)
{
    std::vector<double> kc;
    kc.push_back(kf[0]);
    kc.push_back(kf[1]);
    for (int e = 0; e < nef; e += 2) kc.push_back(kf[2 + e]);  // element boundaries 0, 2, ...
    kc.push_back(kf[2 + nef]);                                  // and the last one
    kc.push_back(kf[nef + 3]);
    kc.push_back(kf[nef + 4]);
    return kc;
}

int prolongation_1d(const std::vector<double> &kf, int nef, const std::vector<double> &kc, int nec, std::vector<double> &P)
{
    const int p = 2, nbf = nef + p, nbc = nec + p;
    const int nkf = nef + 5, nkc = nec + 5;
    // collocation points: Greville abscissae, the two ends moved inside
    std::vector<double> x(nbf);
    for (int a = 0; a < nbf; a++) x[a] = 0.5 * (kf[a + 1] + kf[a + 2]);
    double h0 = kf[3] - kf[2], h1 = kf[nef + 2] - kf[nef + 1];
    x[0] = kf[2] + 0.25 * h0;
    x[nbf - 1] = kf[nef + 2] - 0.25 * h1;
    // B[a][f] = N^f_f(x_a), C[a][A] = N^c_A(x_a); solve B P = C
    std::vector<double> B((size_t)nbf * nbf), C((size_t)nbf * nbc);
    for (int a = 0; a < nbf; a++) {
        for (int f = 0; f < nbf; f++) B[(size_t)a * nbf + f] = evalN(kf.data(), nkf, 0, f, p, x[a]);
        for (int A = 0; A < nbc; A++) C[(size_t)a * nbc + A] = evalN(kc.data(), nkc, 0, A, p, x[a]);
    }
    for (int k = 0; k < nbf; k++) {  // Gaussian elimination, partial pivoting
        int piv = k;
        for (int r = k + 1; r < nbf; r++)
            if (std::fabs(B[(size_t)r * nbf + k]) > std::fabs(B[(size_t)piv * nbf + k])) piv = r;
        if (std::fabs(B[(size_t)piv * nbf + k]) < 1e-300) return -1;
        if (piv != k) {
            for (int c = 0; c < nbf; c++) std::swap(B[(size_t)k * nbf + c], B[(size_t)piv * nbf + c]);
            for (int c = 0; c < nbc; c++) std::swap(C[(size_t)k * nbc + c], C[(size_t)piv * nbc + c]);
        }
        for (int r = 0; r < nbf; r++) {
            if (r == k) continue;
            double f = B[(size_t)r * nbf + k] / B[(size_t)k * nbf + k];
            if (f == 0.0) continue;
            for (int c = k; c < nbf; c++) B[(size_t)r * nbf + c] -= f * B[(size_t)k * nbf + c];
            for (int c = 0; c < nbc; c++) C[(size_t)r * nbc + c] -= f * C[(size_t)k * nbc + c];
        }
    }
    P.assign((size_t)nbf * nbc, 0.0);
    for (int a = 0; a < nbf; a++)
        for (int A = 0; A < nbc; A++) {
            double v = C[(size_t)a * nbc + A] / B[(size_t)a * nbf + a];
            P[(size_t)a * nbc + A] = std::fabs(v) < 1e-13 ? 0.0 : v;
        }
    // nested: the coarse functions are reproduced at a second point set
    std::vector<double> Nf(nbf);
    for (int a = 0; a < nbf; a++) {
        double xm = 0.5 * (x[a] + (a + 1 < nbf ? x[a + 1] : kf[nef + 2]));
        for (int f = 0; f < nbf; f++) Nf[f] = evalN(kf.data(), nkf, 0, f, p, xm);
        for (int A = 0; A < nbc; A++) {
            double s = 0.0;
            for (int f = 0; f < nbf; f++) s += P[(size_t)f * nbc + A] * Nf[f];
            if (std::fabs(s - evalN(kc.data(), nkc, 0, A, p, xm)) > 1e-10) return -1;
        }
    }
    return 0;
}

}  // namespace igap4h
