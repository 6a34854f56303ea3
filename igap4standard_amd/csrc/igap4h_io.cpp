// Host utilities around the assembly: solution files and knot-insertion
// refinement of a solution, the one-rank counterparts of src/ioutil.c:89-158
// and mesh/nurbs/nurbs_knotinsert.c:7-107.  No HIP calls.
#include <cerrno>
#include <cstdio>
#include <cstring>
#include <vector>

#include "igap4h_internal.hpp"

using namespace igap4h;

// write_solution (src/ioutil.c:89-120): slices ihist0..ihist1-1 of U_hist as
// raw doubles in universal order (one rank: the local order), no header.
extern "C" int igap4h_write_solution(const char *fname, const double *U_hist, int64_t ndof, int ihist0, int ihist1)
{
    if (!fname || !U_hist || ndof <= 0 || ihist0 < 0 || ihist1 <= ihist0) return fail(IGAP4H_EINVAL, "bad write_solution arguments");
    FILE *f = std::fopen(fname, "wb");
    if (!f) return fail(IGAP4H_EINVAL, "cannot open %s for writing: %s", fname, std::strerror(errno));
    const size_t n = (size_t)(ihist1 - ihist0) * (size_t)ndof;
    size_t w = std::fwrite(U_hist + (size_t)ihist0 * ndof, sizeof(double), n, f);
    int cl = std::fclose(f);
    if (w != n || cl != 0) return fail(IGAP4H_EINVAL, "short write to %s (%zu of %zu doubles)", fname, w, n);
    return 0;
}

// read_solution (src/ioutil.c:122-158): the inverse; the file must hold
// exactly (ihist1 - ihist0) x ndof doubles.
extern "C" int igap4h_read_solution(const char *fname, double *U_hist, int64_t ndof, int ihist0, int ihist1)
{
    if (!fname || !U_hist || ndof <= 0 || ihist0 < 0 || ihist1 <= ihist0) return fail(IGAP4H_EINVAL, "bad read_solution arguments");
    FILE *f = std::fopen(fname, "rb");
    if (!f) return fail(IGAP4H_EINVAL, "cannot open %s: %s", fname, std::strerror(errno));
    const size_t n = (size_t)(ihist1 - ihist0) * (size_t)ndof;
    size_t r = std::fread(U_hist + (size_t)ihist0 * ndof, sizeof(double), n, f);
    bool extra = std::fgetc(f) != EOF;
    std::fclose(f);
    if (r != n) return fail(IGAP4H_EINVAL, "%s holds %zu doubles, expected %zu", fname, r, n);
    if (extra) return fail(IGAP4H_EINVAL, "%s holds more than the expected %zu doubles", fname, n);
    return 0;
}

// nurbs_knotinsert (mesh/nurbs/nurbs_knotinsert.c:7-107): U_universal holds
// the control values of the mesh with nelem[d] / 2^nref elements (ndof per
// control point, x slowest) and is overwritten with those of nelem[d]
// elements on open uniform knots -- the same spline, refined nref times by
// midpoint insertion.  The transfer per direction is prolongation_1d.
extern "C" int igap4h_knot_insert(double *U_universal, const int nelem[3], int nref, int porder, int ndim, int ndof,
                                  const int bc_periodic[3])
{
    if (!U_universal || !nelem || nref < 0 || ndof < 1) return fail(IGAP4H_EINVAL, "bad knot_insert arguments");
    if (porder != 2 || ndim != 3) return fail(IGAP4H_EINVAL, "knot insertion supports porder 2, ndim 3 (as the reference)");
    for (int d = 0; d < 3; d++) {
        if (bc_periodic && bc_periodic[d]) return fail(IGAP4H_EINVAL, "periodic directions are not supported");
        if (nelem[d] < 1 || nelem[d] % (1 << nref) != 0) return fail(IGAP4H_EINVAL, "nelem[%d] = %d is not a multiple of 2^%d", d, nelem[d], nref);
    }
    auto uniform = [](int ne) {
        std::vector<double> k(ne + 5);
        for (int i = 0; i < ne + 5; i++) k[i] = i < 3 ? 0.0 : (i >= ne + 2 ? 1.0 : (i - 2) * (1.0 / ne));
        return k;
    };
    int ne[3], nb[3];
    for (int d = 0; d < 3; d++) {
        ne[d] = nelem[d] >> nref;
        nb[d] = ne[d] + 2;
    }
    std::vector<double> cur((size_t)nb[0] * nb[1] * nb[2] * ndof), nxt;
    std::memcpy(cur.data(), U_universal, cur.size() * sizeof(double));
    for (int r = 0; r < nref; r++) {
        // one direction at a time: out[..a..] = sum_A P[a][A] in[..A..]
        for (int d = 0; d < 3; d++) {
            std::vector<double> kc = uniform(ne[d]), kf = uniform(2 * ne[d]), P;
            if (prolongation_1d(kf, 2 * ne[d], kc, ne[d], P)) return fail(IGAP4H_EINVAL, "knot insertion failed");
            const int nbc = nb[d], nbf = 2 * ne[d] + 2;
            int nbo[3] = {nb[0], nb[1], nb[2]};
            nbo[d] = nbf;
            nxt.assign((size_t)nbo[0] * nbo[1] * nbo[2] * ndof, 0.0);
            for (int i = 0; i < nbo[0]; i++)
                for (int j = 0; j < nbo[1]; j++)
                    for (int k = 0; k < nbo[2]; k++) {
                        const int a = d == 0 ? i : (d == 1 ? j : k);
                        double *o = &nxt[(((size_t)i * nbo[1] + j) * nbo[2] + k) * ndof];
                        for (int A = 0; A < nbc; A++) {
                            const double w = P[(size_t)a * nbc + A];
                            if (w == 0.0) continue;
                            int ii = d == 0 ? A : i, jj = d == 1 ? A : j, kk = d == 2 ? A : k;
                            const double *in = &cur[(((size_t)ii * nb[1] + jj) * nb[2] + kk) * ndof];
                            for (int c = 0; c < ndof; c++) o[c] += w * in[c];
                        }
                    }
            cur.swap(nxt);
            nb[d] = nbf;
            ne[d] *= 2;
        }
    }
    std::memcpy(U_universal, cur.data(), cur.size() * sizeof(double));
    return 0;
}
