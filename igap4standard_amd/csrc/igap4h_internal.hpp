// Library-internal declarations shared by the assembly (igap4h.hip) and the
// Newton-Krylov driver (igap4h_solver.hip): the context, error reporting.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "../../include/igap4h.h"
#include "igap4h_setup.hpp"

namespace igap4h {
// records the message for igap4h_last_error() and returns code
int fail(int code, const char *fmt, ...);

struct Solver;
}  // namespace igap4h

#define HIPCHK(call)                                                                                   \
    do {                                                                                               \
        hipError_t e_ = (call);                                                                        \
        if (e_ != hipSuccess) return igap4h::fail(IGAP4H_EHIP, "%s: %s", #call, hipGetErrorString(e_)); \
    } while (0)

struct igap4h_ctx {
    igap4h::Mesh mesh;
    int physics, nquad, nhist;
    int device;
    double par_mat[9];
    double cq[4], wq[4];
    int64_t nnz;
    std::vector<int32_t> drows;
    // device
    double *d_knot[3] = {nullptr, nullptr, nullptr};
    double *d_phi = nullptr;     // 1D basis tables [d][element][Gauss point][9], built at create
    double *d_coef = nullptr;
    long long *d_rowptr = nullptr;
    int *d_col = nullptr;
    unsigned char *d_mask = nullptr;
    double *d_g = nullptr;
    double *d_fneu = nullptr;
    int *d_drows = nullptr;
    double *d_Re = nullptr;      // n_elem x 81
    double *d_Uh = nullptr;      // staging for the host API
    double *d_R = nullptr;
    void *d_vals = nullptr;      // staging for the host API
    size_t d_vals_bytes = 0;
    double *d_Pe = nullptr;      // per-element energies / density outputs
    size_t d_Pe_n = 0;
    double *d_part = nullptr;    // block partial sums of the density reductions
    double neu_measure = 0.0;    // neumann_measure() of the current boundary data
    double *U_hist_app = nullptr;
    igap4h_neumann_fn neumann = nullptr;  // traction hook given at create
    igap4h::Solver *solver = nullptr;  // Krylov work space, created on first use
    // incremented whenever the Dirichlet/Neumann device data change (their
    // arrays may be reallocated), so that users of copies (batches) refresh
    int bc_gen = 0;
    // one rank's block of a partitioned patch: mesh.nown < mesh.nb in some
    // direction; gidx = globalIdx of the local window
    bool partitioned = false;
    std::vector<int> gidx;
    int tangent_form = IGAP4H_TANGENT_AUTO;  // igap4h_set_tangent_form
};

namespace igap4h {
// Fault injection: with IGAP4H_DEBUG_ALLOC_LIMIT=<bytes> in the environment,
// device allocations of more than that many bytes are refused as if
// hipMalloc had failed, so that the out-of-memory paths can be exercised.
bool alloc_refused(size_t bytes);

// hipMalloc with the fault-injection check
static inline hipError_t dev_malloc(void **p, size_t bytes)
{
    if (alloc_refused(bytes)) {
        *p = nullptr;
        return hipErrorOutOfMemory;
    }
    return hipMalloc(p, bytes);
}

template <class X>
static inline int dev_alloc(X **p, size_t n)
{
    if (n == 0) n = 1;
    if (dev_malloc((void **)p, n * sizeof(X)) != hipSuccess) {
        *p = nullptr;
        return fail(IGAP4H_ENOMEM, "hipMalloc of %zu bytes failed", n * sizeof(X));
    }
    return 0;
}
void solver_free(igap4h_ctx *c);
void solver_invalidate_jacobian(igap4h_ctx *c);  // a lagged tangent no longer fits the problem
}  // namespace igap4h
