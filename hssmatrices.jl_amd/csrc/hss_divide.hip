// hss_divide.hip -- hssA \ hssB and hssA / hssB as HSS matrices (hss_mat_ldiv,
// hss_mat_rdiv; reference ldiv! / rdiv!, src/ulvdivide.jl:13-54, reached from
// \(hssA, hssB) and /(hssA, hssB), src/hssmatrix.jl:235-237).
//
// The reference's ULV divide works on the generators: it factors A's leaves, applies the
// transforms to B's generators, subtracts an HSS product, recompresses, prunes one tree level
// and recurses, then unpacks the recursion's result into the next level up.  Every step
// changes the tree's shape.  Here the same matrix is formed the way one MI355X does it well:
// A is factored once (hss_ulv_factor, resident in HBM), the dense A^-1 B (rows x cols,
// Float64: 34 GB at n = 65536, inside 288 GB of HBM) is filled block of columns by block of
// columns -- B's columns by the HSS matvec on unit columns, then the ULV solve -- and
// compressed by hss_compress_dev at (atol, rtol) over A's column tree and B's column tree.
// The work is O(n^2 k) instead of O(n k^2); the result meets the same tolerance test
// (test/runtests.jl:69-77).  A / B is (B' \ A')'.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/hss_hip.h"
#include "hss_host.h"
#include "hss_internal.h"
#include "hss_launch.h"
#include "hss_tasks.h"

using namespace hss;

namespace {
constexpr int kBlock = 256;  // columns of B per matvec + solve
}

extern "C" hss_status hss_mat_ldiv(hss_mat_t a, hss_mat_t b, double atol, double rtol, hss_mat_t* out) {
    if (!a || !b || !out) return host_fail(HSS_ERR_ARG, "null argument");
    *out = nullptr;
    if (!(atol >= 0.0) || !(rtol >= 0.0)) return host_fail(HSS_ERR_ARG, "atol and rtol must be >= 0");
    hss_ctx_t ctx = host_mat_ctx(a);
    if (host_mat_ctx(b) != ctx) return host_fail(HSS_ERR_ARG, "the matrices belong to different contexts");
    int na = 0, nb = 0;
    const hss_node_desc* ta = host_mat_nodes(a, &na);
    const hss_node_desc* tb = host_mat_nodes(b, &nb);
    if (na != nb) return host_fail(HSS_ERR_DIM, "the trees of A and B differ");
    for (int i = 0; i < na; ++i)
        if (ta[i].left != tb[i].left || ta[i].right != tb[i].right || ta[i].parent != tb[i].parent ||
            ta[i].n != tb[i].m)
            return host_fail(HSS_ERR_DIM, "A's column tree and B's row tree differ at node " + std::to_string(i));
    const int64_t n = ta[0].n, nc = tb[0].n;
    if (ta[0].m != n) return host_fail(HSS_ERR_DIM, "A is not square");
    if (hipSetDevice(host_ctx_device(ctx)) != hipSuccess) return host_fail(HSS_ERR_HIP, "hipSetDevice");
    hipStream_t st = reinterpret_cast<hipStream_t>(hss_ctx_stream(ctx));
    Buffers buf;
    buf.st = st;
#define CTRY(expr)                                                                                         \
    do {                                                                                                   \
        hipError_t _e = (expr);                                                                            \
        if (_e != hipSuccess)                                                                              \
            return host_fail(_e == hipErrorOutOfMemory ? HSS_ERR_OOM : HSS_ERR_HIP,                        \
                             std::string("hss_mat_ldiv: ") + #expr + ": " + hipGetErrorString(_e));        \
    } while (0)
    hss_fact_t F = nullptr;
    hss_status s = hss_ulv_factor(a, HSS_PREC_FP64, &F);
    if (s != HSS_OK) return s;
    struct FactGuard {
        hss_fact_t f;
        ~FactGuard() { hss_fact_destroy(f); }
    } guard{F};
    const int w = (int)std::min<int64_t>(kBlock, std::max<int64_t>(nc, 1));
    double *C = nullptr, *E = nullptr, *T = nullptr;
    CTRY(buf.get(&C, (size_t)std::max<int64_t>(n, 1) * std::max<int64_t>(nc, 1)));
    CTRY(buf.get(&E, (size_t)std::max<int64_t>(nc, 1) * w));
    CTRY(buf.get(&T, (size_t)std::max<int64_t>(n, 1) * w));
    CTRY(hipMemsetAsync(E, 0, sizeof(double) * (size_t)std::max<int64_t>(nc, 1) * w, st));
    for (int64_t j0 = 0; j0 < nc; j0 += w) {
        const int c = (int)std::min<int64_t>(w, nc - j0);
        CTRY(launch_unit_cols(E, nc, j0, c, nc, 1.0, st));
        // T = B E (columns j0 .. j0 + c of B), then C[:, j0 ..] = A^-1 T
        if ((s = hss_matmul_dev(b, c, 1.0, E, nc, 0.0, T, n)) != HSS_OK) return s;
        if ((s = hss_ulv_solve_dev(F, c, T, n, C + j0 * n, n, nullptr, nullptr)) != HSS_OK) return s;
        CTRY(launch_unit_cols(E, nc, j0, c, nc, 0.0, st));
    }
    buf.release(T);
    buf.release(E);
    // compress over A's column tree (rows of A^-1 B) and B's column tree
    std::vector<hss_node_desc> t(ta, ta + na);
    for (int i = 0; i < na; ++i) {
        t[i].m = ta[i].n;
        t[i].row_off = ta[i].col_off;
        t[i].n = tb[i].n;
        t[i].col_off = tb[i].col_off;
    }
    CTRY(hipStreamSynchronize(st));
    s = hss_compress_dev(ctx, C, std::max<int64_t>(n, 1), t.data(), na, atol, rtol, out);
    return s;
#undef CTRY
}

extern "C" hss_status hss_mat_rdiv(hss_mat_t a, hss_mat_t b, double atol, double rtol, hss_mat_t* out) {
    if (!a || !b || !out) return host_fail(HSS_ERR_ARG, "null argument");
    *out = nullptr;
    // A / B = (B' \ A')'
    hss_mat_t at = nullptr, bt = nullptr, ct = nullptr;
    hss_status s = hss_mat_adjoint(a, &at);
    if (s == HSS_OK) s = hss_mat_adjoint(b, &bt);
    if (s == HSS_OK) s = hss_mat_ldiv(bt, at, atol, rtol, &ct);
    if (s == HSS_OK) s = hss_mat_adjoint(ct, out);
    for (hss_mat_t m : {at, bt, ct})
        if (m) hss_mat_destroy(m);
    return s;
}
