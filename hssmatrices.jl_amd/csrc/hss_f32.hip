// hss_f32.hip -- Float32 at the boundary: HssMatrix{Float32} generators and Float32
// right-hand sides (the reference's hot path is generic in T; its tests run it for Float32,
// test/runtests.jl:16).  The generators are widened to Float64 when the matrix is created and
// the data are widened on the way in and narrowed on the way out, so the device computes
// in Float64 (the results are the Float64 results rounded once to Float32, at least as
// accurate as a Float32 computation).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <vector>

#include "../../include/hss_hip.h"
#include "hss_host.h"

using namespace hss;

namespace {
// column-major rows x cols (ld) <-> dense (ld rows)
std::vector<double> widen(const float* A, int64_t rows, int64_t cols, int64_t ld) {
    std::vector<double> out((size_t)std::max<int64_t>(rows * cols, 1));
    for (int64_t j = 0; j < cols; ++j)
        for (int64_t i = 0; i < rows; ++i) out[(size_t)(i + j * rows)] = A[i + j * ld];
    return out;
}
void narrow(const std::vector<double>& A, int64_t rows, int64_t cols, float* out, int64_t ld) {
    for (int64_t j = 0; j < cols; ++j)
        for (int64_t i = 0; i < rows; ++i) out[i + j * ld] = (float)A[(size_t)(i + j * rows)];
}
}  // namespace

extern "C" hss_status hss_mat_create_f32(hss_ctx_t ctx, const hss_node_desc* nodes, int32_t num_nodes,
                                         const float* blob, int64_t blob_len, hss_mat_t* out) {
    if (!ctx || !nodes || !out || (blob_len > 0 && !blob) || blob_len < 0)
        return host_fail(HSS_ERR_ARG, "null argument");
    std::vector<double> b(blob, blob + blob_len);
    return hss_mat_create(ctx, nodes, num_nodes, b.data(), blob_len, out);
}

extern "C" hss_status hss_matmul_f32(hss_mat_t mat, int32_t nrhs, float alpha, const float* B, int64_t ldb,
                                     float beta, float* C, int64_t ldc) {
    if (!mat || nrhs < 0 || (nrhs > 0 && (!B || !C))) return host_fail(HSS_ERR_ARG, "null argument or nrhs < 0");
    int64_t rows = 0, cols = 0;
    hss_status s = hss_mat_size(mat, &rows, &cols);
    if (s != HSS_OK) return s;
    if (nrhs == 0) return HSS_OK;
    if (ldb < std::max<int64_t>(cols, 1) || ldc < std::max<int64_t>(rows, 1))
        return host_fail(HSS_ERR_DIM, "leading dimension smaller than the matrix size");
    std::vector<double> b = widen(B, cols, nrhs, ldb);
    std::vector<double> c = beta != 0.0f ? widen(C, rows, nrhs, ldc) : std::vector<double>((size_t)std::max<int64_t>(rows * nrhs, 1));
    s = hss_matmul(mat, nrhs, (double)alpha, b.data(), std::max<int64_t>(cols, 1), (double)beta, c.data(),
                   std::max<int64_t>(rows, 1));
    if (s != HSS_OK) return s;
    narrow(c, rows, nrhs, C, ldc);
    return HSS_OK;
}

extern "C" hss_status hss_ulvfactsolve_f32(hss_mat_t mat, int32_t nrhs, const float* B, int64_t ldb, float* X,
                                           int64_t ldx) {
    if (!mat || nrhs < 0 || (nrhs > 0 && (!B || !X))) return host_fail(HSS_ERR_ARG, "null argument or nrhs < 0");
    int64_t rows = 0, cols = 0;
    hss_status s = hss_mat_size(mat, &rows, &cols);
    if (s != HSS_OK) return s;
    if (nrhs == 0) return HSS_OK;
    if (ldb < std::max<int64_t>(rows, 1) || ldx < std::max<int64_t>(cols, 1))
        return host_fail(HSS_ERR_DIM, "leading dimension smaller than the matrix size");
    std::vector<double> b = widen(B, rows, nrhs, ldb), x((size_t)std::max<int64_t>(cols * nrhs, 1));
    s = hss_ulvfactsolve(mat, nrhs, b.data(), std::max<int64_t>(rows, 1), x.data(), std::max<int64_t>(cols, 1));
    if (s != HSS_OK) return s;
    narrow(x, cols, nrhs, X, ldx);
    return HSS_OK;
}

extern "C" hss_status hss_ulv_solve_f32(hss_fact_t fact, int32_t nrhs, const float* B, int64_t ldb, float* X,
                                        int64_t ldx) {
    if (!fact || nrhs < 0 || (nrhs > 0 && (!B || !X))) return host_fail(HSS_ERR_ARG, "null argument or nrhs < 0");
    if (nrhs == 0) return HSS_OK;
    const int64_t n = host_fact_rows(fact);
    if (ldb < std::max<int64_t>(n, 1) || ldx < std::max<int64_t>(n, 1))
        return host_fail(HSS_ERR_DIM, "leading dimension smaller than the matrix size");
    std::vector<double> b = widen(B, n, nrhs, ldb), x((size_t)std::max<int64_t>(n * nrhs, 1));
    hss_status s = hss_ulv_solve(fact, nrhs, b.data(), std::max<int64_t>(n, 1), x.data(), std::max<int64_t>(n, 1),
                                 nullptr, nullptr);
    if (s != HSS_OK) return s;
    narrow(x, n, nrhs, X, ldx);
    return HSS_OK;
}
