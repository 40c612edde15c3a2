/*
 * hss_hip.h -- C ABI of libhss_hip.so: HSS matvec and ULV factor/solve on
 * AMD Instinct MI355X (gfx950).
 *
 * Drop-in boundary for bonevbs/HssMatrices.jl 0.1.6.  The reference has no
 * FFI: its boundary is Julia dispatch on HssMatrix{T}.  Each entry point below
 * replaces one reference method; the Julia binding a maintainer would add is
 * hssmatrices.jl_amd/julia/HssHip.jl.
 *
 *   hss_matmul / hss_matmul_dev   mul!(C, hssA, B, alpha, beta)  src/matmul.jl:18
 *                                 *(hssA, B) / *(hssA, x)        src/matmul.jl:13,15
 *   hss_ulvfactsolve              ulvfactsolve(hssA, b)          src/ulvfactor.jl:10
 *                                 \(hssA, B::Matrix)             src/hssmatrix.jl:234
 *   hss_ulv_factor + hss_ulv_solve   the same, split into a factorization that
 *                                 stays resident in HBM and a solve that reuses
 *                                 it (the reference refactors on every call;
 *                                 src/ulvdivide.jl:29-33 is its only split form)
 *
 * Conventions (mirroring the reference):
 *   - Float64, column-major, leading dimensions in elements.
 *   - Every call returns an hss_status; 0 = success.  HSS_ERR_DIM is the
 *     reference's DimensionMismatch (src/matmul.jl:19-20, src/ulvfactor.jl:25-26),
 *     HSS_ERR_ARG its ArgumentError (src/hssmatrix.jl:41-42), HSS_ERR_SINGULAR
 *     the SingularException of the dense root solve (src/ulvfactor.jl:83).
 *     hss_last_error() returns a thread-local message for the last failure.
 *   - Inputs are never modified; outputs are caller-allocated.  Host buffers
 *     are only borrowed for the duration of a call.  hss_mat_create copies the
 *     generators into library-owned HBM.
 *   - Handles are externally synchronized: one caller at a time per handle.
 */
#ifndef HSS_HIP_H
#define HSS_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum hss_status {
    HSS_OK = 0,
    HSS_ERR_DIM = 1,      /* DimensionMismatch */
    HSS_ERR_ARG = 2,      /* ArgumentError / malformed tree */
    HSS_ERR_SINGULAR = 3, /* SingularException (zero pivot in the dense root solve) */
    HSS_ERR_NOCONV = 4,   /* iterative refinement did not reach the requested residual */
    HSS_ERR_OOM = 5,      /* device allocation failed */
    HSS_ERR_BUDGET = 6,   /* admission refused: over the memory budget */
    HSS_ERR_HIP = 7,      /* HIP runtime error */
    HSS_ERR_UNSUPPORTED = 8,
    HSS_ERR_IO = 9        /* a generator file could not be opened, read or written */
} hss_status;

/*
 * One record per HSS tree node, in breadth-first order (root = 0, then
 * depth 1 left to right, ...).  Generators live in one Float64 blob,
 * column-major, addressed by element offsets (-1 = absent).
 *
 * Mirrors HssMatrix{T} (src/hssmatrix.jl:11-68):
 *   leaf:   D (m x n) at off_D, U (m x ku) at off_U, V (n x kv) at off_V;
 *   branch: B12 (ku(left) x kv(right)) at off_B12, B21 (ku(right) x kv(left)) at off_B21;
 *   every non-root node: its translator into the parent, R (ku x ku(parent))
 *   at off_R and W (kv x kv(parent)) at off_W -- i.e. the reference's R1/W1 are
 *   stored with the left child and R2/W2 with the right child.
 * ku/kv are gensize(node) (src/hssmatrix.jl:254-262); 0 at the root, whose
 * R/W have no columns (src/hssmatrix.jl:46-55).
 */
typedef struct hss_node_desc {
    int32_t parent, left, right; /* node indices, -1 = none */
    int32_t depth, height;       /* root depth 0, leaf height 0 */
    int32_t m, n;                /* block rows / columns */
    int32_t ku, kv;              /* column / row basis ranks */
    int32_t pad;
    int64_t row_off, col_off;    /* first global row / column of the block */
    int64_t off_D, off_U, off_V;
    int64_t off_B12, off_B21;
    int64_t off_R, off_W;
} hss_node_desc; /* 112 bytes */

typedef struct hss_ctx_s* hss_ctx_t;
typedef struct hss_mat_s* hss_mat_t;
typedef struct hss_fact_s* hss_fact_t;

typedef enum hss_prec {
    HSS_PREC_FP64 = 0, /* factor and solve in Float64 */
    HSS_PREC_FP32 = 1  /* factor and inner solve in Float32; Float64 refinement */
} hss_prec;

typedef struct hss_refine_opts {
    double tol;   /* target ||b - A x|| / ||b|| per column (max over columns); <= 0: no refinement */
    int max_iter; /* refinement steps after the first solve */
    int mode;     /* 0 = none, 1 = iterative refinement with the HSS matvec residual,
                     2 = GMRES-IR: each correction by Float64 GMRES preconditioned with the
                     factor (hss_ulv_solve / hss_ulv_solve_dev only; plans and servers
                     return HSS_ERR_ARG) */
} hss_refine_opts;

typedef struct hss_solve_report {
    int iterations;     /* refinement steps taken */
    double rel_residual; /* final max_j ||b_j - A x_j|| / ||b_j|| (computed when refinement is on) */
    double seconds;      /* device time of the call */
} hss_solve_report;

/* Handles may be destroyed in any order: a context, matrix or factor that is destroyed
 * while a matrix, factor, plan or server still uses it is released when the last of those
 * is destroyed. */

/* ---- context: owns the HIP stream used by every call on its handles ---- */
hss_status hss_ctx_create(int device, hss_ctx_t* out);
hss_status hss_ctx_destroy(hss_ctx_t ctx);
/* hipStream_t of the context, for callers that enqueue around hss_*_dev calls */
void* hss_ctx_stream(hss_ctx_t ctx);

/* ---- matrix: copies the generators into HBM (level-major) ----
 * Replaces the HssMatrix{Float64} object itself (src/hssmatrix.jl:11-68). */
hss_status hss_mat_create(hss_ctx_t ctx, const hss_node_desc* nodes, int32_t num_nodes, const double* blob,
                          int64_t blob_len, hss_mat_t* out);
hss_status hss_mat_destroy(hss_mat_t mat);
/* The adjoint A' as a new matrix on the same context (adjoint(hssA), src/hssmatrix.jl:165-171:
 * leaf D', U <-> V; branch B12 <- B21', B21 <- B12'; R <-> W).  With it, hss_matmul gives
 * A' * B and A * hssB = (hssB' * A')' (src/matmul.jl:14), and hss_ulvfactsolve on the adjoint
 * gives A / hssB = (hssB' \ A')' (src/hssmatrix.jl:236). */
hss_status hss_mat_adjoint(hss_mat_t mat, hss_mat_t* out);
/* The HSS x HSS product A * B as a new matrix on the same context (*(hssA, hssB),
 * src/matmul.jl:65-118).  A and B must share one tree (A's column clusters are B's row
 * clusters; else HSS_ERR_DIM).  C's generator ranks are the sums of A's and B's, as in the
 * reference, which does not recompress; the generators are formed on the device. */
hss_status hss_mat_mul(hss_mat_t a, hss_mat_t b, hss_mat_t* out);
/* size(hssA) (src/hssmatrix.jl:94) */
hss_status hss_mat_size(hss_mat_t mat, int64_t* rows, int64_t* cols);
/* bytes of HBM held by the matrix (generators + tables) */
int64_t hss_mat_device_bytes(hss_mat_t mat);

/* ---- generator interchange file ----
 * The reference keeps HssMatrix{T} in memory only (src/hssmatrix.jl:11-68).  This file
 * carries the same generators between processes and languages, e.g. written by the Julia
 * binding from a matrix the reference compressed (src/compression.jl:47-63) and loaded
 * here, or the reverse.  Layout, little-endian: an hss_gen_header, then num_nodes
 * hss_node_desc records (the hss_mat_create table), then blob_len Float64 values (the
 * generator blob). */
typedef struct hss_gen_header {
    char magic[4];        /* "HSSG" */
    uint32_t version;     /* 1 */
    uint32_t elem_size;   /* 8: Float64 generators */
    int32_t num_nodes;
    int64_t blob_len;     /* Float64 elements */
    int64_t reserved[2];  /* 0 */
} hss_gen_header; /* 40 bytes */
/* Writes a table and blob that hss_mat_create accepts (else HSS_ERR_ARG / HSS_ERR_DIM as it
 * would return).  Host only: needs no device. */
hss_status hss_gen_write(const char* path, const hss_node_desc* nodes, int32_t num_nodes, const double* blob,
                         int64_t blob_len);
/* Reads a generator file.  *num_nodes and *blob_len are set whenever the header is valid;
 * nodes and blob are filled when both are non-null (HSS_ERR_DIM if nodes_cap or blob_cap
 * is too small).  A malformed file (magic, version, element size, sizes, truncation) is
 * HSS_ERR_ARG; a table hss_mat_create would refuse returns the status it would return.
 * A file that cannot be opened or read is HSS_ERR_IO.  Host only. */
hss_status hss_gen_read(const char* path, hss_node_desc* nodes, int32_t nodes_cap, int32_t* num_nodes, double* blob,
                        int64_t blob_cap, int64_t* blob_len);
/* hss_mat_create from a generator file */
hss_status hss_mat_load(hss_ctx_t ctx, const char* path, hss_mat_t* out);
/* The matrix's generators (read back from HBM) and node table to a generator file */
hss_status hss_mat_save(hss_mat_t mat, const char* path);

/* ---- direct compression on the device: compress(A, rcl, ccl; atol, rtol), src/compression.jl:47-63 ----
 * dA: the dense rows(tree[0]) x cols(tree[0]) matrix, column-major (lda), in device memory.
 * tree: the cluster trees as a node table in breadth-first order (parent, left, right, m, n,
 * row_off, col_off of every node: the row tree rcl and column tree ccl of the reference,
 * which must be compatible, src/binarytree.jl:68-74); ranks and generator offsets are not
 * read.  Each node's basis is the truncated column-pivoted QR of its HSS block row (column),
 * cut at the first pivot whose norm is not above max(atol, rtol * the first pivot's norm),
 * the reference's rank rule (src/compression.jl:22-30); wide blocks are sampled with
 * Gaussian columns first.  The result is a new matrix (ranks as found) on ctx.  The
 * reference's default tolerances are atol = rtol = 1e-9 (src/HssMatrices.jl:62-63). */
hss_status hss_compress_dev(hss_ctx_t ctx, const double* dA, int64_t lda, const hss_node_desc* tree, int32_t num_nodes,
                            double atol, double rtol, hss_mat_t* out);

/* Randomized adaptive compression of a dense matrix in device memory on one GPU:
 * randcompress_adaptive(A, rcl, ccl; kest, atol, rtol, noversampling, stepsize)
 * (src/compression.jl:312-355, recursion :358-430, interpolative decomposition :433-444).
 * dA is column-major (lda >= rows), tree as for hss_compress_dev.  The matrix is sampled with
 * kest + noversampling Gaussian columns from each side; every node's basis is an
 * interpolative decomposition of its samples (rank = #{|R_ii| > min(atol, rtol |R_11|)}) and
 * the couplings are entries of A at the skeleton rows and columns.  The result is tested on
 * stepsize further columns, ||A Om - hssA Om||_F <= max(atol, rtol ||A Om||_F); on failure
 * the sample grows by stepsize columns and the compression is redone, while kest < columns.
 * samples (may be null) receives the sample width of the returned matrix.  The reference's
 * defaults: kest 10, noversampling 10, stepsize 20 (src/HssMatrices.jl:62-66). */
hss_status hss_randcompress_dev(hss_ctx_t ctx, const double* dA, int64_t lda, const hss_node_desc* tree,
                                int32_t num_nodes, double atol, double rtol, int32_t kest, int32_t noversampling,
                                int32_t stepsize, hss_mat_t* out, int32_t* samples);

/* Recompression: recompress!(hssA; atol, rtol) (src/compression.jl:148-260, with
 * orthonormalize_generators!, src/generators.jl:64-125) as a new matrix on mat's context:
 * the bases are made orthonormal bottom-up, then every branch's couplings and its children's
 * bases are truncated top-down by column-pivoted QR at max(atol, rtol * the first pivot's
 * norm).  Ranks never grow; D is copied.  For products (hss_mat_mul), whose ranks add. */
hss_status hss_mat_recompress(hss_mat_t mat, double atol, double rtol, hss_mat_t* out);

/* HSS right-hand sides: \(hssA, hssB) and /(hssA, hssB) (src/hssmatrix.jl:235-237; the
 * reference's ULV divide, ldiv! / rdiv!, src/ulvdivide.jl:13-54) as new matrices on a's
 * context.  ldiv: A^-1 B, compressed at (atol, rtol) over A's column tree and B's column
 * tree; A square, its column tree the shape of B's row tree (HSS_ERR_DIM otherwise).  A is
 * factored once (Float64, resident) and A^-1 B is formed densely in device memory
 * (8 rows(B) cols(B) bytes, HSS_ERR_OOM if it does not fit) from B's matvec and the ULV
 * solve, then compressed as by hss_compress_dev.  rdiv: A B^-1 = (B' \ A')'. */
hss_status hss_mat_ldiv(hss_mat_t a, hss_mat_t b, double atol, double rtol, hss_mat_t* out);
hss_status hss_mat_rdiv(hss_mat_t a, hss_mat_t b, double atol, double rtol, hss_mat_t* out);

/* The matrix's node table and generator blob in host memory (the layout hss_mat_create
 * takes): the HssMatrix{Float64} itself (src/hssmatrix.jl:11-68), e.g. after
 * hss_compress_dev.  *num_nodes and *blob_len are always set; nodes and blob are each
 * filled when non-null (HSS_ERR_DIM if nodes_cap or blob_cap is too small). */
hss_status hss_mat_generators(hss_mat_t mat, hss_node_desc* nodes, int32_t nodes_cap, int32_t* num_nodes,
                              double* blob, int64_t blob_cap, int64_t* blob_len);

/* ---- matvec: C = alpha*A*B + beta*C,  mul!(C, hssA, B, alpha, beta), src/matmul.jl:18-62 ----
 * Host pointers; B is cols(A) x nrhs (ldb), C is rows(A) x nrhs (ldc). */
hss_status hss_matmul(hss_mat_t mat, int32_t nrhs, double alpha, const double* B, int64_t ldb, double beta,
                      double* C, int64_t ldc);
/* Same with device pointers, enqueued on the context stream; returns without waiting. */
hss_status hss_matmul_dev(hss_mat_t mat, int32_t nrhs, double alpha, const double* dB, int64_t ldb, double beta,
                          double* dC, int64_t ldc);

/* ---- ULV: ulvfactsolve(hssA, b), src/ulvfactor.jl:10-107; \(hssA, B) src/hssmatrix.jl:234 ----
 * Factor + solve + release in one call, as the reference does on every "\". */
hss_status hss_ulvfactsolve(hss_mat_t mat, int32_t nrhs, const double* B, int64_t ldb, double* X, int64_t ldx);

/* Resident factorization (HBM) and solves that reuse it.
 * Node limits of the factor (this and hss_ulvfactsolve, the batched factor, the server; the
 * reference has none): every node's block W = [D; V'] after its children's elimination, m
 * rows of D and n columns, with kv rows of V', must have m <= 256, n <= 256 and
 * (m + kv) * S <= 1024 for S the power of two >= ceil(n / 32), that is
 *   n <= 32: m + kv <= 1024;  n <= 64: m + kv <= 512;  n <= 128: m + kv <= 256;  n <= 256: m + kv <= 128.
 * Leaves of at most 128 points (the reference's default leaf size is 64, src/HssMatrices.jl:64)
 * meet them whenever the ranks stay at or below 128.  A larger node is refused with
 * HSS_ERR_UNSUPPORTED before any device work, naming the node and its sizes. */
hss_status hss_ulv_factor(hss_mat_t mat, hss_prec prec, hss_fact_t* out);
hss_status hss_fact_destroy(hss_fact_t fact);
int64_t hss_fact_device_bytes(hss_fact_t fact);
/* X = A \ B with host pointers.  opts may be NULL (no refinement for FP64, refinement to 1e-10 for FP32);
 * report may be NULL. */
hss_status hss_ulv_solve(hss_fact_t fact, int32_t nrhs, const double* B, int64_t ldb, double* X, int64_t ldx,
                         const hss_refine_opts* opts, hss_solve_report* report);
/* Same with device pointers, on the context stream.  Synchronizes only when refinement is on. */
hss_status hss_ulv_solve_dev(hss_fact_t fact, int32_t nrhs, const double* dB, int64_t ldb, double* dX, int64_t ldx,
                             const hss_refine_opts* opts, hss_solve_report* report);

/* Allocation-free solve.  No counterpart in the reference: it is for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  X = A \ B with device pointers on the context stream, for a Float64
 * factor, without refinement; the same kernels as hss_ulv_solve_dev.  The solve workspace
 * is a fixed 64 MiB device array of the library (one per GPU, held until the process
 * ends); columns run in chunks that fit it.  HSS_ERR_UNSUPPORTED for a Float32 factor or a
 * matrix whose one column needs more workspace than the array holds.  Calls on one GPU share
 * the array: calls that overlap in time (other streams or threads) race. */
hss_status hss_ulv_solve_dev_noalloc(hss_fact_t fact, int32_t nrhs, const double* dB, int64_t ldb, double* dX,
                                     int64_t ldx);

/* ---- batched solves over many (uneven) matrices: one launch per tree level ----
 * Each request i solves fact[i] with nrhs[i] columns, device pointers.  Float64 factors only
 * (HSS_ERR_ARG for a Float32 factor, whose unrefined solve is accurate to about 1e-7: batch
 * Float32 factors in a plan with refinement). */
hss_status hss_ulv_solve_batch_dev(int32_t count, const hss_fact_t* facts, const int32_t* nrhs,
                                   const double* const* dB, const int64_t* ldb, double* const* dX,
                                   const int64_t* ldx);
/* Batch plan: requests i = 0..count-1 solve facts[i] (facts may repeat, sizes may differ) with
 * nrhs[i] columns from dB[i] into dX[i] (device pointers, fixed for the plan's lifetime).  Every
 * execution issues one launch per tree level across all requests; without refinement the launch
 * sequence is captured once in a HIP graph and replayed.  With with_refinement != 0 an execution
 * may refine (opts->mode == 1) in lockstep: one batched Float64 HSS matvec per step, requests
 * leave the batch as they reach opts->tol.  A plan of Float32 factors executes only with
 * refinement (with_refinement != 0 and opts->mode == 1; HSS_ERR_ARG otherwise).  No counterpart
 * in the reference (batched and served solves). */
typedef struct hss_plan_s* hss_plan_t;
hss_status hss_plan_create(int32_t count, const hss_fact_t* facts, const int32_t* nrhs, const double* const* dB,
                           const int64_t* ldb, double* const* dX, const int64_t* ldx, int32_t with_refinement,
                           hss_plan_t* out);
hss_status hss_plan_execute(hss_plan_t plan, const hss_refine_opts* opts, hss_solve_report* report);
hss_status hss_plan_destroy(hss_plan_t plan);
/* Autotuner (no reference counterpart; the reference has no device kernels to tune): times
 * every valid launch choice of each up- and down-sweep level of the plan's solve (kernel form
 * LDS-resident / streamed / chunked, 16 or 32 columns, 256 or 512 threads per workgroup) with
 * reps launches each and keeps the fastest per level.  Choices are cached per level shape
 * for the process.  Results stay those of the same solve up to rounding.  Tuning runs the
 * plan's solve (b -> x) repeatedly, so the plan's solution buffers hold its result afterwards.
 * ms_before / ms_after (may be null): one solve of the plan before and after, in ms.
 * Synchronous. */
hss_status hss_plan_tune(hss_plan_t plan, int32_t reps, double* ms_before, double* ms_after);
/* The choice for level `level` of the up (dir 0, by height) or down (dir 1, by depth)
 * sweep: form 0 resident, 1 streamed, 2 chunked, -1 built-in rules; tw; nt (0 = by width) */
hss_status hss_plan_tune_choice(hss_plan_t plan, int32_t dir, int32_t level, int32_t* form, int32_t* tw, int32_t* nt);
int64_t hss_plan_device_bytes(hss_plan_t plan);

/* ---- serving (no reference counterpart; config C5) ----
 * Factorizations stay resident in HBM, admitted while their bytes fit hbm_budget_bytes
 * (HSS_ERR_BUDGET otherwise).  Submitted solves queue; every hss_server_step runs one round:
 * queued requests take free slots (max_inflight, memory), then all requests in flight advance
 * together -- their first solve or their next refinement correction in one launch per tree
 * level, and one batched Float64 matvec + residual -- and requests at opts->tol retire.
 * Device pointers dB/dX must stay valid until the ticket is reported finished by
 * hss_server_poll.  opts NULL: refinement to 1e-10, at most 10 steps.  A Float32 server needs
 * opts->mode == 1 (HSS_ERR_ARG otherwise).  Not thread-safe: one caller drives a server. */
typedef struct hss_server_s* hss_server_t;
hss_status hss_server_create(hss_ctx_t ctx, int64_t hbm_budget_bytes, hss_prec prec, const hss_refine_opts* opts,
                             int32_t max_inflight, hss_server_t* out);
hss_status hss_server_admit(hss_server_t server, hss_mat_t mat, int32_t* matrix_id);
hss_status hss_server_evict(hss_server_t server, int32_t matrix_id);
int64_t hss_server_used_bytes(hss_server_t server);
hss_status hss_server_submit(hss_server_t server, int32_t matrix_id, int32_t nrhs, const double* dB, int64_t ldb,
                             double* dX, int64_t ldx, int64_t* ticket);
hss_status hss_server_step(hss_server_t server, int32_t* outstanding);
hss_status hss_server_poll(hss_server_t server, int64_t ticket, int32_t* status, double* latency_s,
                           hss_solve_report* report);
hss_status hss_server_destroy(hss_server_t server);

/* Batched factorization of many matrices (one launch per level across all of them). */
hss_status hss_ulv_factor_batch(int32_t count, const hss_mat_t* mats, hss_prec prec, hss_fact_t* out);

/* ---- Float32 at the boundary (the reference's hot path for T = Float32, test/runtests.jl:16) ----
 * HssMatrix{Float32} generators (the same node table, a Float32 blob) and Float32 host data.
 * The device computes in Float64: generators and inputs are widened, results rounded once
 * to Float32.  Otherwise as the Float64 entry points. */
hss_status hss_mat_create_f32(hss_ctx_t ctx, const hss_node_desc* nodes, int32_t num_nodes, const float* blob,
                              int64_t blob_len, hss_mat_t* out);
hss_status hss_matmul_f32(hss_mat_t mat, int32_t nrhs, float alpha, const float* B, int64_t ldb, float beta, float* C,
                          int64_t ldc);
hss_status hss_ulvfactsolve_f32(hss_mat_t mat, int32_t nrhs, const float* B, int64_t ldb, float* X, int64_t ldx);
hss_status hss_ulv_solve_f32(hss_fact_t fact, int32_t nrhs, const float* B, int64_t ldb, float* X, int64_t ldx);

/* ---- diagnostics ---- */
const char* hss_last_error(void);
/* Per-launch timing of the library's kernels (no counterpart in the reference;
 * for profiling and the roofline report): when enabled, every kernel launch is
 * one record.  Its time is the kernel's own duration, from the start and end
 * events the launch records itself (hipExtLaunchKernelGGL); a record that
 * covers several launches is timed by HIP events recorded around them on the
 * context stream.  hss_prof_count() waits for the recorded launches and
 * returns how many there are. */
void hss_prof_enable(int on);
void hss_prof_reset(void);
int32_t hss_prof_count(void);
hss_status hss_prof_get(int32_t i, const char** name, int32_t* level, int32_t* nwork, double* ms);
/* Factor sizes of one node: {m, n, k, nk, mr, nr, pass, nblk, My, Kx} (10 ints).  The node's
 * up-sweep reads Kx input rows and writes My = nk + mr + kv output rows through two operators,
 * nk x Kx and (My - nk) x (Kx + nk); its down-sweep operator is n x n. */
hss_status hss_fact_node_info(hss_fact_t fact, int32_t node, int32_t* info);
/* The matvec's launch structure (no counterpart in the reference): *cut_depth = the depth c
 * whose up and down steps above collapse into one dense product (0: none), *folded = the
 * leaves whose down step is folded into their leaf product. */
hss_status hss_mat_matvec_info(hss_mat_t mat, int32_t* cut_depth, int32_t* folded);
/* Per-call device timing of the last matvec/solve on this thread (milliseconds, -1 if none). */
double hss_last_device_ms(void);

#ifdef __cplusplus
}
#endif
#endif /* HSS_HIP_H */
