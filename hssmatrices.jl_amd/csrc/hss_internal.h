// hss_internal.h -- device-side tables shared by the HIP kernels and the host
// runtime of libhss_hip.so.  Not part of the public ABI (include/hss_hip.h).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace hss {

// Device memory pointers in argument tables: in the device compilation of the kernels they
// carry the global address space, so the compiler emits global (not flat) loads and stores
// through them, also after they pass through inlined helpers taking plain pointers.  The
// layout is the same as a plain pointer's; the host runtime (which only fills the tables)
// sees plain pointers.
#if defined(__HIP_DEVICE_COMPILE__) && defined(HSS_KERNELS_TU)
#define HSS_GLOBAL __attribute__((address_space(1)))
#else
#define HSS_GLOBAL
#endif

// Reflector block size of the compact-WY representation used by the solve.
constexpr int NB = 16;
// Threads of the factor kernel (one workgroup per tree node).
constexpr int FACT_THREADS = 1024;
// Columns of the node matrix held in registers by one factor thread.
constexpr int FACT_CPT = 32;
// Threads of the matvec / solve kernels.
constexpr int SWEEP_THREADS = 256;
// Right-hand-side columns per workgroup in the matvec / solve sweeps.
constexpr int RHS_TILE = 32;

// One tree node of a matrix, derived from hss_node_desc at hss_mat_create.
struct DevNode {
    int32_t m, n, ku, kv;       // generator sizes; ku/kv = basis ranks (0 at the root)
    int32_t left, right, parent;
    int32_t is_leaf, is_left;   // is_left: this node is its parent's first child
    int32_t depth, height;
    int32_t fold;               // leaf: its down step is folded into its leaf product (opC below)
    int64_t row_off, col_off;
    int64_t D, U, V, B12, B21, R, W;  // element offsets into the generator blob (-1 = none)
    int64_t zrow, frow;         // matvec workspace rows: Z (kv rows), F (ku rows)
    // matvec operators (column-major, element offsets into the operator buffer, -1 = none):
    //   opZ kv x (n | kv_1 + kv_2): V' at a leaf, [W_1' W_2'] at a branch
    //   opF ku x (kv_sibling + ku_parent): [B_sibling R] (B12 for a left child, B21 for a right one)
    //   opC m x (n + ku) at a leaf: [D U]; a folded leaf's is m x (n + kv_sibling + ku_parent):
    //       [D, U B_sibling, U R], so C = D x + U (B_sibling Z_sibling + R F_parent) in one product
    int64_t opZ, opF, opC;
};

// Factorization metadata of one node (host-computed before the factor runs).
struct FNode {
    int32_t m, n;      // size of the node's (merged) diagonal block
    int32_t k, nk;     // k = min(ku, m) eliminated-by-QL rank; nk = m - k LQ reflectors
    int32_t mr, nr;    // rows / columns passed to the parent
    int32_t pass;      // 1: k >= m, nothing is eliminated (src/ulvfactor.jl:32-36)
    int32_t S;         // factor threads per matrix row
    int32_t nblk;      // ceil(nk / NB)
    int32_t pad;
    // Solve operators of a non-root node (resident).  The node's input x is the leaf's
    // rows of B, or for a branch [b_1; u_1; b_2; u_2] (both children's reduced b and u,
    // in that order); its output y = [zloc; reduced b; u] (My = nk + mr + kv rows):
    //   Gz (nk x Kx):           zloc = Gz x
    //   Gr (R x (Kx + nk)):     [reduced b; u] = Gr [x; zloc], R = mr + kv; the last nk
    //                           columns are [-L2a; V_top'], applied to the zloc just
    //                           computed (keeps the reduced b consistent with zloc)
    //   Qt (n x n):             z = Qt [zloc; z from the parent] (Q_lq' explicitly;
    //                           -1 = identity)
    // Kx = m (leaf) or mr_1 + kv_1 + mr_2 + kv_2.
    int32_t My, Kx;
    int64_t Gz, Gr, Qt;
    // factor storage (elements of the factor type), -1 = none.  Everything except G, Qt
    // and the root's LU, LUinv, P12, P21 is factorization-time data, released once G and
    // Qt are formed.
    int64_t Yql, Tql, Ylq, taulq, Tlq, L1, L1inv, L2a, Vt, P12, P21, LU, LUinv, piv;
    // factorization-time workspace (reduced D/U/V passed to the parent, merged D scratch)
    int64_t rwD, rwU, rwV, rwM;
    // root: elements of the contiguous block LU | LUinv | P12 | P21
    int64_t rlen;
    // root (branch): the merged block D (m x m), its inverse Dinv (formed from the LU
    // factors) and the merge Mb (m x Kx, b = Mb [b_1; u_1; b_2; u_2]); solves use
    // z0 = Dinv b, z = z0 + Dinv (b - D z0)
    int64_t Dr, Dinv, Mb;
    // solve workspace rows (times nrhs; column-major blocks of nrhs columns):
    //   yrow: the node's y = [zloc; reduced b; u] (My rows, non-root nodes)
    //   zrow: the node's z (n rows, branches and the root)
    int64_t yrow, zrow;
};

struct WorkItem {
    int32_t inst, node;
};

// One work item of a solve sweep with the node's sizes and row offsets resolved on the
// host, so a workgroup reads this record and its instance before loading data (instead
// of walking node, child and parent tables).  Row bases (y, x0, x1) are in units of
// nrhs elements (the workspace blocks are nrhs columns wide).
//   up (k_ulv_gup*):  K = Kx; A = Gz (Gr when nk == 0); y = own yrow;
//                     leaf: x0 = row_off into B; branch: x = rows [add0, add0 + r0) of the
//                     first child's y (x0 = its yrow, pitch l0), then rows from add1 of the
//                     second child's y (x1 = its yrow, pitch l1)
//   down (k_ulv_gdown*): K = n; A = Qt (-1 = identity); x0 = own yrow (zloc, pitch My);
//                     x1 = parent's zrow, rows from add1, pitch l1 = the parent's n;
//                     y = own zrow, or col_off into X at a leaf
//   root: inst and node only
struct SweepRec {
    int32_t inst, node, leaf, nk;
    int32_t K, My, r0, l0;
    int32_t l1, add0, add1, pad;
    int64_t A, y, x0, x1;
};

// One work item of a matvec sweep with the node's operator and segment offsets resolved
// on the host.  The step is y = A [x_0; x_1] with A (M x K, column-major, lda M) at
// element offset A of the operator buffer and rows [0, r0) of x from segment 0:
//   kind 0 (up):   Z_node = opZ [B rows (leaf) | Z_left; Z_right]
//   kind 1 (down): F_node = opF [Z_sibling; F_parent]
//   kind 2 (leaf): C rows = alpha opC [B rows; F_node] + beta C rows, or for a folded leaf
//                  alpha opC [B rows; Z_sibling; F_parent] + beta C rows
// B / C offsets (x0 with from_b, y of kind 2) are row offsets (pitch: the call's ldb /
// ldc); Z / F offsets are row bases in units of nrhs elements, with pitches ld0, ld1, ldy.
// A cut record (k_cut): inst; A = T, x0 = its input row table, x1 = its output row table
// (device addresses), ld0 = T's rows (lda), K = its columns, r0 = the first output row of
// the strip, M = the strip's rows (<= 16); the workspace is Z, with F at Z + zrows * nrhs.
// Rows [r0, r1) of x come from segment 1 and rows [r1, K) from segment 2 (x2, pitch ld2;
// folded leaves only).  A leaf's segments 1 and 2 are row bases into Z, whose F part
// follows at Z + zrows * nrhs.
struct MvRec {
    int32_t inst, M, K, r0;
    int32_t ld0, ld1, ldy, from_b;
    int64_t A, x0, x1, y;
    int32_t r1, ld2;
    int64_t x2;
};

// A tree cut of the matvec: the sweep steps above the nodes of one depth (the cut nodes)
// collapsed into one dense product y = T x (T M x K, column-major, lda M).  Row k of the
// stacked input x is read from W + xrow[k].off * nrhs + xrow[k].inner (column pitch
// xrow[k].ld) and row g of the output y is written to W + yrow[g].off * nrhs +
// yrow[g].inner (pitch yrow[g].ld), where W is the call's matvec workspace (Z followed by F)
// and row bases are in units of nrhs.
struct CutRow {
    int64_t off;
    int32_t inner, ld;
};

// Per-instance arguments of one matvec launch sequence.
struct MvInst {
    const DevNode* nodes;
    const double* op;  // matvec operators (DevNode::opZ/opF/opC)
    const double* B;
    double* C;
    int64_t ldb, ldc;
    double* Z;
    double* F;
    double alpha, beta;
    int32_t nrhs, pad;
};

// Per-instance arguments of a factorization.
template <typename TF>
struct FInst {
    const DevNode HSS_GLOBAL* nodes;
    const FNode HSS_GLOBAL* fn;
    const double HSS_GLOBAL* blob;
    TF HSS_GLOBAL* fac;
    TF HSS_GLOBAL* rw;
    int32_t HSS_GLOBAL* ipiv;
    int32_t HSS_GLOBAL* status;  // first zero pivot (1-based) of the root LU, 0 = none
};

// Per-instance arguments of a solve.  TF = element type of the factor and of the
// solve workspaces; B and X are Float64.
template <typename TF>
struct SInst {
    const DevNode* nodes;
    const FNode* fn;
    const TF* fac;
    const int32_t* ipiv;
    const double* B;
    double* X;
    int64_t ldb, ldx;
    TF* wy;  // y blocks (FNode::yrow)
    TF* wz;  // z blocks (FNode::zrow)
    int32_t nrhs, pad;
};

// One small product of an HSS x HSS product's generator build (src/matmul.jl:65-118):
//   c = op(a) op(b)  (+ c when GT_ACC), op = transpose when GT_TA / GT_TB;
//   GT_COPY: c = op(a) (b, k unused).
// Column-major, absolute device pointers; m x n output.
enum { GT_TA = 1, GT_TB = 2, GT_ACC = 4, GT_COPY = 8, GT_NEG = 16 };  // GT_NEG: c = -op(a) op(b) (+ c)
struct GemmTask {
    int32_t m, n, k, flags;
    int32_t lda, ldb, ldc, pad;
    const double* a;
    const double* b;
    double* c;
};

// One truncated column-pivoted QR of the direct compression (k_rrqr): Y (r x s, ld r) is
// overwritten; the rank goes to *rank and the first rank columns of Q (ld r, capacity
// kmax = min(r, s) columns) to Q.  scale multiplies the column norms in the truncation
// test (1 / sqrt(s) for a Gaussian sample of s columns, 1 for the block itself).
struct RrqrTask {
    double* Y;
    double* Q;        // may be null: no basis formed
    int32_t* rank;
    int32_t* perm;    // may be null; else perm[j] = the original index of column j after pivoting
    int32_t r, s, kmax;
    int32_t id_rule;  // 0: stop when the pivot norm <= max(atol, rtol |R11|) (compression.jl:22-30);
                      // 1: keep |R_ii| > min(atol, rtol |R11|) (_interpolate, compression.jl:433-444)
    double scale;
};

// Interpolation matrix of one interpolative decomposition (compression.jl:441-443): R is the
// r x s output of k_rrqr (R in rows 0..k-1, columns in pivoted order, ld r) and perm its
// pivots; Xt (s x k, ld ldx) = X' with X(:, perm) = [I_k, R11^-1 R12].
struct IdTask {
    const double* R;
    const int32_t* perm;
    double* Xt;
    int32_t r, s, k, ldx;
};

// dst (m x n, ld ldd) = src[ri[i] + ci[j] * ld] (ri / ci null: i / j)
struct GatherTask {
    const double* src;
    const int32_t* ri;
    const int32_t* ci;
    double* dst;
    int64_t ld;
    int32_t m, n, ldd, pad;
};

// Per-request vectors of a batched refinement step.
struct BatchVec {
    const double* b;  // right-hand side (n x nrhs, ldb)
    double* x;        // solution (ldx)
    double* r;        // residual (n x nrhs, ld n)
    double* d;        // correction (n x nrhs, ld n)
    int64_t ldb, ldx, n;
    int32_t nrhs;     // 0 = inactive (converged)
    int32_t pad;
};

}  // namespace hss
