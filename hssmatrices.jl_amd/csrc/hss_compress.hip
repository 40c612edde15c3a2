// hss_compress.hip -- direct HSS compression of a dense matrix in device memory
// (hss_compress_dev; reference compress(A, rcl, ccl; atol, rtol), src/compression.jl:47-133).
//
// The reference walks the tree in post-order: each node compresses its HSS block row
// (the rows of A outside the node's own columns, with the parts that earlier nodes already
// compressed in their compressed form) and block column by a rank-revealing QR, and reads
// the couplings B12 / B21 off its siblings' compressed blocks (Xia, Chandrasekaran, Gu, Li
// 2010).  A post-order walk is sequential, so here every node of one tree height is done in
// the same launches, bottom-up:
//
//   row basis   leaf: the block row A[I, ~J]; branch p = (a, b): [Phi_a; Phi_b][:, ~J_p],
//               the children's compressed block rows restricted to the columns outside p
//               (compression.jl:123: Brow = [Brow1[:, n2+1:end]; Brow2[:, rn1+1:end]]).
//               Wide inputs are sampled, Y = input * Omega / sqrt(s) with s = rows + 32
//               Gaussian columns (rank(input) <= rows < s, so range(Y) = range(input));
//               narrow ones are factored as they are.  k_rrqr's truncated pivoted QR of Y
//               gives the rank k and the orthonormal basis: U at a leaf, [R1; R2] at a
//               branch (compression.jl:126-129).
//   compressed  Phi_leaf = U' A[I, :], Phi_p = R1' Phi_a + R2' Phi_b  (k x all columns)
//   column side the same on A' (V, [W1; W2], Psi)
//   couplings   B12 = Phi_a[:, J_b] Vbig_b and B21 = Phi_b[:, J_a] Vbig_a, where Vbig is the
//               node's expanded column basis (Vbig_leaf = V, Vbig_p = [Vbig_a W1; Vbig_b W2]):
//               B12 = Ubig_a' A[I_a, J_b] Vbig_b, the projection of the block onto both
//               bases, as the reference's B12 = Bcol2[1:rm1, :] is (compression.jl:117-118).
//
// Differences from the reference, both outside its published interface: a node's block
// row holds its sibling's columns uncompressed (the reference's second child sees the
// first child's columns already projected onto V1), and wide blocks are sampled before the
// pivoted QR.  Each basis is the truncation of the same block row, so the result meets the
// same tolerance; the generators are not the reference's bit for bit, which its
// third-party rank-revealing QR (LowRankApprox.pqrfact) leaves unpinned anyway.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/hss_hip.h"
#include "hss_host.h"
#include "hss_internal.h"
#include "hss_launch.h"
#include "hss_tasks.h"

using namespace hss;

namespace {

constexpr int kOversample = 32;                  // Gaussian columns beyond the input's rows
constexpr uint64_t kSeed = 0x4853534350524553ull;  // fixed: compressions are reproducible

// per-node state of one side (rows: U/R and Phi; columns: V/W and Psi)
struct Side {
    int r = 0;             // rows of the node's sample input (m at a leaf, k_a + k_b at a branch)
    int s = 0;             // columns of Y (sample width, or the input's width when factored as is)
    bool sampled = false;
    double* Y = nullptr;   // r x s
    double* Q = nullptr;   // r x min(r, s): the basis (first k columns)
    int k = 0;             // rank
    double* Phi = nullptr; // k x (the other dimension): compressed block row / column
    double* big = nullptr; // expanded basis (size x k), columns only
};

}  // namespace

extern "C" hss_status hss_compress_dev(hss_ctx_t ctx, const double* dA, int64_t lda, const hss_node_desc* tree,
                                       int32_t nn, double atol, double rtol, hss_mat_t* out) {
    if (!ctx || !tree || !out) return host_fail(HSS_ERR_ARG, "null argument");
    *out = nullptr;
    if (!(atol >= 0.0)) return host_fail(HSS_ERR_ARG, "atol must be >= 0");
    if (!(rtol >= 0.0)) return host_fail(HSS_ERR_ARG, "rtol must be >= 0");
    hss_status vs = host_validate_tree(tree, nn);
    if (vs != HSS_OK) return vs;
    if (tree[0].row_off != 0 || tree[0].col_off != 0) return host_fail(HSS_ERR_ARG, "the root must start at (0, 0)");
    const int64_t M = tree[0].m, N = tree[0].n;
    if (M > 0 && N > 0 && !dA) return host_fail(HSS_ERR_ARG, "null matrix");
    if (lda < std::max<int64_t>(M, 1)) return host_fail(HSS_ERR_DIM, "lda smaller than the matrix's rows");
    if (lda > INT32_MAX || std::max(M, N) > INT32_MAX / 2)
        return host_fail(HSS_ERR_UNSUPPORTED, "matrix too large for the compression's 32-bit strides");
    const int dev = host_ctx_device(ctx);
    if (hipSetDevice(dev) != hipSuccess) return host_fail(HSS_ERR_HIP, "hipSetDevice");
    hipStream_t st = reinterpret_cast<hipStream_t>(hss_ctx_stream(ctx));
    Buffers buf;
    buf.st = st;
#define CTRY(expr)                                                                                         \
    do {                                                                                                   \
        hipError_t _e = (expr);                                                                            \
        if (_e != hipSuccess)                                                                              \
            return host_fail(_e == hipErrorOutOfMemory ? HSS_ERR_OOM : HSS_ERR_HIP,                        \
                             std::string("hss_compress_dev: ") + #expr + ": " + hipGetErrorString(_e));    \
    } while (0)

    // heights (leaf 0) and the processing order
    std::vector<int> height(nn, 0);
    for (int i = nn - 1; i >= 0; --i)
        if (tree[i].left >= 0) height[i] = 1 + std::max(height[tree[i].left], height[tree[i].right]);
    const int H = height[0];
    std::vector<Side> row(nn), col(nn);
    std::vector<double*> B12(nn, nullptr), B21(nn, nullptr);
    const int64_t om_rows = std::max(M, N);

    // Gaussian test matrix (om_rows x om_cols, ld om_rows), grown when a level needs more
    double* om = nullptr;
    int om_cols = 0;
    auto ensure_omega = [&](int cols) -> hss_status {
        if (cols <= om_cols) return HSS_OK;
        if (om) buf.release(om);
        om = nullptr;
        om_cols = 0;
        CTRY(buf.get(&om, (size_t)om_rows * cols));
        CTRY(launch_gauss(om, om_rows, cols, kSeed, st));
        om_cols = cols;
        return HSS_OK;
    };
    GemmTask* dT = nullptr;
    size_t dT_cap = 0;
    auto run = [&](TaskList& L) -> hss_status {
        if (L.t.empty()) return HSS_OK;
        if (L.t.size() > dT_cap) {
            if (dT) buf.release(dT);
            dT = nullptr;
            dT_cap = 0;
            CTRY(buf.get(&dT, L.t.size() * 2));
            dT_cap = L.t.size() * 2;
        }
        // the previous launch may still read the table: the copy is ordered on the stream
        CTRY(hipMemcpyAsync(dT, L.t.data(), sizeof(GemmTask) * L.t.size(), hipMemcpyHostToDevice, st));
        CTRY(launch_gemm_tasks(dT, (int)L.t.size(), st));
        CTRY(hipStreamSynchronize(st));  // the host table is reused by the next call
        L.t.clear();
        return HSS_OK;
    };
    hss_status s = HSS_OK;
    // B12 = Phi_a[:, J_b] Vbig_b and B21 = Phi_b[:, J_a] Vbig_a of branch i
    auto couplings = [&](int i, TaskList& c) -> hss_status {
        const hss_node_desc& x = tree[i];
        if (x.left < 0) return HSS_OK;
        const int a = x.left, b = x.right;
        const hss_node_desc &xa = tree[a], &xb = tree[b];
        const int kua = row[a].k, kub = row[b].k, kva = col[a].k, kvb = col[b].k;
        if (kua > 0 && kvb > 0) {
            CTRY(buf.get(&B12[i], (size_t)kua * kvb));
            c.gemm(kua, kvb, xb.n, 0, row[a].Phi + (int64_t)xb.col_off * kua, kua, col[b].big, xb.n, B12[i], kua);
        }
        if (kub > 0 && kva > 0) {
            CTRY(buf.get(&B21[i], (size_t)kub * kva));
            c.gemm(kub, kva, xa.n, 0, row[b].Phi + (int64_t)xa.col_off * kub, kub, col[a].big, xa.n, B21[i], kub);
        }
        return HSS_OK;
    };
    // once a branch's own compressed blocks and couplings are formed, its children's are not
    // read again (their bases Q stay: they become generators)
    auto release_children = [&](int i) {
        for (int c : {tree[i].left, tree[i].right}) {
            for (Side* S : {&row[c], &col[c]}) {
                if (S->Phi) buf.release(S->Phi);
                S->Phi = nullptr;
                if (S->big && S->big != S->Q) buf.release(S->big);
                S->big = nullptr;
            }
        }
    };

    // ---- bottom-up over the heights of the non-root nodes ----
    for (int h = 0; h < H; ++h) {
        std::vector<int> lev;
        for (int i = 1; i < nn; ++i)
            if (height[i] == h) lev.push_back(i);
        if (lev.empty()) continue;
        // sample inputs: sizes, then Y
        int need_cols = 0;
        for (int i : lev) {
            const hss_node_desc& x = tree[i];
            const bool leaf = x.left < 0;
            for (int sd = 0; sd < 2; ++sd) {
                Side& S = sd ? col[i] : row[i];
                const int64_t W = sd ? M - x.m : N - x.n;  // width of the block row / column
                S.r = leaf ? (sd ? x.n : x.m) : (sd ? col[x.left].k + col[x.right].k : row[x.left].k + row[x.right].k);
                S.sampled = W > (int64_t)S.r + kOversample;
                S.s = S.sampled ? S.r + kOversample : (int)W;
                if (S.sampled) need_cols = std::max(need_cols, S.s);
                if (S.r > 0 && S.s > 0) CTRY(buf.get(&S.Y, (size_t)S.r * S.s));
            }
        }
        if ((s = ensure_omega(need_cols)) != HSS_OK) return s;
        TaskList first, second;
        for (int i : lev) {
            const hss_node_desc& x = tree[i];
            const bool leaf = x.left < 0;
            for (int sd = 0; sd < 2; ++sd) {
                Side& S = sd ? col[i] : row[i];
                if (S.r <= 0 || S.s <= 0) continue;
                // the outside range [0, lo) u [hi, total) of the other dimension
                const int64_t lo = sd ? x.row_off : x.col_off;
                const int64_t hi = lo + (sd ? x.m : x.n);
                const int64_t total = sd ? M : N;
                // the input's row blocks: (source, ld, rows, transposed)
                struct Blk { const double* src; int64_t ld; int rows; bool tr; int r0; };
                std::vector<Blk> blks;
                if (leaf) {
                    // rows: A[I, :] (ld lda); columns: A[:, J]' (transposed)
                    const double* src = sd ? dA + (int64_t)x.col_off * lda : dA + x.row_off;
                    blks.push_back(Blk{src, lda, S.r, sd == 1, 0});
                } else {
                    const Side& a = sd ? col[x.left] : row[x.left];
                    const Side& b = sd ? col[x.right] : row[x.right];
                    if (a.k > 0) blks.push_back(Blk{a.Phi, a.k, a.k, false, 0});
                    if (b.k > 0) blks.push_back(Blk{b.Phi, b.k, b.k, false, a.k});
                }
                // element (i, j) of a block in the outside range: src[i + j*ld] or src[j + i*ld]
                for (const Blk& B : blks) {
                    auto at = [&](int64_t j) { return B.tr ? B.src + j : B.src + j * B.ld; };
                    double* y = S.Y + B.r0;
                    if (S.sampled) {
                        // Y rows = block[:, [0, lo)] Om[[0, lo), :] + block[:, [hi, total)] Om[[hi, total), :]
                        first.gemm(B.rows, S.s, (int)lo, B.tr ? GT_TA : 0, at(0), (int)B.ld, om, (int)om_rows, y, S.r);
                        second.gemm(B.rows, S.s, (int)(total - hi), GT_ACC | (B.tr ? GT_TA : 0), at(hi), (int)B.ld,
                                    om + hi, (int)om_rows, y, S.r);
                    } else {
                        first.copy(B.rows, (int)lo, B.tr, at(0), (int)B.ld, y, S.r);
                        first.copy(B.rows, (int)(total - hi), B.tr, at(hi), (int)B.ld, y + (int64_t)lo * S.r, S.r);
                    }
                }
            }
        }
        if ((s = run(first)) != HSS_OK) return s;
        if ((s = run(second)) != HSS_OK) return s;
        // truncated pivoted QR of every sample
        std::vector<RrqrTask> rq;
        int32_t* dk = nullptr;
        CTRY(buf.get(&dk, lev.size() * 2));
        int max_s = 1, max_k = 1;
        for (size_t t = 0; t < lev.size(); ++t) {
            for (int sd = 0; sd < 2; ++sd) {
                Side& S = sd ? col[lev[t]] : row[lev[t]];
                if (S.r <= 0 || S.s <= 0) continue;
                const int kmax = std::min(S.r, S.s);
                CTRY(buf.get(&S.Q, (size_t)S.r * kmax));
                RrqrTask q{};
                q.Y = S.Y; q.Q = S.Q; q.rank = dk + 2 * t + sd;
                q.r = S.r; q.s = S.s; q.kmax = kmax;
                q.scale = S.sampled ? 1.0 / std::sqrt((double)S.s) : 1.0;
                rq.push_back(q);
                max_s = std::max(max_s, S.s);
                max_k = std::max(max_k, kmax);
            }
        }
        if (max_s + max_k + 8 > 160 * 1024 / 8)
            return host_fail(HSS_ERR_UNSUPPORTED, "a node's sample is too wide for the rank-revealing QR kernel");
        std::vector<int32_t> hk(lev.size() * 2, 0);
        if (!rq.empty()) {
            RrqrTask* dq = nullptr;
            CTRY(buf.get(&dq, rq.size()));
            CTRY(hipMemsetAsync(dk, 0, sizeof(int32_t) * lev.size() * 2, st));
            CTRY(hipMemcpyAsync(dq, rq.data(), sizeof(RrqrTask) * rq.size(), hipMemcpyHostToDevice, st));
            CTRY(launch_rrqr(dq, (int)rq.size(), max_s, max_k, atol, rtol, st));
            CTRY(hipMemcpyAsync(hk.data(), dk, sizeof(int32_t) * lev.size() * 2, hipMemcpyDeviceToHost, st));
            CTRY(hipStreamSynchronize(st));
            buf.release(dq);
        }
        buf.release(dk);
        for (size_t t = 0; t < lev.size(); ++t) {
            Side& R = row[lev[t]];
            Side& C = col[lev[t]];
            R.k = hk[2 * t];
            C.k = hk[2 * t + 1];
            if (R.Y) { buf.release(R.Y); R.Y = nullptr; }
            if (C.Y) { buf.release(C.Y); C.Y = nullptr; }
        }
        // compressed block rows / columns, expanded column bases, and the couplings of the
        // branches of this height
        TaskList p1, p2;
        for (int i : lev) {
            const hss_node_desc& x = tree[i];
            const bool leaf = x.left < 0;
            Side& R = row[i];
            Side& C = col[i];
            if (R.k > 0) CTRY(buf.get(&R.Phi, (size_t)R.k * N));
            if (C.k > 0) CTRY(buf.get(&C.Phi, (size_t)C.k * M));
            if (leaf) {
                // Phi = U' A[I, :]; Psi = V' A[:, J]'
                p1.gemm(R.k, (int)N, x.m, GT_TA, R.Q, R.r, dA + x.row_off, (int)lda, R.Phi, R.k);
                p1.gemm(C.k, (int)M, x.n, GT_TA | GT_TB, C.Q, C.r, dA + (int64_t)x.col_off * lda, (int)lda, C.Phi,
                        C.k);
                C.big = C.Q;  // Vbig = V (n x kv, ld n)
            } else {
                const Side &ra = row[x.left], &rb = row[x.right], &ca = col[x.left], &cb = col[x.right];
                // Phi = R1' Phi_a + R2' Phi_b (and the same for Psi); k = 0 children add nothing
                if (ra.k > 0) p1.gemm(R.k, (int)N, ra.k, GT_TA, R.Q, R.r, ra.Phi, ra.k, R.Phi, R.k);
                if (rb.k > 0)
                    (ra.k > 0 ? p2 : p1)
                        .gemm(R.k, (int)N, rb.k, GT_TA | (ra.k > 0 ? GT_ACC : 0), R.Q + ra.k, R.r, rb.Phi, rb.k,
                              R.Phi, R.k);
                if (ca.k > 0) p1.gemm(C.k, (int)M, ca.k, GT_TA, C.Q, C.r, ca.Phi, ca.k, C.Phi, C.k);
                if (cb.k > 0)
                    (ca.k > 0 ? p2 : p1)
                        .gemm(C.k, (int)M, cb.k, GT_TA | (ca.k > 0 ? GT_ACC : 0), C.Q + ca.k, C.r, cb.Phi, cb.k,
                              C.Phi, C.k);
                // Vbig = [Vbig_a W1; Vbig_b W2]; a rank-0 child's rows are zeros (a product over
                // no terms)
                if (C.k > 0) {
                    CTRY(buf.get(&C.big, (size_t)x.n * C.k));
                    const hss_node_desc &xa = tree[x.left], &xb = tree[x.right];
                    p1.gemm(xa.n, C.k, ca.k, 0, ca.big, xa.n, C.Q, C.r, C.big, x.n);
                    p1.gemm(xb.n, C.k, cb.k, 0, cb.big, xb.n, C.Q + ca.k, C.r, C.big + xa.n, x.n);
                }
            }
        }
        // the couplings of this height's branches read only their children's data
        for (int i : lev)
            if ((s = couplings(i, p1)) != HSS_OK) return s;
        if ((s = run(p1)) != HSS_OK) return s;
        if ((s = run(p2)) != HSS_OK) return s;
        for (int i : lev)
            if (tree[i].left >= 0) release_children(i);
    }
    {
        TaskList c;
        if ((s = couplings(0, c)) != HSS_OK) return s;
        if ((s = run(c)) != HSS_OK) return s;
        release_children(0);
    }

    // ---- generator blob in the layout of hss_mat_create (breadth-first node order) ----
    std::vector<hss_node_desc> nd(tree, tree + nn);
    int64_t pos = 0;
    auto take = [&](int64_t count) {
        const int64_t o = pos;
        pos += count;
        return count > 0 ? o : (int64_t)-1;
    };
    for (int i = 0; i < nn; ++i) {
        hss_node_desc& x = nd[i];
        x.ku = i == 0 ? 0 : row[i].k;
        x.kv = i == 0 ? 0 : col[i].k;
        x.height = height[i];
        x.off_D = x.off_U = x.off_V = x.off_B12 = x.off_B21 = x.off_R = x.off_W = -1;
    }
    for (int i = 0; i < nn; ++i) {
        hss_node_desc& x = nd[i];
        if (x.left < 0) {
            x.off_D = take((int64_t)x.m * x.n);
            x.off_U = take((int64_t)x.m * x.ku);
            x.off_V = take((int64_t)x.n * x.kv);
        } else {
            x.off_B12 = take((int64_t)nd[x.left].ku * nd[x.right].kv);
            x.off_B21 = take((int64_t)nd[x.right].ku * nd[x.left].kv);
        }
        if (i > 0) {
            x.off_R = take((int64_t)x.ku * nd[x.parent].ku);
            x.off_W = take((int64_t)x.kv * nd[x.parent].kv);
        }
    }
    const int64_t blob_len = pos;
    double* blob = nullptr;
    CTRY(hipMalloc(reinterpret_cast<void**>(&blob), sizeof(double) * (size_t)std::max<int64_t>(blob_len, 1)));
    {
        TaskList g;
        for (int i = 0; i < nn; ++i) {
            const hss_node_desc& x = nd[i];
            if (x.left < 0) {
                g.copy(x.m, x.n, false, dA + x.row_off + (int64_t)x.col_off * lda, (int)lda, blob + x.off_D, x.m);
                if (x.ku > 0) g.copy(x.m, x.ku, false, row[i].Q, row[i].r, blob + x.off_U, x.m);
                if (x.kv > 0) g.copy(x.n, x.kv, false, col[i].Q, col[i].r, blob + x.off_V, x.n);
            } else {
                const int kua = nd[x.left].ku, kvb = nd[x.right].kv, kub = nd[x.right].ku, kva = nd[x.left].kv;
                if (kua * kvb > 0) g.copy(kua, kvb, false, B12[i], kua, blob + x.off_B12, kua);
                if (kub * kva > 0) g.copy(kub, kva, false, B21[i], kub, blob + x.off_B21, kub);
            }
            if (i > 0) {
                const int p = x.parent;
                const bool second = nd[p].right == i;
                // R = rows of the parent's [R1; R2] (the row basis of the parent's sample)
                const int r0 = second ? nd[nd[p].left].ku : 0;
                const int c0 = second ? nd[nd[p].left].kv : 0;
                if (x.ku > 0 && nd[p].ku > 0) g.copy(x.ku, nd[p].ku, false, row[p].Q + r0, row[p].r, blob + x.off_R, x.ku);
                if (x.kv > 0 && nd[p].kv > 0) g.copy(x.kv, nd[p].kv, false, col[p].Q + c0, col[p].r, blob + x.off_W, x.kv);
            }
        }
        hss_status gs = run(g);
        if (gs != HSS_OK) {
            (void)hipFree(blob);
            return gs;
        }
    }
    hss_status ms = host_mat_from_device(ctx, nd.data(), nn, &blob, blob_len, out);
    if (blob) (void)hipFree(blob);
    return ms;
#undef CTRY
}
