// hss_recompress.hip -- recompression of an HSS matrix on the device (hss_mat_recompress;
// reference recompress!, src/compression.jl:148-203, its recursion _recompress!, :205-260, and
// orthonormalize_generators!, src/generators.jl:64-125).
//
// Two passes over the generators, each one tree level at a time:
//   1. bottom-up, orthonormal bases: a leaf's U (V) becomes Q of U = Q T; a branch's stacked
//      translators [R1; R2] ([W1; W2]) become Q of [R1; R2] = Q T; every T moves into the
//      parent's data (R_c <- T_c R_c, B12 <- T_u(a) B12 T_v(b)', B21 <- T_u(b) B21 T_v(a)').
//   2. top-down, truncation: a branch p with children a, b, and the block row Brow and
//      block column Bcol its parent hands it (none at the root), compresses
//        [B12, R_a Brow]  = P1 S2    [B12', W_b Bcol] = Q2 T1
//        [B21, R_b Brow]  = P2 S1    [B21', W_a Bcol] = Q1 T2
//      by truncated column-pivoted QR (the reference's rank rule, compression.jl:22-30), so
//      B12 <- S2[:, 1:kv_b] Q2, B21 <- S1[:, 1:kv_a] Q1, R_a <- P1' R_a, W_b <- Q2' W_b,
//      R_b <- P2' R_b, W_a <- Q1' W_a; child a's bases take P1 / Q1 (U_a P1 and V_a Q1 at a
//      leaf, the translators of a's children times P1 / Q1 at a branch), child b's P2 / Q2;
//      a then works on Brow = [B12, S2[:, kv_b+1:end]], Bcol = [B21', T2[:, ku_b+1:end]], and
//      b on [B21, S1[:, kv_a+1:end]], [B12', T1[:, ku_a+1:end]].
// The result is a new matrix (the ranks change; D is copied).  The QR of pass 1 is the
// library's pivoted QR without truncation, with T = Q' (input).  The reference's
// LowRankApprox.pqrfact is a third-party routine, so the generators are not the
// reference's bit for bit; the result is pinned by its tolerance.
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

extern "C" hss_status hss_mat_recompress(hss_mat_t mat, double atol, double rtol, hss_mat_t* out) {
    if (!mat || !out) return host_fail(HSS_ERR_ARG, "null argument");
    *out = nullptr;
    if (!(atol >= 0.0)) return host_fail(HSS_ERR_ARG, "atol must be >= 0");
    if (!(rtol >= 0.0)) return host_fail(HSS_ERR_ARG, "rtol must be >= 0");
    hss_ctx_t ctx = host_mat_ctx(mat);
    int nn = 0;
    const hss_node_desc* src = host_mat_nodes(mat, &nn);
    const double* blob0 = host_mat_blob(mat);
    if (hipSetDevice(host_ctx_device(ctx)) != hipSuccess) return host_fail(HSS_ERR_HIP, "hipSetDevice");
    hipStream_t st = reinterpret_cast<hipStream_t>(hss_ctx_stream(ctx));
    Buffers buf;
    buf.st = st;
#define CTRY(expr)                                                                                         \
    do {                                                                                                   \
        hipError_t _e = (expr);                                                                            \
        if (_e != hipSuccess)                                                                              \
            return host_fail(_e == hipErrorOutOfMemory ? HSS_ERR_OOM : HSS_ERR_HIP,                        \
                             std::string("hss_mat_recompress: ") + #expr + ": " + hipGetErrorString(_e));  \
    } while (0)
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
        CTRY(hipMemcpyAsync(dT, L.t.data(), sizeof(GemmTask) * L.t.size(), hipMemcpyHostToDevice, st));
        CTRY(launch_gemm_tasks(dT, (int)L.t.size(), st));
        CTRY(hipStreamSynchronize(st));
        L.t.clear();
        return HSS_OK;
    };
    hss_status s = HSS_OK;
    // a new m x n buffer (null when empty)
    auto alloc = [&](double** p, int m, int n) -> hss_status {
        *p = nullptr;
        if ((int64_t)m * n > 0) CTRY(buf.get(p, (size_t)m * n));
        return HSS_OK;
    };
    // buffers replaced by a task list, released once it has run
    std::vector<void*> pending;
    auto retire = [&](double*& p, double* np) {
        if (p) pending.push_back(p);
        p = np;
    };
    auto flush = [&]() {
        for (void* p : pending) buf.release(p);
        pending.clear();
    };
    // column-pivoted QRs (k_rrqr, Q formed) of r x c inputs, each copied first (the kernel
    // works in place): Q (r x k) and the rank k of each
    struct QrIn {
        const double* A;
        int r, c;
        double* Q;
        int k;
    };
    auto qr_batch = [&](std::vector<QrIn>& in, double tatol, double trtol) -> hss_status {
        std::vector<RrqrTask> rq;
        std::vector<double*> ys;
        std::vector<size_t> idx;
        TaskList cp;
        int32_t* dk = nullptr;
        CTRY(buf.get(&dk, std::max<size_t>(in.size(), 1)));
        int max_s = 1, max_k = 1;
        for (size_t t = 0; t < in.size(); ++t) {
            QrIn& q = in[t];
            q.Q = nullptr;
            q.k = 0;
            if (q.r <= 0 || q.c <= 0) continue;
            double* Y = nullptr;
            CTRY(buf.get(&Y, (size_t)q.r * q.c));
            const int kmax = std::min(q.r, q.c);
            CTRY(buf.get(&q.Q, (size_t)q.r * kmax));
            cp.copy(q.r, q.c, false, q.A, q.r, Y, q.r);
            RrqrTask r{};
            r.Y = Y; r.Q = q.Q; r.rank = dk + t;
            r.r = q.r; r.s = q.c; r.kmax = kmax; r.scale = 1.0;
            rq.push_back(r);
            ys.push_back(Y);
            idx.push_back(t);
            max_s = std::max(max_s, q.c);
            max_k = std::max(max_k, kmax);
        }
        if (!rq.empty()) {
            if (max_s + max_k + 8 > 160 * 1024 / 8)
                return host_fail(HSS_ERR_UNSUPPORTED, "a block is too wide for the rank-revealing QR kernel");
            if ((s = run(cp)) != HSS_OK) return s;
            RrqrTask* dq = nullptr;
            CTRY(buf.get(&dq, rq.size()));
            CTRY(hipMemcpyAsync(dq, rq.data(), sizeof(RrqrTask) * rq.size(), hipMemcpyHostToDevice, st));
            CTRY(launch_rrqr(dq, (int)rq.size(), max_s, max_k, tatol, trtol, st));
            std::vector<int32_t> hk(in.size(), 0);
            CTRY(hipMemcpyAsync(hk.data(), dk, sizeof(int32_t) * in.size(), hipMemcpyDeviceToHost, st));
            CTRY(hipStreamSynchronize(st));
            for (size_t t : idx) {
                if (hk[t] < 0 || hk[t] > std::min(in[t].r, in[t].c))
                    return host_fail(HSS_ERR_HIP, "rank out of range");
                in[t].k = hk[t];
            }
            buf.release(dq);
            for (double* y : ys) buf.release(y);
        }
        buf.release(dk);
        return HSS_OK;
    };

    // ---- the generators as separate device matrices (copied from the matrix's blob) ----
    std::vector<int> ku(nn), kv(nn), height(nn, 0), depth(nn, 0);
    std::vector<double*> U(nn, nullptr), V(nn, nullptr), R(nn, nullptr), W(nn, nullptr), B12(nn, nullptr),
        B21(nn, nullptr);
    for (int i = nn - 1; i >= 0; --i)
        if (src[i].left >= 0) height[i] = 1 + std::max(height[src[i].left], height[src[i].right]);
    int maxd = 0;
    for (int i = 1; i < nn; ++i) {
        depth[i] = depth[src[i].parent] + 1;
        maxd = std::max(maxd, depth[i]);
    }
    const int H = height[0];
    for (int i = 0; i < nn; ++i) {
        ku[i] = i == 0 ? 0 : src[i].ku;
        kv[i] = i == 0 ? 0 : src[i].kv;
    }
    {
        TaskList L;
        auto piece = [&](double** dst, int m, int n, int64_t off) -> hss_status {
            if ((s = alloc(dst, m, n)) != HSS_OK) return s;
            if (*dst) L.copy(m, n, false, blob0 + off, m, *dst, m);
            return HSS_OK;
        };
        for (int i = 0; i < nn; ++i) {
            const hss_node_desc& x = src[i];
            if (x.left < 0) {
                if ((s = piece(&U[i], x.m, ku[i], x.off_U)) != HSS_OK) return s;
                if ((s = piece(&V[i], x.n, kv[i], x.off_V)) != HSS_OK) return s;
            } else {
                if ((s = piece(&B12[i], ku[x.left], kv[x.right], x.off_B12)) != HSS_OK) return s;
                if ((s = piece(&B21[i], ku[x.right], kv[x.left], x.off_B21)) != HSS_OK) return s;
            }
            if (i > 0) {
                if ((s = piece(&R[i], ku[i], ku[x.parent], x.off_R)) != HSS_OK) return s;
                if ((s = piece(&W[i], kv[i], kv[x.parent], x.off_W)) != HSS_OK) return s;
            }
        }
        if ((s = run(L)) != HSS_OK) return s;
    }

    // ---- pass 1: orthonormal bases, bottom-up (orthonormalize_generators!) ----
    // after node c's QR: Tu[c] / Tv[c] (new rank x old rank), old_ku[c] / old_kv[c]
    std::vector<double*> Tu(nn, nullptr), Tv(nn, nullptr);
    std::vector<int> old_ku(ku), old_kv(kv);
    for (int h = 0; h <= H; ++h) {
        std::vector<int> lev;
        for (int i = 0; i < nn; ++i)
            if (height[i] == h) lev.push_back(i);
        // (a) the children's T into this level's branches
        TaskList a1, a2;
        std::vector<double*> tmp;
        for (int i : lev) {
            const hss_node_desc& x = src[i];
            if (x.left < 0) continue;
            const int a = x.left, b = x.right;
            if (i != 0)
                for (int c : {a, b}) {
                    double *nr = nullptr, *nw = nullptr;
                    if ((s = alloc(&nr, ku[c], ku[i])) != HSS_OK) return s;
                    if ((s = alloc(&nw, kv[c], kv[i])) != HSS_OK) return s;
                    if (nr) a1.gemm(ku[c], ku[i], old_ku[c], 0, Tu[c], ku[c], R[c], old_ku[c], nr, ku[c]);
                    if (nw) a1.gemm(kv[c], kv[i], old_kv[c], 0, Tv[c], kv[c], W[c], old_kv[c], nw, kv[c]);
                    retire(R[c], nr);
                    retire(W[c], nw);
                }
            // B12 <- Tu[a] B12 Tv[b]', B21 <- Tu[b] B21 Tv[a]'
            double *t12 = nullptr, *t21 = nullptr, *n12 = nullptr, *n21 = nullptr;
            if ((s = alloc(&t12, ku[a], old_kv[b])) != HSS_OK) return s;
            if ((s = alloc(&t21, ku[b], old_kv[a])) != HSS_OK) return s;
            if ((s = alloc(&n12, ku[a], kv[b])) != HSS_OK) return s;
            if ((s = alloc(&n21, ku[b], kv[a])) != HSS_OK) return s;
            if (t12) a1.gemm(ku[a], old_kv[b], old_ku[a], 0, Tu[a], ku[a], B12[i], old_ku[a], t12, ku[a]);
            if (t21) a1.gemm(ku[b], old_kv[a], old_ku[b], 0, Tu[b], ku[b], B21[i], old_ku[b], t21, ku[b]);
            if (n12) a2.gemm(ku[a], kv[b], old_kv[b], GT_TB, t12, ku[a], Tv[b], kv[b], n12, ku[a]);
            if (n21) a2.gemm(ku[b], kv[a], old_kv[a], GT_TB, t21, ku[b], Tv[a], kv[a], n21, ku[b]);
            retire(B12[i], n12);
            retire(B21[i], n21);
            for (double* p : {t12, t21})
                if (p) tmp.push_back(p);
        }
        if ((s = run(a1)) != HSS_OK) return s;
        if ((s = run(a2)) != HSS_OK) return s;
        flush();
        for (double* p : tmp) buf.release(p);
        for (int i : lev)
            if (src[i].left >= 0)
                for (int c : {src[i].left, src[i].right})
                    for (double** p : {&Tu[c], &Tv[c]})
                        if (*p) { buf.release(*p); *p = nullptr; }
        if (h == H) break;  // the root has no basis
        // (b) this level's bases: QR of U / V at a leaf, of [R_a; R_b] / [W_a; W_b] at a branch
        std::vector<double*> stk;
        TaskList sk;
        std::vector<QrIn> qin;
        for (int i : lev) {
            const hss_node_desc& x = src[i];
            for (int sd = 0; sd < 2; ++sd) {
                const int k = sd ? kv[i] : ku[i];
                if (x.left < 0) {
                    qin.push_back(QrIn{sd ? V[i] : U[i], sd ? x.n : x.m, k, nullptr, 0});
                    continue;
                }
                const int a = x.left, b = x.right;
                const int ra = sd ? kv[a] : ku[a], rb = sd ? kv[b] : ku[b];
                double* S = nullptr;
                if ((s = alloc(&S, ra + rb, k)) != HSS_OK) return s;
                if (S) {
                    if (ra > 0) sk.copy(ra, k, false, sd ? W[a] : R[a], ra, S, ra + rb);
                    if (rb > 0) sk.copy(rb, k, false, sd ? W[b] : R[b], rb, S + ra, ra + rb);
                    stk.push_back(S);
                }
                qin.push_back(QrIn{S, ra + rb, k, nullptr, 0});
            }
        }
        if ((s = run(sk)) != HSS_OK) return s;
        if ((s = qr_batch(qin, 0.0, 0.0)) != HSS_OK) return s;
        TaskList tt, wb;
        for (size_t t = 0; t < lev.size(); ++t) {
            const int i = lev[t];
            const hss_node_desc& x = src[i];
            for (int sd = 0; sd < 2; ++sd) {
                const QrIn& q = qin[2 * t + sd];
                const int kn = q.k, k = sd ? kv[i] : ku[i];
                double*& T = sd ? Tv[i] : Tu[i];
                if ((s = alloc(&T, kn, k)) != HSS_OK) return s;
                if (T) tt.gemm(kn, k, q.r, GT_TA, q.Q, q.r, q.A, q.r, T, kn);  // T = Q' (input)
                if (x.left < 0) {
                    double* nb = nullptr;
                    if ((s = alloc(&nb, q.r, kn)) != HSS_OK) return s;
                    if (nb) wb.copy(q.r, kn, false, q.Q, q.r, nb, q.r);
                    retire(sd ? V[i] : U[i], nb);
                } else {
                    const int a = x.left, b = x.right;
                    const int ra = sd ? kv[a] : ku[a], rb = sd ? kv[b] : ku[b];
                    double *na = nullptr, *nb = nullptr;
                    if ((s = alloc(&na, ra, kn)) != HSS_OK) return s;
                    if ((s = alloc(&nb, rb, kn)) != HSS_OK) return s;
                    if (na) wb.copy(ra, kn, false, q.Q, ra + rb, na, ra);
                    if (nb) wb.copy(rb, kn, false, q.Q + ra, ra + rb, nb, rb);
                    retire(sd ? W[a] : R[a], na);
                    retire(sd ? W[b] : R[b], nb);
                }
                (sd ? old_kv : old_ku)[i] = k;
                (sd ? kv : ku)[i] = kn;
            }
        }
        if ((s = run(tt)) != HSS_OK) return s;
        if ((s = run(wb)) != HSS_OK) return s;
        flush();
        for (const QrIn& q : qin)
            if (q.Q) buf.release(q.Q);
        for (double* p : stk) buf.release(p);
    }

    // ---- pass 2: truncation, top-down (recompress!, _recompress!) ----
    // Brow[p] (ku_p x cb[p]) and Bcol[p] (kv_p x cc[p]) handed to branch p by its parent
    std::vector<double*> Brow(nn, nullptr), Bcol(nn, nullptr);
    std::vector<int> cb(nn, 0), cc(nn, 0);
    for (int d = 0; d < maxd; ++d) {
        std::vector<int> lev;
        for (int i = 0; i < nn; ++i)
            if (depth[i] == d && src[i].left >= 0) lev.push_back(i);
        if (lev.empty()) continue;
        const size_t L = lev.size();
        // the four blocks of each branch: [B12, R_a Brow] (ua x (vb + cb)), [B12', W_b Bcol]
        // (vb x (ua + cc)), [B21, R_b Brow] (ub x (va + cb)), [B21', W_a Bcol] (va x (ub + cc))
        std::vector<double*> blk(4 * L, nullptr);
        std::vector<QrIn> qin(4 * L);
        TaskList f1;
        for (size_t t = 0; t < L; ++t) {
            const int p = lev[t], a = src[p].left, b = src[p].right;
            const int ua = ku[a], va = kv[a], ub = ku[b], vb = kv[b];
            const bool root = p == 0;
            const int bc = root ? 0 : cb[p], bcc = root ? 0 : cc[p];
            struct Spec { int rows, left, extra; const double* B; bool tr; const double* T; const double* X; int xr; };
            const Spec sp[4] = {
                {ua, vb, bc, B12[p], false, R[a], Brow[p], ku[p]},
                {vb, ua, bcc, B12[p], true, W[b], Bcol[p], kv[p]},
                {ub, va, bc, B21[p], false, R[b], Brow[p], ku[p]},
                {va, ub, bcc, B21[p], true, W[a], Bcol[p], kv[p]},
            };
            for (int q = 0; q < 4; ++q) {
                const Spec& z = sp[q];
                const int cols = z.left + (root ? 0 : z.extra);
                double*& Bk = blk[4 * t + q];
                if ((s = alloc(&Bk, z.rows, cols)) != HSS_OK) return s;
                if (Bk) {
                    if (z.left > 0 && z.B) f1.copy(z.rows, z.left, z.tr, z.B, z.tr ? z.left : z.rows, Bk, z.rows);
                    if (!root && z.extra > 0)
                        f1.gemm(z.rows, z.extra, z.xr, 0, z.T, z.rows, z.X, z.xr, Bk + (int64_t)z.left * z.rows,
                                z.rows);
                }
                qin[4 * t + q] = QrIn{Bk, z.rows, cols, nullptr, 0};
            }
        }
        if ((s = run(f1)) != HSS_OK) return s;
        if ((s = qr_batch(qin, atol, rtol)) != HSS_OK) return s;
        // S2 = P1' [..], T1 = Q2' [..], S1 = P2' [..], T2 = Q1' [..]
        std::vector<double*> prj(4 * L, nullptr);
        TaskList f2;
        for (size_t t = 0; t < 4 * L; ++t) {
            const QrIn& q = qin[t];
            if ((s = alloc(&prj[t], q.k, q.c)) != HSS_OK) return s;
            if (prj[t]) f2.gemm(q.k, q.c, q.r, GT_TA, q.Q, q.r, q.A, q.r, prj[t], q.k);
        }
        if ((s = run(f2)) != HSS_OK) return s;
        // new couplings, translators, children's bases
        TaskList f3;
        std::vector<int> nk1(L), nk2(L), nk3(L), nk4(L);
        for (size_t t = 0; t < L; ++t) {
            const int p = lev[t], a = src[p].left, b = src[p].right;
            const int va = kv[a], vb = kv[b];
            const QrIn &q1 = qin[4 * t], &q2 = qin[4 * t + 1], &q3 = qin[4 * t + 2], &q4 = qin[4 * t + 3];
            const int k1 = q1.k, k2 = q2.k, k3 = q3.k, k4 = q4.k;
            const double *P1 = q1.Q, *Q2 = q2.Q, *P2 = q3.Q, *Q1 = q4.Q;
            const double *S2 = prj[4 * t], *S1 = prj[4 * t + 2];
            // B12 = S2[:, 1:vb] Q2 (k1 x k2), B21 = S1[:, 1:va] Q1 (k3 x k4)
            double *n12 = nullptr, *n21 = nullptr;
            if ((s = alloc(&n12, k1, k2)) != HSS_OK) return s;
            if ((s = alloc(&n21, k3, k4)) != HSS_OK) return s;
            if (n12) f3.gemm(k1, k2, vb, 0, S2, k1, Q2, vb, n12, k1);
            if (n21) f3.gemm(k3, k4, va, 0, S1, k3, Q1, va, n21, k3);
            retire(B12[p], n12);
            retire(B21[p], n21);
            if (p != 0) {
                // R_a <- P1' R_a, W_b <- Q2' W_b, R_b <- P2' R_b, W_a <- Q1' W_a
                struct Up { double** X; int k; const double* P; int rows; int cols; };
                const Up up[4] = {{&R[a], k1, P1, ku[a], ku[p]}, {&W[b], k2, Q2, kv[b], kv[p]},
                                  {&R[b], k3, P2, ku[b], ku[p]}, {&W[a], k4, Q1, kv[a], kv[p]}};
                for (const Up& z : up) {
                    double* nx = nullptr;
                    if ((s = alloc(&nx, z.k, z.cols)) != HSS_OK) return s;
                    if (nx) f3.gemm(z.k, z.cols, z.rows, GT_TA, z.P, z.rows, *z.X, z.rows, nx, z.k);
                    retire(*z.X, nx);
                }
            }
            // the children's bases: U P / V Q at a leaf, the grandchildren's translators times
            // P / Q at a branch
            for (int side = 0; side < 2; ++side) {
                const int c = side ? b : a;
                const double* Pu = side ? P2 : P1;
                const double* Pv = side ? Q2 : Q1;
                const int kup = side ? k3 : k1, kvp = side ? k2 : k4;
                const hss_node_desc& xc = src[c];
                if (xc.left < 0) {
                    double *nu = nullptr, *nv = nullptr;
                    if ((s = alloc(&nu, xc.m, kup)) != HSS_OK) return s;
                    if ((s = alloc(&nv, xc.n, kvp)) != HSS_OK) return s;
                    if (nu) f3.gemm(xc.m, kup, ku[c], 0, U[c], xc.m, Pu, ku[c], nu, xc.m);
                    if (nv) f3.gemm(xc.n, kvp, kv[c], 0, V[c], xc.n, Pv, kv[c], nv, xc.n);
                    retire(U[c], nu);
                    retire(V[c], nv);
                } else {
                    for (int g : {xc.left, xc.right}) {
                        double *nr = nullptr, *nw = nullptr;
                        if ((s = alloc(&nr, ku[g], kup)) != HSS_OK) return s;
                        if ((s = alloc(&nw, kv[g], kvp)) != HSS_OK) return s;
                        if (nr) f3.gemm(ku[g], kup, ku[c], 0, R[g], ku[g], Pu, ku[c], nr, ku[g]);
                        if (nw) f3.gemm(kv[g], kvp, kv[c], 0, W[g], kv[g], Pv, kv[c], nw, kv[g]);
                        retire(R[g], nr);
                        retire(W[g], nw);
                    }
                }
            }
            nk1[t] = k1; nk2[t] = k2; nk3[t] = k3; nk4[t] = k4;
        }
        if ((s = run(f3)) != HSS_OK) return s;
        flush();
        // the children's block row / column for the next depth:
        //   a: [B12, S2[:, vb+1:end]], [B21', T2[:, ub+1:end]];  b: [B21, S1[:, va+1:end]], [B12', T1[:, ua+1:end]]
        TaskList f4;
        for (size_t t = 0; t < L; ++t) {
            const int p = lev[t], a = src[p].left, b = src[p].right;
            const int ua = ku[a], va = kv[a], ub = ku[b], vb = kv[b];  // ranks before this step
            const int k1 = nk1[t], k2 = nk2[t], k3 = nk3[t], k4 = nk4[t];
            const bool root = p == 0;
            const int bc = root ? 0 : cb[p], bcc = root ? 0 : cc[p];
            const double *S2 = prj[4 * t], *T1 = prj[4 * t + 1], *S1 = prj[4 * t + 2], *T2 = prj[4 * t + 3];
            struct Nx { int c; int rows; int left; const double* B; bool tr; int bld; const double* tail; int tail_off; int extra; };
            const Nx nx[4] = {
                {a, k1, k2, B12[p], false, k1, S2, vb, bc},   // Brow[a]
                {a, k4, k3, B21[p], true, k3, T2, ub, bcc},   // Bcol[a]
                {b, k3, k4, B21[p], false, k3, S1, va, bc},   // Brow[b]
                {b, k2, k1, B12[p], true, k1, T1, ua, bcc},   // Bcol[b]
            };
            for (int q = 0; q < 4; ++q) {
                const Nx& z = nx[q];
                if (src[z.c].left < 0) continue;  // leaves hand nothing down
                double*& dst = (q & 1) ? Bcol[z.c] : Brow[z.c];
                int& w = (q & 1) ? cc[z.c] : cb[z.c];
                w = z.left + z.extra;
                if ((s = alloc(&dst, z.rows, w)) != HSS_OK) return s;
                if (!dst) continue;
                if (z.left > 0 && z.B) f4.copy(z.rows, z.left, z.tr, z.B, z.bld, dst, z.rows);
                if (z.extra > 0 && z.tail)
                    f4.copy(z.rows, z.extra, false, z.tail + (int64_t)z.tail_off * z.rows, z.rows,
                            dst + (int64_t)z.left * z.rows, z.rows);
            }
            ku[a] = k1; kv[a] = k4; ku[b] = k3; kv[b] = k2;
        }
        if ((s = run(f4)) != HSS_OK) return s;
        for (int p : lev)
            for (double** q : {&Brow[p], &Bcol[p]})
                if (*q) { buf.release(*q); *q = nullptr; }
        for (double* q : blk)
            if (q) buf.release(q);
        for (double* q : prj)
            if (q) buf.release(q);
        for (const QrIn& q : qin)
            if (q.Q) buf.release(q.Q);
    }

    // ---- generators in the layout of hss_mat_create ----
    std::vector<hss_node_desc> nd(src, src + nn);
    int64_t pos = 0;
    auto take = [&](int64_t count) {
        const int64_t o = pos;
        pos += count;
        return count > 0 ? o : (int64_t)-1;
    };
    for (int i = 0; i < nn; ++i) {
        hss_node_desc& x = nd[i];
        x.ku = i == 0 ? 0 : ku[i];
        x.kv = i == 0 ? 0 : kv[i];
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
                if ((int64_t)x.m * x.n > 0) g.copy(x.m, x.n, false, blob0 + src[i].off_D, x.m, blob + x.off_D, x.m);
                if (x.ku > 0) g.copy(x.m, x.ku, false, U[i], x.m, blob + x.off_U, x.m);
                if (x.kv > 0) g.copy(x.n, x.kv, false, V[i], x.n, blob + x.off_V, x.n);
            } else {
                const int kua = nd[x.left].ku, kvb = nd[x.right].kv, kub = nd[x.right].ku, kva = nd[x.left].kv;
                if (kua * kvb > 0) g.copy(kua, kvb, false, B12[i], kua, blob + x.off_B12, kua);
                if (kub * kva > 0) g.copy(kub, kva, false, B21[i], kub, blob + x.off_B21, kub);
            }
            if (i > 0) {
                const int kp = nd[x.parent].ku, kq = nd[x.parent].kv;
                if (x.ku > 0 && kp > 0) g.copy(x.ku, kp, false, R[i], x.ku, blob + x.off_R, x.ku);
                if (x.kv > 0 && kq > 0) g.copy(x.kv, kq, false, W[i], x.kv, blob + x.off_W, x.kv);
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
