// hss_randcompress.hip -- randomized adaptive HSS compression of a dense matrix in device
// memory (hss_randcompress_dev; reference randcompress_adaptive, src/compression.jl:312-355,
// with its recursion _randcompress!, :358-430, and _interpolate, :433-444).
//
// The reference samples the whole matrix once, Scol = A Om_c and Srow = A' Om_r with k + r
// Gaussian columns, and walks the tree in post-order:
//   leaf      the samples of the block row outside the diagonal block are
//             Su = Scol[I, :] - D Om_c[J, :]; an interpolative decomposition of Su'
//             (column-pivoted QR, rank = #{|R_ii| > min(atol, rtol |R_11|)}) gives U = X'
//             and the skeleton rows J_u; the leaf passes on its skeleton samples Su[J_u, :]
//             and its reduced test matrix U' Om_r[I, :].  The column side is the same on
//             Srow and D' (V, J_v, V' Om_c[J, :]).
//   branch    the couplings are entries of A: B12 = A[J_u(a), J_v(b)], B21 = A[J_u(b), J_v(a)];
//             the branch's samples are its children's skeleton samples less the couplings
//             times the other child's reduced test matrix,
//               Su = [Su_a - B12 Om_v(b); Su_b - B21 Om_v(a)],
//               Sv = [Sv_a - B21' Om_u(b); Sv_b - B12' Om_u(a)],
//             and their decompositions give [R1; R2] and [W1; W2]; the skeleton indices are
//             the children's, selected, and the reduced test matrices R1' Om_u(a) + R2' Om_u(b)
//             and W1' Om_v(a) + W2' Om_v(b).
//   root      its couplings only.
// randcompress_adaptive then checks the result on stepsize fresh columns:
// ||A Om_t - hssA Om_t|| <= max(atol, rtol ||A Om_t||) (Frobenius norms over sqrt(stepsize));
// on failure the test columns join the sample (k grows by stepsize) and the compression is
// redone, while k < n.
//
// Here every node of one tree height is processed in the same launches, bottom-up (the
// post-order only orders a node after its children).  The test matrices come from the
// hashed Gaussian generator (k_gauss): an entry depends only on (seed, row, column), so the
// test columns are the next stepsize columns of the same matrices and "append them to the
// sample" is "take more columns".  The reference draws them from Julia's global RNG, so
// the generators are not the reference's bit for bit; the tolerance is the same.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
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

constexpr uint64_t kSeedCol = 0x5241434f4c554d4eull;  // Om_c (fixed: compressions are reproducible)
constexpr uint64_t kSeedRow = 0x52414e44524f5753ull;  // Om_r
constexpr int kRowBlock = 128;                        // rows per task of the sampling products

// one side of a node: the U side (samples from Scol, basis U / [R1; R2], skeleton rows) or
// the V side (samples from Srow, basis V / [W1; W2], skeleton columns)
struct RSide {
    int r = 0;                 // rows of the node's sample (m or n at a leaf, k_a + k_b at a branch)
    double* S = nullptr;       // r x s samples (ld r)
    double* Y = nullptr;       // s x r: S', factored in place by k_rrqr
    int32_t* perm = nullptr;   // r pivots
    int k = 0;                 // rank
    double* X = nullptr;       // r x k: the interpolation basis (U at a leaf, [R1; R2] at a branch)
    double* Sk = nullptr;      // k x s skeleton samples
    double* Om = nullptr;      // k x s reduced test matrix (basis' Om)
    std::vector<int32_t> J;    // k skeleton indices (rows of A on the U side, columns on the V side)
    int32_t* dJ = nullptr;     // the same on the device
};

}  // namespace

extern "C" hss_status hss_randcompress_dev(hss_ctx_t ctx, const double* dA, int64_t lda, const hss_node_desc* tree,
                                           int32_t nn, double atol, double rtol, int32_t kest, int32_t noversampling,
                                           int32_t stepsize, hss_mat_t* out, int32_t* samples) {
    if (!ctx || !tree || !out) return host_fail(HSS_ERR_ARG, "null argument");
    *out = nullptr;
    if (!(atol >= 0.0)) return host_fail(HSS_ERR_ARG, "atol must be >= 0");
    if (!(rtol >= 0.0)) return host_fail(HSS_ERR_ARG, "rtol must be >= 0");
    if (kest < 0 || noversampling < 0 || stepsize < 1)
        return host_fail(HSS_ERR_ARG, "kest and noversampling must be >= 0 and stepsize >= 1");
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
                             std::string("hss_randcompress_dev: ") + #expr + ": " + hipGetErrorString(_e)); \
    } while (0)

    std::vector<int> height(nn, 0);
    for (int i = nn - 1; i >= 0; --i)
        if (tree[i].left >= 0) height[i] = 1 + std::max(height[tree[i].left], height[tree[i].right]);
    const int H = height[0];

    // ---- task runners (each waits for its launch: the host tables are reused) ----
    GemmTask* dT = nullptr;
    size_t dT_cap = 0;
    GatherTask* dG = nullptr;
    size_t dG_cap = 0;
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
    auto run_g = [&](GatherList& L) -> hss_status {
        if (L.t.empty()) return HSS_OK;
        if (L.t.size() > dG_cap) {
            if (dG) buf.release(dG);
            dG = nullptr;
            dG_cap = 0;
            CTRY(buf.get(&dG, L.t.size() * 2));
            dG_cap = L.t.size() * 2;
        }
        CTRY(hipMemcpyAsync(dG, L.t.data(), sizeof(GatherTask) * L.t.size(), hipMemcpyHostToDevice, st));
        CTRY(launch_gather_tasks(dG, (int)L.t.size(), st));
        CTRY(hipStreamSynchronize(st));
        L.t.clear();
        return HSS_OK;
    };
    hss_status s = HSS_OK;

    // ---- samples: Scol = A Om_c (M x cap), Srow = A' Om_r (N x cap), columns added on demand ----
    int cap = 0, have = 0;  // allocated and computed sample columns
    double *om_c = nullptr, *om_r = nullptr, *Scol = nullptr, *Srow = nullptr;
    auto ensure_cols = [&](int cols) -> hss_status {
        if (cols > cap) {
            const int nc = std::max(cols, cap + cap / 2);
            double *oc = nullptr, *orr = nullptr, *sc = nullptr, *sr = nullptr;
            CTRY(buf.get(&oc, (size_t)N * nc));
            CTRY(buf.get(&orr, (size_t)M * nc));
            CTRY(buf.get(&sc, (size_t)M * nc));
            CTRY(buf.get(&sr, (size_t)N * nc));
            // entries depend on (seed, row, column) only: regenerating gives the same columns
            if (N > 0) CTRY(launch_gauss(oc, N, nc, kSeedCol, st));
            if (M > 0) CTRY(launch_gauss(orr, M, nc, kSeedRow, st));
            if (have > 0) {
                CTRY(hipMemcpyAsync(sc, Scol, sizeof(double) * (size_t)M * have, hipMemcpyDeviceToDevice, st));
                CTRY(hipMemcpyAsync(sr, Srow, sizeof(double) * (size_t)N * have, hipMemcpyDeviceToDevice, st));
            }
            for (double* p : {om_c, om_r, Scol, Srow})
                if (p) buf.release(p);
            om_c = oc; om_r = orr; Scol = sc; Srow = sr;
            cap = nc;
        }
        if (cols > have) {
            TaskList L;
            const int c0 = have, w = cols - have;
            for (int64_t rb = 0; rb < M; rb += kRowBlock)
                L.gemm((int)std::min<int64_t>(kRowBlock, M - rb), w, (int)N, 0, dA + rb, (int)lda, om_c + (int64_t)c0 * N,
                       (int)N, Scol + rb + (int64_t)c0 * M, (int)M);
            for (int64_t cb = 0; cb < N; cb += kRowBlock)
                L.gemm((int)std::min<int64_t>(kRowBlock, N - cb), w, (int)M, GT_TA, dA + cb * lda, (int)lda,
                       om_r + (int64_t)c0 * M, (int)M, Srow + cb + (int64_t)c0 * N, (int)N);
            if ((s = run(L)) != HSS_OK) return s;
            have = cols;
        }
        return HSS_OK;
    };

    const int64_t maxrank = N;
    int k = kest;
    hss_mat_t result = nullptr;
    for (;;) {
        const int smp = k + noversampling;  // sample columns of this attempt
        if ((s = ensure_cols(smp + stepsize)) != HSS_OK) return s;

        // ---- one compression with smp sample columns (_randcompress!) ----
        std::vector<RSide> su(nn), sv(nn);
        std::vector<double*> B12(nn, nullptr), B21(nn, nullptr);
        std::vector<int32_t*> jbufs;  // skeleton index buffers, one per height
        auto couplings = [&](int i, GatherList& g) -> hss_status {
            const hss_node_desc& x = tree[i];
            const int a = x.left, b = x.right;
            if (su[a].k > 0 && sv[b].k > 0) {
                CTRY(buf.get(&B12[i], (size_t)su[a].k * sv[b].k));
                g.add(su[a].k, sv[b].k, dA, lda, su[a].dJ, sv[b].dJ, B12[i], su[a].k);
            }
            if (su[b].k > 0 && sv[a].k > 0) {
                CTRY(buf.get(&B21[i], (size_t)su[b].k * sv[a].k));
                g.add(su[b].k, sv[a].k, dA, lda, su[b].dJ, sv[a].dJ, B21[i], su[b].k);
            }
            return HSS_OK;
        };
        auto release_children = [&](int i) {
            for (int c : {tree[i].left, tree[i].right})
                for (RSide* S : {&su[c], &sv[c]}) {
                    for (double** p : {&S->Sk, &S->Om})
                        if (*p) { buf.release(*p); *p = nullptr; }
                }
        };
        for (int h = 0; h < H; ++h) {
            std::vector<int> lev;
            for (int i = 1; i < nn; ++i)
                if (height[i] == h) lev.push_back(i);
            if (lev.empty()) continue;
            GatherList g1;
            TaskList c1, p1, tr;
            for (int i : lev) {
                const hss_node_desc& x = tree[i];
                const bool leaf = x.left < 0;
                RSide& U = su[i];
                RSide& V = sv[i];
                if (leaf) {
                    U.r = x.m;
                    V.r = x.n;
                } else {
                    if ((s = couplings(i, g1)) != HSS_OK) return s;
                    U.r = su[x.left].k + su[x.right].k;
                    V.r = sv[x.left].k + sv[x.right].k;
                }
                for (RSide* S : {&U, &V}) {
                    if (S->r <= 0 || smp <= 0) continue;
                    CTRY(buf.get(&S->S, (size_t)S->r * smp));
                    CTRY(buf.get(&S->Y, (size_t)S->r * smp));
                }
                if (smp <= 0) continue;
                if (leaf) {
                    // Su = Scol[I, :] - D Om_c[J, :];  Sv = Srow[J, :] - D' Om_r[I, :]
                    const double* D = dA + x.row_off + x.col_off * lda;
                    if (U.r > 0) {
                        c1.copy(U.r, smp, false, Scol + x.row_off, (int)M, U.S, U.r);
                        p1.gemm(U.r, smp, x.n, GT_NEG | GT_ACC, D, (int)lda, om_c + x.col_off, (int)N, U.S, U.r);
                    }
                    if (V.r > 0) {
                        c1.copy(V.r, smp, false, Srow + x.col_off, (int)N, V.S, V.r);
                        p1.gemm(V.r, smp, x.m, GT_TA | GT_NEG | GT_ACC, D, (int)lda, om_r + x.row_off, (int)M, V.S,
                                V.r);
                    }
                } else {
                    const int a = x.left, b = x.right;
                    const RSide &ua = su[a], &ub = su[b], &va = sv[a], &vb = sv[b];
                    // Su = [Su_a - B12 Om_v(b); Su_b - B21 Om_v(a)]
                    if (ua.k > 0) c1.copy(ua.k, smp, false, ua.Sk, ua.k, U.S, U.r);
                    if (ub.k > 0) c1.copy(ub.k, smp, false, ub.Sk, ub.k, U.S + ua.k, U.r);
                    if (ua.k > 0 && vb.k > 0)
                        p1.gemm(ua.k, smp, vb.k, GT_NEG | GT_ACC, B12[i], ua.k, vb.Om, vb.k, U.S, U.r);
                    if (ub.k > 0 && va.k > 0)
                        p1.gemm(ub.k, smp, va.k, GT_NEG | GT_ACC, B21[i], ub.k, va.Om, va.k, U.S + ua.k, U.r);
                    // Sv = [Sv_a - B21' Om_u(b); Sv_b - B12' Om_u(a)]
                    if (va.k > 0) c1.copy(va.k, smp, false, va.Sk, va.k, V.S, V.r);
                    if (vb.k > 0) c1.copy(vb.k, smp, false, vb.Sk, vb.k, V.S + va.k, V.r);
                    if (va.k > 0 && ub.k > 0)
                        p1.gemm(va.k, smp, ub.k, GT_TA | GT_NEG | GT_ACC, B21[i], ub.k, ub.Om, ub.k, V.S, V.r);
                    if (vb.k > 0 && ua.k > 0)
                        p1.gemm(vb.k, smp, ua.k, GT_TA | GT_NEG | GT_ACC, B12[i], ua.k, ua.Om, ua.k, V.S + va.k, V.r);
                }
                for (RSide* S : {&U, &V})
                    if (S->r > 0) tr.copy(smp, S->r, true, S->S, S->r, S->Y, smp);  // Y = S'
            }
            if ((s = run_g(g1)) != HSS_OK) return s;
            if ((s = run(c1)) != HSS_OK) return s;
            if ((s = run(p1)) != HSS_OK) return s;
            if ((s = run(tr)) != HSS_OK) return s;

            // interpolative decompositions of S' (rank and pivots, R left in Y); the level's
            // pivots share one buffer (one read-back)
            std::vector<RrqrTask> rq;
            std::vector<RSide*> rs;
            std::vector<int64_t> poff;
            int64_t ptotal = 0;
            for (int i : lev)
                for (RSide* S : {&su[i], &sv[i]}) {
                    S->k = 0;
                    if (S->r <= 0 || smp <= 0) continue;
                    rs.push_back(S);
                    poff.push_back(ptotal);
                    ptotal += S->r;
                }
            int32_t *dk = nullptr, *dperm = nullptr;
            CTRY(buf.get(&dk, rs.size()));
            CTRY(buf.get(&dperm, (size_t)ptotal));
            int max_r = 1, max_k = 1;
            for (size_t t = 0; t < rs.size(); ++t) {
                RSide* S = rs[t];
                S->perm = dperm + poff[t];
                {
                    RrqrTask q{};
                    q.Y = S->Y;
                    q.Q = nullptr;
                    q.rank = dk + t;
                    q.perm = S->perm;
                    q.r = smp;
                    q.s = S->r;
                    q.kmax = std::min(smp, S->r);
                    q.id_rule = 1;
                    q.scale = 1.0;
                    rq.push_back(q);
                    max_r = std::max(max_r, S->r);
                    max_k = std::max(max_k, q.kmax);
                }
            }
            if (max_r + max_k + 8 > 160 * 1024 / 8)
                return host_fail(HSS_ERR_UNSUPPORTED, "a node is too large for the rank-revealing QR kernel");
            std::vector<int32_t> hk(rs.size(), 0), hp((size_t)ptotal);
            if (!rq.empty()) {
                RrqrTask* dq = nullptr;
                CTRY(buf.get(&dq, rq.size()));
                CTRY(hipMemcpyAsync(dq, rq.data(), sizeof(RrqrTask) * rq.size(), hipMemcpyHostToDevice, st));
                CTRY(launch_rrqr(dq, (int)rq.size(), max_r, max_k, atol, rtol, st));
                CTRY(hipMemcpyAsync(hk.data(), dk, sizeof(int32_t) * rs.size(), hipMemcpyDeviceToHost, st));
                CTRY(hipMemcpyAsync(hp.data(), dperm, sizeof(int32_t) * (size_t)ptotal, hipMemcpyDeviceToHost, st));
                CTRY(hipStreamSynchronize(st));
                buf.release(dq);
            }
            buf.release(dk);
            auto piv = [&](size_t t, int c) { return hp[(size_t)poff[t] + c]; };
            // interpolation bases, skeleton samples and indices, reduced test matrices
            std::vector<IdTask> idt;
            GatherList g2;
            for (size_t t = 0; t < rs.size(); ++t) {
                RSide& S = *rs[t];
                S.k = hk[t];
                if (S.k < 0 || S.k > std::min(smp, S.r)) return host_fail(HSS_ERR_HIP, "rank out of range");
                for (int c = 0; c < S.r; ++c)
                    if (piv(t, c) < 0 || piv(t, c) >= S.r) return host_fail(HSS_ERR_HIP, "pivot out of range");
                if (S.k == 0) continue;
                CTRY(buf.get(&S.X, (size_t)S.r * S.k));
                CTRY(buf.get(&S.Sk, (size_t)S.k * smp));
                CTRY(buf.get(&S.Om, (size_t)S.k * smp));
                IdTask it{};
                it.R = S.Y;
                it.perm = S.perm;
                it.Xt = S.X;
                it.r = smp;
                it.s = S.r;
                it.k = S.k;
                it.ldx = S.r;
                idt.push_back(it);
                g2.add(S.k, smp, S.S, S.r, S.perm, nullptr, S.Sk, S.k);  // Sk = S[perm[0:k], :]
            }
            // skeleton indices: a leaf's rows / columns, or its children's, selected; the level's
            // indices share one device buffer (kept until this attempt ends)
            std::vector<int32_t> hJ;
            std::vector<std::pair<RSide*, size_t>> joff;
            for (size_t t = 0; t < rs.size(); ++t) {
                RSide& S = *rs[t];
                if (S.k == 0) continue;
                const bool colside = &S >= sv.data() && &S < sv.data() + nn;  // a V side (else U)
                const int i = (int)(colside ? &S - sv.data() : &S - su.data());
                const hss_node_desc& x = tree[i];
                S.J.resize(S.k);
                if (x.left < 0) {
                    const int64_t off = colside ? x.col_off : x.row_off;
                    for (int c = 0; c < S.k; ++c) S.J[c] = (int32_t)(off + piv(t, c));
                } else {
                    const RSide& a = colside ? sv[x.left] : su[x.left];
                    const RSide& b = colside ? sv[x.right] : su[x.right];
                    for (int c = 0; c < S.k; ++c) {
                        const int q = piv(t, c);
                        S.J[c] = q < a.k ? a.J[q] : b.J[q - a.k];
                    }
                }
                joff.emplace_back(&S, hJ.size());
                hJ.insert(hJ.end(), S.J.begin(), S.J.end());
            }
            if (!hJ.empty()) {
                int32_t* dJ = nullptr;
                CTRY(buf.get(&dJ, hJ.size()));
                CTRY(hipMemcpyAsync(dJ, hJ.data(), sizeof(int32_t) * hJ.size(), hipMemcpyHostToDevice, st));
                for (auto& pr : joff) pr.first->dJ = dJ + pr.second;
                jbufs.push_back(dJ);
            }
            if (!idt.empty()) {
                IdTask* di = nullptr;
                CTRY(buf.get(&di, idt.size()));
                CTRY(hipMemcpyAsync(di, idt.data(), sizeof(IdTask) * idt.size(), hipMemcpyHostToDevice, st));
                CTRY(launch_id_form(di, (int)idt.size(), st));
                CTRY(hipStreamSynchronize(st));
                buf.release(di);
            }
            if ((s = run_g(g2)) != HSS_OK) return s;
            // reduced test matrices: leaf U' Om_r[I, :], V' Om_c[J, :]; branch
            // R1' Om_u(a) + R2' Om_u(b), W1' Om_v(a) + W2' Om_v(b)
            TaskList o1, o2;
            for (int i : lev) {
                const hss_node_desc& x = tree[i];
                for (int sd = 0; sd < 2; ++sd) {
                    RSide& S = sd ? sv[i] : su[i];
                    if (S.k == 0) continue;
                    if (x.left < 0) {
                        const double* om = sd ? om_c + x.col_off : om_r + x.row_off;
                        o1.gemm(S.k, smp, S.r, GT_TA, S.X, S.r, om, (int)(sd ? N : M), S.Om, S.k);
                    } else {
                        const RSide& a = sd ? sv[x.left] : su[x.left];
                        const RSide& b = sd ? sv[x.right] : su[x.right];
                        if (a.k > 0) o1.gemm(S.k, smp, a.k, GT_TA, S.X, S.r, a.Om, a.k, S.Om, S.k);
                        if (b.k > 0)
                            (a.k > 0 ? o2 : o1)
                                .gemm(S.k, smp, b.k, GT_TA | (a.k > 0 ? GT_ACC : 0), S.X + a.k, S.r, b.Om, b.k, S.Om, S.k);
                    }
                }
            }
            if ((s = run(o1)) != HSS_OK) return s;
            if ((s = run(o2)) != HSS_OK) return s;
            for (int i : lev) {
                for (RSide* S : {&su[i], &sv[i]})
                    for (void* p : {(void*)S->S, (void*)S->Y})
                        if (p) buf.release(p);
                su[i].S = su[i].Y = sv[i].S = sv[i].Y = nullptr;
                su[i].perm = sv[i].perm = nullptr;
                if (tree[i].left >= 0) release_children(i);
            }
            buf.release(dperm);
        }
        if (nn > 1) {
            GatherList g;
            if ((s = couplings(0, g)) != HSS_OK) return s;
            if ((s = run_g(g)) != HSS_OK) return s;
            release_children(0);
        }

        // ---- generators in the layout of hss_mat_create (as hss_compress_dev) ----
        std::vector<hss_node_desc> nd(tree, tree + nn);
        int64_t pos = 0;
        auto take = [&](int64_t count) {
            const int64_t o = pos;
            pos += count;
            return count > 0 ? o : (int64_t)-1;
        };
        for (int i = 0; i < nn; ++i) {
            hss_node_desc& x = nd[i];
            x.ku = i == 0 ? 0 : su[i].k;
            x.kv = i == 0 ? 0 : sv[i].k;
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
            TaskList gt;
            for (int i = 0; i < nn; ++i) {
                const hss_node_desc& x = nd[i];
                if (x.left < 0) {
                    gt.copy(x.m, x.n, false, dA + x.row_off + (int64_t)x.col_off * lda, (int)lda, blob + x.off_D, x.m);
                    if (x.ku > 0) gt.copy(x.m, x.ku, false, su[i].X, su[i].r, blob + x.off_U, x.m);
                    if (x.kv > 0) gt.copy(x.n, x.kv, false, sv[i].X, sv[i].r, blob + x.off_V, x.n);
                } else {
                    const int kua = nd[x.left].ku, kvb = nd[x.right].kv, kub = nd[x.right].ku, kva = nd[x.left].kv;
                    if (kua * kvb > 0) gt.copy(kua, kvb, false, B12[i], kua, blob + x.off_B12, kua);
                    if (kub * kva > 0) gt.copy(kub, kva, false, B21[i], kub, blob + x.off_B21, kub);
                }
                if (i > 0) {
                    const int p = x.parent;
                    const bool second = nd[p].right == i;
                    // R / W = this child's rows of the parent's [R1; R2] / [W1; W2]
                    const int r0 = second ? nd[nd[p].left].ku : 0;
                    const int c0 = second ? nd[nd[p].left].kv : 0;
                    if (x.ku > 0 && nd[p].ku > 0)
                        gt.copy(x.ku, nd[p].ku, false, su[p].X + r0, su[p].r, blob + x.off_R, x.ku);
                    if (x.kv > 0 && nd[p].kv > 0)
                        gt.copy(x.kv, nd[p].kv, false, sv[p].X + c0, sv[p].r, blob + x.off_W, x.kv);
                }
            }
            hss_status gs = run(gt);
            if (gs != HSS_OK) {
                (void)hipFree(blob);
                return gs;
            }
        }
        hss_mat_t mat = nullptr;
        hss_status ms = host_mat_from_device(ctx, nd.data(), nn, &blob, blob_len, &mat);
        if (blob) (void)hipFree(blob);
        if (ms != HSS_OK) return ms;
        // this attempt's device state (bases, couplings, indices) is not needed any more
        for (int i = 0; i < nn; ++i) {
            for (RSide* S : {&su[i], &sv[i]})
                for (void* p : {(void*)S->X, (void*)S->Sk, (void*)S->Om})
                    if (p) buf.release(p);
            for (double* p : {B12[i], B21[i]})
                if (p) buf.release(p);
        }
        for (int32_t* p : jbufs) buf.release(p);

        // ---- the adaptive test on columns [smp, smp + stepsize) ----
        bool failed = false;
        if (M > 0 && N > 0) {
            double *Yt = nullptr, *nrm = nullptr;
            CTRY(buf.get(&Yt, (size_t)M * stepsize));
            CTRY(buf.get(&nrm, (size_t)2 * stepsize));
            hss_status mm = hss_matmul_dev(mat, stepsize, 1.0, om_c + (int64_t)smp * N, N, 0.0, Yt, M);
            if (mm != HSS_OK) {
                hss_mat_destroy(mat);
                return mm;
            }
            CTRY(launch_colnorm2(Scol + (int64_t)smp * M, M, M, stepsize, nrm, st));
            CTRY(launch_residual(Scol + (int64_t)smp * M, M, Yt, M, M, stepsize, st));  // Yt = A Om_t - hssA Om_t
            CTRY(launch_colnorm2(Yt, M, M, stepsize, nrm + stepsize, st));
            std::vector<double> hn(2 * (size_t)stepsize);
            CTRY(hipMemcpyAsync(hn.data(), nrm, sizeof(double) * hn.size(), hipMemcpyDeviceToHost, st));
            CTRY(hipStreamSynchronize(st));
            buf.release(Yt);
            buf.release(nrm);
            double a2 = 0.0, e2 = 0.0;
            for (int j = 0; j < stepsize; ++j) {
                a2 += hn[j];
                e2 += hn[stepsize + j];
            }
            const double est = std::sqrt(e2 / stepsize), ref = std::sqrt(a2 / stepsize);
            failed = !(est <= atol) && !(est <= rtol * ref);  // a NaN estimate fails
        }
        if (result) hss_mat_destroy(result);
        result = mat;
        if (!failed || (int64_t)k + stepsize >= maxrank) break;
        k += stepsize;  // the test columns join the sample
    }
    if (samples) *samples = k + noversampling;
    *out = result;
    return HSS_OK;
#undef CTRY
}
