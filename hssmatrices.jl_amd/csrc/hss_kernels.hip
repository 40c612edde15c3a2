// hss_kernels.hip -- gfx950 kernels of libhss_hip.so.
//
// Matvec (reference src/matmul.jl:18-62) as level-synchronous sweeps:
//   k_mv_up      Z_leaf = V' B_seg ; Z = W1' Z1 + W2' Z2        (_matmatup, matmul.jl:32-42)
//   k_mv_down    F1 = B12 Z2 + R1 F ; F2 = B21 Z1 + R2 F       (_matmatdown!, matmul.jl:51-57)
//   k_mv_leaf    C_seg = alpha (D B_seg + U F) + beta C_seg      (matmul.jl:46-47)
//
// ULV (reference src/ulvfactor.jl:10-107), split into a resident factor and a solve:
//   k_ulv_factor  per node: merge of the children (ulvfactor.jl:74-79), QL of U,
//                 Q'D, LQ of the top rows with the reflectors applied to the
//                 bottom rows and to V' (ulvfactor.jl:39-54); stores the
//                 reflectors in compact-WY blocks for the solve
//   k_ulv_root    merge + LU with partial pivoting of the root block (ulvfactor.jl:82-86)
//   k_ulv_sup     solve up-sweep: Q_ql' b, blocked forward substitution, b and u
//                 updates (ulvfactor.jl:42,48-51,74,89)
//   k_ulv_sroot   root LU solve (ulvfactor.jl:83)
//   k_ulv_sdown   solve down-sweep: z <- Q_lq' z (ulvfactor.jl:98-107)
//
// Every kernel takes a list of (instance, node) work items, so one launch can
// cover the same tree level of many HSS matrices.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "hss_internal.h"
#include "hss_launch.h"

namespace hss {

// ---------------------------------------------------------------------------
// small helpers
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ T shfl_xor(T v, int mask) {
    return __shfl_xor(v, mask, 64);
}

// sum over the S adjacent lanes of a row group (S a power of two <= 64)
template <typename T>
__device__ __forceinline__ T row_reduce(T v, int S) {
    for (int o = 1; o < S; o <<= 1) v += shfl_xor(v, o);
    return v;
}

template <typename T>
__device__ __forceinline__ T wave_reduce(T v) {
    for (int o = 1; o < 64; o <<= 1) v += shfl_xor(v, o);
    return v;
}

// Householder reflector as LAPACK dlarfg: given alpha and ||x||^2, returns
// beta, tau and the scale applied to x (v = [1; x*scale]).  xnorm == 0 gives
// tau = 0 (H = I).
template <typename T>
__device__ __forceinline__ void larfg(T alpha, T xnorm2, T& beta, T& tau, T& scale) {
    if (xnorm2 == T(0)) {
        beta = alpha;
        tau = T(0);
        scale = T(0);
        return;
    }
    T nrm = sqrt(alpha * alpha + xnorm2);
    beta = (alpha >= T(0)) ? -nrm : nrm;
    tau = (beta - alpha) / beta;
    scale = T(1) / (alpha - beta);
}

// Y[i + j*ldy] = alpha * sum_l opA(i,l) X[l + j*ldx] (+ Y if acc), i<M, j<N.
// A lives in global memory (column-major, lda), X and Y anywhere the calling
// block can address.  All threads of the block participate (no barrier inside).
template <bool TA, typename TA_t, typename T>
__device__ void blk_gemm(int M, int N, int K, const TA_t* __restrict__ A, int64_t lda, const T* X, int64_t ldx,
                         T* Y, int64_t ldy, T alpha, bool acc) {
    const int total = M * N;
    for (int o = threadIdx.x; o < total; o += blockDim.x) {
        const int i = o % M, j = o / M;
        T s0 = T(0), s1 = T(0);
        const T* x = X + (int64_t)j * ldx;
        int l = 0;
        if (TA) {
            const TA_t* a = A + (int64_t)i * lda;
            for (; l + 1 < K; l += 2) {
                s0 += (T)a[l] * x[l];
                s1 += (T)a[l + 1] * x[l + 1];
            }
            if (l < K) s0 += (T)a[l] * x[l];
        } else {
            const TA_t* a = A + i;
            for (; l + 1 < K; l += 2) {
                s0 += (T)a[(int64_t)l * lda] * x[l];
                s1 += (T)a[(int64_t)(l + 1) * lda] * x[l + 1];
            }
            if (l < K) s0 += (T)a[(int64_t)l * lda] * x[l];
        }
        const T s = alpha * (s0 + s1);
        T* y = Y + i + (int64_t)j * ldy;
        *y = acc ? *y + s : s;
    }
}

// ---------------------------------------------------------------------------
// matvec
// ---------------------------------------------------------------------------
// Up-sweep of one level.  Leaf: Z = V' B_seg.  Branch: Z = W1' Z1 + W2' Z2.
__global__ __launch_bounds__(SWEEP_THREADS) void k_mv_up(const MvInst* __restrict__ insts,
                                                        const WorkItem* __restrict__ work) {
    const WorkItem wi = work[blockIdx.x];
    const MvInst& I = insts[wi.inst];
    const int j0 = blockIdx.y * RHS_TILE;
    if (j0 >= I.nrhs) return;
    const int nt = min(RHS_TILE, I.nrhs - j0);
    const DevNode nd = I.nodes[wi.node];
    extern __shared__ double lds_mv[];
    double* Zt = I.Z + nd.zrow * I.nrhs + (int64_t)j0 * nd.kv;  // kv x nt, ld kv
    if (nd.is_leaf) {
        // stage B_seg (n x nt) in LDS
        double* xs = lds_mv;
        const double* Bs = I.B + nd.col_off + (int64_t)j0 * I.ldb;
        for (int e = threadIdx.x; e < nd.n * nt; e += blockDim.x) {
            const int r = e % nd.n, j = e / nd.n;
            xs[e] = Bs[r + (int64_t)j * I.ldb];
        }
        __syncthreads();
        blk_gemm<true>(nd.kv, nt, nd.n, I.blob + nd.V, nd.n, xs, nd.n, Zt, nd.kv, 1.0, false);
    } else {
        const DevNode L = I.nodes[nd.left];
        const DevNode R = I.nodes[nd.right];
        const double* ZL = I.Z + L.zrow * I.nrhs + (int64_t)j0 * L.kv;
        const double* ZR = I.Z + R.zrow * I.nrhs + (int64_t)j0 * R.kv;
        blk_gemm<true>(nd.kv, nt, L.kv, I.blob + L.W, L.kv, ZL, L.kv, Zt, nd.kv, 1.0, false);
        __syncthreads();
        blk_gemm<true>(nd.kv, nt, R.kv, I.blob + R.W, R.kv, ZR, R.kv, Zt, nd.kv, 1.0, true);
    }
}

// Down-sweep of one level: for a branch node p, F of both children.
__global__ __launch_bounds__(SWEEP_THREADS) void k_mv_down(const MvInst* __restrict__ insts,
                                                          const WorkItem* __restrict__ work) {
    const WorkItem wi = work[blockIdx.x];
    const MvInst& I = insts[wi.inst];
    const int j0 = blockIdx.y * RHS_TILE;
    if (j0 >= I.nrhs) return;
    const int nt = min(RHS_TILE, I.nrhs - j0);
    const DevNode p = I.nodes[wi.node];
    const DevNode L = I.nodes[p.left];
    const DevNode R = I.nodes[p.right];
    const bool has_f = p.parent >= 0;
    const double* ZL = I.Z + L.zrow * I.nrhs + (int64_t)j0 * L.kv;
    const double* ZR = I.Z + R.zrow * I.nrhs + (int64_t)j0 * R.kv;
    double* FL = I.F + L.frow * I.nrhs + (int64_t)j0 * L.ku;
    double* FR = I.F + R.frow * I.nrhs + (int64_t)j0 * R.ku;
    // F1 = B12 * Z2 ; F2 = B21 * Z1
    blk_gemm<false>(L.ku, nt, R.kv, I.blob + p.B12, L.ku, ZR, R.kv, FL, L.ku, 1.0, false);
    blk_gemm<false>(R.ku, nt, L.kv, I.blob + p.B21, R.ku, ZL, L.kv, FR, R.ku, 1.0, false);
    if (has_f) {
        const double* Fp = I.F + p.frow * I.nrhs + (int64_t)j0 * p.ku;
        __syncthreads();
        // F1 += R1 * F ; F2 += R2 * F
        blk_gemm<false>(L.ku, nt, p.ku, I.blob + L.R, L.ku, Fp, p.ku, FL, L.ku, 1.0, true);
        blk_gemm<false>(R.ku, nt, p.ku, I.blob + R.R, R.ku, Fp, p.ku, FR, R.ku, 1.0, true);
    }
}

// Leaves: C_seg = alpha (D B_seg + U F) + beta C_seg.
__global__ __launch_bounds__(SWEEP_THREADS) void k_mv_leaf(const MvInst* __restrict__ insts,
                                                          const WorkItem* __restrict__ work) {
    const WorkItem wi = work[blockIdx.x];
    const MvInst& I = insts[wi.inst];
    const int j0 = blockIdx.y * RHS_TILE;
    if (j0 >= I.nrhs) return;
    const int nt = min(RHS_TILE, I.nrhs - j0);
    const DevNode nd = I.nodes[wi.node];
    const bool has_f = nd.parent >= 0 && nd.ku > 0;
    extern __shared__ double lds_mv[];
    double* xs = lds_mv;                         // n x nt
    double* fs = lds_mv + nd.n * nt;             // ku x nt
    const double* Bs = I.B + nd.col_off + (int64_t)j0 * I.ldb;
    for (int e = threadIdx.x; e < nd.n * nt; e += blockDim.x) {
        const int r = e % nd.n, j = e / nd.n;
        xs[e] = Bs[r + (int64_t)j * I.ldb];
    }
    if (has_f) {
        const double* Fn = I.F + nd.frow * I.nrhs + (int64_t)j0 * nd.ku;
        for (int e = threadIdx.x; e < nd.ku * nt; e += blockDim.x) fs[e] = Fn[e];
    }
    __syncthreads();
    const double* D = I.blob + nd.D;
    const double* U = I.blob + (nd.U < 0 ? 0 : nd.U);
    double* Cs = I.C + nd.row_off + (int64_t)j0 * I.ldc;
    const double alpha = I.alpha, beta = I.beta;
    const int M = nd.m, N = nd.n;
    for (int o = threadIdx.x; o < M * nt; o += blockDim.x) {
        const int i = o % M, j = o / M;
        double s0 = 0.0, s1 = 0.0;
        const double* x = xs + j * N;
        int l = 0;
        for (; l + 1 < N; l += 2) {
            s0 += D[i + (int64_t)l * M] * x[l];
            s1 += D[i + (int64_t)(l + 1) * M] * x[l + 1];
        }
        if (l < N) s0 += D[i + (int64_t)l * M] * x[l];
        double s = s0 + s1;
        if (has_f) {
            const double* fj = fs + j * nd.ku;
            double t = 0.0;
            for (int q = 0; q < nd.ku; ++q) t += U[i + (int64_t)q * M] * fj[q];
            s += t;
        }
        double* c = Cs + i + (int64_t)j * I.ldc;
        *c = (beta == 0.0) ? alpha * s : alpha * s + beta * (*c);
    }
}

// ---------------------------------------------------------------------------
// ULV factorization
// ---------------------------------------------------------------------------
// LDS layout of k_ulv_factor (elements of T):
//   Us    m x ku    row generators, QL-factored in place
//   vb    2 x n     current LQ reflector (double buffered)
//   red   nwaves x n partial column sums of the QL application
//   tot   n         column sums
//   sc    8         scalars
//   taus  max(k, nk)
//   gram  k x k     (QL T-factor scratch)
template <typename T>
struct FactLds {
    T *Us, *vb, *red, *tot, *sc, *taus, *gram;
};

template <typename T>
__device__ __forceinline__ T gen(const double* blob, int64_t off, int64_t idx) {
    return (T)blob[off + idx];
}

// reduced-block accessors of a child (rw workspace)
template <typename T>
struct ChildRed {
    const T *D, *U, *V;
    int mr, nr, ku, kv;
};

template <typename T>
__device__ ChildRed<T> child_red(const FInst<T>& I, int c) {
    const FNode& f = I.fn[c];
    const DevNode& d = I.nodes[c];
    ChildRed<T> r;
    r.D = I.rw + f.rwD;
    r.U = I.rw + f.rwU;
    r.V = I.rw + f.rwV;
    r.mr = f.mr;
    r.nr = f.nr;
    r.ku = d.ku;
    r.kv = d.kv;
    return r;
}

// Merged diagonal block entry (r, c) of a branch node (ulvfactor.jl:75):
//   [ D1            U1 B12 V2' ]
//   [ U2 B21 V1'    D2         ]
// with P12 = U1 B12 and P21 = U2 B21 precomputed.
template <typename T>
__device__ __forceinline__ T merged_D(const ChildRed<T>& a, const ChildRed<T>& b, const T* P12, const T* P21, int r,
                                      int c) {
    if (r < a.mr) {
        if (c < a.nr) return a.D[r + (int64_t)c * a.mr];
        const int cc = c - a.nr;
        T s = T(0);
        for (int l = 0; l < b.kv; ++l) s += P12[r + (int64_t)l * a.mr] * b.V[cc + (int64_t)l * b.nr];
        return s;
    }
    const int rr = r - a.mr;
    if (c >= a.nr) return b.D[rr + (int64_t)(c - a.nr) * b.mr];
    T s = T(0);
    for (int l = 0; l < a.kv; ++l) s += P21[rr + (int64_t)l * b.mr] * a.V[c + (int64_t)l * a.nr];
    return s;
}

// P12 = U1 B12 (mr1 x kv2), P21 = U2 B21 (mr2 x kv1) into global factor storage.
template <typename T>
__device__ void merge_products(const FInst<T>& I, const DevNode& nd, const FNode& f, const ChildRed<T>& a,
                               const ChildRed<T>& b) {
    T* P12 = I.fac + f.P12;
    T* P21 = I.fac + f.P21;
    for (int e = threadIdx.x; e < a.mr * b.kv; e += blockDim.x) {
        const int r = e % a.mr, c = e / a.mr;
        T s = T(0);
        for (int l = 0; l < a.ku; ++l) s += a.U[r + (int64_t)l * a.mr] * gen<T>(I.blob, nd.B12, l + (int64_t)c * a.ku);
        P12[e] = s;
    }
    for (int e = threadIdx.x; e < b.mr * a.kv; e += blockDim.x) {
        const int r = e % b.mr, c = e / b.mr;
        T s = T(0);
        for (int l = 0; l < b.ku; ++l) s += b.U[r + (int64_t)l * b.mr] * gen<T>(I.blob, nd.B21, l + (int64_t)c * b.ku);
        P21[e] = s;
    }
}

template <typename T, int CPT>
__global__ __launch_bounds__(FACT_THREADS) void k_ulv_factor(const FInst<T>* __restrict__ insts,
                                                             const WorkItem* __restrict__ work) {
    const WorkItem wi = work[blockIdx.x];
    const FInst<T>& I = insts[wi.inst];
    const DevNode nd = I.nodes[wi.node];
    const FNode f = I.fn[wi.node];
    const int m = f.m, n = f.n, k = f.k, nk = f.nk, ku = nd.ku, kv = nd.kv;
    const int S = f.S;
    const int tid = threadIdx.x;
    const int r = tid / S, s = tid % S;
    const int c0 = s * CPT;
    const int R = m + kv;
    const bool row_on = r < R;
    const int nwaves = blockDim.x / 64;

    extern __shared__ unsigned char lds_raw[];
    T* lds = reinterpret_cast<T*>(lds_raw);
    FactLds<T> L;
    L.Us = lds;
    L.vb = L.Us + (size_t)m * ku;
    L.red = L.vb + 2 * n;
    L.tot = L.red + (size_t)nwaves * n;
    L.sc = L.tot + n;
    L.taus = L.sc + 8;
    L.gram = L.taus + (k > nk ? k : nk) + 1;

    // ---- phase 0: assemble W = [D; V'] (rows in registers) and U (LDS) ----
    T w[CPT];
#pragma unroll
    for (int q = 0; q < CPT; ++q) w[q] = T(0);

    if (nd.is_leaf) {
        if (row_on) {
#pragma unroll
            for (int q = 0; q < CPT; ++q) {
                const int c = c0 + q;
                if (c < n) w[q] = (r < m) ? gen<T>(I.blob, nd.D, r + (int64_t)c * m)
                                          : gen<T>(I.blob, nd.V, c + (int64_t)(r - m) * n);
            }
        }
        for (int e = tid; e < m * ku; e += blockDim.x) L.Us[e] = gen<T>(I.blob, nd.U, e);
    } else {
        const ChildRed<T> a = child_red(I, nd.left);
        const ChildRed<T> b = child_red(I, nd.right);
        const DevNode dl = I.nodes[nd.left];
        const DevNode dr = I.nodes[nd.right];
        merge_products(I, nd, f, a, b);
        __threadfence_block();
        __syncthreads();
        const T* P12 = I.fac + f.P12;
        const T* P21 = I.fac + f.P21;
        // merged rows of W = [D; V'] into the row-major scratch rwM ((m + kv) x n)
        T* Wm = I.rw + f.rwM;
        for (int e = tid; e < R * n; e += blockDim.x) {
            const int rr = e / n, c = e % n;
            T val;
            if (rr < m) {
                val = merged_D(a, b, P12, P21, rr, c);
            } else {
                // V = [V1 W1; V2 W2]  (ulvfactor.jl:79), row j of V'
                const int j = rr - m;
                T sum = T(0);
                if (c < a.nr) {
                    for (int l = 0; l < a.kv; ++l)
                        sum += a.V[c + (int64_t)l * a.nr] * gen<T>(I.blob, dl.W, l + (int64_t)j * a.kv);
                } else {
                    const int cc = c - a.nr;
                    for (int l = 0; l < b.kv; ++l)
                        sum += b.V[cc + (int64_t)l * b.nr] * gen<T>(I.blob, dr.W, l + (int64_t)j * b.kv);
                }
                val = sum;
            }
            Wm[e] = val;
        }
        // U = [U1 R1; U2 R2]  (ulvfactor.jl:78)
        for (int e = tid; e < m * ku; e += blockDim.x) {
            const int rr = e % m, j = e / m;
            T sum = T(0);
            if (rr < a.mr) {
                for (int l = 0; l < a.ku; ++l)
                    sum += a.U[rr + (int64_t)l * a.mr] * gen<T>(I.blob, dl.R, l + (int64_t)j * a.ku);
            } else {
                const int r2 = rr - a.mr;
                for (int l = 0; l < b.ku; ++l)
                    sum += b.U[r2 + (int64_t)l * b.mr] * gen<T>(I.blob, dr.R, l + (int64_t)j * b.ku);
            }
            L.Us[e] = sum;
        }
        __threadfence_block();
        __syncthreads();
        if (row_on) {
            const T* wr = Wm + (int64_t)r * n;
#pragma unroll
            for (int q = 0; q < CPT; ++q)
                if (c0 + q < n) w[q] = wr[c0 + q];
        }
    }
    __syncthreads();

    if (f.pass) {
        // k >= m: hand D, U, V to the parent unchanged (ulvfactor.jl:32-36)
        T* Dr = I.rw + f.rwD;
        T* Vr = I.rw + f.rwV;
        if (row_on) {
#pragma unroll
            for (int q = 0; q < CPT; ++q) {
                const int c = c0 + q;
                if (c < n) {
                    if (r < m) Dr[r + (int64_t)c * m] = w[q];
                    else Vr[c + (int64_t)(r - m) * n] = w[q];
                }
            }
        }
        T* Ur = I.rw + f.rwU;
        for (int e = tid; e < m * ku; e += blockDim.x) Ur[e] = L.Us[e];
        return;
    }

    // ---- phase 1: QL of U (m x k) in LDS, LAPACK dgeqlf semantics (ulvfactor.jl:39) ----
    const int wave = tid / 64, lane = tid % 64;
    for (int jj = k - 1; jj >= 0; --jj) {
        const int p = m - k + jj;
        T* col = L.Us + (int64_t)jj * m;
        if (wave == 0) {
            T part = T(0);
            for (int i = lane; i < p; i += 64) part += col[i] * col[i];
            part = wave_reduce(part);
            if (lane == 0) {
                T beta, tau, scale;
                larfg(col[p], part, beta, tau, scale);
                L.sc[0] = beta;
                L.sc[1] = tau;
                L.sc[2] = scale;
                L.taus[jj] = tau;
            }
        }
        __syncthreads();
        const T tau = L.sc[1], scale = L.sc[2];
        if (tau != T(0))
            for (int i = tid; i < p; i += blockDim.x) col[i] *= scale;
        __syncthreads();
        // apply H = I - tau v v' to U[0:p+1, 0:jj], v = [col[0:p]; 1]
        if (tau != T(0)) {
            for (int c = wave; c < jj; c += nwaves) {
                T* uc = L.Us + (int64_t)c * m;
                T part = T(0);
                for (int i = lane; i < p; i += 64) part += col[i] * uc[i];
                part = wave_reduce(part) + uc[p];
                const T t = tau * part;
                for (int i = lane; i < p; i += 64) uc[i] -= t * col[i];
                if (lane == 0) uc[p] -= t;
            }
        }
        __syncthreads();
        if (tid == 0) col[p] = L.sc[0];
        __syncthreads();
    }

    // Y_ql (explicit, m x k), T_ql (k x k), reduced U = tril(U[m-k:m, :]) for the parent
    {
        T* Y = I.fac + f.Yql;
        for (int e = tid; e < m * k; e += blockDim.x) {
            const int i = e % m, j = e / m;
            const int p = m - k + j;
            Y[e] = (i < p) ? L.Us[e] : (i == p ? T(1) : T(0));
        }
        T* Ur = I.rw + f.rwU;  // k x ku (ku == k here)
        for (int e = tid; e < k * ku; e += blockDim.x) {
            const int i = e % k, j = e / k;
            Ur[e] = (i >= j) ? L.Us[(m - k + i) + (int64_t)j * m] : T(0);
        }
        // Gram G = Y'Y (k x k) for the T factor (dlarft, forward columnwise)
        for (int e = tid; e < k * k; e += blockDim.x) {
            const int a = e % k, bcol = e / k;
            if (a < bcol) {
                // y_a has its unit at p_a < p_b; y_b is zero below p_b
                const int pa = m - k + a, pb = m - k + bcol;
                const T* ya = L.Us + (int64_t)a * m;
                const T* yb = L.Us + (int64_t)bcol * m;
                T sum = yb[pa];  // ya[pa] = 1
                for (int i = 0; i < pa; ++i) sum += ya[i] * yb[i];
                (void)pb;
                L.gram[e] = sum;
            }
        }
        __syncthreads();
        if (tid == 0) {
            T* Tq = I.fac + f.Tql;
            for (int e = 0; e < k * k; ++e) Tq[e] = T(0);
            for (int i = 0; i < k; ++i) {
                const T tau = L.taus[i];
                Tq[i + (int64_t)i * k] = tau;
                // T[0:i, i] = -tau T[0:i,0:i] G[0:i, i]
                for (int a = 0; a < i; ++a) {
                    T sum = T(0);
                    for (int bb = a; bb < i; ++bb) sum += Tq[a + (int64_t)bb * k] * L.gram[bb + (int64_t)i * k];
                    Tq[a + (int64_t)i * k] = -tau * sum;
                }
            }
        }
    }

    // ---- phase 2: D <- Q_ql' D = H(1)...H(k) D  (ulvfactor.jl:41) ----
    for (int jj = k - 1; jj >= 0; --jj) {
        const int p = m - k + jj;
        const T tau = L.taus[jj];
        if (tau == T(0)) continue;  // uniform across the block
        T vr = T(0);
        if (r < m) vr = (r < p) ? L.Us[r + (int64_t)jj * m] : (r == p ? T(1) : T(0));
        // column sums of vr * D[r, c] over the rows of this wave (lanes with the same s)
#pragma unroll
        for (int q = 0; q < CPT; ++q) {
            T pq = vr * w[q];
            for (int o = S; o < 64; o <<= 1) pq += shfl_xor(pq, o);
            if (lane < S && c0 + q < n) L.red[(int64_t)wave * n + c0 + q] = pq;
        }
        __syncthreads();
        for (int c = tid; c < n; c += blockDim.x) {
            T sum = T(0);
            for (int wv = 0; wv < nwaves; ++wv) sum += L.red[(int64_t)wv * n + c];
            L.tot[c] = sum;
        }
        __syncthreads();
        if (vr != T(0)) {
            const T t = tau * vr;
#pragma unroll
            for (int q = 0; q < CPT; ++q)
                if (c0 + q < n) w[q] -= t * L.tot[c0 + q];
        }
        __syncthreads();
    }

    // ---- phase 3: LQ of the first nk rows of [D; V'], reflectors applied to all
    //      rows below (L2 = D_bot Q', V <- Q V; ulvfactor.jl:44,47,50) ----
    for (int i = 0; i < nk; ++i) {
        T* vb = L.vb + (i & 1) * n;
        if (r == i) {
            T part = T(0), alpha_p = T(0);
#pragma unroll
            for (int q = 0; q < CPT; ++q) {
                const int c = c0 + q;
                if (c > i && c < n) part += w[q] * w[q];
                if (c == i) alpha_p = w[q];
            }
            part = row_reduce(part, S);
            const T alpha = row_reduce(alpha_p, S);
            T beta, tau, scale;
            larfg(alpha, part, beta, tau, scale);
#pragma unroll
            for (int q = 0; q < CPT; ++q) {
                const int c = c0 + q;
                if (c < n) {
                    T v;
                    if (c < i) v = T(0);
                    else if (c == i) { v = T(1); w[q] = beta; }
                    else { v = (tau != T(0)) ? w[q] * scale : T(0); if (tau != T(0)) w[q] = v; }
                    vb[c] = v;
                }
            }
            if (s == 0) L.taus[i] = tau;
        }
        __syncthreads();
        if (r > i && row_on) {
            // all S lanes of a row take this branch together (same r), so the
            // S-lane reduction below is convergent
            const T tau = L.taus[i];
            if (tau != T(0)) {
                T part = T(0);
#pragma unroll
                for (int q = 0; q < CPT; ++q) {
                    const int c = c0 + q;
                    if (c >= i && c < n) part += w[q] * vb[c];
                }
                const T dot = row_reduce(part, S) * tau;
#pragma unroll
                for (int q = 0; q < CPT; ++q) {
                    const int c = c0 + q;
                    if (c >= i && c < n) w[q] -= dot * vb[c];
                }
            }
        }
    }
    __syncthreads();

    // ---- phase 4: outputs ----
    if (row_on) {
        if (r < nk) {
            T* L1 = I.fac + f.L1;
            T* Y = I.fac + f.Ylq;
#pragma unroll
            for (int q = 0; q < CPT; ++q) {
                const int c = c0 + q;
                if (c < n) {
                    if (c < nk) L1[r + (int64_t)c * nk] = (c <= r) ? w[q] : T(0);
                    Y[c + (int64_t)r * n] = (c < r) ? T(0) : (c == r ? T(1) : w[q]);
                }
            }
        } else if (r < m) {
            T* L2a = I.fac + f.L2a;
            T* Dr = I.rw + f.rwD;
            const int rr = r - nk;
#pragma unroll
            for (int q = 0; q < CPT; ++q) {
                const int c = c0 + q;
                if (c < n) {
                    if (c < nk) L2a[rr + (int64_t)c * k] = w[q];
                    else Dr[rr + (int64_t)(c - nk) * f.mr] = w[q];
                }
            }
        } else {
            T* Vt = I.fac + f.Vt;
            T* Vr = I.rw + f.rwV;
            const int j = r - m;
#pragma unroll
            for (int q = 0; q < CPT; ++q) {
                const int c = c0 + q;
                if (c < n) {
                    if (c < nk) Vt[j + (int64_t)c * kv] = w[q];
                    else Vr[(c - nk) + (int64_t)j * f.nr] = w[q];
                }
            }
        }
    }
    __threadfence();
    __syncthreads();

    // LQ reflector scalars for the compact-WY blocks (built by k_ulv_wy)
    for (int e = tid; e < nk; e += blockDim.x) I.fac[f.taulq + e] = L.taus[e];
}

// Compact-WY T blocks of the LQ reflectors (dlarft, forward columnwise) and the
// inverted diagonal blocks of L1, from the factor data in global memory: one
// wave per block of NB reflectors.
template <typename T>
__global__ __launch_bounds__(256) void k_ulv_wy(const FInst<T>* __restrict__ insts, const WorkItem* __restrict__ work) {
    const WorkItem wi = work[blockIdx.x];
    const FInst<T>& I = insts[wi.inst];
    const FNode f = I.fn[wi.node];
    if (f.pass || f.nk == 0) return;
    const int n = f.n, nk = f.nk;
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    const int nwaves = blockDim.x / 64;
    const T* taus = I.fac + f.taulq;
    for (int bblk = wave; bblk < f.nblk; bblk += nwaves) {
        const int b0 = bblk * NB;
        const int nb = min(NB, nk - b0);
        const T* Y = I.fac + f.Ylq;
        const T* L1 = I.fac + f.L1;
        T* Tb = I.fac + f.Tlq + (int64_t)bblk * NB * NB;
        T* Li = I.fac + f.L1inv + (int64_t)bblk * NB * NB;
        // Gram of the block (strict upper part): G[a][b] = y_{b0+a} . y_{b0+b}
        T g[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const int e = lane + 64 * t;  // 0..255 = NB*NB entries
            const int a = e % NB, bb = e / NB;
            T sum = T(0);
            if (a < bb && bb < nb) {
                const T* ya = Y + (int64_t)(b0 + a) * n;
                const T* yb = Y + (int64_t)(b0 + bb) * n;
                for (int c = b0 + a; c < n; ++c) sum += ya[c] * yb[c];
            }
            g[t] = sum;
        }
        // T recurrence, lane a owns row a of T
        T trow[NB];
#pragma unroll
        for (int q = 0; q < NB; ++q) trow[q] = T(0);
        for (int i = 0; i < nb; ++i) {
            const T tau = taus[b0 + i];
            // G[a][i] for lane a: entry a + NB*i lives in lane (a + NB*i) % 64, slot (a + NB*i) / 64
            T gai = T(0);
            const int srcl = (lane + NB * i) % 64;
            const int slot = (lane + NB * i) / 64;
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const T v = __shfl(g[t], srcl, 64);
                if (slot == t) gai = v;
            }
            const T tmpa = (lane < i) ? -tau * gai : T(0);
            T acc = T(0);
#pragma unroll
            for (int bb = 0; bb < NB; ++bb) {
                const T tb = __shfl(tmpa, bb, 64);
                if (bb < i) acc += trow[bb] * tb;
            }
#pragma unroll
            for (int q = 0; q < NB; ++q)
                if (q == i) trow[q] = (lane < i) ? acc : (lane == i ? tau : T(0));
        }
        if (lane < NB) {
#pragma unroll
            for (int q = 0; q < NB; ++q) Tb[lane + q * NB] = (lane < nb && q < nb) ? trow[q] : T(0);
        }
        // inverse of the lower-triangular diagonal block of L1: lane j solves L x = e_j
        if (lane < NB) {
            T x[NB];
#pragma unroll
            for (int q = 0; q < NB; ++q) x[q] = T(0);
            if (lane < nb) {
#pragma unroll
                for (int i = 0; i < NB; ++i) {
                    if (i < nb) {
                        T sum = (i == lane) ? T(1) : T(0);
#pragma unroll
                        for (int q = 0; q < NB; ++q)
                            if (q < i) sum -= L1[(b0 + i) + (int64_t)(b0 + q) * nk] * x[q];
                        x[i] = sum / L1[(b0 + i) + (int64_t)(b0 + i) * nk];
                    }
                }
            }
#pragma unroll
            for (int q = 0; q < NB; ++q) Li[q + lane * NB] = x[q];
        }
    }
}

// Root: merge the two children (or take the leaf D) and LU-factor with partial pivoting.
template <typename T>
__global__ __launch_bounds__(1024) void k_ulv_root(const FInst<T>* __restrict__ insts,
                                                   const WorkItem* __restrict__ work) {
    const WorkItem wi = work[blockIdx.x];
    const FInst<T>& I = insts[wi.inst];
    const DevNode nd = I.nodes[wi.node];
    const FNode f = I.fn[wi.node];
    const int m = f.m;
    T* A = I.fac + f.LU;
    int32_t* piv = I.ipiv + f.piv;
    const int tid = threadIdx.x;
    if (nd.is_leaf) {
        for (int e = tid; e < m * m; e += blockDim.x) A[e] = gen<T>(I.blob, nd.D, e);
    } else {
        const ChildRed<T> a = child_red(I, nd.left);
        const ChildRed<T> b = child_red(I, nd.right);
        merge_products(I, nd, f, a, b);
        __threadfence_block();
        __syncthreads();
        const T* P12 = I.fac + f.P12;
        const T* P21 = I.fac + f.P21;
        for (int e = tid; e < m * m; e += blockDim.x) A[e] = merged_D(a, b, P12, P21, e % m, e / m);
    }
    __threadfence_block();
    __syncthreads();
    __shared__ int s_piv;
    __shared__ T s_val;
    const int wave = tid / 64, lane = tid % 64;
    int info = 0;
    for (int i = 0; i < m; ++i) {
        if (wave == 0) {
            // idamax over A[i:m, i] (first index of the largest magnitude)
            T best = T(-1);
            int bi = m;
            for (int rr = i + lane; rr < m; rr += 64) {
                const T v = fabs(A[rr + (int64_t)i * m]);
                if (v > best) { best = v; bi = rr; }
            }
            for (int o = 1; o < 64; o <<= 1) {
                const T ob = __shfl_xor(best, o, 64);
                const int oi = __shfl_xor(bi, o, 64);
                if (ob > best || (ob == best && oi < bi)) { best = ob; bi = oi; }
            }
            if (lane == 0) {
                s_piv = bi;
                s_val = A[bi + (int64_t)i * m];
                piv[i] = bi;
            }
        }
        __syncthreads();
        const int p = s_piv;
        const T pv = s_val;
        if (pv == T(0)) {
            if (info == 0) info = i + 1;
        } else {
            if (p != i)
                for (int c = tid; c < m; c += blockDim.x) {
                    const T t = A[i + (int64_t)c * m];
                    A[i + (int64_t)c * m] = A[p + (int64_t)c * m];
                    A[p + (int64_t)c * m] = t;
                }
            __syncthreads();
            const T rcp = T(1) / pv;
            for (int rr = i + 1 + tid; rr < m; rr += blockDim.x) A[rr + (int64_t)i * m] *= rcp;
        }
        __syncthreads();
        const int mm = m - i - 1;
        for (int e = tid; e < mm * mm; e += blockDim.x) {
            const int rr = i + 1 + e % mm, c = i + 1 + e / mm;
            A[rr + (int64_t)c * m] -= A[rr + (int64_t)i * m] * A[i + (int64_t)c * m];
        }
        __syncthreads();
    }
    if (tid == 0) I.status[0] = info;
}

// ---------------------------------------------------------------------------
// ULV solve
// ---------------------------------------------------------------------------
// Assemble the node's right-hand side b (m x nt) in LDS:
//   leaf:   B_seg;  branch: [b1 - P12 u2; b2 - P21 u1]  (ulvfactor.jl:74)
template <typename TF, typename TV>
__device__ void solve_form_b(const SInst<TF, TV>& I, const DevNode& nd, const FNode& f, int j0, int nt, TV* bv,
                             TV* tmp) {
    const int m = f.m;
    const int nrhs = I.nrhs;
    if (nd.is_leaf) {
        const double* Bs = I.B + nd.row_off + (int64_t)j0 * I.ldb;
        for (int e = threadIdx.x; e < m * nt; e += blockDim.x) {
            const int r = e % m, j = e / m;
            bv[e] = (TV)Bs[r + (int64_t)j * I.ldb];
        }
        __syncthreads();
        return;
    }
    const FNode fl = I.fn[nd.left], fr = I.fn[nd.right];
    const DevNode dl = I.nodes[nd.left], dr = I.nodes[nd.right];
    const TV* b1 = I.wb + fl.brow * nrhs + (int64_t)j0 * fl.mr;
    const TV* b2 = I.wb + fr.brow * nrhs + (int64_t)j0 * fr.mr;
    const TV* u1 = I.wu + fl.urow * nrhs + (int64_t)j0 * dl.kv;
    const TV* u2 = I.wu + fr.urow * nrhs + (int64_t)j0 * dr.kv;
    for (int e = threadIdx.x; e < m * nt; e += blockDim.x) {
        const int r = e % m, j = e / m;
        bv[e] = (r < fl.mr) ? b1[r + (int64_t)j * fl.mr] : b2[(r - fl.mr) + (int64_t)j * fr.mr];
    }
    __syncthreads();
    // tmp = P12 u2 (mr1 x nt), then P21 u1 (mr2 x nt) below it
    blk_gemm<false>(fl.mr, nt, dr.kv, I.fac + f.P12, fl.mr, u2, dr.kv, tmp, m, TV(1), false);
    blk_gemm<false>(fr.mr, nt, dl.kv, I.fac + f.P21, fr.mr, u1, dl.kv, tmp + fl.mr, m, TV(1), false);
    __syncthreads();
    for (int e = threadIdx.x; e < m * nt; e += blockDim.x) bv[e] -= tmp[e];
    __syncthreads();
}

// u = Vt zloc + W1' u1 + W2' u2  (ulvfactor.jl:51, 89) -> global u workspace
template <typename TF, typename TV>
__device__ void solve_u(const SInst<TF, TV>& I, const DevNode& nd, const FNode& f, int j0, int nt, const TV* zl,
                        TV* ug) {
    const int kv = nd.kv;
    const int nrhs = I.nrhs;
    if (f.nk > 0) blk_gemm<false>(kv, nt, f.nk, I.fac + f.Vt, kv, zl, f.nk, ug, kv, TV(1), false);
    else
        for (int e = threadIdx.x; e < kv * nt; e += blockDim.x) ug[e] = TV(0);
    if (!nd.is_leaf) {
        const FNode fl = I.fn[nd.left], fr = I.fn[nd.right];
        const DevNode dl = I.nodes[nd.left], dr = I.nodes[nd.right];
        const TV* u1 = I.wu + fl.urow * nrhs + (int64_t)j0 * dl.kv;
        const TV* u2 = I.wu + fr.urow * nrhs + (int64_t)j0 * dr.kv;
        __syncthreads();
        blk_gemm<true>(kv, nt, dl.kv, I.blob + dl.W, dl.kv, u1, dl.kv, ug, kv, TV(1), true);
        __syncthreads();
        blk_gemm<true>(kv, nt, dr.kv, I.blob + dr.W, dr.kv, u2, dr.kv, ug, kv, TV(1), true);
    }
}

template <typename TF, typename TV>
__global__ __launch_bounds__(SWEEP_THREADS) void k_ulv_sup(const SInst<TF, TV>* __restrict__ insts,
                                                          const WorkItem* __restrict__ work) {
    const WorkItem wi = work[blockIdx.x];
    const SInst<TF, TV>& I = insts[wi.inst];
    const int j0 = blockIdx.y * RHS_TILE;
    if (j0 >= I.nrhs) return;
    const int nt = min(RHS_TILE, I.nrhs - j0);
    const DevNode nd = I.nodes[wi.node];
    const FNode f = I.fn[wi.node];
    const int m = f.m, k = f.k, nk = f.nk;
    const int nrhs = I.nrhs;
    extern __shared__ unsigned char lds_raw[];
    const int mp = m > NB ? m : NB;
    TV* bv = reinterpret_cast<TV*>(lds_raw);  // m x nt
    TV* tmp = bv + mp * nt;                   // max(m, NB) x nt
    TV* t2 = tmp + mp * nt;                   // max(m, NB) x nt
    solve_form_b(I, nd, f, j0, nt, bv, tmp);
    TV* wbn = I.wb + f.brow * nrhs + (int64_t)j0 * f.mr;
    TV* ug = I.wu + f.urow * nrhs + (int64_t)j0 * nd.kv;
    if (f.pass) {
        for (int e = threadIdx.x; e < m * nt; e += blockDim.x) wbn[e] = bv[e];
        solve_u(I, nd, f, j0, nt, (const TV*)nullptr, ug);
        return;
    }
    // b <- Q_ql' b = b - Y (T (Y' b))   (ulvfactor.jl:42)
    if (k > 0) {
        blk_gemm<true>(k, nt, m, I.fac + f.Yql, m, bv, m, tmp, k, TV(1), false);
        __syncthreads();
        // t2 = T tmp, T upper triangular k x k
        blk_gemm<false>(k, nt, k, I.fac + f.Tql, k, tmp, k, t2, k, TV(1), false);
        __syncthreads();
        blk_gemm<false>(m, nt, k, I.fac + f.Yql, m, t2, k, bv, m, TV(-1), true);
        __syncthreads();
    }
    // zloc = L1 \ b[0:nk]  (ulvfactor.jl:48), blocked with inverted diagonal blocks
    TV* zl = tmp;  // nk x nt
    for (int bb = 0; bb < f.nblk; ++bb) {
        const int b0 = bb * NB, nb = min(NB, nk - b0);
        // t2[0:nb] = b[b0:b0+nb] - L1[b0:b0+nb, 0:b0] zl[0:b0]
        for (int e = threadIdx.x; e < nb * nt; e += blockDim.x) {
            const int i = e % nb, j = e / nb;
            TV sacc = bv[b0 + i + j * m];
            const TF* Lr = I.fac + f.L1 + b0 + i;
            for (int q = 0; q < b0; ++q) sacc -= (TV)Lr[(int64_t)q * nk] * zl[q + j * nk];
            t2[i + j * NB] = sacc;
        }
        __syncthreads();
        const TF* Li = I.fac + f.L1inv + (int64_t)bb * NB * NB;
        for (int e = threadIdx.x; e < nb * nt; e += blockDim.x) {
            const int i = e % nb, j = e / nb;
            TV sacc = TV(0);
            for (int q = 0; q <= i; ++q) sacc += (TV)Li[i + q * NB] * t2[q + j * NB];
            zl[b0 + i + j * nk] = sacc;
        }
        __syncthreads();
    }
    // b_r = b[nk:m] - L2a zloc  (ulvfactor.jl:49) -> parent's input
    for (int e = threadIdx.x; e < k * nt; e += blockDim.x) {
        const int i = e % k, j = e / k;
        TV sacc = bv[nk + i + j * m];
        const TF* Lr = I.fac + f.L2a + i;
        for (int q = 0; q < nk; ++q) sacc -= (TV)Lr[(int64_t)q * k] * zl[q + j * nk];
        wbn[i + j * f.mr] = sacc;
    }
    // store zloc into the node's z vector (rows 0..nk-1)
    TV* zn = I.wz + f.zrow * nrhs + (int64_t)j0 * f.n;
    for (int e = threadIdx.x; e < nk * nt; e += blockDim.x) {
        const int i = e % nk, j = e / nk;
        zn[i + j * f.n] = zl[e];
    }
    solve_u(I, nd, f, j0, nt, zl, ug);
}

// Root: b assembled like a branch, z = LU \ b
template <typename TF, typename TV>
__global__ __launch_bounds__(SWEEP_THREADS) void k_ulv_sroot(const SInst<TF, TV>* __restrict__ insts,
                                                            const WorkItem* __restrict__ work) {
    const WorkItem wi = work[blockIdx.x];
    const SInst<TF, TV>& I = insts[wi.inst];
    const int j0 = blockIdx.y * RHS_TILE;
    if (j0 >= I.nrhs) return;
    const int nt = min(RHS_TILE, I.nrhs - j0);
    const DevNode nd = I.nodes[wi.node];
    const FNode f = I.fn[wi.node];
    const int m = f.m;
    extern __shared__ unsigned char lds_raw[];
    TV* bv = reinterpret_cast<TV*>(lds_raw);
    TV* tmp = bv + m * nt;
    solve_form_b(I, nd, f, j0, nt, bv, tmp);
    const TF* A = I.fac + f.LU;
    const int32_t* piv = I.ipiv + f.piv;
    // row interchanges, then L (unit) and U solves, one thread per column
    for (int j = threadIdx.x; j < nt; j += blockDim.x) {
        TV* x = bv + j * m;
        for (int i = 0; i < m; ++i) {
            const int p = piv[i];
            if (p != i) { const TV t = x[i]; x[i] = x[p]; x[p] = t; }
        }
        for (int i = 0; i < m; ++i) {
            const TV xi = x[i];
            for (int rr = i + 1; rr < m; ++rr) x[rr] -= (TV)A[rr + (int64_t)i * m] * xi;
        }
        for (int i = m - 1; i >= 0; --i) {
            x[i] /= (TV)A[i + (int64_t)i * m];
            const TV xi = x[i];
            for (int rr = 0; rr < i; ++rr) x[rr] -= (TV)A[rr + (int64_t)i * m] * xi;
        }
    }
    __syncthreads();
    if (nd.is_leaf) {
        double* Xs = I.X + nd.col_off + (int64_t)j0 * I.ldx;
        for (int e = threadIdx.x; e < m * nt; e += blockDim.x) Xs[e % m + (int64_t)(e / m) * I.ldx] = (double)bv[e];
    } else {
        TV* zn = I.wz + f.zrow * I.nrhs + (int64_t)j0 * f.n;
        for (int e = threadIdx.x; e < m * nt; e += blockDim.x) zn[e] = bv[e];
    }
}

// Down-sweep: z = [zloc; z_from_parent], z <- Q_lq' z  (ulvfactor.jl:98-107)
template <typename TF, typename TV>
__global__ __launch_bounds__(SWEEP_THREADS) void k_ulv_sdown(const SInst<TF, TV>* __restrict__ insts,
                                                            const WorkItem* __restrict__ work) {
    const WorkItem wi = work[blockIdx.x];
    const SInst<TF, TV>& I = insts[wi.inst];
    const int j0 = blockIdx.y * RHS_TILE;
    if (j0 >= I.nrhs) return;
    const int nt = min(RHS_TILE, I.nrhs - j0);
    const DevNode nd = I.nodes[wi.node];
    const FNode f = I.fn[wi.node];
    const int n = f.n, nk = f.nk;
    const int nrhs = I.nrhs;
    extern __shared__ unsigned char lds_raw[];
    TV* z = reinterpret_cast<TV*>(lds_raw);  // n x nt
    TV* t1 = z + n * nt;                      // NB x nt
    TV* t2 = t1 + NB * nt;                    // NB x nt
    const FNode fp = I.fn[nd.parent];
    const int off = nd.is_left ? 0 : I.fn[I.nodes[nd.parent].left].nr;
    const TV* zp = I.wz + fp.zrow * nrhs + (int64_t)j0 * fp.n;
    const TV* zn = I.wz + f.zrow * nrhs + (int64_t)j0 * n;
    for (int e = threadIdx.x; e < n * nt; e += blockDim.x) {
        const int r = e % n, j = e / n;
        z[e] = (r < nk) ? zn[r + j * n] : zp[off + (r - nk) + j * fp.n];
    }
    __syncthreads();
    // Q' = H(1)...H(nk) = P_0 ... P_{B-1}; apply P_{B-1} first
    for (int bb = f.nblk - 1; bb >= 0; --bb) {
        const int b0 = bb * NB, nb = min(NB, nk - b0);
        const TF* Yb = I.fac + f.Ylq + (int64_t)b0 * n;  // n x nb
        const TF* Tb = I.fac + f.Tlq + (int64_t)bb * NB * NB;
        blk_gemm<true>(nb, nt, n - b0, Yb + b0, n, z + b0, n, t1, NB, TV(1), false);
        __syncthreads();
        blk_gemm<false>(nb, nt, nb, Tb, NB, t1, NB, t2, NB, TV(1), false);
        __syncthreads();
        blk_gemm<false>(n - b0, nt, nb, Yb + b0, n, t2, NB, z + b0, n, TV(-1), true);
        __syncthreads();
    }
    if (nd.is_leaf) {
        double* Xs = I.X + nd.col_off + (int64_t)j0 * I.ldx;
        for (int e = threadIdx.x; e < n * nt; e += blockDim.x) Xs[e % n + (int64_t)(e / n) * I.ldx] = (double)z[e];
    } else {
        TV* zo = I.wz + f.zrow * nrhs + (int64_t)j0 * n;
        for (int e = threadIdx.x; e < n * nt; e += blockDim.x) zo[e] = z[e];
    }
}

// ---------------------------------------------------------------------------
// vector helpers for refinement
// ---------------------------------------------------------------------------
// R = B - R  (R holds A*X on entry), per column
__global__ void k_residual(const double* __restrict__ B, int64_t ldb, double* __restrict__ Rr, int64_t ldr, int64_t n,
                           int nrhs) {
    const int64_t total = n * nrhs;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = e % n, j = e / n;
        Rr[i + j * ldr] = B[i + j * ldb] - Rr[i + j * ldr];
    }
}

// X += D
__global__ void k_axpy(double* __restrict__ X, int64_t ldx, const double* __restrict__ Dd, int64_t ldd, int64_t n,
                       int nrhs) {
    const int64_t total = n * nrhs;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = e % n, j = e / n;
        X[i + j * ldx] += Dd[i + j * ldd];
    }
}

// per-column sum of squares of n x nrhs (one block per column)
__global__ void k_colnorm2(const double* __restrict__ A, int64_t lda, int64_t n, double* __restrict__ out) {
    const int j = blockIdx.x;
    double s = 0.0;
    for (int64_t i = threadIdx.x; i < n; i += blockDim.x) {
        const double v = A[i + j * lda];
        s += v * v;
    }
    __shared__ double red[16];
    s = wave_reduce(s);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += red[w];
        out[j] = t;
    }
}

// ---------------------------------------------------------------------------
// host-side launchers
// ---------------------------------------------------------------------------
static inline int rhs_tiles(int nrhs) { return (nrhs + RHS_TILE - 1) / RHS_TILE; }

hipError_t launch_mv_up(const MvInst* insts, const WorkItem* work, int nwork, int max_nrhs, size_t lds,
                        hipStream_t st) {
    if (nwork <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_mv_up, dim3(nwork, rhs_tiles(max_nrhs)), dim3(SWEEP_THREADS), lds, st, insts, work);
    return hipGetLastError();
}
hipError_t launch_mv_down(const MvInst* insts, const WorkItem* work, int nwork, int max_nrhs, hipStream_t st) {
    if (nwork <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_mv_down, dim3(nwork, rhs_tiles(max_nrhs)), dim3(SWEEP_THREADS), 0, st, insts, work);
    return hipGetLastError();
}
hipError_t launch_mv_leaf(const MvInst* insts, const WorkItem* work, int nwork, int max_nrhs, size_t lds,
                          hipStream_t st) {
    if (nwork <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_mv_leaf, dim3(nwork, rhs_tiles(max_nrhs)), dim3(SWEEP_THREADS), lds, st, insts, work);
    return hipGetLastError();
}

template <typename T>
static hipError_t set_lds_attr(const void* fn, size_t lds) {
    if (lds > 65536) return hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    return hipSuccess;
}

template <typename T>
hipError_t launch_ulv_factor(const FInst<T>* insts, const WorkItem* work, int nwork, size_t lds, hipStream_t st) {
    if (nwork <= 0) return hipSuccess;
    const void* fn = reinterpret_cast<const void*>(&k_ulv_factor<T, FACT_CPT>);
    hipError_t e = set_lds_attr<T>(fn, lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((k_ulv_factor<T, FACT_CPT>), dim3(nwork), dim3(FACT_THREADS), lds, st, insts, work);
    return hipGetLastError();
}
template <typename T>
hipError_t launch_ulv_wy(const FInst<T>* insts, const WorkItem* work, int nwork, hipStream_t st) {
    if (nwork <= 0) return hipSuccess;
    hipLaunchKernelGGL((k_ulv_wy<T>), dim3(nwork), dim3(256), 0, st, insts, work);
    return hipGetLastError();
}
template <typename T>
hipError_t launch_ulv_root(const FInst<T>* insts, const WorkItem* work, int nwork, hipStream_t st) {
    if (nwork <= 0) return hipSuccess;
    hipLaunchKernelGGL((k_ulv_root<T>), dim3(nwork), dim3(1024), 0, st, insts, work);
    return hipGetLastError();
}
template <typename TF, typename TV>
hipError_t launch_ulv_sup(const SInst<TF, TV>* insts, const WorkItem* work, int nwork, int max_nrhs, size_t lds,
                          hipStream_t st) {
    if (nwork <= 0) return hipSuccess;
    const void* fn = reinterpret_cast<const void*>(&k_ulv_sup<TF, TV>);
    hipError_t e = set_lds_attr<TV>(fn, lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((k_ulv_sup<TF, TV>), dim3(nwork, rhs_tiles(max_nrhs)), dim3(SWEEP_THREADS), lds, st, insts,
                       work);
    return hipGetLastError();
}
template <typename TF, typename TV>
hipError_t launch_ulv_sroot(const SInst<TF, TV>* insts, const WorkItem* work, int nwork, int max_nrhs, size_t lds,
                            hipStream_t st) {
    if (nwork <= 0) return hipSuccess;
    const void* fn = reinterpret_cast<const void*>(&k_ulv_sroot<TF, TV>);
    hipError_t e = set_lds_attr<TV>(fn, lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((k_ulv_sroot<TF, TV>), dim3(nwork, rhs_tiles(max_nrhs)), dim3(SWEEP_THREADS), lds, st,
                       insts, work);
    return hipGetLastError();
}
template <typename TF, typename TV>
hipError_t launch_ulv_sdown(const SInst<TF, TV>* insts, const WorkItem* work, int nwork, int max_nrhs, size_t lds,
                            hipStream_t st) {
    if (nwork <= 0) return hipSuccess;
    const void* fn = reinterpret_cast<const void*>(&k_ulv_sdown<TF, TV>);
    hipError_t e = set_lds_attr<TV>(fn, lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((k_ulv_sdown<TF, TV>), dim3(nwork, rhs_tiles(max_nrhs)), dim3(SWEEP_THREADS), lds, st,
                       insts, work);
    return hipGetLastError();
}

hipError_t launch_residual(const double* B, int64_t ldb, double* R, int64_t ldr, int64_t n, int nrhs, hipStream_t st) {
    const int64_t total = n * nrhs;
    int blocks = (int)((total + 255) / 256);
    if (blocks > 4096) blocks = 4096;
    if (blocks < 1) return hipSuccess;
    hipLaunchKernelGGL(k_residual, dim3(blocks), dim3(256), 0, st, B, ldb, R, ldr, n, nrhs);
    return hipGetLastError();
}
hipError_t launch_axpy(double* X, int64_t ldx, const double* D, int64_t ldd, int64_t n, int nrhs, hipStream_t st) {
    const int64_t total = n * nrhs;
    int blocks = (int)((total + 255) / 256);
    if (blocks > 4096) blocks = 4096;
    if (blocks < 1) return hipSuccess;
    hipLaunchKernelGGL(k_axpy, dim3(blocks), dim3(256), 0, st, X, ldx, D, ldd, n, nrhs);
    return hipGetLastError();
}
hipError_t launch_colnorm2(const double* A, int64_t lda, int64_t n, int nrhs, double* out, hipStream_t st) {
    if (nrhs <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_colnorm2, dim3(nrhs), dim3(256), 0, st, A, lda, n, out);
    return hipGetLastError();
}

// explicit instantiations
template hipError_t launch_ulv_factor<double>(const FInst<double>*, const WorkItem*, int, size_t, hipStream_t);
template hipError_t launch_ulv_factor<float>(const FInst<float>*, const WorkItem*, int, size_t, hipStream_t);
template hipError_t launch_ulv_wy<double>(const FInst<double>*, const WorkItem*, int, hipStream_t);
template hipError_t launch_ulv_wy<float>(const FInst<float>*, const WorkItem*, int, hipStream_t);
template hipError_t launch_ulv_root<double>(const FInst<double>*, const WorkItem*, int, hipStream_t);
template hipError_t launch_ulv_root<float>(const FInst<float>*, const WorkItem*, int, hipStream_t);
template hipError_t launch_ulv_sup<double, double>(const SInst<double, double>*, const WorkItem*, int, int, size_t,
                                                   hipStream_t);
template hipError_t launch_ulv_sup<float, float>(const SInst<float, float>*, const WorkItem*, int, int, size_t,
                                                 hipStream_t);
template hipError_t launch_ulv_sroot<double, double>(const SInst<double, double>*, const WorkItem*, int, int, size_t,
                                                     hipStream_t);
template hipError_t launch_ulv_sroot<float, float>(const SInst<float, float>*, const WorkItem*, int, int, size_t,
                                                   hipStream_t);
template hipError_t launch_ulv_sdown<double, double>(const SInst<double, double>*, const WorkItem*, int, int, size_t,
                                                     hipStream_t);
template hipError_t launch_ulv_sdown<float, float>(const SInst<float, float>*, const WorkItem*, int, int, size_t,
                                                   hipStream_t);

}  // namespace hss
