// hss_kernels.hip -- gfx950 kernels of libhss_hip.so.
//
// Matvec (reference src/matmul.jl:18-62) as level-synchronous sweeps over per-node
// operators, one MFMA product per node and step:
//   k_mv_s<0> / k_mv_t<0>   Z = V' B_seg (leaf) | [W1' W2'] [Z1; Z2]     (_matmatup, matmul.jl:32-42)
//   k_mv_s<1> / k_mv_t<1>   F = [B_sibling R] [Z_sibling; F_parent]      (_matmatdown!, matmul.jl:51-57)
//   k_mv_s<2> / k_mv_t<2>   C_seg = alpha [D U] [B_seg; F] + beta C_seg  (matmul.jl:46-47), or with
//                           the leaf's down step folded in: alpha [D, U B_sib, U R] [B_seg; Z_sib; F_parent]
//   k_cut                   the steps above one depth as one dense product F_cut = T Z_cut (tree cut)
//
// ULV (reference src/ulvfactor.jl:10-107), split into a resident factor and a solve:
//   k_ulv_factor  per node: merge of the children (ulvfactor.jl:74-79), QL of U,
//                 Q'D, LQ of the top rows with the reflectors applied to the
//                 bottom rows and to V' (ulvfactor.jl:39-54)
//   k_ulv_wy      compact-WY blocks of the LQ reflectors
//   k_ulv_root    merge + LU with partial pivoting of the root block (ulvfactor.jl:82-86)
//   k_ulv_rinv    the root block's inverse and merge operator
//   k_ulv_gform / k_ulv_qform   explicit per-node solve operators (Gz, Gr, Qt), formed
//                 by running the node's reduction on identity columns
//   k_ulv_gup*    solve up-sweep: zloc = Gz x, [b; u] = Gr [x; zloc] (ulvfactor.jl:42,48-51,74,89)
//   k_ulv_groot*  root solve (ulvfactor.jl:83)
//   k_ulv_gdown*  solve down-sweep: z = Qt [zloc; z_parent] (ulvfactor.jl:98-107)
// Sweep kernels come in three forms: _s with operator and input staged in LDS,
// _s<.., false> with the operator streamed by the waves, and chunked tile kernels.
//
// Every kernel takes a list of (instance, node) work items, so one launch can
// cover the same tree level of many HSS matrices.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>
#include <math.h>

#include <algorithm>
#include <utility>
#include <stdint.h>

#define HSS_KERNELS_TU
#include "hss_internal.h"
#include "hss_launch.h"

namespace hss {

// Every kernel of this file is launched here.  Inside a timed scope (g_prof_launch set) the
// scope's first launch also records the scope's kernel events at its own start and end.
template <typename... KArgs, typename... Args>
static inline void klaunch(void (*k)(KArgs...), dim3 grid, dim3 block, size_t lds, hipStream_t st, Args... args) {
    static_assert(sizeof...(KArgs) == sizeof...(Args), "kernel argument count");
    ProfLaunch* p = g_prof_launch;
    if (p && p->launches++ == 0) {
        hipExtLaunchKernelGGL(k, grid, block, (std::uint32_t)lds, st, p->k0, p->k1, 0u, static_cast<KArgs>(args)...);
        return;
    }
    hipLaunchKernelGGL(k, grid, block, lds, st, static_cast<KArgs>(args)...);
}

// ---------------------------------------------------------------------------
// small helpers
// ---------------------------------------------------------------------------
// Device pointers read from argument tables are generic to the compiler, which then
// emits flat loads and stores.  Flat operations also count against the LDS wait
// counter, so a wait for an LDS read would wait for every outstanding global
// prefetch.  Cast to the global address space, they become global_load / global_store.
// (The host compilation pass only parses the device code; there the cast is a no-op.)
#if defined(__HIP_DEVICE_COMPILE__)
template <typename T>
using gptr = __attribute__((address_space(1))) T*;
#else
template <typename T>
using gptr = T*;
#endif
template <typename T>
__device__ __forceinline__ gptr<T> gbl(T* p) {
    return (gptr<T>)p;
}
#if defined(__HIP_DEVICE_COMPILE__)
// Raw buffer descriptor over [p, p + bytes) (gfx9 format word 0x00020000) and loads through
// it: the descriptor and a per-instruction offset live in scalar registers.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const void* p, int bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, bytes, 0x00020000);
}
template <typename T>
__device__ __forceinline__ T buf_load(__amdgpu_buffer_rsrc_t r, int voff, int soff);
template <>
__device__ __forceinline__ double buf_load<double>(__amdgpu_buffer_rsrc_t r, int voff, int soff) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
}
template <>
__device__ __forceinline__ float buf_load<float>(__amdgpu_buffer_rsrc_t r, int voff, int soff) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
}
#else
struct HostRsrc {};
__device__ inline HostRsrc buf_rsrc(const void*, int) { return {}; }
template <typename T>
__device__ inline T buf_load(HostRsrc, int, int) {
    return T(0);
}
#endif
template <typename T>
__device__ __forceinline__ T shfl_xor(T v, int mask) {
    return __shfl_xor(v, mask, 64);
}

// v from another lane by a DPP lane move (no LDS round trip); 64-bit values move as two
// 32-bit halves.  CTRL: 0xB1 = lane ^ 1, 0x4E = lane ^ 2 (quad permutes), 0x141 = mirror
// within each half row of 8 lanes, 0x140 = mirror within each row of 16.
template <int CTRL>
__device__ __forceinline__ double dpp_mov(double v) {
#if defined(__HIP_DEVICE_COMPILE__)
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xF, 0xF, false);
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
#else
    return v;
#endif
}
template <int CTRL>
__device__ __forceinline__ float dpp_mov(float v) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
#else
    return v;
#endif
}

// sum over the S adjacent lanes of a row group (S a power of two <= 64, groups aligned to
// S, all lanes of a group active): the first two steps are quad permutes
template <typename T>
__device__ __forceinline__ T row_reduce(T v, int S) {
    if (S > 1) v += dpp_mov<0xB1>(v);
    if (S > 2) v += dpp_mov<0x4E>(v);
    for (int o = 4; o < S; o <<= 1) v += shfl_xor(v, o);
    return v;
}

// sum over the wave (all lanes active), as the butterfly lane ^ 1, 2, 4, 8, 16, 32: once
// the values are equal within quads (then within half rows), the half-row mirror (then the
// row mirror) pairs lanes exactly as ^ 4 (then ^ 8) does, so every lane ends with the same
// sum, added in the same order
template <typename T>
__device__ __forceinline__ T wave_reduce(T v) {
    v += dpp_mov<0xB1>(v);
    v += dpp_mov<0x4E>(v);
    v += dpp_mov<0x141>(v);
    v += dpp_mov<0x140>(v);
    v += shfl_xor(v, 16);
    v += shfl_xor(v, 32);
    return v;
}

// Sum over the four 16-lane rows of a wave (lanes l, l^16, l^32, l^48), on the VALU:
// v_permlane16_swap then v_permlane32_swap (gfx950), each adding the value the swap brings
// to the lane's own; every lane ends with the same sum, added in the same order.
__device__ __forceinline__ float rows_sum(float v) {
#if defined(__HIP_DEVICE_COMPILE__)
    const int x = __float_as_int(v);
    auto a = __builtin_amdgcn_permlane16_swap(x, x, false, false);
    v = __int_as_float(a[0]) + __int_as_float(a[1]);
    const int y = __float_as_int(v);
    auto b = __builtin_amdgcn_permlane32_swap(y, y, false, false);
    v = __int_as_float(b[0]) + __int_as_float(b[1]);
#endif
    return v;
}
__device__ __forceinline__ double rows_sum(double v) {
#if defined(__HIP_DEVICE_COMPILE__)
    auto pair = [](int lo, int hi) {
        return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
    };
    long long bits = __double_as_longlong(v);
    auto a = __builtin_amdgcn_permlane16_swap((int)bits, (int)bits, false, false);
    auto c = __builtin_amdgcn_permlane16_swap((int)(bits >> 32), (int)(bits >> 32), false, false);
    v = pair(a[0], c[0]) + pair(a[1], c[1]);
    bits = __double_as_longlong(v);
    a = __builtin_amdgcn_permlane32_swap((int)bits, (int)bits, false, false);
    c = __builtin_amdgcn_permlane32_swap((int)(bits >> 32), (int)(bits >> 32), false, false);
    v = pair(a[0], c[0]) + pair(a[1], c[1]);
#endif
    return v;
}
// lane l's value (l uniform across the wave) to every lane (v_readlane)
__device__ __forceinline__ float lane_read(float v, int l) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
__device__ __forceinline__ double lane_read(double v, int l) {
    const long long b = __double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)b, l);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
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

// ---------------------------------------------------------------------------
// workgroup GEMM on small blocks
// ---------------------------------------------------------------------------
// Right-hand-side tiles live in LDS row-major: element (r, j) at X[r * NTP + j],
// NTP = RHS_TILE, so the CB columns a thread needs are contiguous.
constexpr int NTP = RHS_TILE;

// out(i, j, v) for v = sum_l opA(i, l) X(l, j), i < M, j < N.
//   op(A) = A (M x K) when !TA, A^T with A stored K x M when TA; A column-major
//   in global memory with leading dimension lda.
// Threads form a (TR x TC) grid, TC = ceil(N / CB) (column groups fastest), each
// owning an RB x CB register tile; the TC lanes of one thread-row read the same
// A element in the same instruction, so A streams from memory once per
// workgroup.  Outputs with j >= N are computed from padding and never stored.
template <int RB, int CB, bool TA, typename TAe, typename T, typename Store>
__device__ __forceinline__ void wg_gemm(int M, int N, int K, const TAe* __restrict__ A, int64_t lda, const T* X,
                                        Store store) {
    const int TC = (N + CB - 1) / CB;
    const int TR = blockDim.x / TC;
    const int tc = threadIdx.x % TC, tr = threadIdx.x / TC;
    if (tr >= TR || M <= 0 || N <= 0) return;
    const int j0 = tc * CB;
    for (int i0 = tr * RB; i0 < M; i0 += TR * RB) {
        T acc[RB][CB];
#pragma unroll
        for (int r = 0; r < RB; ++r)
#pragma unroll
            for (int c = 0; c < CB; ++c) acc[r][c] = T(0);
#pragma unroll 4
        for (int l = 0; l < K; ++l) {
            T a[RB];
#pragma unroll
            for (int r = 0; r < RB; ++r) {
                const int i = i0 + r;
                a[r] = (i < M) ? (T)(TA ? A[l + (int64_t)i * lda] : A[i + (int64_t)l * lda]) : T(0);
            }
            const T* xr = X + l * NTP + j0;
            T x[CB];
#pragma unroll
            for (int c = 0; c < CB; ++c) x[c] = xr[c];
#pragma unroll
            for (int r = 0; r < RB; ++r)
#pragma unroll
                for (int c = 0; c < CB; ++c) acc[r][c] += a[r] * x[c];
        }
#pragma unroll
        for (int r = 0; r < RB; ++r)
#pragma unroll
            for (int c = 0; c < CB; ++c)
                if (i0 + r < M && j0 + c < N) store(i0 + r, j0 + c, acc[r][c]);
    }
}

// Tile shape by the row count, so that short products (the 16-row reflector and
// substitution blocks) still keep every thread busy: 1 x 2 per thread up to 16 rows, 1 x 4
// up to 32, 1 x 8 up to 64, else 2 x 8 (256 threads cover 128 x 32).  Every output sums over
// K in the same order whatever the shape.
template <bool TA, typename TAe, typename T, typename Store>
__device__ __forceinline__ void gemm(int M, int N, int K, const TAe* __restrict__ A, int64_t lda, const T* X,
                                     Store store) {
    if (M <= 16) wg_gemm<1, 2, TA>(M, N, K, A, lda, X, store);
    else if (M <= 32) wg_gemm<1, 4, TA>(M, N, K, A, lda, X, store);
    else if (M <= 64) wg_gemm<1, 8, TA>(M, N, K, A, lda, X, store);
    else wg_gemm<2, 8, TA>(M, N, K, A, lda, X, store);
}

// global column-major (ld) -> LDS row-major rows [r0, r0 + rows)
// for e in [0, n): st(e, ld(e)), with U loads per thread issued before their stores
// (a load and a store through generic pointers may alias, so a plain loop would wait
// for every load in turn)
template <int U, typename T, typename Ld, typename St>
__device__ __forceinline__ void batched(int n, Ld ld, St st) {
    const int bd = blockDim.x;
    for (int e0 = threadIdx.x; e0 < n; e0 += U * bd) {
        T r[U];
#pragma unroll
        for (int q = 0; q < U; ++q) {
            const int e = e0 + q * bd;
            r[q] = e < n ? ld(e) : T{};
        }
#pragma unroll
        for (int q = 0; q < U; ++q) {
            const int e = e0 + q * bd;
            if (e < n) st(e, r[q]);
        }
    }
}

template <typename Tsrc, typename T>
__device__ __forceinline__ void stage(T* dst, int r0, const Tsrc* src, int64_t ld, int rows, int nt) {
    batched<8, T>(
        rows * nt, [&](int e) { return (T)src[(e % rows) + (int64_t)(e / rows) * ld]; },
        [&](int e, T v) { dst[(r0 + e % rows) * NTP + e / rows] = v; });
}

// dst[0, n) = src[0, n) with direct global -> LDS loads (global_load_lds_dwordx4: lane l
// of a wave writes 16 bytes at the wave's LDS base + 16 l), so the whole copy is in
// flight at once without passing through registers.  src and dst 16-byte aligned; may
// read up to 15 bytes past n.  The data is in LDS after s_waitcnt vmcnt(0) and, for
// other waves, a barrier.
template <typename TF>
__device__ __forceinline__ void lds_dma(TF* dst, const TF* __restrict__ src, int64_t n) {
    const int nv = (int)((n * (int64_t)sizeof(TF) + 15) / 16);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    const char* s = reinterpret_cast<const char*>(src);
    char* d = reinterpret_cast<char*>(dst);
    for (int c0 = wave * 64; c0 < nv; c0 += nw * 64) {
        if (c0 + lane < nv) {
#if defined(__HIP_DEVICE_COMPILE__)
            __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void*)(s + (int64_t)(c0 + lane) * 16),
                                             (__attribute__((address_space(3))) void*)(d + (int64_t)c0 * 16), 16, 0, 0);
#endif
        }
    }
}
// wait for this wave's outstanding memory operations (direct LDS loads included)
__device__ __forceinline__ void wait_all() {
#if defined(__HIP_DEVICE_COMPILE__)
    __builtin_amdgcn_s_waitcnt(0);
#endif
}

// dst[0, n) = src[0, n) (lds_dma and the wait; the caller's barrier publishes it)
template <typename TF>
__device__ __forceinline__ void lds_copy(TF* dst, const TF* __restrict__ src, int64_t n) {
    lds_dma(dst, src, n);
    wait_all();
}

// One round of loads for an LDS-resident step: the operator (dst[0, na) = src[0, na) by
// direct LDS loads) and the input gather (st(e, ld(e)) for e < nx, through registers,
// eight loads per thread issued before their stores); the caller's barrier publishes both.
template <typename TF, typename Ld, typename St>
__device__ __forceinline__ void stage2(TF* dst, const TF* __restrict__ src, int64_t na, int nx, Ld ld, St st) {
    lds_dma(dst, src, na);
    batched<8, TF>(nx, ld, st);
    wait_all();
}

// LDS row-major -> global column-major
template <typename T, typename Tdst>
__device__ __forceinline__ void unstage(Tdst* dst, int64_t ld, const T* src, int r0, int rows, int nt) {
    for (int e = threadIdx.x; e < rows * nt; e += blockDim.x) {
        const int r = e % rows, j = e / rows;
        dst[r + (int64_t)j * ld] = (Tdst)src[(r0 + r) * NTP + j];
    }
}

// 16x16x4 MFMA of the element type: lane l supplies A(l%16, l/16) and B(l/16, l%16) and
// receives 4 outputs of column l%16, rows row(l/16, r) in register r (4r + l/16 for the
// Float64 instruction, 4(l/16) + r for the Float32 one).
template <typename TF>
struct Mfma;
template <>
struct Mfma<double> {
    typedef double v4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ v4 mac(double a, double b, v4 c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lk, int r) { return 4 * r + lk; }
};
template <>
struct Mfma<float> {
    typedef float v4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ v4 mac(float a, float b, v4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lk, int r) { return 4 * lk + r; }
};

// ---------------------------------------------------------------------------
// matvec
// ---------------------------------------------------------------------------
// ---------------------------------------------------------------------------
// ULV factorization
// ---------------------------------------------------------------------------
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

// Padded index of column c in the LQ reflector buffer: the CPT-column chunks of
// the S threads of a row start in different LDS banks.
template <int CPT>
__device__ __forceinline__ int vpad(int c) { return c + (c / CPT) * 2; }

// Shared front of the factor kernels (one workgroup per node, any thread count): the merged
// U in LDS and the source of W = [D; V'] (phase 0), the exit of a pass-through node, the QL
// of U with its outputs (phase 1), and M = T (Y' D) in LDS (phase 2), so that
// Q_ql' D = D - Y M.  LDS: Us (m x ku) | G (k x k) | Tq (k x k) | M (k rows of vstride) |
// vb (2 x vstride) | taus | sc, with vstride = S (CPT + 2) + 2.  Wm is set to the
// row-major merge buffer of a branch (null at a leaf, whose D and V are read from the
// generator blob).  Returns false for a pass-through node, whose outputs it has written.
template <typename T, int CPT>
__device__ __forceinline__ bool ulv_front(const FInst<T>& I, const DevNode& nd, const FNode& f,
                                          unsigned char* lds_raw, const T*& Wm_out, int64_t lds_elems) {
    const int m = f.m, n = f.n, k = f.k, nk = f.nk, ku = nd.ku, kv = nd.kv;
    const int S = f.S;
    const int tid = threadIdx.x;
    const int R = m + kv;
    const int wave = tid / 64, lane = tid % 64;
    const int nwaves = blockDim.x / 64;

    // LDS: Us (m x ku) | G (k x k) | Tq (k x k) | M (k rows of vstride) | vb (2 x vstride) | taus | sc
    // every row chunk of the S threads is CPT wide (columns >= n are zero padding)
    const int vstride = S * (CPT + 2) + 2;
    T* Us = reinterpret_cast<T*>(lds_raw);
    T* G = Us + (size_t)m * ku;
    T* Tq = G + (size_t)k * k;
    T* Mq = Tq + (size_t)k * k;  // row j of M at Mq + j * vstride, column c at vpad(c)
    T* vbuf = Mq + (size_t)k * vstride;
    T* taus = vbuf + 2 * vstride;
    T* sc = taus + (k > nk ? k : nk) + 1;

    // ---- phase 0: U (LDS); the source of W = [D; V'] ----
    //   leaf:   D and V from the generator blob (column-major)
    //   branch: merged rows written row-major into rwM (ulvfactor.jl:74-79)
    const T* Wm = nullptr;
    if (nd.is_leaf) {
        for (int e = tid; e < m * ku; e += blockDim.x) Us[e] = gen<T>(I.blob, nd.U, e);
    } else {
        const ChildRed<T> a = child_red(I, nd.left);
        const ChildRed<T> b = child_red(I, nd.right);
        const DevNode dl = I.nodes[nd.left];
        const DevNode dr = I.nodes[nd.right];
        T* Wmw = I.rw + f.rwM;
        // The merge's inner products run over the children's ranks, so they are staged:
        // the LDS after Us (used by phases 1-5 only later) holds the rank-sized operands,
        // loaded with many loads in flight, and the products read LDS.  The launch's
        // dynamic LDS (lds_elems elements, at least this node's own formula) bounds what
        // fits; a merge that does not fit reads its operands from memory.
        const int64_t own = 2 * (int64_t)k * k + (int64_t)(k + 2) * vstride + (k > nk ? k : nk) + 1 + 8;
        const int64_t cap = lds_elems - (int64_t)m * ku > own ? lds_elems - (int64_t)m * ku : own;
        auto mx = [](int64_t x, int64_t y) { return x > y ? x : y; };
        const int64_t nP12 = (int64_t)a.mr * b.kv, nP21 = (int64_t)b.mr * a.kv;
        const int64_t nU1 = (int64_t)a.mr * a.ku, nU2 = (int64_t)b.mr * b.ku;
        const int64_t nV1 = (int64_t)a.nr * a.kv, nV2 = (int64_t)b.nr * b.kv;
        // the most any one step below holds at once
        const int64_t need = mx(mx(nP12 + mx(nU1 + (int64_t)a.ku * b.kv, nV2), nP21 + mx(nU2 + (int64_t)b.ku * a.kv, nV1)),
                                mx(mx(nV1 + (int64_t)a.kv * kv, nV2 + (int64_t)b.kv * kv),
                                   mx(nU1 + (int64_t)a.ku * ku, nU2 + (int64_t)b.ku * ku)));
        if (need <= cap) {
            auto stage_to = [&](T* dst, int cnt, auto ld) { batched<8, T>(cnt, ld, [&](int e, T v) { dst[e] = v; }); };
            // rows of the merged D (ulvfactor.jl:75): [D1, U1 B12 V2'] then [U2 B21 V1', D2].
            // P = U1 B12 (or U2 B21) is kept in the factor too: the solve operators use it.
            for (int half = 0; half < 2; ++half) {
                const ChildRed<T>& x = half ? b : a;  // rows from this child
                const ChildRed<T>& o = half ? a : b;  // the other child (columns of V')
                const int64_t nP = (int64_t)x.mr * o.kv;
                T* sP = G;
                T* sU = sP + nP;
                T* sB = sU + (int64_t)x.mr * x.ku;
                stage_to(sU, x.mr * x.ku, [&](int e) { return x.U[e]; });
                stage_to(sB, x.ku * o.kv, [&](int e) { return gen<T>(I.blob, half ? nd.B21 : nd.B12, e); });
                __syncthreads();
                T* Pg = I.fac + (half ? f.P21 : f.P12);
                for (int e = tid; e < nP; e += blockDim.x) {
                    const int rr = e % x.mr, c = e / x.mr;
                    T sum = T(0);
                    for (int l = 0; l < x.ku; ++l) sum += sU[rr + l * x.mr] * sB[l + c * x.ku];
                    sP[e] = sum;
                    Pg[e] = sum;
                }
                __syncthreads();
                T* sV = sU;  // the other child's V, over U and B
                stage_to(sV, o.nr * o.kv, [&](int e) { return o.V[e]; });
                __syncthreads();
                const int r0 = half ? a.mr : 0, own0 = half ? a.nr : 0;  // own columns start
                for (int e = tid; e < x.mr * n; e += blockDim.x) {
                    const int rr = e / n, c = e % n;
                    T val = T(0);
                    if (c >= own0 && c < own0 + x.nr) {
                        val = x.D[rr + (int64_t)(c - own0) * x.mr];
                    } else {
                        const int cc = half ? c : c - a.nr;  // column of the other child's V'
                        for (int l = 0; l < o.kv; ++l) val += sP[rr + l * x.mr] * sV[cc + l * o.nr];
                    }
                    Wmw[(int64_t)(r0 + rr) * n + c] = val;
                }
                __syncthreads();
            }
            // rows of V' (ulvfactor.jl:79): V = [V1 W1; V2 W2]; columns from each child in turn
            for (int half = 0; half < 2; ++half) {
                const ChildRed<T>& x = half ? b : a;
                T* sV = G;
                T* sW = sV + (int64_t)x.nr * x.kv;
                stage_to(sV, x.nr * x.kv, [&](int e) { return x.V[e]; });
                stage_to(sW, x.kv * kv, [&](int e) { return gen<T>(I.blob, half ? dr.W : dl.W, e); });
                __syncthreads();
                const int c0x = half ? a.nr : 0;
                for (int e = tid; e < kv * x.nr; e += blockDim.x) {
                    const int j = e / x.nr, c = e % x.nr;
                    T val = T(0);
                    for (int l = 0; l < x.kv; ++l) val += sV[c + l * x.nr] * sW[l + j * x.kv];
                    Wmw[(int64_t)(m + j) * n + c0x + c] = val;
                }
                __syncthreads();
            }
            // U = [U1 R1; U2 R2]  (ulvfactor.jl:78)
            for (int half = 0; half < 2; ++half) {
                const ChildRed<T>& x = half ? b : a;
                T* sU = G;
                T* sR = sU + (int64_t)x.mr * x.ku;
                stage_to(sU, x.mr * x.ku, [&](int e) { return x.U[e]; });
                stage_to(sR, x.ku * ku, [&](int e) { return gen<T>(I.blob, half ? dr.R : dl.R, e); });
                __syncthreads();
                const int r0 = half ? a.mr : 0;
                for (int e = tid; e < x.mr * ku; e += blockDim.x) {
                    const int rr = e % x.mr, j = e / x.mr;
                    T sum = T(0);
                    for (int l = 0; l < x.ku; ++l) sum += sU[rr + l * x.mr] * sR[l + j * x.ku];
                    Us[(r0 + rr) + (int64_t)j * m] = sum;
                }
                __syncthreads();
            }
            Wm = Wmw;
            __threadfence_block();
        } else {
            merge_products(I, nd, f, a, b);
            __threadfence_block();
            __syncthreads();
            const T* P12 = I.fac + f.P12;
            const T* P21 = I.fac + f.P21;
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
                Wmw[e] = val;
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
                Us[e] = sum;
            }
            Wm = Wmw;
            __threadfence_block();
        }
    }
    __syncthreads();
    // entry (rr, c) of W = [D; V']
    auto Wsrc = [&](int rr, int c) -> T {
        if (Wm) return Wm[(int64_t)rr * n + c];
        return rr < m ? gen<T>(I.blob, nd.D, rr + (int64_t)c * m) : gen<T>(I.blob, nd.V, c + (int64_t)(rr - m) * n);
    };

    if (f.pass) {
        // k >= m: hand D, U, V to the parent unchanged (ulvfactor.jl:32-36)
        T* Dr = I.rw + f.rwD;
        T* Vr = I.rw + f.rwV;
        for (int e = tid; e < m * n; e += blockDim.x) Dr[e] = Wsrc(e % m, e / m);
        for (int e = tid; e < n * kv; e += blockDim.x) Vr[e] = Wsrc(m + e / n, e % n);
        T* Ur = I.rw + f.rwU;
        for (int e = tid; e < m * ku; e += blockDim.x) Ur[e] = Us[e];
        return false;
    }

    // ---- phase 1: QL of U (m x k) in LDS, LAPACK dgeqlf semantics (ulvfactor.jl:39) ----
    // Two barriers per reflector.  While reflector jj is applied its vector stays unscaled in
    // LDS and each use forms col[i] * scale, the value dgeqlf stores; the stored vector is
    // scaled in place during the next reflector's norm (another column).  beta goes to col[p]
    // at once: the update never reads it (the vector's unit entry is implicit).  The scalars
    // alternate between sc[0..2] and sc[4..6], so the deferred scaling reads the previous
    // reflector's while the current one's are written.
    for (int jj = k - 1; jj >= 0; --jj) {
        const int p = m - k + jj;
        T* col = Us + (int64_t)jj * m;
        T* scj = sc + 4 * (jj & 1);
        if (wave == 0) {
            T part = T(0);
            for (int i = lane; i < p; i += 64) part += col[i] * col[i];
            part = wave_reduce(part);
            if (lane == 0) {
                T beta, tau, scale;
                larfg(col[p], part, beta, tau, scale);
                scj[0] = beta;
                scj[1] = tau;
                scj[2] = scale;
                taus[jj] = tau;
                col[p] = beta;
            }
        }
        if (jj + 1 < k) {  // reflector jj + 1's vector: rows [0, p + 1) of the next column
            const T* scp = sc + 4 * ((jj + 1) & 1);
            if (scp[1] != T(0))
                for (int i = tid; i < p + 1; i += blockDim.x) col[m + i] *= scp[2];
        }
        __syncthreads();
        // apply H = I - tau v v' to U[0:p+1, 0:jj], v = [col[0:p] * scale; 1]
        const T tau = scj[1], scale = scj[2];
        if (tau != T(0)) {
            for (int c = wave; c < jj; c += nwaves) {
                T* uc = Us + (int64_t)c * m;
                T part = T(0);
                for (int i = lane; i < p; i += 64) part += (col[i] * scale) * uc[i];
                part = wave_reduce(part) + uc[p];
                const T t = tau * part;
                for (int i = lane; i < p; i += 64) uc[i] -= t * (col[i] * scale);
                if (lane == 0) uc[p] -= t;
            }
        }
        __syncthreads();
    }
    if (k > 0) {  // reflector 0's vector: rows [0, m - k) of column 0
        if (sc[1] != T(0))
            for (int i = tid; i < m - k; i += blockDim.x) Us[i] *= sc[2];
        __syncthreads();
    }
    // reflector j: y_j[i] = Us[i, j] (i < p_j), 1 (i = p_j), 0 below
    auto yql = [&](int i, int j) -> T {
        const int p = m - k + j;
        return i < p ? Us[i + (int64_t)j * m] : (i == p ? T(1) : T(0));
    };
    {
        T* Y = I.fac + f.Yql;
        for (int e = tid; e < m * k; e += blockDim.x) Y[e] = yql(e % m, e / m);
        T* Ur = I.rw + f.rwU;  // reduced U = tril(U[m-k:m, :]), k x ku (ku == k here)
        for (int e = tid; e < k * ku; e += blockDim.x) {
            const int i = e % k, j = e / k;
            Ur[e] = (i >= j) ? Us[(m - k + i) + (int64_t)j * m] : T(0);
        }
        // Gram G[a][b] = y_a . y_b for a < b (y_a has its unit above y_b's)
        for (int e = tid; e < k * k; e += blockDim.x) {
            const int a = e % k, bcol = e / k;
            T sum = T(0);
            if (a < bcol) {
                const int pa = m - k + a;
                const T* ya = Us + (int64_t)a * m;
                const T* yb = Us + (int64_t)bcol * m;
                sum = yb[pa];
                for (int i = 0; i < pa; ++i) sum += ya[i] * yb[i];
            }
            G[e] = sum;
        }
    }
    __syncthreads();
    // padding columns of M (n <= c < S * CPT) must read as zero in phase 3
    for (int e = tid; e < k * vstride; e += blockDim.x) Mq[e] = T(0);
    // T factor of H(1)...H(k) = I - Y T Y' (dlarft, forward columnwise); row a of T
    // depends only on row a itself, so thread a builds it alone
    for (int a = tid; a < k; a += blockDim.x) {
        for (int q = 0; q < k; ++q) Tq[a + (int64_t)q * k] = T(0);
        Tq[a + (int64_t)a * k] = taus[a];
        for (int i = a + 1; i < k; ++i) {
            T sum = T(0);
            for (int bb = a; bb < i; ++bb) sum += Tq[a + (int64_t)bb * k] * G[bb + (int64_t)i * k];
            Tq[a + (int64_t)i * k] = -taus[i] * sum;
        }
    }
    __syncthreads();
    {
        T* Tg = I.fac + f.Tql;
        for (int e = tid; e < k * k; e += blockDim.x) Tg[e] = Tq[e];
    }

    // ---- phase 2: M = T (Y' D), so that Q_ql' D = D - Y M  (ulvfactor.jl:41) ----
    // thread (column c, row group g): rows g, g+16, ... ; reduction over the 16 lanes of g
    {
        constexpr int RG = 16, MAXR = 16;
        const int g = tid % RG;
        const int cols_per_pass = blockDim.x / RG;
        for (int cb = tid / RG; cb < ((n + cols_per_pass - 1) / cols_per_pass) * cols_per_pass; cb += cols_per_pass) {
            const int c = cb;
            T d[MAXR];
#pragma unroll
            for (int q = 0; q < MAXR; ++q) {
                const int i = g + RG * q;
                d[q] = (c < n && i < m) ? Wsrc(i, c) : T(0);
            }
            for (int j0 = 0; j0 < k; j0 += 8) {
                T acc[8];
#pragma unroll
                for (int t = 0; t < 8; ++t) acc[t] = T(0);
#pragma unroll
                for (int q = 0; q < MAXR; ++q) {
                    const int i = g + RG * q;
                    if (i < m) {
#pragma unroll
                        for (int t = 0; t < 8; ++t)
                            if (j0 + t < k) acc[t] += yql(i, j0 + t) * d[q];
                    }
                }
#pragma unroll
                for (int t = 0; t < 8; ++t) {
                    T v = acc[t];
                    v += shfl_xor(v, 1);
                    v += shfl_xor(v, 2);
                    v += shfl_xor(v, 4);
                    v += shfl_xor(v, 8);
                    if (g == 0 && c < n && j0 + t < k) Mq[(j0 + t) * vstride + vpad<CPT>(c)] = v;
                }
            }
        }
    }
    __syncthreads();
    // M <- T M in place (T upper triangular: ascending j reads only rows >= j)
    for (int c = tid; c < n; c += blockDim.x) {
        T* mc = Mq + vpad<CPT>(c);
        for (int j = 0; j < k; ++j) {
            T sum = T(0);
            for (int l = j; l < k; ++l) sum += Tq[j + (int64_t)l * k] * mc[l * vstride];
            mc[j * vstride] = sum;
        }
    }
    __syncthreads();
    Wm_out = Wm;
    return true;
}

// NT = threads per workgroup (the launch bound).  FP64 levels whose nodes need at most 768
// threads run 768: the register budget per thread grows from 128 to 168 VGPRs and the
// 1024-thread build's spills go (the leaves' LQ 28 % faster).  FP32 fits 1024 threads
// without spills and keeps them.
template <typename T, int CPT, int NT>
__global__ __launch_bounds__(NT) void k_ulv_factor(const FInst<T>* __restrict__ insts,
                                                             const WorkItem* __restrict__ work, int64_t lds_elems) {
    const WorkItem wi = work[blockIdx.x];
    const FInst<T>& I = insts[wi.inst];
    const DevNode nd = I.nodes[wi.node];
    const FNode f = I.fn[wi.node];
    const int m = f.m, n = f.n, k = f.k, nk = f.nk, kv = nd.kv;
    const int S = f.S;
    const int tid = threadIdx.x;
    const int r = tid / S, s = tid % S;
    const int c0 = s * CPT;
    const int R = m + kv;
    const bool row_on = r < R;
    extern __shared__ unsigned char lds_raw[];
    const T* Wm = nullptr;
    if (!ulv_front<T, CPT>(I, nd, f, lds_raw, Wm, lds_elems)) return;
    const int vstride = S * (CPT + 2) + 2;
    T* Us = reinterpret_cast<T*>(lds_raw);
    T* Mq = Us + (size_t)m * nd.ku + 2 * (size_t)k * k;
    T* vbuf = Mq + (size_t)k * vstride;
    T* taus = vbuf + 2 * vstride;
    auto Wsrc = [&](int rr, int c) -> T {
        if (Wm) return Wm[(int64_t)rr * n + c];
        return rr < m ? gen<T>(I.blob, nd.D, rr + (int64_t)c * m) : gen<T>(I.blob, nd.V, c + (int64_t)(rr - m) * n);
    };
    auto yql = [&](int i, int j) -> T {
        const int p = m - k + j;
        return i < p ? Us[i + (int64_t)j * m] : (i == p ? T(1) : T(0));
    };

    // ---- phase 3: rows of W in registers: D rows updated by -Y M, V' rows as is ----
    T w[CPT];
#pragma unroll
    for (int q = 0; q < CPT; ++q) w[q] = T(0);
    if (row_on) {
#pragma unroll
        for (int q = 0; q < CPT; ++q)
            if (c0 + q < n) w[q] = Wsrc(r, c0 + q);
        if (r < m) {
            for (int j = 0; j < k; ++j) {
                const T y = yql(r, j);
                if (y != T(0)) {
                    const T* mj = Mq + j * vstride + vpad<CPT>(c0);
#pragma unroll
                    for (int q = 0; q < CPT; ++q) w[q] -= y * mj[q];
                }
            }
        }
    }

    // ---- phase 4: LQ of the first nk rows of [D; V'], reflectors applied to all
    //      rows below (L2 = D_bot Q', V <- Q V; ulvfactor.jl:44,47,50) ----
    for (int i = 0; i < nk; ++i) {
        T* vb = vbuf + (i & 1) * vstride;
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
            // v = [0 .. 0, 1, x * scale]; row i keeps L (c < i), beta (c = i) and v (c > i).
            // tau = 0 only when x = 0, and then x * scale = 0 = x as well.
            T* vs = vb + vpad<CPT>(c0);
#pragma unroll
            for (int q = 0; q < CPT; ++q) {
                const int c = c0 + q;
                const T v = (c == i) ? T(1) : ((c > i) ? w[q] * scale : T(0));
                w[q] = (c == i) ? beta : ((c > i) ? v : w[q]);
                vs[q] = v;
            }
            if (s == 0) taus[i] = tau;
        }
        __syncthreads();
        if (r > i && row_on) {
            // all S lanes of a row take this branch together (same r), so the
            // S-lane reduction below is convergent
            const T tau = taus[i];
            if (tau != T(0)) {
                // vs is zero for c < i and beyond n, so whole chunks can be used
                const T* vs = vb + vpad<CPT>(c0);
                T part = T(0);
#pragma unroll
                for (int q = 0; q < CPT; ++q) part += w[q] * vs[q];
                const T dot = row_reduce(part, S) * tau;
                // re-read v from LDS instead of keeping CPT more values live across the
                // reduction (the row itself already holds CPT registers)
                asm volatile("" ::: "memory");
#pragma unroll
                for (int q = 0; q < CPT; ++q) w[q] -= dot * vs[q];
            }
        }
    }
    __syncthreads();

    // ---- phase 5: outputs ----
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
                    if (c < nk) Vt[c + (int64_t)j * nk] = w[q];  // V_top (nk x kv)
                    else Vr[(c - nk) + (int64_t)j * f.nr] = w[q];
                }
            }
        }
    }
    // LQ reflector scalars for the compact-WY blocks (built by k_ulv_wy)
    for (int e = tid; e < nk; e += blockDim.x) I.fac[f.taulq + e] = taus[e];
}

// Blocked factor of one node per workgroup, on MFMA (NT threads, 16 rows of W = [D; V']
// per wave, n <= 16 NTILE).  Phases 0-2 as ulv_front.  Then each wave keeps its 16 rows of W
// in registers, transposed, in the accumulator layout of the 16x16x4 MFMA: lane l holds row
// 16 w + l%16 of W, columns 16 t + Mfma<T>::row(l/16, r) in a[t][r].  In that layout the
// rows serve directly as the B operand of a product over W's columns, so every update is
// an MFMA chain with the small operand (M, Y, V, T) read from LDS:
//   Q_ql' D = D - Y M           (ulvfactor.jl:41; M = T (Y' D) from ulv_front)
//   LQ of the first nk rows in panels of 16 reflectors (ulvfactor.jl:44,47,50): the panel's
//   wave factors its own rows (dlarfg reflectors, each applied to the panel's later rows;
//   cross-lane shuffles, no workgroup barrier), writes the reflectors V (16 x n) to LDS and
//   forms their compact-WY T (dlarft forward, from the Gram V V' on MFMA); then every wave
//   below applies the block, W <- W (I - V' T V) = W - ((W V') T) V, as three MFMA chains.
//   One workgroup barrier per panel (V and T double-buffered).
// Outputs as k_ulv_factor (L1, the LQ reflectors, L2a, V_top, the reduced D and V, taus)
// plus the T blocks.
template <typename T, int NTILE, int NT>
__global__ __launch_bounds__(NT) void k_ulv_factor_mf(const FInst<T>* __restrict__ insts,
                                                      const WorkItem* __restrict__ work, int64_t lds_elems) {
    using MF = Mfma<T>;
    using V4 = typename MF::v4;
    constexpr int NC = 16 * NTILE;  // columns held per row (n <= NC)
    const WorkItem wi = work[blockIdx.x];
    const FInst<T>& I = insts[wi.inst];
    const DevNode nd = I.nodes[wi.node];
    const FNode f = I.fn[wi.node];
    extern __shared__ unsigned char lds_raw[];
    const T* Wm = nullptr;
    if (!ulv_front<T, FACT_CPT>(I, nd, f, lds_raw, Wm, lds_elems)) return;
    const int m = f.m, n = f.n, k = f.k, nk = f.nk, kv = nd.kv, R = m + kv;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int li = lane & 15, lk = lane >> 4;
    const int row = 16 * wave + li;
    const int vstride = f.S * (FACT_CPT + 2) + 2;
    const T* Us = reinterpret_cast<const T*>(lds_raw);
    const T* Mq = Us + (size_t)m * nd.ku + 2 * (size_t)k * k;
    auto colof = [&](int t, int r) { return 16 * t + MF::row(lk, r); };

    // ---- rows of W; Q_ql' D = D - Y M as Wt -= M' Y' ----
    V4 a[NTILE];
#pragma unroll
    for (int t = 0; t < NTILE; ++t)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int c = colof(t, r);
            T x = T(0);
            if (row < R && c < n) {
                if (Wm) x = Wm[(int64_t)row * n + c];
                else x = row < m ? gen<T>(I.blob, nd.D, row + (int64_t)c * m) : gen<T>(I.blob, nd.V, c + (int64_t)(row - m) * n);
            }
            a[t][r] = x;
        }
    for (int s4 = 0; s4 < (k + 3) / 4; ++s4) {
        const int kk = 4 * s4 + lk;  // the k index this lane supplies in the step
        T yb = T(0);                 // B(kk, i) = Y(row, kk), reflector kk of the QL
        if (kk < k && row < m) {
            const int pp = m - k + kk;
            yb = row < pp ? Us[row + (int64_t)kk * m] : (row == pp ? T(1) : T(0));
        }
#pragma unroll
        for (int t = 0; t < NTILE; ++t) {
            const int c = 16 * t + li;  // A(c, kk) = -M(kk, c)
            const T am = (kk < k && c < n) ? -Mq[kk * vstride + vpad<FACT_CPT>(c)] : T(0);
            a[t] = MF::mac(am, yb, a[t]);
        }
    }
    __syncthreads();  // Us and M are dead from here; the panels' buffers reuse the LDS

    // ---- blocked LQ ----
    // [2][16][NCP]: the panel's reflectors (row-major).  The row pitch NC + 2 doubles puts the
    // 16 rows a column read touches (lanes li) in distinct LDS banks (a pitch of NC = 64 or
    // 128 maps them all to one) and keeps every row 16-byte aligned for the vector reads.
    constexpr int NCP = NC + 2;
    T* Vs = reinterpret_cast<T*>(lds_raw);
    // [2][16][TPP]: T(a, q) at a * TPP + q; the pitch 17 puts the 16 rows (lanes li) in
    // distinct LDS banks
    constexpr int TPP = 17;
    T* Ts = Vs + 2 * 16 * NCP;
    T* tl = Ts + 2 * 16 * TPP;              // nk taus (nk <= n <= 128)
    T* Gs = tl + 128;                       // [16]: the panel step's Gram column v_b . v_j
    const int npan = (nk + 15) / 16;
    for (int p = 0; p < npan; ++p) {
        const int i0 = 16 * p, nb = min(16, nk - i0), buf = p & 1;
        T* Vp = Vs + buf * 16 * NCP;
        T* Tp = Ts + buf * 16 * TPP;
        if (wave == p) {
            // the panel's reflectors, one row at a time (a loop, not unrolled), each applied
            // to the wave's later rows.  Rows live in the 4 lanes l%16 of the wave's 16-lane
            // rows, so a row's sums are rows_sum; row j's norm and pivot reach every lane by
            // readlane, and its reflector v_j reaches the other rows through row j of the
            // panel's V buffer in LDS (written by row j's lanes, read back by every lane; one
            // wave, so LDS order is program order).  Columns left of the panel (tiles t < p)
            // are already reduced and skipped.  The same dot loop gives v_b . v_j for the
            // panel's earlier reflectors b < j (their masked rows), which builds column j of
            // T (dlarft forward) through the 16-entry Gram column Gs.  Row ra of T lives in
            // LDS (Tp[ra * TPP + q]); the 4 lanes of row ra write the same values.
            const int ra = li;
#pragma unroll
            for (int q = 0; q < 16; ++q) Tp[ra * TPP + q] = T(0);
            // Only tile p (the panel's own 16 columns) needs column masks; in tiles t > p every
            // column is right of the reflector (c > i) and of every panel row's diagonal.  In
            // tile p, column 16 p + cr with cr = Mfma<T>::row(lk, r).  Row j's reflector is
            // stored lane-contiguously (slot 16 t + 4 lk + r) so each lane reads its 4 values
            // of a tile as one vector.
            for (int j = 0; j < nb; ++j) {
                const int i = i0 + j;
                T part = T(0), al = T(0);
#pragma unroll
                for (int t = 0; t < NTILE; ++t) {
                    if (t < p) continue;
                    if (t == p) {
#pragma unroll
                        for (int r = 0; r < 4; ++r) {
                            const int cr = MF::row(lk, r);
                            const T x = a[t][r];
                            part += cr > j ? x * x : T(0);
                            al += cr == j ? x : T(0);
                        }
                    } else {
#pragma unroll
                        for (int r = 0; r < 4; ++r) part += a[t][r] * a[t][r];
                    }
                }
                const T xn2 = lane_read(rows_sum(part), j);
                const T alpha = lane_read(rows_sum(al), j);
                T beta, tau, scl;
                larfg(alpha, xn2, beta, tau, scl);
                // row j becomes [L, beta, v_j]; its lanes also store v_j = [0, 1, x scl]
                const bool own = li == j;
                T* vrow = Vp + j * NCP;
#pragma unroll
                for (int t = 0; t < NTILE; ++t) {
                    if (t < p) continue;
                    V4 v;
                    if (t == p) {
#pragma unroll
                        for (int r = 0; r < 4; ++r) {
                            const int cr = MF::row(lk, r);
                            const T x = a[t][r];
                            v[r] = cr > j ? x * scl : (cr == j ? T(1) : T(0));
                            a[t][r] = own ? (cr == j ? beta : (cr > j ? v[r] : x)) : x;
                        }
                    } else {
#pragma unroll
                        for (int r = 0; r < 4; ++r) {
                            v[r] = a[t][r] * scl;
                            a[t][r] = own ? v[r] : a[t][r];
                        }
                    }
                    if (own) *reinterpret_cast<V4*>(vrow + 16 * t + 4 * lk) = v;
                }
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                // d = row . v_j for the rows below j; v_li . v_j for the reflectors above
                // (row li < j reads as [0, 1, v] about its diagonal column 16 p + li)
                T d = T(0);
#pragma unroll
                for (int t = 0; t < NTILE; ++t) {
                    if (t < p) continue;
                    const V4 v = *reinterpret_cast<const V4*>(vrow + 16 * t + 4 * lk);
                    if (t == p) {
#pragma unroll
                        for (int r = 0; r < 4; ++r) {
                            const int cr = MF::row(lk, r);
                            const T x = a[t][r];
                            const T left = li > j ? x : (cr < li ? T(0) : (cr == li ? T(1) : x));
                            d += left * v[r];
                        }
                    } else {
#pragma unroll
                        for (int r = 0; r < 4; ++r) d += a[t][r] * v[r];
                    }
                }
                d = rows_sum(d);
                const T dt = li > j ? d * tau : T(0);
#pragma unroll
                for (int t = 0; t < NTILE; ++t) {
                    if (t < p) continue;
                    const V4 v = *reinterpret_cast<const V4*>(vrow + 16 * t + 4 * lk);
#pragma unroll
                    for (int r = 0; r < 4; ++r) a[t][r] -= dt * v[r];
                }
                // T(a, j) = -tau_j sum_{a <= b < j} T(a, b) (v_b . v_j) for a < j
                Gs[li] = d;
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                T acc = T(0);
                for (int b2 = 0; b2 < j; ++b2) acc += Tp[ra * TPP + b2] * Gs[b2];
                Tp[ra * TPP + j] = ra < j ? -tau * acc : (ra == j ? tau : T(0));
                if (lane == 0) tl[i] = tau;
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            }
            // V (rows beyond nb zero) and T to LDS; T also to the factor
#pragma unroll
            for (int t = 0; t < NTILE; ++t)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int c = colof(t, r), ii = i0 + li;
                    Vp[li * NCP + c] = li < nb ? (c < ii ? T(0) : (c == ii ? T(1) : a[t][r])) : T(0);
                }
            if (lane < 16) {
                T* Tg = I.fac + f.Tlq + (int64_t)p * NB * NB;
#pragma unroll
                for (int q = 0; q < 16; ++q) Tg[ra + q * NB] = (q < nb && ra < nb) ? Tp[ra * TPP + q] : T(0);
            }
        }
        __syncthreads();
        if (wave > p) {
            // W <- W - ((W V') T) V for this wave's rows, as Wt -= V' (T' (V Wt))
            V4 P = {T(0), T(0), T(0), T(0)};
#pragma unroll
            for (int t = 0; t < NTILE; ++t)
#pragma unroll
                for (int r = 0; r < 4; ++r) P = MF::mac(Vp[li * NCP + colof(t, r)], a[t][r], P);
            V4 Q = {T(0), T(0), T(0), T(0)};
#pragma unroll
            for (int r = 0; r < 4; ++r) Q = MF::mac(Tp[MF::row(lk, r) * TPP + li], P[r], Q);
#pragma unroll
            for (int t = 0; t < NTILE; ++t)
#pragma unroll
                for (int r = 0; r < 4; ++r) a[t] = MF::mac(-Vp[MF::row(lk, r) * NCP + 16 * t + li], Q[r], a[t]);
        }
    }
    __syncthreads();

    // ---- outputs (the same blocks and layouts as k_ulv_factor's phase 5) ----
    // (the row index passes through an empty asm so that the output addresses are formed
    // here, not hoisted to the top of the kernel and kept live through the panels)
    int orow = row;
    asm volatile("" : "+v"(orow));
    if (orow < R) {
        const int row = orow;
#pragma unroll
        for (int t = 0; t < NTILE; ++t)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int c = colof(t, r);
                if (c >= n) continue;
                const T x = a[t][r];
                if (row < nk) {
                    if (c < nk) I.fac[f.L1 + row + (int64_t)c * nk] = (c <= row) ? x : T(0);
                    I.fac[f.Ylq + c + (int64_t)row * n] = (c < row) ? T(0) : (c == row ? T(1) : x);
                } else if (row < m) {
                    const int rr = row - nk;
                    if (c < nk) I.fac[f.L2a + rr + (int64_t)c * k] = x;
                    else I.rw[f.rwD + rr + (int64_t)(c - nk) * f.mr] = x;
                } else {
                    const int j = row - m;
                    if (c < nk) I.fac[f.Vt + c + (int64_t)j * nk] = x;
                    else I.rw[f.rwV + (c - nk) + (int64_t)j * f.nr] = x;
                }
            }
    }
    for (int e = tid; e < nk; e += blockDim.x) I.fac[f.taulq + e] = tl[e];
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
    if (f.Dr >= 0)  // the merged block itself, for the refinement step of every root solve
        for (int e = tid; e < m * m; e += blockDim.x) I.fac[f.Dr + e] = A[e];
    __shared__ int s_piv;
    __shared__ T s_val;
    const int wave = tid / 64, lane = tid % 64;
    int info = 0;
    for (int i = 0; i < m; ++i) {
        if (wave == 0) {
            // idamax over A[i:m, i] (first index of the largest magnitude).  bi starts at
            // i, so an all-NaN column pivots in range (and is reported below)
            T best = T(-1);
            int bi = i;
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
        if (pv == T(0) || pv != pv) {  // a zero or NaN pivot: singular (status i + 1)
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
    // inverted diagonal NB x NB blocks of the unit-lower L and the upper U, for
    // blocked triangular solves in every root solve; thread (block, column j)
    __threadfence_block();
    __syncthreads();
    if (info == 0) {
        const int nblk = (m + NB - 1) / NB;
        T* Li = I.fac + f.LUinv;                           // nblk x NB x NB
        T* Ui = I.fac + f.LUinv + (int64_t)nblk * NB * NB;  // nblk x NB x NB
        for (int t = tid; t < nblk * NB; t += blockDim.x) {
            const int bb = t / NB, jj = t % NB;
            const int b0 = bb * NB, nb = min(NB, m - b0);
            T x[NB];
            // L_bb x = e_j (unit diagonal)
#pragma unroll
            for (int i2 = 0; i2 < NB; ++i2) {
                T sum = (i2 == jj) ? T(1) : T(0);
#pragma unroll
                for (int q = 0; q < NB; ++q)
                    if (q < i2 && i2 < nb) sum -= A[(b0 + i2) + (int64_t)(b0 + q) * m] * x[q];
                x[i2] = (i2 < nb && jj < nb) ? sum : T(0);
            }
#pragma unroll
            for (int q = 0; q < NB; ++q) Li[(int64_t)bb * NB * NB + q + jj * NB] = x[q];
            // U_bb x = e_j
#pragma unroll
            for (int i2 = NB - 1; i2 >= 0; --i2) {
                T sum = (i2 == jj) ? T(1) : T(0);
#pragma unroll
                for (int q = 0; q < NB; ++q)
                    if (q > i2 && q < nb) sum -= A[(b0 + i2) + (int64_t)(b0 + q) * m] * x[q];
                x[i2] = (i2 < nb && jj < nb) ? sum / A[(b0 + i2) + (int64_t)(b0 + i2) * m] : T(0);
            }
#pragma unroll
            for (int q = 0; q < NB; ++q) Ui[(int64_t)bb * NB * NB + q + jj * NB] = x[q];
        }
    }
}

// ---------------------------------------------------------------------------
// ULV solve operators, formed once per factorization
// ---------------------------------------------------------------------------
// The reduction of a node (ulvfactor.jl:42,48-51,74,89: Q_ql' b, the forward
// substitution with L1, the reduced b and u, the children's merges) is linear in
// its input x = the leaf's rows of B, or [b_1; u_1; b_2; u_2] for a branch.
template <bool TA, typename TF, typename Store>
__device__ __forceinline__ void mfma_gemm(int M, int K, const TF* __restrict__ A, int64_t lda, const TF* xs, int ncols,
                                          Store store, int xp);

// k_ulv_gform runs its steps on columns of the identity:
//   Gz (nk x Kx):   column c = zloc for x = e_c
//   Gr (R x (Kx + nk)), R = mr + kv: column c < Kx = [(Q_ql' b)[nk:m]; W_1'u_1 + W_2'u_2]
//                   for x = e_c (the reduced b and u before their zloc terms); the last
//                   nk columns are [-L2a; V_top'], so [reduced b; u] = Gr [x; zloc].
// Grid columns run over Kx + nk; the last nk are copies of factor blocks.
//
// LDS (row-major, NTP columns): bv m rows | zl max(nk,1) | us kv_1 + kv_2 | t1 tr | t2 tr,
// tr = max(k, NB, kv)
template <typename TF>
__global__ __launch_bounds__(SWEEP_THREADS) void k_ulv_gform(const FInst<TF>* __restrict__ insts,
                                                            const WorkItem* __restrict__ work) {
    const WorkItem wi = work[blockIdx.x];
    const FInst<TF>& I = insts[wi.inst];
    const DevNode nd = I.nodes[wi.node];
    const FNode f = I.fn[wi.node];
    const int j0 = blockIdx.y * RHS_TILE;
    const int Kx = f.Kx, nk = f.nk, mr = f.mr, kv = nd.kv, R = mr + kv;
    if (j0 >= Kx + nk) return;
    TF* Gr = I.fac + f.Gr;  // ld R
    // columns [Kx, Kx + nk) of Gr: [-L2a; V_top']
    for (int c = max(j0, Kx); c < min(j0 + RHS_TILE, Kx + nk); ++c) {
        const int l = c - Kx;
        for (int i = threadIdx.x; i < R; i += blockDim.x)
            Gr[i + (int64_t)c * R] = i < mr ? -I.fac[f.L2a + i + (int64_t)l * f.k] : I.fac[f.Vt + l + (int64_t)(i - mr) * nk];
    }
    if (j0 >= Kx) return;
    const int nt = min(RHS_TILE, Kx - j0);
    const int m = f.m, k = f.k;
    int kvl = 0, kvr = 0, mrl = 0;
    if (!nd.is_leaf) {
        kvl = I.nodes[nd.left].kv;
        kvr = I.nodes[nd.right].kv;
        mrl = I.fn[nd.left].mr;
    }
    const int tr = max(max(k, NB), kv);
    extern __shared__ unsigned char lds_raw[];
    TF* bv = reinterpret_cast<TF*>(lds_raw);
    TF* zl = bv + m * NTP;
    TF* us = zl + max(nk, 1) * NTP;
    TF* t1 = us + (kvl + kvr) * NTP;
    TF* t2 = t1 + tr * NTP;
    // identity columns j0.. scattered over b (bv) and the children's u (us):
    // input index of bv row r is r (b_1) or r + kv_1 (b_2); of us row r, mr_1 + r (u_1) or m + r (u_2)
    for (int e = threadIdx.x; e < m * NTP; e += blockDim.x) {
        const int r = e / NTP, j = e % NTP;
        const int c = (nd.is_leaf || r < mrl) ? r : r + kvl;
        bv[e] = (j < nt && c == j0 + j) ? TF(1) : TF(0);
    }
    for (int e = threadIdx.x; e < (kvl + kvr) * NTP; e += blockDim.x) {
        const int r = e / NTP, j = e % NTP;
        const int c = r < kvl ? mrl + r : m + r;
        us[e] = (j < nt && c == j0 + j) ? TF(1) : TF(0);
    }
    __syncthreads();
    if (!nd.is_leaf) {
        // b = [b_1 - P12 u_2; b_2 - P21 u_1]   (ulvfactor.jl:74)
        const int mrr = m - mrl;
        mfma_gemm<false, TF>(mrl, kvr, I.fac + f.P12, mrl, us + kvl * NTP, nt,
                         [&](int i, int j, TF v) { bv[i * NTP + j] -= v; }, NTP);
        mfma_gemm<false, TF>(mrr, kvl, I.fac + f.P21, mrr, us, nt,
                         [&](int i, int j, TF v) { bv[(mrl + i) * NTP + j] -= v; }, NTP);
        __syncthreads();
    }
    TF* Grc = Gr + (int64_t)j0 * R;  // column j0 of Gr
    if (!f.pass) {
        // b <- Q_ql' b = b - Y (T (Y' b))   (ulvfactor.jl:42)
        if (k > 0) {
            mfma_gemm<true, TF>(k, m, I.fac + f.Yql, m, bv, nt, [&](int i, int j, TF v) { t1[i * NTP + j] = v; }, NTP);
            __syncthreads();
            mfma_gemm<false, TF>(k, k, I.fac + f.Tql, k, t1, nt, [&](int i, int j, TF v) { t2[i * NTP + j] = v; }, NTP);
            __syncthreads();
            mfma_gemm<false, TF>(m, k, I.fac + f.Yql, m, t2, nt, [&](int i, int j, TF v) { bv[i * NTP + j] -= v; }, NTP);
            __syncthreads();
        }
        // zloc = L1 \ b[0:nk]  (ulvfactor.jl:48), blocked with inverted diagonal blocks
        for (int bb = 0; bb < f.nblk; ++bb) {
            const int b0 = bb * NB, nb = min(NB, nk - b0);
            mfma_gemm<false, TF>(nb, b0, I.fac + f.L1 + b0, nk, zl, nt,
                             [&](int i, int j, TF v) { t1[i * NTP + j] = bv[(b0 + i) * NTP + j] - v; }, NTP);
            __syncthreads();
            mfma_gemm<false, TF>(nb, nb, I.fac + f.L1inv + (int64_t)bb * NB * NB, NB, t1, nt,
                             [&](int i, int j, TF v) { zl[(b0 + i) * NTP + j] = v; }, NTP);
            __syncthreads();
        }
        unstage(I.fac + f.Gz + (int64_t)j0 * nk, nk, zl, 0, nk, nt);
        unstage(Grc, R, bv, nk, mr, nt);  // (Q_ql' b)[nk:m]
    } else {
        unstage(Grc, R, bv, 0, m, nt);  // nothing eliminated: reduced b = b
    }
    // u part: W_1' u_1 + W_2' u_2 (ulvfactor.jl:89); a leaf has none
    if (kv > 0) {
        TF* Gu = Grc + mr;
        if (nd.is_leaf) {
            for (int e = threadIdx.x; e < kv * nt; e += blockDim.x) Gu[(e % kv) + (int64_t)(e / kv) * R] = TF(0);
        } else {
            const DevNode dl = I.nodes[nd.left], dr = I.nodes[nd.right];
            gemm<true>(kv, nt, kvl, I.blob + dl.W, kvl, us, [&](int i, int j, TF v) { t2[i * NTP + j] = v; });
            __syncthreads();
            gemm<true>(kv, nt, kvr, I.blob + dr.W, kvr, us + kvl * NTP,
                       [&](int i, int j, TF v) { Gu[i + (int64_t)j * R] = t2[i * NTP + j] + v; });
        }
    }
}

// Qt = Q_lq' explicitly (ulvfactor.jl:102): column c is H(1)...H(nk) e_c, applied
// as blocks of NB reflectors z - Y_b T_b Y_b' z, the last block first.
// LDS: z n x NTP | t1, t2 NB x NTP
template <typename TF>
__global__ __launch_bounds__(SWEEP_THREADS) void k_ulv_qform(const FInst<TF>* __restrict__ insts,
                                                            const WorkItem* __restrict__ work) {
    const WorkItem wi = work[blockIdx.x];
    const FInst<TF>& I = insts[wi.inst];
    const FNode f = I.fn[wi.node];
    if (f.Qt < 0) return;
    const int n = f.n;
    const int j0 = blockIdx.y * RHS_TILE;
    if (j0 >= n) return;
    const int nt = min(RHS_TILE, n - j0);
    const int nk = f.nk;
    extern __shared__ unsigned char lds_raw[];
    TF* z = reinterpret_cast<TF*>(lds_raw);
    TF* t1 = z + n * NTP;
    TF* t2 = t1 + NB * NTP;
    for (int e = threadIdx.x; e < n * NTP; e += blockDim.x) {
        const int r = e / NTP, j = e % NTP;
        z[e] = (j < nt && r == j0 + j) ? TF(1) : TF(0);
    }
    __syncthreads();
    for (int bb = f.nblk - 1; bb >= 0; --bb) {
        const int b0 = bb * NB, nb = min(NB, nk - b0);
        const TF* Yb = I.fac + f.Ylq + (int64_t)b0 * n + b0;  // rows b0.., columns b0..b0+nb
        mfma_gemm<true, TF>(nb, n - b0, Yb, n, z + b0 * NTP, nt, [&](int i, int j, TF v) { t1[i * NTP + j] = v; }, NTP);
        __syncthreads();
        mfma_gemm<false, TF>(nb, nb, I.fac + f.Tlq + (int64_t)bb * NB * NB, NB, t1, nt,
                         [&](int i, int j, TF v) { t2[i * NTP + j] = v; }, NTP);
        __syncthreads();
        mfma_gemm<false, TF>(n - b0, nb, Yb, n, t2, nt, [&](int i, int j, TF v) { z[(b0 + i) * NTP + j] -= v; }, NTP);
        __syncthreads();
    }
    unstage(I.fac + f.Qt + (int64_t)j0 * n, n, z, 0, n, nt);
}

// ---------------------------------------------------------------------------
// ULV solve
// ---------------------------------------------------------------------------
// bv (m rows, NTP columns in LDS) <- A^-1 bv with the LU factors of A (partial
// pivoting, unit-lower L and U in LU, m x m) and the inverted NB x NB diagonal blocks of
// L then U in LUinv: row interchanges, then blocked forward and backward substitution.
// t1: NB x NTP scratch.
template <typename TF>
__device__ __forceinline__ void lu_solve(TF* bv, TF* t1, int m, int nt, const TF* LU, const TF* LUinv, const int32_t* piv) {
    for (int j = threadIdx.x; j < nt; j += blockDim.x)
        for (int i = 0; i < m; ++i) {
            const int p = piv[i];
            if (p != i) { const TF t = bv[i * NTP + j]; bv[i * NTP + j] = bv[p * NTP + j]; bv[p * NTP + j] = t; }
        }
    __syncthreads();
    const int nblk = (m + NB - 1) / NB;
    const TF* Li = LUinv;
    const TF* Ui = Li + (int64_t)nblk * NB * NB;
    // forward: y_b = L_bb^-1 (b_b - L[b, 0:b0] y[0:b0])
    for (int bb = 0; bb < nblk; ++bb) {
        const int b0 = bb * NB, nb = min(NB, m - b0);
        gemm<false>(nb, nt, b0, LU + b0, m, bv, [&](int i, int j, TF v) { t1[i * NTP + j] = bv[(b0 + i) * NTP + j] - v; });
        __syncthreads();
        gemm<false>(nb, nt, nb, Li + (int64_t)bb * NB * NB, NB, t1, [&](int i, int j, TF v) { bv[(b0 + i) * NTP + j] = v; });
        __syncthreads();
    }
    // backward: x_b = U_bb^-1 (y_b - U[b, b1:m] x[b1:m])
    for (int bb = nblk - 1; bb >= 0; --bb) {
        const int b0 = bb * NB, nb = min(NB, m - b0), b1 = b0 + nb;
        gemm<false>(nb, nt, m - b1, LU + b0 + (int64_t)b1 * m, m, bv + b1 * NTP,
                    [&](int i, int j, TF v) { t1[i * NTP + j] = bv[(b0 + i) * NTP + j] - v; });
        __syncthreads();
        gemm<false>(nb, nt, nb, Ui + (int64_t)bb * NB * NB, NB, t1, [&](int i, int j, TF v) { bv[(b0 + i) * NTP + j] = v; });
        __syncthreads();
    }
}

// Root (ulvfactor.jl:74,83): b = [b_1 - P12 u_2; b_2 - P21 u_1] from the children's
// y blocks, then the LU solve with partial pivoting (blocked triangular solves with
// the inverted diagonal blocks of L and U).
// LDS (row-major, NTP columns): bv m rows | us kv_1 + kv_2 | t1 NB | with resident != 0,
// the root factor block LU | LUinv | P12 | P21 (staged whole, so the solve's many
// small dependent steps read LDS rather than memory)
template <typename TF>
__global__ __launch_bounds__(SWEEP_THREADS) void k_ulv_groot(const SInst<TF>* __restrict__ insts,
                                                            const SweepRec* __restrict__ work, int resident) {
    const SweepRec wi = work[blockIdx.x];
    const SInst<TF>& I = insts[wi.inst];
    const int j0 = blockIdx.y * RHS_TILE;
    if (j0 >= I.nrhs) return;
    const int nt = min(RHS_TILE, I.nrhs - j0);
    const int nrhs = I.nrhs;
    const DevNode nd = I.nodes[wi.node];
    const FNode f = I.fn[wi.node];
    const int m = f.m;
    int kvl = 0, kvr = 0;
    if (!nd.is_leaf) {
        kvl = I.nodes[nd.left].kv;
        kvr = I.nodes[nd.right].kv;
    }
    extern __shared__ unsigned char lds_raw[];
    TF* bv = reinterpret_cast<TF*>(lds_raw);
    TF* us = bv + m * NTP;
    TF* t1 = us + (kvl + kvr) * NTP;
    // the root factor block, in LDS or in memory
    const TF* fb = I.fac + f.LU;
    if (resident) {
        TF* Ls = t1 + NB * NTP;
        Ls += (16 - ((Ls - reinterpret_cast<TF*>(lds_raw)) & 15)) & 15;  // 16-element aligned
        lds_copy(Ls, fb, f.rlen);
        fb = Ls;
    }
    const TF* P12 = fb + (f.P12 - f.LU);
    const TF* P21 = fb + (f.P21 - f.LU);
    if (nd.is_leaf) {
        stage(bv, 0, I.B + nd.row_off + (int64_t)j0 * I.ldb, I.ldb, m, nt);
        __syncthreads();
    } else {
        const FNode fl = I.fn[nd.left], fr = I.fn[nd.right];
        const TF* yl = I.wy + fl.yrow * nrhs + (int64_t)j0 * fl.My;
        const TF* yr = I.wy + fr.yrow * nrhs + (int64_t)j0 * fr.My;
        stage(bv, 0, yl + fl.nk, fl.My, fl.mr, nt);
        stage(bv, fl.mr, yr + fr.nk, fr.My, fr.mr, nt);
        stage(us, 0, yl + fl.nk + fl.mr, fl.My, kvl, nt);
        stage(us, kvl, yr + fr.nk + fr.mr, fr.My, kvr, nt);
        __syncthreads();
        gemm<false>(fl.mr, nt, kvr, P12, fl.mr, us + kvl * NTP, [&](int i, int j, TF v) { bv[i * NTP + j] -= v; });
        const int mr1 = fl.mr;
        gemm<false>(fr.mr, nt, kvl, P21, fr.mr, us, [&](int i, int j, TF v) { bv[(mr1 + i) * NTP + j] -= v; });
        __syncthreads();
    }
    lu_solve(bv, t1, m, nt, fb, fb + (f.LUinv - f.LU), I.ipiv + f.piv);
    if (nd.is_leaf) unstage(I.X + nd.col_off + (int64_t)j0 * I.ldx, I.ldx, bv, 0, m, nt);
    else unstage(I.wz + f.zrow * nrhs + (int64_t)j0 * f.n, f.n, bv, 0, m, nt);
}


// Root operators for the inverse-and-refine solve (factor time, after k_ulv_root):
// columns [j0, j0 + NTP) of Dinv (LU solves of identity columns) and of Mb.
// LDS: bv m x NTP | t1 NB x NTP
template <typename TF>
__global__ __launch_bounds__(SWEEP_THREADS) void k_ulv_rinv(const FInst<TF>* __restrict__ insts,
                                                           const WorkItem* __restrict__ work) {
    const WorkItem wi = work[blockIdx.x];
    const FInst<TF>& I = insts[wi.inst];
    const DevNode nd = I.nodes[wi.node];
    const FNode f = I.fn[wi.node];
    if (f.Dinv < 0) return;
    const int m = f.m, Kx = f.Kx;
    const int j0 = blockIdx.y * RHS_TILE;
    // Mb columns: input order [b_1 (mr1); u_1 (kv1); b_2 (mr2); u_2 (kv2)]
    const int mr1 = I.fn[nd.left].mr, mr2 = I.fn[nd.right].mr;
    const int kv1 = I.nodes[nd.left].kv;
    for (int e = threadIdx.x; e < m * RHS_TILE; e += blockDim.x) {
        const int i = e % m, c = j0 + e / m;
        if (c >= Kx) continue;
        TF v = TF(0);
        if (c < mr1) v = (i == c) ? TF(1) : TF(0);
        else if (c < mr1 + kv1) v = i >= mr1 ? -I.fac[f.P21 + (i - mr1) + (int64_t)(c - mr1) * mr2] : TF(0);
        else if (c < mr1 + kv1 + mr2) v = (i - mr1 == c - mr1 - kv1) ? TF(1) : TF(0);
        else v = i < mr1 ? -I.fac[f.P12 + i + (int64_t)(c - mr1 - kv1 - mr2) * mr1] : TF(0);
        I.fac[f.Mb + i + (int64_t)c * m] = v;
    }
    if (j0 >= m) return;
    const int nt = min(RHS_TILE, m - j0);
    extern __shared__ unsigned char lds_raw[];
    TF* bv = reinterpret_cast<TF*>(lds_raw);
    TF* t1 = bv + m * NTP;
    for (int e = threadIdx.x; e < m * NTP; e += blockDim.x) {
        const int r = e / NTP, j = e % NTP;
        bv[e] = (j < nt && r == j0 + j) ? TF(1) : TF(0);
    }
    __syncthreads();
    lu_solve<TF>(bv, t1, m, nt, I.fac + f.LU, I.fac + f.LUinv, I.ipiv + f.piv);
    unstage(I.fac + f.Dinv + (int64_t)j0 * m, m, bv, 0, m, nt);
}

// ---- MFMA tile GEMM of the solve sweeps ----
// y(0:M, 0:N) = A(0:M, 0:K) x(0:K, 0:N) over one workgroup tile of TM x TN outputs.
// A is column-major (lda = M) in the factor type, or the identity when null.  x is
// two stacked column-major segments (rows [0, r0) from seg 0, the rest from seg 1),
// each Float64 (the caller's B) or the factor type (workspace blocks); y likewise.
// The 4 waves form a WM x WN grid; each wave owns 2 x 2 blocks of 16 x 16 outputs
// computed with v_mfma_*_16x16x4 (lane l supplies A(l%16, l/16) and x(l/16, l%16)
// and receives 4 outputs of column l%16: rows 4r + l/16 in register r for the
// Float64 instruction, rows 4(l/16) + r for the Float32 one).  K advances in chunks
// of GK, staged through double-buffered LDS with one barrier per chunk.

// x operand: rows [0, r0) from segment 0, [r0, r1) from segment 1 (global,
// column-major, Float64 or the factor type), rows [r1, K) from an LDS block zs
// (row-major, pitch zp, columns local to the workgroup tile).
// Rows [0, r0) of x from p0, [r0, r1) from p1, then from p2 (Float64, pitch ld2) when it
// is set, else from the LDS rows zs.
struct SegOp {
    const void* p0;
    const void* p1;
    const void* p2;
    int64_t ld0, ld1, ld2;
    int r0, r1, dbl0, dbl1;
};

constexpr int GK = 16;  // K chunk of the MFMA tile GEMM

template <typename TF>
__device__ __forceinline__ TF seg_load(const SegOp& x, const TF* zs, int zp, int r, int64_t j, int jl) {
    if (r < x.r0)
        return x.dbl0 ? (TF)gbl(static_cast<const double*>(x.p0))[r + j * x.ld0]
                      : gbl(static_cast<const TF*>(x.p0))[r + j * x.ld0];
    if (r < x.r1) {
        r -= x.r0;
        return x.dbl1 ? (TF)gbl(static_cast<const double*>(x.p1))[r + j * x.ld1]
                      : gbl(static_cast<const TF*>(x.p1))[r + j * x.ld1];
    }
    if (x.p2) return (TF)gbl(static_cast<const double*>(x.p2))[(r - x.r1) + j * x.ld2];
    return zs[(r - x.r1) * zp + jl];
}

// LDS staging of one tile GEMM: two buffers of A (GK x (TM + 1)) and x (GK x (TN + 1)).
template <typename TF, int WM, int WN>
struct TileLds {
    static constexpr int TM = 32 * WM, TN = 32 * WN, AP = TM + 1, XP = TN + 1;
    TF As[2][GK * AP];
    TF Xs[2][GK * XP];
};

// y(i0:i0+TM, j0:j0+TN) = A(., 0:K) x(0:K, .), outputs passed to store(i, j, v) for
// i < M, j < N.
template <typename TF, int WM, int WN, typename Store>
__device__ __forceinline__ void tile_gemm(TileLds<TF, WM, WN>& L, int M, int N, int K, const TF* __restrict__ A,
                                          int64_t lda, const SegOp& x, const TF* zs, int zp, int i0, int j0,
                                          Store store) {
    constexpr int TM = 32 * WM, TN = 32 * WN;
    constexpr int AP = TM + 1, XP = TN + 1;  // padded LDS rows
    constexpr int NA = TM * GK / 256, NX = GK * TN / 256;
    auto& As = L.As;
    auto& Xs = L.Xs;
    using V4 = typename Mfma<TF>::v4;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave % WM, wn = wave / WM;
    TF ra[NA], rx[NX];
    auto load = [&](int k0) {
#pragma unroll
        for (int q = 0; q < NA; ++q) {
            const int e = tid + q * 256;
            const int i = e % TM, kk = e / TM;
            const int gi = i0 + i, gk = k0 + kk;
            TF v = TF(0);
            if (gi < M && gk < K) v = A ? gbl(A)[gi + (int64_t)gk * lda] : (gi == gk ? TF(1) : TF(0));
            ra[q] = v;
        }
#pragma unroll
        for (int q = 0; q < NX; ++q) {
            const int e = tid + q * 256;
            const int kk = e % GK, j = e / GK;
            const int gk = k0 + kk, gj = j0 + j;
            rx[q] = (gk < K && gj < N) ? seg_load<TF>(x, zs, zp, gk, gj, j) : TF(0);
        }
    };
    auto put = [&](int b) {
#pragma unroll
        for (int q = 0; q < NA; ++q) {
            const int e = tid + q * 256;
            As[b][(e / TM) * AP + e % TM] = ra[q];
        }
#pragma unroll
        for (int q = 0; q < NX; ++q) {
            const int e = tid + q * 256;
            Xs[b][(e % GK) * XP + e / GK] = rx[q];
        }
    };
    V4 acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = V4{TF(0), TF(0), TF(0), TF(0)};
    const int nch = (K + GK - 1) / GK;
    load(0);
    put(0);
    __syncthreads();
    const int lr = lane & 15, lk = lane >> 4;
    for (int c = 0; c < nch; ++c) {
        const int b = c & 1;
        if (c + 1 < nch) load((c + 1) * GK);
#pragma unroll
        for (int kk = 0; kk < GK; kk += 4) {
            TF av[2], xv[2];
#pragma unroll
            for (int t = 0; t < 2; ++t) {
                av[t] = As[b][(kk + lk) * AP + wm * 32 + t * 16 + lr];
                xv[t] = Xs[b][(kk + lk) * XP + wn * 32 + t * 16 + lr];
            }
#pragma unroll
            for (int ra_ = 0; ra_ < 2; ++ra_)
#pragma unroll
                for (int cb = 0; cb < 2; ++cb) acc[ra_][cb] = Mfma<TF>::mac(av[ra_], xv[cb], acc[ra_][cb]);
        }
        if (c + 1 < nch) put(b ^ 1);
        __syncthreads();
    }
#pragma unroll
    for (int ra_ = 0; ra_ < 2; ++ra_)
#pragma unroll
        for (int cb = 0; cb < 2; ++cb) {
            const int gj = j0 + wn * 32 + cb * 16 + lr;
            if (gj >= N) continue;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int gi = i0 + wm * 32 + ra_ * 16 + Mfma<TF>::row(lk, r);
                if (gi < M) store(gi, gj, acc[ra_][cb][r]);
            }
        }
}

// Flat tile index -> (work item, row tile, column tile).  Launches are padded to a
// multiple of 8 workgroups; consecutive workgroups go to different XCDs, so the
// tiles of one node are gathered onto one XCD (and its L2).
__device__ __forceinline__ bool tile_of(int total, int tm, int tn, int& item, int& ti, int& tj) {
    const int per = gridDim.x / 8;
    const int logical = (blockIdx.x % 8) * per + blockIdx.x / 8;
    if (logical >= total) return false;
    item = logical / (tm * tn);
    const int rem = logical % (tm * tn);
    ti = rem % tm;
    tj = rem / tm;
    return true;
}

// Up-sweep of one height (ulvfactor.jl:42,48-51,74,89 for every node of the level),
// one workgroup per node and TN-column tile, in two phases:
//   zloc = Gz x                   (stored to y and kept in LDS)
//   [reduced b; u] = Gr [x; zloc]
// x = the leaf's rows of B, or the children's reduced b and u (rows nk.. of their y).
// Dynamic LDS: zloc, (max nk of the level) x (TN + 1).
template <typename TF>
__global__ __launch_bounds__(256) void k_ulv_gup(const SInst<TF>* __restrict__ insts,
                                                 const SweepRec* __restrict__ work, int nwork, int tn) {
    constexpr int WM = 4, WN = 1, TM = 32 * WM, TN = 32 * WN, ZP = TN + 1;
    int item, ti, tj;
    if (!tile_of(nwork * tn, 1, tn, item, ti, tj)) return;
    const SweepRec rc = work[item];
    const SInst<TF>& I = insts[rc.inst];
    const int nrhs = I.nrhs;
    const int j0 = tj * TN;
    if (j0 >= nrhs) return;
    extern __shared__ unsigned char lds_raw[];
    TF* zs = reinterpret_cast<TF*>(lds_raw);
    const int My = rc.My, nk = rc.nk, Kx = rc.K, R = My - nk;
    SegOp x;
    x.p2 = nullptr;
    x.ld2 = 0;
    if (rc.leaf) {
        x.p0 = I.B + rc.x0;
        x.ld0 = I.ldb;
        x.dbl0 = 1;
        x.r0 = Kx;
        x.p1 = nullptr;
        x.ld1 = 0;
        x.dbl1 = 0;
    } else {
        x.p0 = I.wy + rc.x0 * nrhs + rc.add0;
        x.ld0 = rc.l0;
        x.dbl0 = 0;
        x.r0 = rc.r0;
        x.p1 = I.wy + rc.x1 * nrhs + rc.add1;
        x.ld1 = rc.l1;
        x.dbl1 = 0;
    }
    x.r1 = Kx;
    const gptr<TF> y = gbl(I.wy + rc.y * nrhs);
    __shared__ TileLds<TF, WM, WN> L;
    // phase 0 rows [0, nk) of y from Gz, phase 1 rows [nk, My) from Gr (stored after Gz)
    for (int ph = 0; ph < 2; ++ph) {
        const int M = ph ? R : nk, K = ph ? Kx + nk : Kx;
        const TF* A = I.fac + rc.A + (ph ? (int64_t)nk * Kx : 0);
        const gptr<TF> yo = y + (ph ? nk : 0);
        for (int p0 = 0; p0 < M; p0 += TM)
            tile_gemm<TF, WM, WN>(L, M, nrhs, K, A, M, x, zs, ZP, p0, j0, [&](int i, int j, TF v) {
                yo[i + (int64_t)j * My] = v;
                if (!ph) zs[i * ZP + (j - j0)] = v;
            });
        __syncthreads();
    }
}

// Down-sweep of one depth: z = Qt [zloc; z from the parent] (ulvfactor.jl:98-107);
// leaves write the solution X.
template <typename TF, int WM, int WN>
__global__ __launch_bounds__(256) void k_ulv_gdown(const SInst<TF>* __restrict__ insts,
                                                   const SweepRec* __restrict__ work, int nwork, int tm, int tn) {
    int item, ti, tj;
    if (!tile_of(nwork * tm * tn, tm, tn, item, ti, tj)) return;
    const SweepRec rc = work[item];
    const SInst<TF>& I = insts[rc.inst];
    const int nrhs = I.nrhs, n = rc.K;
    const int i0 = ti * 32 * WM, j0 = tj * 32 * WN;
    if (i0 >= n || j0 >= nrhs) return;
    SegOp x;
    x.p2 = nullptr;
    x.ld2 = 0;
    x.p0 = I.wy + rc.x0 * nrhs;
    x.ld0 = rc.My;
    x.dbl0 = 0;
    x.r0 = rc.nk;
    x.p1 = I.wz + rc.x1 * nrhs + rc.add1;
    x.ld1 = rc.l1;
    x.dbl1 = 0;
    x.r1 = n;
    const TF* Q = rc.A >= 0 ? I.fac + rc.A : nullptr;
    __shared__ TileLds<TF, WM, WN> L;
    const bool leaf = rc.leaf;
    const gptr<double> X = gbl(I.X + rc.y);
    const int64_t ldx = I.ldx;
    const gptr<TF> z = gbl(I.wz + rc.y * nrhs);
    tile_gemm<TF, WM, WN>(L, n, nrhs, n, Q, n, x, nullptr, 0, i0, j0, [&](int i, int j, TF v) {
        if (leaf) X[i + (int64_t)j * ldx] = (double)v;
        else z[i + (int64_t)j * n] = v;
    });
}
// ---- solve sweeps for nodes whose operators fit in LDS ----
// Upper tree levels have small operators and few nodes: a streamed K loop would pay
// one memory latency per chunk.  These variants stage the node's whole operator and
// input into LDS with batched 16-byte loads (one or two latencies), then run every
// MFMA from LDS.  Column tile TN = 32; the 4 waves share the 16 x 16 output blocks.
constexpr int XS = 33;  // LDS pitch of the staged input (32 columns + 1)

// out(i, j) = sum_{k < K} A(i, k) xs(k, j), A column-major in LDS (lda = M); the
// (16-row, 16-column) blocks of an M x 32 tile are dealt to the waves, each block
// accumulated in two independent MFMA chains (even and odd k-steps).
template <typename TF, typename Store>
__device__ __forceinline__ void lds_gemm(int M, int K, const TF* As, const TF* xs, int ncols, Store store,
                                         int xp = XS) {
    using V4 = typename Mfma<TF>::v4;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    const int lr = lane & 15, lk = lane >> 4;
    const int nrb = (M + 15) / 16, ncb = (ncols + 15) / 16;
    for (int t = wave; t < nrb * ncb; t += nw) {
        const int rb = t % nrb, cb = t / nrb;
        const int i = rb * 16 + lr, jc = cb * 16 + lr;
        auto av = [&](int kk) { return (i < M && kk < K) ? As[i + kk * M] : TF(0); };
        auto xv = [&](int kk) { return kk < K ? xs[kk * xp + jc] : TF(0); };
        V4 acc0 = V4{TF(0), TF(0), TF(0), TF(0)}, acc1 = acc0;
        int k = 0;
        for (; k + 4 < K; k += 8) {
            const TF a0 = av(k + lk), b0 = xv(k + lk), a1 = av(k + 4 + lk), b1 = xv(k + 4 + lk);
            acc0 = Mfma<TF>::mac(a0, b0, acc0);
            acc1 = Mfma<TF>::mac(a1, b1, acc1);
        }
        if (k < K) acc0 = Mfma<TF>::mac(av(k + lk), xv(k + lk), acc0);
        if (jc < ncols) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int ii = rb * 16 + Mfma<TF>::row(lk, r);
                if (ii < M) store(ii, jc, acc0[r] + acc1[r]);
            }
        }
    }
}

// Strip products: acc += A(16-row strip rb, 0:K) xs(0:K, 0:16 or 0:32), A column-major
// (lda) in LDS (AL) or in memory, xs row-major in LDS (pitch xp).  A strip covers all
// (one or two) 16-column blocks, so every A element is read by one wave and feeds one
// MFMA per column block; with one block the k-steps alternate between acc0 and acc1.
// TA (A in memory only): the strip's rows are columns of A, i.e. the product is A' xs with
// A (K x M) column-major (lda).
// A in memory is loaded straight into registers U k-steps ahead of the MFMAs; the loads
// are unconditional (columns past K are clamped in bounds and k past K is zeroed at the
// MFMA), so no wait is forced before a value is used.  Rows past M read row M - 1 and
// are never stored.
template <bool AL, bool two, typename TF, bool TA = false>
__device__ __forceinline__ void strip_acc(int M, int K, const TF* __restrict__ A, int64_t lda, const TF* xs, int xp,
                                          int rb, typename Mfma<TF>::v4& acc0, typename Mfma<TF>::v4& acc1) {
    static_assert(!(AL && TA), "a transposed A is read from memory only");
    if (K <= 0) return;
    const int lane = threadIdx.x & 63, lr = lane & 15, lk = lane >> 4;
    const int i = rb * 16 + lr;
    const int ic = i < M ? i : M - 1;
    const int S = (K + 3) / 4, Sf = K / 4;  // k-steps; steps [0, Sf) lie entirely inside K
    auto mac = [&](int u, TF av, const TF* xk) {
        if (two) {
            acc0 = Mfma<TF>::mac(av, xk[0], acc0);
            acc1 = Mfma<TF>::mac(av, xk[16], acc1);
        } else if (u & 1) {
            acc1 = Mfma<TF>::mac(av, xk[0], acc1);
        } else {
            acc0 = Mfma<TF>::mac(av, xk[0], acc0);
        }
    };
    // a step that may pass K: clamped x row, A value zeroed past K
    auto edge = [&](int u, int st, TF aval) {
        const int kk = 4 * st + lk;
        mac(u, kk < K ? aval : TF(0), xs + (kk < K ? kk : K - 1) * xp + lr);
    };
    if (AL) {
        constexpr int U = 4;
        const TF* Ai = A + ic;
        int st = 0;
        for (; st + U <= Sf; st += U) {
            TF av[U], b0[U], b1[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int kk = 4 * (st + u) + lk;
                av[u] = Ai[(int64_t)kk * lda];
                b0[u] = xs[kk * xp + lr];
                if (two) b1[u] = xs[kk * xp + 16 + lr];
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                if (two) {
                    acc0 = Mfma<TF>::mac(av[u], b0[u], acc0);
                    acc1 = Mfma<TF>::mac(av[u], b1[u], acc1);
                } else if (u & 1) {
                    acc1 = Mfma<TF>::mac(av[u], b0[u], acc1);
                } else {
                    acc0 = Mfma<TF>::mac(av[u], b0[u], acc0);
                }
            }
        }
        for (int u = 0; st < S; ++st, ++u) {
            const int kk = 4 * st + lk;
            edge(u, st, Ai[(int64_t)(kk < K ? kk : K - 1) * lda]);
        }
        return;
    }
    constexpr int U = 8;
    // A through buffer loads: the descriptor and the per-step offset live in scalar
    // registers, the lane's row/column offset in one VGPR, so the prefetch costs two
    // VGPRs per step.  Every access is inside A: the lane's column is clamped to K - 1
    // (full steps never need it) and steps at or past Sf are loaded by the edge code.
    const int kl = lk < K ? lk : K - 1;
    // element (row i, column k) of the strip's operator: A[i + k lda], or A[k + i lda] (TA)
    auto at = [&](int i, int k) { return TA ? i * (int)lda + k : i + k * (int)lda; };
    const int voff = at(ic, kl) * (int)sizeof(TF);
    const int step = 4 * (TA ? 1 : (int)lda) * (int)sizeof(TF);
    const auto rsrc = buf_rsrc(A, (int)((TA ? lda * (M - 1) + K : lda * (K - 1) + M) * (int64_t)sizeof(TF)));
    auto ld = [&](int st) { return buf_load<TF>(rsrc, voff, (st < Sf ? st : (Sf > 0 ? Sf - 1 : 0)) * step); };
    auto ld_edge = [&](int st) {
        const int kk = 4 * st + lk;
        return buf_load<TF>(rsrc, at(ic, kk < K ? kk : K - 1) * (int)sizeof(TF), 0);
    };
    TF a[U];
#pragma unroll
    for (int u = 0; u < U; ++u) a[u] = ld(u);
    int s0 = 0;
    // blocks of U full steps; the two register sets alternate, so no prefetched value is
    // copied (a copy would wait for its load)
    auto blk = [&](TF(&cur)[U], TF(&nxt)[U], int b0s) {
#pragma unroll
        for (int u = 0; u < U; ++u) nxt[u] = ld(b0s + U + u);
        const TF* xb = xs + (4 * b0s + lk) * xp + lr;
        // x from LDS in groups of G steps (registers go to the A prefetch)
        constexpr int G = 4;
#pragma unroll
        for (int g = 0; g < U; g += G) {
            TF b0[G], b1[G];
#pragma unroll
            for (int u = 0; u < G; ++u) {
                b0[u] = xb[4 * (g + u) * xp];
                if (two) b1[u] = xb[4 * (g + u) * xp + 16];
            }
#pragma unroll
            for (int u = 0; u < G; ++u) {
                if (two) {
                    acc0 = Mfma<TF>::mac(cur[g + u], b0[u], acc0);
                    acc1 = Mfma<TF>::mac(cur[g + u], b1[u], acc1);
                } else if (u & 1) {
                    acc1 = Mfma<TF>::mac(cur[g + u], b0[u], acc1);
                } else {
                    acc0 = Mfma<TF>::mac(cur[g + u], b0[u], acc0);
                }
            }
        }
    };
    TF an[U];
    for (; s0 + 2 * U <= Sf; s0 += 2 * U) {
        blk(a, an, s0);
        blk(an, a, s0 + U);
    }
    if (s0 + U <= Sf) {
        blk(a, an, s0);
        s0 += U;
#pragma unroll
        for (int u = 0; u < U; ++u) a[u] = an[u];
    }
    // the remaining steps: full ones from the prefetch, the partial one loaded here
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const int st = s0 + u;
        if (st < S) edge(u, st, st < Sf ? a[u] : ld_edge(st));
    }
}

// store(i, j, v) for the rows i < M of strip rb held by lane ln
template <bool two, typename TF, typename Store>
__device__ __forceinline__ void strip_put(int M, int ncols, int rb, int ln, const typename Mfma<TF>::v4& a0,
                                          const typename Mfma<TF>::v4& a1, Store store) {
    const int lr = ln & 15, lk = ln >> 4;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int ii = rb * 16 + Mfma<TF>::row(lk, r);
        if (ii >= M) continue;
        if (two) {
            store(ii, lr, a0[r]);
            if (16 + lr < ncols) store(ii, 16 + lr, a1[r]);
        } else if (lr < ncols) {
            store(ii, lr, a0[r] + a1[r]);
        }
    }
}

// out = A xs (M x ncols, ncols <= 32) with A in memory, one strip per wave at a time
template <bool two, typename TF, typename Store>
__device__ __forceinline__ void glb_gemm_cb(int M, int K, const TF* __restrict__ A, int64_t lda, const TF* xs,
                                            int ncols, Store store, int xp) {
    using V4 = typename Mfma<TF>::v4;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    for (int rb = wave; rb < (M + 15) / 16; rb += nw) {
        V4 a0 = V4{TF(0), TF(0), TF(0), TF(0)}, a1 = a0;
        strip_acc<false, two>(M, K, A, lda, xs, xp, rb, a0, a1);
        strip_put<two, TF>(M, ncols, rb, lane, a0, a1, store);
    }
}
template <typename TF, typename Store>
__device__ __forceinline__ void glb_gemm(int M, int K, const TF* __restrict__ A, int64_t lda, const TF* xs, int ncols,
                                         Store store, int xp = XS) {
    if (ncols > 16) glb_gemm_cb<true>(M, K, A, lda, xs, ncols, store, xp);
    else glb_gemm_cb<false>(M, K, A, lda, xs, ncols, store, xp);
}

// out = op(A) xs on MFMA (M x ncols, ncols <= 32; K terms): op(A) = A (M x K, lda), or A'
// with A (K x M, lda) when TA; xs row-major in LDS (pitch xp).  No barrier inside.
template <bool TA, typename TF, typename Store>
__device__ __forceinline__ void mfma_gemm(int M, int K, const TF* __restrict__ A, int64_t lda, const TF* xs, int ncols,
                                          Store store, int xp) {
    using V4 = typename Mfma<TF>::v4;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    for (int rb = wave; rb < (M + 15) / 16; rb += nw) {
        V4 a0 = V4{TF(0), TF(0), TF(0), TF(0)}, a1 = a0;
        if (ncols > 16) {
            strip_acc<false, true, TF, TA>(M, K, A, lda, xs, xp, rb, a0, a1);
            strip_put<true, TF>(M, ncols, rb, lane, a0, a1, store);
        } else {
            strip_acc<false, false, TF, TA>(M, K, A, lda, xs, xp, rb, a0, a1);
            strip_put<false, TF>(M, ncols, rb, lane, a0, a1, store);
        }
    }
}

// The two products of a node's up-sweep step with the input x in xs rows [0, Kx):
//   zloc = Gz x           (nk x Kx, lda nk)  -> st0, which also writes xs rows [Kx, Kx + nk)
//   [b; u] = Gr [x; zloc] (R x (Kx + nk), lda R) -> st1
// The strips of Gr start on the x part (its first Kx columns) while other waves compute
// zloc, hold their sums across the barrier that publishes zloc, then add the zloc part.
// Waves [0, n2) own the Gr strips, the Gz strips go to the other waves.  (Unless every
// wave owns a Gr strip; then the two products run one after the other.)  All threads of the
// workgroup call it; it contains a barrier.
template <bool AL, bool two, typename TF, typename S0, typename S1>
__device__ __forceinline__ void up_gemms_cb(int nk, int Kx, int R, const TF* Gz, const TF* Gr, const TF* xs, int xp,
                                            int ncols, S0 st0, S1 st1) {
    using V4 = typename Mfma<TF>::v4;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    const int n1 = (nk + 15) / 16, n2 = (R + 15) / 16;
    const V4 z4 = V4{TF(0), TF(0), TF(0), TF(0)};
    if (n2 >= nw) {
        for (int rb = wave; rb < n1; rb += nw) {
            V4 a0 = z4, a1 = z4;
            strip_acc<AL, two>(nk, Kx, Gz, nk, xs, xp, rb, a0, a1);
            strip_put<two, TF>(nk, ncols, rb, lane, a0, a1, st0);
        }
        __syncthreads();
        for (int rb = wave; rb < n2; rb += nw) {
            V4 a0 = z4, a1 = z4;
            strip_acc<AL, two>(R, Kx + nk, Gr, R, xs, xp, rb, a0, a1);
            strip_put<two, TF>(R, ncols, rb, lane, a0, a1, st1);
        }
        return;
    }
    // Gz strips go to the waves that own no Gr strip (round robin over them), so the
    // MFMA work of each SIMD's waves stays balanced
    const int nf = nw - n2;
    for (int rb = wave - n2; wave >= n2 && rb < n1; rb += nf) {
        V4 a0 = z4, a1 = z4;
        strip_acc<AL, two>(nk, Kx, Gz, nk, xs, xp, rb, a0, a1);
        strip_put<two, TF>(nk, ncols, rb, lane, a0, a1, st0);
    }
    V4 b0 = z4, b1 = z4;
    if (wave < n2) strip_acc<AL, two>(R, Kx, Gr, R, xs, xp, wave, b0, b1);
    __syncthreads();  // zloc is in xs rows [Kx, Kx + nk)
    if (wave < n2) {
        strip_acc<AL, two>(R, nk, Gr + (int64_t)Kx * R, R, xs + (int64_t)Kx * xp, xp, wave, b0, b1);
        strip_put<two, TF>(R, ncols, wave, lane, b0, b1, st1);
    }
}
template <bool AL, typename TF, typename S0, typename S1>
__device__ __forceinline__ void up_gemms(int nk, int Kx, int R, const TF* Gz, const TF* Gr, const TF* xs, int xp,
                                         int ncols, S0 st0, S1 st1) {
    if (ncols > 16) up_gemms_cb<AL, true>(nk, Kx, R, Gz, Gr, xs, xp, ncols, st0, st1);
    else up_gemms_cb<AL, false>(nk, Kx, R, Gz, Gr, xs, xp, ncols, st0, st1);
}

// Up-sweep with the input staged in LDS; tw = columns per workgroup (16 or 32), xs pitch
// tw + 1.  AR: the operator [Gz Gr] is staged too (LDS = [Gz Gr] (nk Kx + R (Kx + nk),
// rounded to 16) | xs (Kx + nk) rows); otherwise the waves stream it from memory
// (LDS = xs), for operators too large to stage whose rows fill few 128-row tiles.
template <typename TF, bool AR, int NT>
__global__ __launch_bounds__(NT, 4) void k_ulv_gup_s(const SInst<TF>* __restrict__ insts,
                                                   const SweepRec* __restrict__ work, int nwork, int tn, int tw) {
    int item, ti, tj;
    if (!tile_of(nwork * tn, 1, tn, item, ti, tj)) return;
    const SweepRec rc = work[item];
    const SInst<TF>& I = insts[rc.inst];
    const int nrhs = I.nrhs;
    const int j0 = tj * tw;
    if (j0 >= nrhs) return;
    const int nc = min(tw, nrhs - j0);
    const int xp = tw + 1;  // LDS pitch of xs
    const int nk = rc.nk, Kx = rc.K, My = rc.My, R = My - nk;
    const int64_t na = (int64_t)nk * Kx + (int64_t)R * (Kx + nk);
    extern __shared__ unsigned char lds_raw[];
    TF* As = reinterpret_cast<TF*>(lds_raw);
    TF* xs = AR ? As + ((na + 15) & ~int64_t(15)) : As;
    const int64_t nst = AR ? na : 0;  // operator elements staged
    // operator and x rows [0, Kx) (the leaf's rows of B, or the children's reduced b and u)
    // in one round of loads; columns past nc are zero
    const TF* Ag = I.fac + rc.A;  // Gz, then Gr
    auto put = [&](int e, TF v) { xs[(e % Kx) * xp + e / Kx] = v; };
    if (rc.leaf) {
        const gptr<const double> B = gbl(I.B + rc.x0 + (int64_t)j0 * I.ldb);
        const int64_t ldb = I.ldb;
        stage2(
            As, Ag, nst, Kx * tw, [&](int e) { return e / Kx < nc ? (TF)B[(e % Kx) + (int64_t)(e / Kx) * ldb] : TF(0); },
            put);
    } else {
        const int r0 = rc.r0, l0 = rc.l0, l1 = rc.l1;
        const gptr<const TF> y0 = gbl(I.wy + rc.x0 * nrhs + rc.add0 + (int64_t)j0 * l0);
        const gptr<const TF> y1 = gbl(I.wy + rc.x1 * nrhs + rc.add1 + (int64_t)j0 * l1);
        stage2(
            As, Ag, nst, Kx * tw,
            [&](int e) {
                const int k = e % Kx, j = e / Kx;
                if (j >= nc) return TF(0);
                return k < r0 ? y0[k + (int64_t)j * l0] : y1[(k - r0) + (int64_t)j * l1];
            },
            put);
    }
    __syncthreads();
    const gptr<TF> y = gbl(I.wy + rc.y * nrhs + (int64_t)j0 * My);
    // zloc = Gz x  -> y rows [0, nk) and xs rows [Kx, Kx + nk)
    auto st0 = [&](int i, int j, TF v) {
        y[i + (int64_t)j * My] = v;
        xs[(Kx + i) * xp + j] = v;
    };
    // [reduced b; u] = Gr [x; zloc]  -> y rows [nk, My)
    auto st1 = [&](int i, int j, TF v) { y[nk + i + (int64_t)j * My] = v; };
    const TF* Gz = AR ? As : Ag;
    up_gemms<AR>(nk, Kx, R, Gz, Gz + (int64_t)nk * Kx, xs, xp, nc, st0, st1);
}

// Down-sweep with the input staged in LDS.  AR: Qt is staged too (LDS = Qt (n x n,
// rounded to 16; empty for a pass-through node) | xs n rows); otherwise streamed.
template <typename TF, bool AR, int NT>
__global__ __launch_bounds__(NT, 4) void k_ulv_gdown_s(const SInst<TF>* __restrict__ insts,
                                                     const SweepRec* __restrict__ work, int nwork, int tn, int tw) {
    int item, ti, tj;
    if (!tile_of(nwork * tn, 1, tn, item, ti, tj)) return;
    const SweepRec rc = work[item];
    const SInst<TF>& I = insts[rc.inst];
    const int nrhs = I.nrhs, n = rc.K, nk = rc.nk;
    const int j0 = tj * tw;
    if (j0 >= nrhs) return;
    const int nc = min(tw, nrhs - j0);
    const int xp = tw + 1;  // LDS pitch of xs
    const int64_t na = rc.A >= 0 ? (int64_t)n * n : 0;
    extern __shared__ unsigned char lds_raw[];
    TF* As = reinterpret_cast<TF*>(lds_raw);
    TF* xs = AR ? As + ((na + 15) & ~int64_t(15)) : As;
    // Qt and x = [zloc (rows 0..nk-1 of the node's y); z from the parent] in one round
    const int l0 = rc.My, l1 = rc.l1;
    const gptr<const TF> y0 = gbl(I.wy + rc.x0 * nrhs + (int64_t)j0 * l0);
    const gptr<const TF> z1 = gbl(I.wz + rc.x1 * nrhs + rc.add1 + (int64_t)j0 * l1);
    stage2(
        As, na > 0 ? I.fac + rc.A : I.fac, AR ? na : 0, n * tw,
        [&](int e) {
            const int k = e % n, j = e / n;
            if (j >= nc) return TF(0);
            return k < nk ? y0[k + (int64_t)j * l0] : z1[(k - nk) + (int64_t)j * l1];
        },
        [&](int e, TF v) { xs[(e % n) * xp + e / n] = v; });
    __syncthreads();
    const bool leaf = rc.leaf;
    const gptr<double> X = gbl(I.X + rc.y + (int64_t)j0 * I.ldx);
    const int64_t ldx = I.ldx;
    const gptr<TF> z = gbl(I.wz + rc.y * nrhs + (int64_t)j0 * n);
    auto store = [&](int i, int j, TF v) {
        if (leaf) X[i + (int64_t)j * ldx] = (double)v;
        else z[i + (int64_t)j * n] = v;
    };
    if (na > 0) {
        if (AR) lds_gemm(n, n, As, xs, nc, store, xp);
        else glb_gemm(n, n, I.fac + rc.A, n, xs, nc, store, xp);
    } else {
        for (int e = threadIdx.x; e < n * nc; e += blockDim.x) store(e % n, e / n, xs[(e % n) * xp + e / n]);
    }
}

// Root solve by inverse and one refinement step (ulvfactor.jl:74,83): with the merge
// b = Mb x (x = the children's reduced b and u), z0 = Dinv b, r = b - D z0,
// z = z0 + Dinv r.  One refinement step restores the accuracy of the LU solve.
// ST (all staged): LDS = [Dinv D Mb] (2 m^2 + m Kx, rounded to 16; contiguous in the
// factor) | xs max(Kx, 2m) x XS | bm m x XS, everything loaded in one round;
// otherwise LDS = Dinv (m x m, rounded to 16) | xs | bm and D, Mb are read from memory.
template <typename TF, bool ST>
__global__ __launch_bounds__(512) void k_ulv_groot2(const SInst<TF>* __restrict__ insts,
                                                    const SweepRec* __restrict__ work, int tn) {
    const int item = blockIdx.x / tn, tj = blockIdx.x % tn;
    const SweepRec wi = work[item];
    const SInst<TF>& I = insts[wi.inst];
    const int nrhs = I.nrhs;
    const int j0 = tj * 32;
    if (j0 >= nrhs) return;
    const int nc = min(32, nrhs - j0);
    const DevNode nd = gbl(I.nodes)[wi.node];
    const FNode f = gbl(I.fn)[wi.node];
    const int m = f.m, Kx = f.Kx;
    const int64_t mm = (int64_t)m * m;
    extern __shared__ unsigned char lds_raw[];
    TF* Di = reinterpret_cast<TF*>(lds_raw);
    const int64_t nops = ST ? 2 * mm + (int64_t)m * Kx : mm;
    TF* xs = Di + ((nops + 15) & ~int64_t(15));
    TF* bm = xs + max(Kx, 2 * m) * XS;
    lds_dma(Di, I.fac + f.Dinv, nops);  // Dinv, then (ST) D and Mb, which follow it
    const FNode fl = gbl(I.fn)[nd.left], fr = gbl(I.fn)[nd.right];
    const int r0 = fl.My - fl.nk;
    const gptr<const TF> y0 = gbl(I.wy + fl.yrow * nrhs + fl.nk + (int64_t)j0 * fl.My);
    const gptr<const TF> y1 = gbl(I.wy + fr.yrow * nrhs + fr.nk + (int64_t)j0 * fr.My);
    const int l0 = fl.My, l1 = fr.My;
    batched<16, TF>(
        Kx * 32,
        [&](int e) {
            const int k = e % Kx, j = e / Kx;
            if (j >= nc) return TF(0);
            return k < r0 ? y0[k + (int64_t)j * l0] : y1[(k - r0) + (int64_t)j * l1];
        },
        [&](int e, TF v) { xs[(e % Kx) * XS + e / Kx] = v; });
    wait_all();
    __syncthreads();
    auto st_b = [&](int i, int j, TF v) { bm[i * XS + j] = v; };
    if (ST) lds_gemm(m, Kx, Di + 2 * mm, xs, nc, st_b);
    else glb_gemm(m, Kx, I.fac + f.Mb, m, xs, nc, st_b);
    __syncthreads();
    lds_gemm(m, m, Di, bm, nc, [&](int i, int j, TF v) { xs[i * XS + j] = v; });  // z0
    __syncthreads();
    auto st_r = [&](int i, int j, TF v) { xs[(m + i) * XS + j] = bm[i * XS + j] - v; };  // r
    if (ST) lds_gemm(m, m, Di + mm, xs, nc, st_r);
    else glb_gemm(m, m, I.fac + f.Dr, m, xs, nc, st_r);
    __syncthreads();
    const gptr<TF> z = gbl(I.wz + f.zrow * nrhs + (int64_t)j0 * f.n);
    lds_gemm(m, m, Di, xs + m * XS, nc, [&](int i, int j, TF v) { z[i + (int64_t)j * f.n] = xs[i * XS + j] + v; });
}

// ---------------------------------------------------------------------------
// matvec (reference src/matmul.jl:18-62) over per-node operators
// ---------------------------------------------------------------------------
// Each sweep step of a node is one product y = A [x_0; x_1] (A = the node's operator):
//   up   (per height):  Z = V' B_seg (leaf) | [W_1' W_2'] [Z_1; Z_2]       (_matmatup, matmul.jl:32-42)
//   down (per depth):   F = [B_sibling R] [Z_sibling; F_parent]             (_matmatdown!, matmul.jl:51-57)
//   leaf:               C_seg = alpha [D U] [B_seg; F] + beta C_seg         (matmul.jl:46-47)
//                       or, down step folded in, alpha [D, U B_sibling, U R] [B_seg; Z_sibling; F_parent]
struct MvOp {
    const double* A;
    int M, K;
    SegOp x;
    double* y;
    int64_t ldy;
};

// segment pointers of one matvec record (Z / F bases scaled by nrhs)
template <int KIND>
__device__ __forceinline__ void mv_op(const MvInst& I, const MvRec& rc, MvOp& o) {
    const int nrhs = I.nrhs;
    o.A = I.op + rc.A;
    o.M = rc.M;
    o.K = rc.K;
    o.x.dbl0 = o.x.dbl1 = 1;
    o.x.r0 = rc.r0;
    o.x.r1 = rc.r1;
    if (rc.from_b) {
        o.x.p0 = I.B + rc.x0;
        o.x.ld0 = I.ldb;
    } else {
        o.x.p0 = I.Z + rc.x0 * nrhs;
        o.x.ld0 = rc.ld0;
    }
    o.x.p1 = (KIND == 1 ? I.F : I.Z) + rc.x1 * nrhs;  // a leaf's rows are Z-based (F follows Z)
    o.x.ld1 = rc.ld1;
    o.x.p2 = (KIND == 2 && rc.r1 < rc.K) ? I.Z + rc.x2 * nrhs : nullptr;
    o.x.ld2 = rc.ld2;
    if (KIND == 2) {
        o.y = I.C + rc.y;
        o.ldy = I.ldc;
    } else {
        o.y = (KIND == 0 ? I.Z : I.F) + rc.y * nrhs;
        o.ldy = rc.ldy;
    }
}

template <int KIND>
__device__ __forceinline__ void mv_store(const MvInst& I, const MvOp& o, int i, int j, double v) {
    const gptr<double> c = gbl(o.y + i + (int64_t)j * o.ldy);
    if (KIND == 2) *c = (I.beta == 0.0) ? I.alpha * v : I.alpha * v + I.beta * (*c);
    else *c = v;
}

// One matvec step per (node, tw-column tile) with the input staged in LDS (xs, K rows,
// pitch tw + 1).  AR: the operator is staged too (LDS = A (M x K, rounded to 16) | xs);
// otherwise the waves stream it from memory (LDS = xs).  Strip products as in the solve.
template <int KIND, bool AR, int NT>
__global__ __launch_bounds__(NT, 4) void k_mv_s(const MvInst* __restrict__ insts, const MvRec* __restrict__ work,
                                              int nwork, int tn, int tw) {
    int item, ti, tj;
    if (!tile_of(nwork * tn, 1, tn, item, ti, tj)) return;
    const MvRec rc = work[item];
    const MvInst& I = insts[rc.inst];
    const int j0 = tj * tw;
    if (j0 >= I.nrhs) return;
    const int nc = min(tw, I.nrhs - j0);
    const int xp = tw + 1;  // LDS pitch of xs
    MvOp o;
    mv_op<KIND>(I, rc, o);
    const int M = o.M, K = o.K;
    const int64_t na = (int64_t)M * K;
    extern __shared__ unsigned char lds_raw[];
    double* As = reinterpret_cast<double*>(lds_raw);
    double* xs = AR ? As + ((na + 15) & ~int64_t(15)) : As;
    const int r0 = o.x.r0;
    const gptr<const double> x0 = gbl(static_cast<const double*>(o.x.p0) + (int64_t)j0 * o.x.ld0);
    const gptr<const double> x1 = gbl(static_cast<const double*>(o.x.p1) + (int64_t)j0 * o.x.ld1);
    const gptr<const double> x2 = gbl(static_cast<const double*>(o.x.p2) + (int64_t)j0 * o.x.ld2);
    const int64_t l0 = o.x.ld0, l1 = o.x.ld1, l2 = o.x.ld2;
    const int r1 = o.x.r1;
    stage2(
        As, o.A, AR ? na : 0, K * tw,
        [&](int e) {
            const int k = e % K, j = e / K;
            if (j >= nc) return 0.0;
            return k < r0 ? x0[k + j * l0] : k < r1 ? x1[(k - r0) + j * l1] : x2[(k - r1) + j * l2];
        },
        [&](int e, double v) { xs[(e % K) * xp + e / K] = v; });
    __syncthreads();
    auto store = [&](int i, int j, double v) { mv_store<KIND>(I, o, i, j0 + j, v); };
    using V4 = typename Mfma<double>::v4;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    const double* Ap = AR ? As : o.A;
    for (int rb = wave; rb < (M + 15) / 16; rb += nw) {
        V4 a0 = V4{0.0, 0.0, 0.0, 0.0}, a1 = a0;
        if (nc > 16) {
            strip_acc<AR, true>(M, K, Ap, M, xs, xp, rb, a0, a1);
            strip_put<true, double>(M, nc, rb, lane, a0, a1, store);
        } else {
            strip_acc<AR, false>(M, K, Ap, M, xs, xp, rb, a0, a1);
            strip_put<false, double>(M, nc, rb, lane, a0, a1, store);
        }
    }
}

// Streamed through the chunked MFMA tile GEMM (nodes whose inputs do not fit in LDS).
template <int KIND, int WM, int WN>
__global__ __launch_bounds__(256) void k_mv_t(const MvInst* __restrict__ insts, const MvRec* __restrict__ work,
                                              int nwork, int tm, int tn) {
    int item, ti, tj;
    if (!tile_of(nwork * tm * tn, tm, tn, item, ti, tj)) return;
    const MvRec rc = work[item];
    const MvInst& I = insts[rc.inst];
    const int i0 = ti * 32 * WM, j0 = tj * 32 * WN;
    MvOp o;
    mv_op<KIND>(I, rc, o);
    if (i0 >= o.M || j0 >= I.nrhs) return;
    __shared__ TileLds<double, WM, WN> L;
    tile_gemm<double, WM, WN>(L, o.M, I.nrhs, o.K, o.A, o.M, o.x, nullptr, 0, i0, j0,
                              [&](int i, int j, double v) { mv_store<KIND>(I, o, i, j, v); });
}

// ---------------------------------------------------------------------------
// tree cut: one dense product in place of the sweep steps above a depth
// ---------------------------------------------------------------------------
struct CutArgs {
    int row0, rows, K, lda;
    int64_t T, xrow, yrow;
};
__device__ __forceinline__ CutArgs cut_args(const MvRec& r) { return CutArgs{r.r0, r.M, r.K, r.ld0, r.A, r.x0, r.x1}; }
__device__ __forceinline__ double* cut_ws(const MvInst& I) { return I.Z; }

constexpr int CUT_TW = 16;       // right-hand-side columns per workgroup
constexpr int CUT_THREADS = 512;
constexpr int CUT_MAXK = 1024;   // largest K (the stacked input is held in LDS)
constexpr int CUT_MAXS = (CUT_MAXK / 4 + CUT_THREADS / 64 - 1) / (CUT_THREADS / 64);  // k-steps per wave
constexpr int CUT_RPT = (CUT_MAXK + CUT_THREADS - 1) / CUT_THREADS;                   // input rows per thread

// One workgroup per (16-row strip of T, 16-column tile).  Every wave owns a contiguous
// share of the k-steps of the strip: its slice of T goes into registers at once, in the
// same round of loads as the row tables, then the stacked input (K x 16) is gathered into
// LDS through the input row table; after the barrier each wave accumulates its share on
// MFMA and the partial sums meet in LDS.  The record carries T and both row tables, so
// the loads start from the record alone.  LDS: 1024 x 17 input (rows past K unused), then
// 8 waves x 256 sums.
template <typename TF, typename Inst, typename Rec>
__global__ __launch_bounds__(CUT_THREADS, 1) void k_cut(const Inst* __restrict__ insts, const Rec* __restrict__ work,
                                                     int nwork, int tn) {
    int item, ti, tj;
    if (!tile_of(nwork * tn, 1, tn, item, ti, tj)) return;
    const Rec rc = work[item];
    const CutArgs c = cut_args(rc);
    const int K = c.K;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    const int lr = lane & 15, lk = lane >> 4;
    using V4 = typename Mfma<TF>::v4;
    // this wave's k-steps [s0, s0 + ns) of the strip; lane (lr, lk) holds row lr, column
    // 4 s + lk of each (rows past the strip repeat its last row and are never stored)
    const int S = (K + 3) / 4;
    const int s0 = S * wave / nw, ns = S * (wave + 1) / nw - s0;
    const gptr<const TF> T = gbl(reinterpret_cast<const TF*>(c.T) + c.row0 + (lr < c.rows ? lr : c.rows - 1));
    // Every global load below is unconditional, its index clamped into range, and no branch
    // depends on its value, so the loads of one round are all in flight at once; the row
    // table goes first, as the input gather waits for it.  Steps past the wave's share or
    // past K get a zero x row at the MFMA instead; columns past nrhs repeat the last one and
    // are never stored.
    const gptr<const CutRow> xrow = gbl(reinterpret_cast<const CutRow*>(c.xrow));
    CutRow xr[CUT_RPT];
#pragma unroll
    for (int q = 0; q < CUT_RPT; ++q) xr[q] = xrow[min((int)threadIdx.x + q * CUT_THREADS, K - 1)];
    TF a[CUT_MAXS];
#pragma unroll
    for (int u = 0; u < CUT_MAXS; ++u) a[u] = T[(int64_t)min(4 * (s0 + u) + lk, K - 1) * c.lda];
    const gptr<const CutRow> yrow = gbl(reinterpret_cast<const CutRow*>(c.yrow));
    CutRow orow[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) orow[r] = yrow[c.row0 + min(Mfma<TF>::row(lk, r), c.rows - 1)];
    const Inst& I = insts[rc.inst];
    const int nrhs = I.nrhs, j0 = tj * CUT_TW;
    if (j0 >= nrhs) return;  // the same for every thread of the workgroup
    const int nc = min(CUT_TW, nrhs - j0);
    const gptr<TF> W = gbl(cut_ws(I));
    constexpr int xp = CUT_TW + 1;
    extern __shared__ unsigned char lds_raw[];
    TF* xs = reinterpret_cast<TF*>(lds_raw);  // CUT_RPT * CUT_THREADS rows (those past K unused)
    TF* red = xs + CUT_RPT * CUT_THREADS * xp;
    TF v[CUT_RPT][CUT_TW];
#pragma unroll
    for (int q = 0; q < CUT_RPT; ++q) {
        const int64_t base = xr[q].off * nrhs + xr[q].inner + (int64_t)j0 * xr[q].ld;
#pragma unroll
        for (int j = 0; j < CUT_TW; ++j) v[q][j] = W[base + (int64_t)min(j, nc - 1) * xr[q].ld];
    }
#pragma unroll
    for (int q = 0; q < CUT_RPT; ++q)
#pragma unroll
        for (int j = 0; j < CUT_TW; ++j) xs[(threadIdx.x + q * CUT_THREADS) * xp + j] = v[q][j];
    __syncthreads();
    // groups of 8 k-steps: the group's LDS reads first, then its MFMAs
    V4 a0 = V4{TF(0), TF(0), TF(0), TF(0)}, a1 = a0;
#pragma unroll
    for (int g = 0; g < CUT_MAXS; g += 8) {
        if (g < ns) {
            TF b[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int k = 4 * (s0 + g + u) + lk;
                const TF t = xs[min(k, K - 1) * xp + lr];
                b[u] = (g + u < ns && k < K) ? t : TF(0);
            }
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                if (u & 1) a1 = Mfma<TF>::mac(a[g + u], b[u], a1);
                else a0 = Mfma<TF>::mac(a[g + u], b[u], a0);
            }
        }
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) red[wave * 256 + r * 64 + lane] = a0[r] + a1[r];
    __syncthreads();
    if (wave != 0) return;
    V4 sum = V4{TF(0), TF(0), TF(0), TF(0)};
    for (int w = 0; w < nw; ++w)
#pragma unroll
        for (int r = 0; r < 4; ++r) sum[r] += red[w * 256 + r * 64 + lane];
    if (lr >= nc) return;
#pragma unroll
    for (int r = 0; r < 4; ++r)
        if (Mfma<TF>::row(lk, r) < c.rows)
            W[orow[r].off * nrhs + orow[r].inner + (int64_t)(j0 + lr) * orow[r].ld] = sum[r];
}

// The identity seed of a cut's build (nrhs columns): row k < K of a stacked input gets a 1
// in column col0 + k.
template <typename TF>
__global__ void k_cut_seed(TF* __restrict__ W, const CutRow* __restrict__ map, int K, int64_t nrhs, int col0) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= K) return;
    const CutRow r = map[k];
    W[r.off * nrhs + r.inner + (int64_t)(col0 + k) * r.ld] = TF(1);
}

// T (M x cols, column-major) = columns [0, cols) of the stacked rows of a cut's build
// (nrhs columns).
template <typename TF>
__global__ void k_cut_gather(TF* __restrict__ T, int M, int cols, const CutRow* __restrict__ map,
                             const TF* __restrict__ W, int64_t nrhs) {
    const int64_t total = (int64_t)M * cols;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        const int g = (int)(e % M);
        const int64_t k = e / M;
        const CutRow r = map[g];
        T[e] = W[r.off * nrhs + r.inner + k * r.ld];
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

// Batched refinement helpers: block (request, column).
// r = b - r (r holds A x on entry); norms[2*(req*maxc + col)] = ||b||^2, [+1] = ||r||^2
__global__ void k_resid_batch(const BatchVec* __restrict__ tab, int maxc, double* __restrict__ norms) {
    const BatchVec v = tab[blockIdx.x];
    const int j = blockIdx.y;
    if (j >= v.nrhs) return;
    double sb = 0.0, sr = 0.0;
    for (int64_t i = threadIdx.x; i < v.n; i += blockDim.x) {
        const double bi = v.b[i + j * v.ldb];
        const double ri = bi - v.r[i + j * v.n];
        v.r[i + j * v.n] = ri;
        sb += bi * bi;
        sr += ri * ri;
    }
    __shared__ double red[2][16];
    sb = wave_reduce(sb);
    sr = wave_reduce(sr);
    if ((threadIdx.x & 63) == 0) {
        red[0][threadIdx.x >> 6] = sb;
        red[1][threadIdx.x >> 6] = sr;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double tb = 0.0, trr = 0.0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) {
            tb += red[0][w];
            trr += red[1][w];
        }
        norms[2 * ((int64_t)blockIdx.x * maxc + j)] = tb;
        norms[2 * ((int64_t)blockIdx.x * maxc + j) + 1] = trr;
    }
}

// x += d for active requests
__global__ void k_axpy_batch(const BatchVec* __restrict__ tab) {
    const BatchVec v = tab[blockIdx.x];
    const int j = blockIdx.y;
    if (j >= v.nrhs) return;
    for (int64_t i = threadIdx.x; i < v.n; i += blockDim.x) v.x[i + j * v.ldx] += v.d[i + j * v.n];
}

// ---------------------------------------------------------------------------
// host-side launchers
// ---------------------------------------------------------------------------
static inline int rhs_tiles(int nrhs) { return (nrhs + RHS_TILE - 1) / RHS_TILE; }

template <typename T>
static hipError_t set_lds_attr(const void* fn, size_t lds) {
    if (lds > 65536) return hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    return hipSuccess;
}

template <typename T, int NT>
static hipError_t launch_factor_nt(const FInst<T>* insts, const WorkItem* work, int nwork, size_t lds,
                                   hipStream_t st) {
    const void* fn = reinterpret_cast<const void*>(&k_ulv_factor<T, FACT_CPT, NT>);
    hipError_t e = set_lds_attr<T>(fn, lds);
    if (e != hipSuccess) return e;
    klaunch((k_ulv_factor<T, FACT_CPT, NT>), dim3(nwork), dim3(NT), lds, st, insts, work, (int64_t)(lds / sizeof(T)));
    return hipGetLastError();
}
template <typename T, int NTILE, int NT>
static hipError_t launch_factor_mf_nt(const FInst<T>* insts, const WorkItem* work, int nwork, size_t lds,
                                      hipStream_t st) {
    const void* fn = reinterpret_cast<const void*>(&k_ulv_factor_mf<T, NTILE, NT>);
    hipError_t e = set_lds_attr<T>(fn, lds);
    if (e != hipSuccess) return e;
    klaunch((k_ulv_factor_mf<T, NTILE, NT>), dim3(nwork), dim3(NT), lds, st, insts, work, (int64_t)(lds / sizeof(T)));
    return hipGetLastError();
}
template <typename T, int NTILE>
static hipError_t launch_factor_mf_tile(const FInst<T>* insts, const WorkItem* work, int nwork, size_t lds, int waves,
                                        hipStream_t st) {
    if (waves <= 8) return launch_factor_mf_nt<T, NTILE, 512>(insts, work, nwork, lds, st);
    if (waves <= 12) return launch_factor_mf_nt<T, NTILE, 768>(insts, work, nwork, lds, st);
    if (sizeof(T) == 4) return launch_factor_mf_nt<T, NTILE, 1024>(insts, work, nwork, lds, st);
    return hipErrorInvalidValue;
}
int factor_mf_max_rows(int elem_bytes) { return elem_bytes == 8 ? 192 : 256; }
template <typename T>
hipError_t launch_ulv_factor_mf(const FInst<T>* insts, const WorkItem* work, int nwork, size_t lds, int max_rows,
                                int max_n, hipStream_t st) {
    if (nwork <= 0) return hipSuccess;
    // nodes whose children eliminated everything have an empty (0 x 0) block: one wave
    if (max_rows < 0 || max_rows > factor_mf_max_rows(sizeof(T)) || max_n < 0 || max_n > 128 || lds > 160 * 1024)
        return hipErrorInvalidValue;
    // at least 8 waves: the rows need ceil(max_rows / 16), and the phases before the
    // blocked LQ (merge, QL, M) are parallel over all threads
    const int waves = std::max(8, (max_rows + 15) / 16);
    return max_n <= 64 ? launch_factor_mf_tile<T, 4>(insts, work, nwork, lds, waves, st)
                       : launch_factor_mf_tile<T, 8>(insts, work, nwork, lds, waves, st);
}
template <typename T>
hipError_t launch_ulv_factor(const FInst<T>* insts, const WorkItem* work, int nwork, size_t lds, int threads,
                             hipStream_t st) {
    if (nwork <= 0) return hipSuccess;
    if (threads <= 0 || threads > FACT_THREADS) return hipErrorInvalidValue;
    return (sizeof(T) == 8 && threads <= 768) ? launch_factor_nt<T, 768>(insts, work, nwork, lds, st)
                                              : launch_factor_nt<T, FACT_THREADS>(insts, work, nwork, lds, st);
}
template <typename T>
hipError_t launch_ulv_wy(const FInst<T>* insts, const WorkItem* work, int nwork, hipStream_t st) {
    if (nwork <= 0) return hipSuccess;
    klaunch((k_ulv_wy<T>), dim3(nwork), dim3(256), 0, st, insts, work);
    return hipGetLastError();
}
template <typename T>
hipError_t launch_ulv_root(const FInst<T>* insts, const WorkItem* work, int nwork, hipStream_t st) {
    if (nwork <= 0) return hipSuccess;
    klaunch((k_ulv_root<T>), dim3(nwork), dim3(1024), 0, st, insts, work);
    return hipGetLastError();
}
template <typename T>
hipError_t launch_ulv_gform(const FInst<T>* insts, const WorkItem* work, int nwork, int max_kx, size_t lds,
                            hipStream_t st) {
    if (nwork <= 0 || max_kx <= 0) return hipSuccess;
    const void* fn = reinterpret_cast<const void*>(&k_ulv_gform<T>);
    hipError_t e = set_lds_attr<T>(fn, lds);
    if (e != hipSuccess) return e;
    klaunch((k_ulv_gform<T>), dim3(nwork, rhs_tiles(max_kx)), dim3(SWEEP_THREADS), lds, st, insts, work);
    return hipGetLastError();
}
template <typename T>
hipError_t launch_ulv_qform(const FInst<T>* insts, const WorkItem* work, int nwork, int max_n, size_t lds,
                            hipStream_t st) {
    if (nwork <= 0 || max_n <= 0) return hipSuccess;
    const void* fn = reinterpret_cast<const void*>(&k_ulv_qform<T>);
    hipError_t e = set_lds_attr<T>(fn, lds);
    if (e != hipSuccess) return e;
    klaunch((k_ulv_qform<T>), dim3(nwork, rhs_tiles(max_n)), dim3(SWEEP_THREADS), lds, st, insts, work);
    return hipGetLastError();
}
template <typename T>
hipError_t launch_ulv_rinv(const FInst<T>* insts, const WorkItem* work, int nwork, int max_cols, size_t lds,
                           hipStream_t st) {
    if (nwork <= 0 || max_cols <= 0) return hipSuccess;
    const void* fn = reinterpret_cast<const void*>(&k_ulv_rinv<T>);
    hipError_t e = set_lds_attr<T>(fn, lds);
    if (e != hipSuccess) return e;
    klaunch((k_ulv_rinv<T>), dim3(nwork, rhs_tiles(max_cols)), dim3(SWEEP_THREADS), lds, st, insts, work);
    return hipGetLastError();
}
template <typename T>
hipError_t launch_ulv_groot2(const SInst<T>* insts, const SweepRec* work, int nwork, int max_nrhs, size_t lds,
                             int staged, hipStream_t st) {
    if (nwork <= 0 || max_nrhs <= 0) return hipSuccess;
    if (lds > 160 * 1024) return hipErrorInvalidValue;
    const int tn = (max_nrhs + 31) / 32;
    const void* fn = staged ? reinterpret_cast<const void*>(&k_ulv_groot2<T, true>)
                            : reinterpret_cast<const void*>(&k_ulv_groot2<T, false>);
    hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    if (staged) klaunch((k_ulv_groot2<T, true>), dim3(nwork * tn), dim3(512), lds, st, insts, work, tn);
    else klaunch((k_ulv_groot2<T, false>), dim3(nwork * tn), dim3(512), lds, st, insts, work, tn);
    return hipGetLastError();
}
template <typename T>
hipError_t launch_ulv_groot(const SInst<T>* insts, const SweepRec* work, int nwork, int max_nrhs, size_t lds,
                            int resident, hipStream_t st) {
    if (nwork <= 0) return hipSuccess;
    if (lds > 160 * 1024) return hipErrorInvalidValue;
    const void* fn = reinterpret_cast<const void*>(&k_ulv_groot<T>);
    hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    klaunch((k_ulv_groot<T>), dim3(nwork, rhs_tiles(max_nrhs)), dim3(SWEEP_THREADS), lds, st, insts,
                       work, resident);
    return hipGetLastError();
}

// Tile shapes of the MFMA sweeps: 64 x 64 outputs per workgroup for more than 32
// right-hand sides, 128 x 32 otherwise.
template <typename T>
static hipError_t sweep_grid(int nwork, int max_m, int max_nrhs, int tm_rows, int tn_cols, dim3& grid, int& tm,
                             int& tn) {
    tm = (max_m + tm_rows - 1) / tm_rows;
    tn = (max_nrhs + tn_cols - 1) / tn_cols;
    const int64_t total = (int64_t)nwork * tm * tn;
    if (total > (int64_t)1 << 30) return hipErrorInvalidValue;
    grid = dim3((unsigned)(8 * ((total + 7) / 8)));
    return hipSuccess;
}

template <typename T>
hipError_t launch_ulv_gup(const SInst<T>* insts, const SweepRec* work, int nwork, int max_nk, int max_nrhs,
                          hipStream_t st) {
    if (nwork <= 0 || max_nrhs <= 0) return hipSuccess;
    const int tn = (max_nrhs + 31) / 32;
    const int64_t total = (int64_t)nwork * tn;
    if (total > (int64_t)1 << 30) return hipErrorInvalidValue;
    const size_t lds = sizeof(T) * (size_t)(max_nk > 0 ? max_nk : 1) * 33;
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_ulv_gup<T>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    klaunch((k_ulv_gup<T>), dim3((unsigned)(8 * ((total + 7) / 8))), dim3(256), lds, st, insts, work, nwork,
                       tn);
    return hipGetLastError();
}
// LDS sweep kernels run 256-thread workgroups on wide levels (wide_level), 512 otherwise.
template <typename T, bool AR, int NT>
static hipError_t launch_gup_s_nt(const SInst<T>* insts, const SweepRec* work, int nwork, int tn, int tw, size_t lds,
                                  hipStream_t st) {
    const int64_t total = (int64_t)nwork * tn;
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_ulv_gup_s<T, AR, NT>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    klaunch((k_ulv_gup_s<T, AR, NT>), dim3((unsigned)(8 * ((total + 7) / 8))), dim3(NT), lds, st, insts,
                       work, nwork, tn, tw);
    return hipGetLastError();
}
template <typename T, bool AR>
static hipError_t launch_gup_s(const SInst<T>* insts, const SweepRec* work, int nwork, int tn, int tw, size_t lds,
                               bool wide, hipStream_t st) {
    return wide ? launch_gup_s_nt<T, AR, 256>(insts, work, nwork, tn, tw, lds, st)
                : launch_gup_s_nt<T, AR, 512>(insts, work, nwork, tn, tw, lds, st);
}
template <typename T>
hipError_t launch_ulv_gup_s(const SInst<T>* insts, const SweepRec* work, int nwork, int max_nrhs, int tw,
                            size_t lds, int resident, int nt, hipStream_t st) {
    if (nwork <= 0 || max_nrhs <= 0) return hipSuccess;
    if ((tw != 16 && tw != 32) || (nt != 0 && nt != 256 && nt != 512)) return hipErrorInvalidValue;
    const int tn = (max_nrhs + tw - 1) / tw;
    if ((int64_t)nwork * tn > (int64_t)1 << 30 || lds > 160 * 1024) return hipErrorInvalidValue;
    const bool wide = nt ? nt == 256 : wide_level(nwork, max_nrhs);
    return resident ? launch_gup_s<T, true>(insts, work, nwork, tn, tw, lds, wide, st)
                    : launch_gup_s<T, false>(insts, work, nwork, tn, tw, lds, wide, st);
}
template <typename T, bool AR, int NT>
static hipError_t launch_gdown_s_nt(const SInst<T>* insts, const SweepRec* work, int nwork, int tn, int tw,
                                    size_t lds, hipStream_t st) {
    const int64_t total = (int64_t)nwork * tn;
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_ulv_gdown_s<T, AR, NT>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    klaunch((k_ulv_gdown_s<T, AR, NT>), dim3((unsigned)(8 * ((total + 7) / 8))), dim3(NT), lds, st, insts,
                       work, nwork, tn, tw);
    return hipGetLastError();
}
template <typename T, bool AR>
static hipError_t launch_gdown_s(const SInst<T>* insts, const SweepRec* work, int nwork, int tn, int tw, size_t lds,
                                 bool wide, hipStream_t st) {
    return wide ? launch_gdown_s_nt<T, AR, 256>(insts, work, nwork, tn, tw, lds, st)
                : launch_gdown_s_nt<T, AR, 512>(insts, work, nwork, tn, tw, lds, st);
}
template <typename T>
hipError_t launch_ulv_gdown_s(const SInst<T>* insts, const SweepRec* work, int nwork, int max_nrhs, int tw,
                              size_t lds, int resident, int nt, hipStream_t st) {
    if (nwork <= 0 || max_nrhs <= 0) return hipSuccess;
    if ((tw != 16 && tw != 32) || (nt != 0 && nt != 256 && nt != 512)) return hipErrorInvalidValue;
    const int tn = (max_nrhs + tw - 1) / tw;
    if ((int64_t)nwork * tn > (int64_t)1 << 30 || lds > 160 * 1024) return hipErrorInvalidValue;
    const bool wide = nt ? nt == 256 : wide_level(nwork, max_nrhs);
    return resident ? launch_gdown_s<T, true>(insts, work, nwork, tn, tw, lds, wide, st)
                    : launch_gdown_s<T, false>(insts, work, nwork, tn, tw, lds, wide, st);
}
template <typename T>
hipError_t launch_ulv_gdown(const SInst<T>* insts, const SweepRec* work, int nwork, int max_m, int max_nrhs,
                            hipStream_t st) {
    if (nwork <= 0 || max_m <= 0 || max_nrhs <= 0) return hipSuccess;
    dim3 grid;
    int tm, tn;
    if (max_nrhs > 32) {
        if (hipError_t e = sweep_grid<T>(nwork, max_m, max_nrhs, 64, 64, grid, tm, tn)) return e;
        klaunch((k_ulv_gdown<T, 2, 2>), grid, dim3(256), 0, st, insts, work, nwork, tm, tn);
    } else {
        if (hipError_t e = sweep_grid<T>(nwork, max_m, max_nrhs, 128, 32, grid, tm, tn)) return e;
        klaunch((k_ulv_gdown<T, 4, 1>), grid, dim3(256), 0, st, insts, work, nwork, tm, tn);
    }
    return hipGetLastError();
}

// The convergence test of one lockstep refinement step, on the device (one workgroup):
// request i is active while tab[i].nrhs > 0; its relative residual is
// max_j ||r_j|| / ||b_j|| from norms (squared, as k_resid_batch leaves them), written to rel[i];
// at or below tol the request is retired by setting its column count to 0 in every table
// the step's launches read (the two solve tables, the matvec table and tab), so they skip it.
// out[0] = the requests still active.
__global__ __launch_bounds__(256) void k_refine_check(BatchVec* __restrict__ tab, char* __restrict__ sinst0,
                                                       char* __restrict__ sinst1, int64_t sinst_size, int nrhs_off,
                                                       MvInst* __restrict__ mv, int cnt, int maxc,
                                                       const double* __restrict__ norms, double tol,
                                                       double* __restrict__ rel, int* __restrict__ out) {
    int active = 0;
    for (int i = threadIdx.x; i < cnt; i += blockDim.x) {
        const int nr = tab[i].nrhs;
        if (nr <= 0) continue;
        double worst = 0.0;
        for (int j = 0; j < nr; ++j) {
            const double nb = sqrt(norms[2 * ((int64_t)i * maxc + j)]);
            const double rr = sqrt(norms[2 * ((int64_t)i * maxc + j) + 1]);
            const double q = nb > 0.0 ? rr / nb : rr;
            worst = q > worst || q != q ? q : worst;  // a NaN residual stays the worst
        }
        rel[i] = worst;
        if (worst <= tol) {
            tab[i].nrhs = 0;
            mv[i].nrhs = 0;
            if (sinst0) *reinterpret_cast<int32_t*>(sinst0 + (int64_t)i * sinst_size + nrhs_off) = 0;
            if (sinst1) *reinterpret_cast<int32_t*>(sinst1 + (int64_t)i * sinst_size + nrhs_off) = 0;
        } else {
            ++active;
        }
    }
    __shared__ int red[16];
    for (int o = 32; o >= 1; o >>= 1) active += __shfl_xor(active, o, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = active;
    __syncthreads();
    if (threadIdx.x == 0) {
        int t = 0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += red[w];
        out[0] = t;
    }
}
hipError_t launch_refine_check(BatchVec* tab, char* sinst0, char* sinst1, int64_t sinst_size, int nrhs_off,
                               MvInst* mv, int cnt, int maxc, const double* norms, double tol, double* rel, int* out,
                               hipStream_t st) {
    if (cnt <= 0) return hipSuccess;
    klaunch(k_refine_check, dim3(1), dim3(256), 0, st, tab, sinst0, sinst1, sinst_size, nrhs_off, mv, cnt, maxc,
            norms, tol, rel, out);
    return hipGetLastError();
}
hipError_t launch_resid_batch(const BatchVec* tab, int cnt, int maxc, double* norms, hipStream_t st) {
    if (cnt <= 0 || maxc <= 0) return hipSuccess;
    klaunch(k_resid_batch, dim3(cnt, maxc), dim3(256), 0, st, tab, maxc, norms);
    return hipGetLastError();
}
hipError_t launch_axpy_batch(const BatchVec* tab, int cnt, int maxc, hipStream_t st) {
    if (cnt <= 0 || maxc <= 0) return hipSuccess;
    klaunch(k_axpy_batch, dim3(cnt, maxc), dim3(256), 0, st, tab);
    return hipGetLastError();
}

constexpr int GT_SPLIT = 4;
constexpr int GT_TILE = 64;  // output tile of the LDS path (four waves of 16 x 64)
constexpr int GT_KC = 32;    // k-chunk staged per step

// A large product task (m, n >= 64, k >= 32) on FP64 MFMA: the task's 64 x 64 output tiles
// go round robin over its gridDim.y workgroups; each stages 64 x 32 chunks of op(a) and
// op(b) in LDS, and wave w accumulates rows [16 w, 16 w + 16) of the tile in four 16 x 16
// column blocks (16x16x4 steps).  Loads past the task's rows, columns or k are clamped in
// bounds; a chunk's rows past k are staged but enter the MFMAs as zeros.
__device__ __forceinline__ void gemm_task_tiled(const GemmTask& g, bool ta, bool tb, bool acc, double sg,
                                                gptr<const double> a, gptr<const double> b, gptr<double> c) {
    using V4 = Mfma<double>::v4;
    __shared__ double As[GT_KC][GT_TILE + 1];  // As[l][i] = op(a)(i0 + i, k0 + l)
    __shared__ double Bs[GT_KC][GT_TILE + 1];  // Bs[l][j] = op(b)(k0 + l, j0 + j)
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, lr = lane & 15, lk = lane >> 4;
    const int mt = (g.m + GT_TILE - 1) / GT_TILE, nt = (g.n + GT_TILE - 1) / GT_TILE;
    for (int tile = blockIdx.y; tile < mt * nt; tile += gridDim.y) {
        const int i0 = (tile % mt) * GT_TILE, j0 = (tile / mt) * GT_TILE;
        V4 s[4];
#pragma unroll
        for (int cb = 0; cb < 4; ++cb) s[cb] = V4{0.0, 0.0, 0.0, 0.0};
        for (int k0 = 0; k0 < g.k; k0 += GT_KC) {
            // 64 x 32 elements of each operand, 8 per thread
            for (int e = tid; e < GT_TILE * GT_KC; e += blockDim.x) {
                const int x = e % GT_TILE, l = e / GT_TILE;  // x: row of op(a) / column of op(b)
                const int kk = min(k0 + l, g.k - 1);
                const int ii = min(i0 + x, g.m - 1), jj = min(j0 + x, g.n - 1);
                As[l][x] = ta ? a[kk + (int64_t)ii * g.lda] : a[ii + (int64_t)kk * g.lda];
                Bs[l][x] = tb ? b[jj + (int64_t)kk * g.ldb] : b[kk + (int64_t)jj * g.ldb];
            }
            __syncthreads();
            const int kn = min(GT_KC, g.k - k0);
            for (int st = 0; st < (kn + 3) / 4; ++st) {
                const int l = 4 * st + lk;  // < GT_KC
                const bool in = l < kn;
                const double av = in ? As[l][16 * wave + lr] : 0.0;
#pragma unroll
                for (int cb = 0; cb < 4; ++cb) s[cb] = Mfma<double>::mac(av, in ? Bs[l][16 * cb + lr] : 0.0, s[cb]);
            }
            __syncthreads();
        }
#pragma unroll
        for (int cb = 0; cb < 4; ++cb) {
            const int j = j0 + 16 * cb + lr;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int i = i0 + 16 * wave + Mfma<double>::row(lk, r);
                if (i < g.m && j < g.n) {
                    const int64_t o = i + (int64_t)j * g.ldc;
                    c[o] = acc ? c[o] + sg * s[cb][r] : sg * s[cb][r];
                }
            }
        }
    }
}

// The small products that build an HSS x HSS product's generators (src/matmul.jl:65-118),
// and the copies that gather a matrix's matvec operators.  Task blockIdx.x runs on
// gridDim.y workgroups.  Products of at least 64 rows and columns over at least 32 terms go
// through gemm_task_tiled (FP64 MFMA); any other gives each thread 4 x 4 outputs
// (consecutive threads take consecutive 4-row groups), each one fma chain over l = 0..k-1
// in order.  Loads past the task's rows or columns are clamped to its last row or column and
// their sums are not stored.
__global__ __launch_bounds__(256) void k_gemm_tasks(const GemmTask* __restrict__ tasks) {
    const GemmTask g = tasks[blockIdx.x];
    const bool ta = g.flags & GT_TA, tb = g.flags & GT_TB, acc = g.flags & GT_ACC, cp = g.flags & GT_COPY;
    const double sg = (g.flags & GT_NEG) ? -1.0 : 1.0;
    const gptr<const double> a = gbl(g.a);
    const gptr<const double> b = gbl(g.b);
    const gptr<double> c = gbl(g.c);
    const int t0 = blockIdx.y * blockDim.x + threadIdx.x, ts = gridDim.y * blockDim.x;
    if (cp) {
        for (int e = t0; e < g.m * g.n; e += ts) {
            const int i = e % g.m, j = e / g.m;
            const double s = sg * (ta ? a[j + (int64_t)i * g.lda] : a[i + (int64_t)j * g.lda]);
            const int64_t o = i + (int64_t)j * g.ldc;
            c[o] = acc ? c[o] + s : s;
        }
        return;
    }
    if (g.m >= GT_TILE && g.n >= GT_TILE && g.k >= GT_KC) {
        gemm_task_tiled(g, ta, tb, acc, sg, a, b, c);
        return;
    }
    const int mb = (g.m + 3) / 4, nb = (g.n + 3) / 4;
    for (int e = t0; e < mb * nb; e += ts) {
        const int i0 = (e % mb) * 4, j0 = (e / mb) * 4;
        int ii[4], jj[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            ii[r] = min(i0 + r, g.m - 1);
            jj[r] = min(j0 + r, g.n - 1);
        }
        double s[4][4];
#pragma unroll
        for (int r = 0; r < 4; ++r)
#pragma unroll
            for (int q = 0; q < 4; ++q) s[r][q] = 0.0;
        for (int l = 0; l < g.k; ++l) {
            double av[4], bv[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                av[r] = ta ? a[l + (int64_t)ii[r] * g.lda] : a[ii[r] + (int64_t)l * g.lda];
                bv[r] = tb ? b[jj[r] + (int64_t)l * g.ldb] : b[l + (int64_t)jj[r] * g.ldb];
            }
#pragma unroll
            for (int r = 0; r < 4; ++r)
#pragma unroll
                for (int q = 0; q < 4; ++q) s[r][q] = fma(av[r], bv[q], s[r][q]);
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            if (j0 + q >= g.n) break;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                if (i0 + r >= g.m) break;
                const int64_t o = (i0 + r) + (int64_t)(j0 + q) * g.ldc;
                c[o] = acc ? c[o] + sg * s[r][q] : sg * s[r][q];
            }
        }
    }
}

hipError_t launch_gemm_tasks(const GemmTask* tasks, int count, hipStream_t st) {
    if (count <= 0) return hipSuccess;
    klaunch(k_gemm_tasks, dim3((unsigned)count, GT_SPLIT), dim3(256), 0, st, tasks);
    return hipGetLastError();
}

hipError_t launch_residual(const double* B, int64_t ldb, double* R, int64_t ldr, int64_t n, int nrhs, hipStream_t st) {
    const int64_t total = n * nrhs;
    int blocks = (int)((total + 255) / 256);
    if (blocks > 4096) blocks = 4096;
    if (blocks < 1) return hipSuccess;
    klaunch(k_residual, dim3(blocks), dim3(256), 0, st, B, ldb, R, ldr, n, nrhs);
    return hipGetLastError();
}
hipError_t launch_axpy(double* X, int64_t ldx, const double* D, int64_t ldd, int64_t n, int nrhs, hipStream_t st) {
    const int64_t total = n * nrhs;
    int blocks = (int)((total + 255) / 256);
    if (blocks > 4096) blocks = 4096;
    if (blocks < 1) return hipSuccess;
    klaunch(k_axpy, dim3(blocks), dim3(256), 0, st, X, ldx, D, ldd, n, nrhs);
    return hipGetLastError();
}
// E[(j0 + c) + c * ld] = v for c < w (the unit columns j0 .. j0 + w - 1 of an identity,
// written into a zeroed rows x w block, or cleared again with v = 0)
__global__ void k_unit_cols(double* __restrict__ E, int64_t ld, int64_t j0, int w, int64_t rows, double v) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < w && j0 + c < rows) E[(j0 + c) + (int64_t)c * ld] = v;
}
hipError_t launch_unit_cols(double* E, int64_t ld, int64_t j0, int w, int64_t rows, double v, hipStream_t st) {
    if (w <= 0) return hipSuccess;
    klaunch(k_unit_cols, dim3((unsigned)((w + 255) / 256)), dim3(256), 0, st, E, ld, j0, w, rows, v);
    return hipGetLastError();
}
int cut_max_k() { return CUT_MAXK; }
size_t cut_lds(size_t w, int max_k) {
    (void)max_k;
    return w * (size_t)(CUT_RPT * CUT_THREADS * (CUT_TW + 1) + (CUT_THREADS / 64) * 256);
}
template <typename TF, typename Inst, typename Rec>
hipError_t launch_cut(const Inst* insts, const Rec* work, int nwork, int max_nrhs, int max_k, hipStream_t st) {
    if (nwork <= 0 || max_nrhs <= 0 || max_k <= 0) return hipSuccess;
    if (max_k > CUT_MAXK) return hipErrorInvalidValue;
    const size_t lds = cut_lds(sizeof(TF), max_k);
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_cut<TF, Inst, Rec>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    const int tn = (max_nrhs + CUT_TW - 1) / CUT_TW;
    const int64_t total = (int64_t)nwork * tn;
    klaunch((k_cut<TF, Inst, Rec>), dim3((unsigned)(8 * ((total + 7) / 8))), dim3(CUT_THREADS), lds, st, insts, work,
            nwork, tn);
    return hipGetLastError();
}
template hipError_t launch_cut<double, MvInst, MvRec>(const MvInst*, const MvRec*, int, int, int, hipStream_t);
template <typename TF>
hipError_t launch_cut_seed(TF* W, const CutRow* map, int K, int64_t nrhs, int col0, hipStream_t st) {
    if (K <= 0) return hipSuccess;
    klaunch(k_cut_seed<TF>, dim3((unsigned)((K + 255) / 256)), dim3(256), 0, st, W, map, K, nrhs, col0);
    return hipGetLastError();
}
template <typename TF>
hipError_t launch_cut_gather(TF* T, int M, int cols, const CutRow* map, const TF* W, int64_t nrhs, hipStream_t st) {
    const int64_t total = (int64_t)M * cols;
    if (total <= 0) return hipSuccess;
    const unsigned blocks = (unsigned)std::min<int64_t>((total + 255) / 256, 4096);
    klaunch(k_cut_gather<TF>, dim3(blocks), dim3(256), 0, st, T, M, cols, map, W, nrhs);
    return hipGetLastError();
}
template hipError_t launch_cut_seed<double>(double*, const CutRow*, int, int64_t, int, hipStream_t);
template hipError_t launch_cut_gather<double>(double*, int, int, const CutRow*, const double*, int64_t, hipStream_t);

hipError_t launch_colnorm2(const double* A, int64_t lda, int64_t n, int nrhs, double* out, hipStream_t st) {
    if (nrhs <= 0) return hipSuccess;
    klaunch(k_colnorm2, dim3(nrhs), dim3(256), 0, st, A, lda, n, out);
    return hipGetLastError();
}

// Gram-Schmidt steps of a lockstep GMRES with one Krylov space per column.  Basis block j
// sits at V + j * vs (n x nrhs, ld n); W is n x nrhs (ld n); coefficients h[j + c * hld].
//   k_gs_dots:   h[j + c * hld] = V_j(:, c)' W(:, c), j < kc   (one workgroup per (c, j))
//   k_gs_update: W(:, c) = (acc ? W(:, c) : 0) + sg * sum_{j < kc} h[j + c * hld] V_j(:, c)
//   k_col_scale: Y(:, c) = W(:, c) / sqrt(s[c]), or 0 where s[c] == 0
__global__ void k_gs_dots(const double* __restrict__ V, int64_t vs, const double* __restrict__ W, int64_t n,
                          double* __restrict__ h, int hld) {
    const int c = blockIdx.x, j = blockIdx.y;
    const double* v = V + j * vs + (int64_t)c * n;
    const double* w = W + (int64_t)c * n;
    double s = 0.0;
    for (int64_t i = threadIdx.x; i < n; i += blockDim.x) s += v[i] * w[i];
    __shared__ double red[16];
    s = wave_reduce(s);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w8 = 0; w8 < (int)(blockDim.x >> 6); ++w8) t += red[w8];
        h[j + (int64_t)c * hld] = t;
    }
}
__global__ void k_gs_update(double* __restrict__ W, const double* __restrict__ V, int64_t vs, int kc, int64_t n,
                            int nrhs, const double* __restrict__ h, int hld, double sg, int acc) {
    const int64_t total = n * nrhs;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t c = e / n;
        double t = 0.0;
        for (int j = 0; j < kc; ++j) t += h[j + c * hld] * V[j * vs + e];
        W[e] = (acc ? W[e] : 0.0) + sg * t;
    }
}
__global__ void k_col_scale(double* __restrict__ Y, const double* __restrict__ W, int64_t n, int nrhs,
                            const double* __restrict__ s) {
    const int64_t total = n * nrhs;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        const double q = s[e / n];
        Y[e] = q > 0.0 ? W[e] / sqrt(q) : 0.0;
    }
}
static inline int stream_blocks(int64_t total) {
    return (int)std::min<int64_t>(std::max<int64_t>((total + 255) / 256, 1), 4096);
}
hipError_t launch_gs_dots(const double* V, int64_t vs, int kc, const double* W, int64_t n, int nrhs, double* h, int hld,
                          hipStream_t st) {
    if (nrhs <= 0 || kc <= 0) return hipSuccess;
    klaunch(k_gs_dots, dim3(nrhs, kc), dim3(256), 0, st, V, vs, W, n, h, hld);
    return hipGetLastError();
}
hipError_t launch_gs_update(double* W, const double* V, int64_t vs, int kc, int64_t n, int nrhs, const double* h,
                            int hld, double sg, int acc, hipStream_t st) {
    if (nrhs <= 0 || n <= 0) return hipSuccess;
    klaunch(k_gs_update, dim3(stream_blocks(n * nrhs)), dim3(256), 0, st, W, V, vs, kc, n, nrhs, h, hld, sg, acc);
    return hipGetLastError();
}
hipError_t launch_col_scale(double* Y, const double* W, int64_t n, int nrhs, const double* s, hipStream_t st) {
    if (nrhs <= 0 || n <= 0) return hipSuccess;
    klaunch(k_col_scale, dim3(stream_blocks(n * nrhs)), dim3(256), 0, st, Y, W, n, nrhs, s);
    return hipGetLastError();
}

template <int KIND, bool AR, int NT>
static hipError_t launch_mv_s_nt(const MvInst* insts, const MvRec* work, int nwork, int tn, int tw, size_t lds,
                                 hipStream_t st) {
    const int64_t total = (int64_t)nwork * tn;
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_mv_s<KIND, AR, NT>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    klaunch((k_mv_s<KIND, AR, NT>), dim3((unsigned)(8 * ((total + 7) / 8))), dim3(NT), lds, st, insts,
                       work, nwork, tn, tw);
    return hipGetLastError();
}
template <int KIND, bool AR>
static hipError_t launch_mv_s(const MvInst* insts, const MvRec* work, int nwork, int tn, int tw, size_t lds,
                              bool wide, hipStream_t st) {
    return wide ? launch_mv_s_nt<KIND, AR, 256>(insts, work, nwork, tn, tw, lds, st)
                : launch_mv_s_nt<KIND, AR, 512>(insts, work, nwork, tn, tw, lds, st);
}
template <int KIND>
static hipError_t launch_mv_kind(const MvInst* insts, const MvRec* work, int nwork, int max_m, int max_nrhs, int tw,
                                 size_t lds, int resident, hipStream_t st) {
    if (nwork <= 0 || max_nrhs <= 0 || max_m <= 0) return hipSuccess;
    if (lds > 0) {
        if (lds > 160 * 1024 || (tw != 16 && tw != 32)) return hipErrorInvalidValue;
        const int tn = (max_nrhs + tw - 1) / tw;
        if ((int64_t)nwork * tn > (int64_t)1 << 30) return hipErrorInvalidValue;
        const bool wide = wide_level(nwork, max_nrhs);
        return resident ? launch_mv_s<KIND, true>(insts, work, nwork, tn, tw, lds, wide, st)
                        : launch_mv_s<KIND, false>(insts, work, nwork, tn, tw, lds, wide, st);
    }
    dim3 grid;
    int tm, tn;
    if (max_nrhs > 32) {
        if (hipError_t e = sweep_grid<double>(nwork, max_m, max_nrhs, 64, 64, grid, tm, tn)) return e;
        klaunch((k_mv_t<KIND, 2, 2>), grid, dim3(256), 0, st, insts, work, nwork, tm, tn);
    } else {
        if (hipError_t e = sweep_grid<double>(nwork, max_m, max_nrhs, 128, 32, grid, tm, tn)) return e;
        klaunch((k_mv_t<KIND, 4, 1>), grid, dim3(256), 0, st, insts, work, nwork, tm, tn);
    }
    return hipGetLastError();
}
hipError_t launch_mv(int kind, const MvInst* insts, const MvRec* work, int nwork, int max_m, int max_nrhs, int tw,
                     size_t lds, int resident, hipStream_t st) {
    switch (kind) {
        case 0: return launch_mv_kind<0>(insts, work, nwork, max_m, max_nrhs, tw, lds, resident, st);
        case 1: return launch_mv_kind<1>(insts, work, nwork, max_m, max_nrhs, tw, lds, resident, st);
        default: return launch_mv_kind<2>(insts, work, nwork, max_m, max_nrhs, tw, lds, resident, st);
    }
}

// ---------------------------------------------------------------------------
// direct compression (reference src/compression.jl:22-30, 47-133): Gaussian test
// matrices and the truncated column-pivoted QR of each node's sampled block row
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    // SplitMix64 finalizer (Steele, Lea, Flood 2014)
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// om[i + j * rows] ~ N(0, 1) for i < rows, j < cols: Box-Muller on two uniforms hashed from
// (seed, i, j), so an entry does not depend on the matrix's size or on the launch
__global__ __launch_bounds__(256) void k_gauss(double* __restrict__ om, int64_t rows, int cols, uint64_t seed) {
    const int64_t total = rows * cols;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = e % rows, j = e / rows;
        const uint64_t h1 = mix64(seed ^ mix64((uint64_t)i * 0x100000001B3ull + (uint64_t)j));
        const uint64_t h2 = mix64(h1 ^ 0xD1B54A32D192ED03ull);
        const double u1 = ((double)(h1 >> 11) + 0.5) * 0x1.0p-53;  // (0, 1)
        const double u2 = (double)(h2 >> 11) * 0x1.0p-53;          // [0, 1)
        gbl(om)[e] = sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
    }
}
hipError_t launch_gauss(double* om, int64_t rows, int cols, uint64_t seed, hipStream_t st) {
    const int64_t total = rows * cols;
    if (total <= 0) return hipSuccess;
    const int64_t blocks = std::min<int64_t>((total + 255) / 256, 8192);
    klaunch(k_gauss, dim3((unsigned)blocks), dim3(256), 0, st, om, rows, cols, seed);
    return hipGetLastError();
}

// Truncated column-pivoted Householder QR of one r x s block Y per workgroup (LAPACK
// dgeqp3 pivoting with dlarfg reflectors), in place in global memory: step i takes the
// column of largest remaining norm; the factorization stops before step i when that norm
// (times t.scale) is not above max(atol, rtol * the first pivot's norm (times t.scale)), the
// truncation rule of the reference's rank-revealing QR (compression.jl:22-30), or, with
// t.id_rule, min(atol, rtol * the first pivot's norm), the rule of its interpolative
// decomposition (compression.jl:439-440).  The rank k goes to *t.rank, Q's first k columns
// (orthonormal, r x k, ld r) to t.Q (if not null), R stays in Y's first k rows (columns in
// pivoted order) and the pivots to t.perm (if not null).  The column
// norms below the current row are recomputed at every step (no downdating), so small
// trailing norms keep full accuracy.  A NaN norm ends the factorization.
// LDS: s + kmax + 8 doubles.
__global__ __launch_bounds__(256) void k_rrqr(const RrqrTask* __restrict__ tasks, double atol, double rtol) {
    const RrqrTask t = tasks[blockIdx.x];
    const int r = t.r, s = t.s, kmax = t.kmax;
    const gptr<double> Y = gbl(t.Y);
    const gptr<double> Q = gbl(t.Q);
    extern __shared__ double lds_d[];
    double* nrm2 = lds_d;          // s: squared norms of the columns below the current row
    double* tauv = nrm2 + s;       // kmax
    double* sc = tauv + kmax;      // scalars
    int* sci = reinterpret_cast<int*>(sc + 6);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, nw = blockDim.x >> 6;
    const gptr<int32_t> perm = gbl(t.perm);
    if (t.perm)
        for (int j = tid; j < s; j += blockDim.x) perm[j] = j;
    for (int j = wave; j < s; j += nw) {
        double part = 0.0;
        for (int i = lane; i < r; i += 64) {
            const double v = Y[i + (int64_t)j * r];
            part += v * v;
        }
        part = wave_reduce(part);
        if (lane == 0) nrm2[j] = part;
    }
    __syncthreads();
    int k = 0;
    double tol = 0.0;
    for (int i = 0; i < kmax; ++i) {
        if (wave == 0) {
            double best = -1.0;
            int bj = s;
            for (int j = i + lane; j < s; j += 64) {
                const double v = nrm2[j];
                if (v > best) { best = v; bj = j; }
            }
            for (int off = 32; off >= 1; off >>= 1) {
                const double ob = shfl_xor(best, off);
                const int oj = __shfl_xor(bj, off, 64);
                if (ob > best || (ob == best && oj < bj)) { best = ob; bj = oj; }
            }
            if (lane == 0) {
                sci[0] = bj < s ? bj : i;  // all norms NaN: stays in range, the test below stops
                sc[0] = best;
            }
        }
        __syncthreads();
        const int p = sci[0];
        const double nrm = sqrt(fmax(sc[0], 0.0)) * t.scale;
        if (i == 0) tol = t.id_rule ? fmin(atol, rtol * nrm) : fmax(atol, rtol * nrm);
        if (!(nrm > tol)) break;  // uniform across the workgroup (LDS values, same arithmetic)
        if (p != i) {
            for (int e = tid; e < r; e += blockDim.x) {
                const double a = Y[e + (int64_t)i * r];
                Y[e + (int64_t)i * r] = Y[e + (int64_t)p * r];
                Y[e + (int64_t)p * r] = a;
            }
            if (tid == 0) {
                const double a = nrm2[i];
                nrm2[i] = nrm2[p];
                nrm2[p] = a;
                if (t.perm) {
                    const int32_t q = perm[i];
                    perm[i] = perm[p];
                    perm[p] = q;
                }
            }
        }
        __syncthreads();
        if (wave == 0) {
            double part = 0.0;
            for (int l = i + 1 + lane; l < r; l += 64) {
                const double v = Y[l + (int64_t)i * r];
                part += v * v;
            }
            part = wave_reduce(part);
            if (lane == 0) {
                double beta, tau, scl;
                larfg(Y[i + (int64_t)i * r], part, beta, tau, scl);
                sc[1] = beta;
                sc[2] = tau;
                sc[3] = scl;
                tauv[i] = tau;
            }
        }
        __syncthreads();
        const double tau = sc[2], scl = sc[3];
        if (tau != 0.0)
            for (int l = i + 1 + tid; l < r; l += blockDim.x) Y[l + (int64_t)i * r] *= scl;
        __syncthreads();
        if (tid == 0) Y[i + (int64_t)i * r] = sc[1];
        // H = I - tau v v', v = [1; Y[i+1:r, i]], on the remaining columns; then each
        // column's squared norm below row i
        for (int j = i + 1 + wave; j < s; j += nw) {
            const gptr<double> yj = Y + (int64_t)j * r;
            double part = 0.0;
            for (int l = i + 1 + lane; l < r; l += 64) part += Y[l + (int64_t)i * r] * yj[l];
            const double w = tau * (wave_reduce(part) + yj[i]);
            double n2 = 0.0;
            for (int l = i + 1 + lane; l < r; l += 64) {
                const double v = yj[l] - w * Y[l + (int64_t)i * r];
                yj[l] = v;
                n2 += v * v;
            }
            n2 = wave_reduce(n2);
            if (lane == 0) {
                yj[i] -= w;
                nrm2[j] = n2;
            }
        }
        __syncthreads();
        k = i + 1;
    }
    if (!t.Q) {
        if (tid == 0) *gbl(t.rank) = k;
        return;
    }
    // Q[:, 0:k] = H_0 ... H_{k-1} [I_k; 0] (LAPACK dorg2r order)
    for (int e = tid; e < r * k; e += blockDim.x) Q[e] = (e % r == e / r) ? 1.0 : 0.0;
    __syncthreads();
    for (int i = k - 1; i >= 0; --i) {
        const double tau = tauv[i];
        for (int j = i + wave; j < k; j += nw) {
            const gptr<double> qj = Q + (int64_t)j * r;
            double part = 0.0;
            for (int l = i + 1 + lane; l < r; l += 64) part += Y[l + (int64_t)i * r] * qj[l];
            const double w = tau * (wave_reduce(part) + qj[i]);
            for (int l = i + 1 + lane; l < r; l += 64) qj[l] -= w * Y[l + (int64_t)i * r];
            if (lane == 0) qj[i] -= w;
        }
        __syncthreads();
    }
    if (tid == 0) *gbl(t.rank) = k;
}
hipError_t launch_rrqr(const RrqrTask* tasks, int count, int max_s, int max_kmax, double atol, double rtol,
                       hipStream_t st) {
    if (count <= 0) return hipSuccess;
    const size_t lds = sizeof(double) * ((size_t)max_s + (size_t)max_kmax + 8);
    if (lds > 160 * 1024) return hipErrorInvalidValue;
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_rrqr),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    klaunch(k_rrqr, dim3((unsigned)count), dim3(256), lds, st, tasks, atol, rtol);
    return hipGetLastError();
}

// Interpolation matrices (one IdTask per workgroup; compression.jl:441-443): column j of
// the pivoted R is column perm[j] of the original, so Xt[perm[j], :] = e_j for j < k and
// Xt[perm[j], :] = (R11^-1 R(:, j))' for j >= k, by back substitution (one thread per
// column; the solution is built in place in Xt, which only that thread touches).
__global__ __launch_bounds__(256) void k_id_form(const IdTask* __restrict__ tasks) {
    const IdTask t = tasks[blockIdx.x];
    const gptr<const double> R = gbl(t.R);
    const gptr<const int32_t> perm = gbl(t.perm);
    const gptr<double> Xt = gbl(t.Xt);
    for (int j = threadIdx.x; j < t.s; j += blockDim.x) {
        const int64_t row = perm[j];
        if (j < t.k) {
            for (int c = 0; c < t.k; ++c) Xt[row + (int64_t)c * t.ldx] = c == j ? 1.0 : 0.0;
            continue;
        }
        for (int c = t.k - 1; c >= 0; --c) {
            double v = R[c + (int64_t)j * t.r];
            for (int l = c + 1; l < t.k; ++l) v -= R[c + (int64_t)l * t.r] * Xt[row + (int64_t)l * t.ldx];
            Xt[row + (int64_t)c * t.ldx] = v / R[c + (int64_t)c * t.r];
        }
    }
}
hipError_t launch_id_form(const IdTask* tasks, int count, hipStream_t st) {
    if (count <= 0) return hipSuccess;
    klaunch(k_id_form, dim3((unsigned)count), dim3(256), 0, st, tasks);
    return hipGetLastError();
}

// Gathers (one GatherTask per blockIdx.x, gridDim.y workgroups each): the skeleton rows of a
// sample and the couplings A[J_rows, J_cols] of randomized compression.
__global__ __launch_bounds__(256) void k_gather_tasks(const GatherTask* __restrict__ tasks) {
    const GatherTask g = tasks[blockIdx.x];
    const gptr<const double> src = gbl(g.src);
    const gptr<const int32_t> ri = gbl(g.ri);
    const gptr<const int32_t> ci = gbl(g.ci);
    const gptr<double> dst = gbl(g.dst);
    const int64_t total = (int64_t)g.m * g.n;
    for (int64_t e = blockIdx.y * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.y * blockDim.x) {
        const int i = (int)(e % g.m), j = (int)(e / g.m);
        const int64_t r = g.ri ? ri[i] : i, c = g.ci ? ci[j] : j;
        dst[i + (int64_t)j * g.ldd] = src[r + c * g.ld];
    }
}
hipError_t launch_gather_tasks(const GatherTask* tasks, int count, hipStream_t st) {
    if (count <= 0) return hipSuccess;
    klaunch(k_gather_tasks, dim3((unsigned)count, GT_SPLIT), dim3(256), 0, st, tasks);
    return hipGetLastError();
}

// A solve record written on the device from the kernel's by-value argument: the
// allocation-free solve places its record in the library's fixed device array without a
// host-to-device copy (stream-ordered with the sweeps that read it).
__global__ void k_put_sinst(SInst<double>* dst, SInst<double> v) {
    if (threadIdx.x == 0) *gbl(dst) = v;
}
hipError_t launch_put_sinst(SInst<double>* dst, const SInst<double>& v, hipStream_t st) {
    klaunch(k_put_sinst, dim3(1), dim3(64), 0, st, dst, v);
    return hipGetLastError();
}

// explicit instantiations
template hipError_t launch_ulv_factor<double>(const FInst<double>*, const WorkItem*, int, size_t, int, hipStream_t);
template hipError_t launch_ulv_factor<float>(const FInst<float>*, const WorkItem*, int, size_t, int, hipStream_t);
template hipError_t launch_ulv_factor_mf<double>(const FInst<double>*, const WorkItem*, int, size_t, int, int,
                                                 hipStream_t);
template hipError_t launch_ulv_factor_mf<float>(const FInst<float>*, const WorkItem*, int, size_t, int, int,
                                                hipStream_t);
template hipError_t launch_ulv_wy<double>(const FInst<double>*, const WorkItem*, int, hipStream_t);
template hipError_t launch_ulv_wy<float>(const FInst<float>*, const WorkItem*, int, hipStream_t);
template hipError_t launch_ulv_root<double>(const FInst<double>*, const WorkItem*, int, hipStream_t);
template hipError_t launch_ulv_root<float>(const FInst<float>*, const WorkItem*, int, hipStream_t);
#define HSS_INST_SOLVE(T)                                                                                      \
    template hipError_t launch_ulv_gform<T>(const FInst<T>*, const WorkItem*, int, int, size_t, hipStream_t);       \
    template hipError_t launch_ulv_qform<T>(const FInst<T>*, const WorkItem*, int, int, size_t, hipStream_t);       \
    template hipError_t launch_ulv_gup<T>(const SInst<T>*, const SweepRec*, int, int, int, hipStream_t);            \
    template hipError_t launch_ulv_groot<T>(const SInst<T>*, const SweepRec*, int, int, size_t, int, hipStream_t);   \
    template hipError_t launch_ulv_rinv<T>(const FInst<T>*, const WorkItem*, int, int, size_t, hipStream_t);       \
    template hipError_t launch_ulv_groot2<T>(const SInst<T>*, const SweepRec*, int, int, size_t, int, hipStream_t);  \
    template hipError_t launch_ulv_gdown<T>(const SInst<T>*, const SweepRec*, int, int, int, hipStream_t);      \
    template hipError_t launch_ulv_gup_s<T>(const SInst<T>*, const SweepRec*, int, int, int, size_t, int, int,       \
                                            hipStream_t);                                                              \
    template hipError_t launch_ulv_gdown_s<T>(const SInst<T>*, const SweepRec*, int, int, int, size_t, int, int,       \
                                              hipStream_t);
HSS_INST_SOLVE(double)
HSS_INST_SOLVE(float)
#undef HSS_INST_SOLVE

}  // namespace hss
