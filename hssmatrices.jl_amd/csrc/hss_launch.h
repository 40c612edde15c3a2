// hss_launch.h -- host launchers of the kernels in hss_kernels.hip.
#pragma once
#include <hip/hip_runtime.h>

#include "hss_internal.h"

namespace hss {

// Profiling hook.  While the runtime times a scope, g_prof_launch points at the scope's
// pair of kernel events: the scope's first launch records them at its own start and end
// (hipExtLaunchKernelGGL), and every launch in the scope counts itself.
struct ProfLaunch {
    hipEvent_t k0 = nullptr, k1 = nullptr;
    int launches = 0;
};
extern ProfLaunch* g_prof_launch;

// Column width of the LDS-resident steps (16 or 32): 16 when a level has too few 32-column
// tiles to occupy the CUs (halves each workgroup's MFMA chain) or at most 16 columns.
// A level is wide when it has at least 2048 (node, 32-column) tiles, eight per CU: its
// LDS kernels then run 256-thread workgroups with the operators streamed (more
// independent workgroups per CU hide each one's staging); narrower levels run 512
// threads per workgroup (more waves per node).
inline bool wide_level(int64_t nwork, int max_nrhs) { return nwork * ((max_nrhs + 31) / 32) >= 2048; }
inline int resident_tw(int nwork, int max_nrhs) {
    if (max_nrhs <= 16) return 16;
    return ((int64_t)nwork * ((max_nrhs + 31) / 32) < 128) ? 16 : 32;
}

// matvec sweep step over nodes (kind 0 = up: Z, 1 = down: F, 2 = leaves: C); lds > 0
// selects the LDS kernel with that many bytes and tw columns per workgroup, with the
// operator staged (resident != 0) or streamed by the waves; lds == 0 the chunked one
hipError_t launch_mv(int kind, const MvInst* insts, const MvRec* work, int nwork, int max_m, int max_nrhs, int tw,
                     size_t lds, int resident, hipStream_t st);

// threads = the most any node of the launch needs ((m + kv) * S, at most FACT_THREADS)
template <typename T>
hipError_t launch_ulv_factor(const FInst<T>* insts, const WorkItem* work, int nwork, size_t lds, int threads,
                             hipStream_t st);
// blocked factor on MFMA: every node of the launch has n <= 128 and m + kv <= max_rows, at
// most factor_mf_max_rows(sizeof(T)) (192 rows for Float64, 256 for Float32)
template <typename T>
hipError_t launch_ulv_factor_mf(const FInst<T>* insts, const WorkItem* work, int nwork, size_t lds, int max_rows,
                                int max_n, hipStream_t st);
int factor_mf_max_rows(int elem_bytes);
template <typename T>
hipError_t launch_ulv_wy(const FInst<T>* insts, const WorkItem* work, int nwork, hipStream_t st);
template <typename T>
hipError_t launch_ulv_root(const FInst<T>* insts, const WorkItem* work, int nwork, hipStream_t st);
// solve operators, once per factorization (every non-root node in one launch)
template <typename T>
hipError_t launch_ulv_gform(const FInst<T>* insts, const WorkItem* work, int nwork, int max_kx, size_t lds,
                            hipStream_t st);
template <typename T>
hipError_t launch_ulv_qform(const FInst<T>* insts, const WorkItem* work, int nwork, int max_n, size_t lds,
                            hipStream_t st);
// solve sweeps: max_nk = largest nk (up), max_m = largest n (down) of the level's nodes
template <typename T>
hipError_t launch_ulv_gup(const SInst<T>* insts, const SweepRec* work, int nwork, int max_nk, int max_nrhs,
                          hipStream_t st);
// root operators of the inverse-and-refine solve (factor time; max_cols = max(m, Kx))
template <typename T>
hipError_t launch_ulv_rinv(const FInst<T>* insts, const WorkItem* work, int nwork, int max_cols, size_t lds,
                           hipStream_t st);
// root solve by inverse and one refinement step (branch roots whose Dinv fits in LDS)
template <typename T>
hipError_t launch_ulv_groot2(const SInst<T>* insts, const SweepRec* work, int nwork, int max_nrhs, size_t lds,
                             int staged, hipStream_t st);
// resident != 0: the root factor is staged in LDS (lds includes it)
template <typename T>
hipError_t launch_ulv_groot(const SInst<T>* insts, const SweepRec* work, int nwork, int max_nrhs, size_t lds,
                            int resident, hipStream_t st);
template <typename T>
hipError_t launch_ulv_gdown(const SInst<T>* insts, const SweepRec* work, int nwork, int max_m, int max_nrhs,
                            hipStream_t st);
// Variants with the input staged in LDS (tw = 16 or 32 columns per workgroup, lds = bytes
// per workgroup); resident != 0: the node operators are staged too, else streamed;
// nt = threads per workgroup (256 or 512; 0: 256 on wide levels, else 512)
template <typename T>
hipError_t launch_ulv_gup_s(const SInst<T>* insts, const SweepRec* work, int nwork, int max_nrhs, int tw, size_t lds,
                            int resident, int nt, hipStream_t st);
template <typename T>
hipError_t launch_ulv_gdown_s(const SInst<T>* insts, const SweepRec* work, int nwork, int max_nrhs, int tw,
                              size_t lds, int resident, int nt, hipStream_t st);

// direct compression: Gaussian test matrix (rows x cols, ld rows) and truncated pivoted QRs
hipError_t launch_gauss(double* om, int64_t rows, int cols, uint64_t seed, hipStream_t st);
hipError_t launch_rrqr(const RrqrTask* tasks, int count, int max_s, int max_kmax, double atol, double rtol,
                       hipStream_t st);
// interpolation matrices of k_rrqr outputs (IdTask) and gathers by index lists (GatherTask)
hipError_t launch_id_form(const IdTask* tasks, int count, hipStream_t st);
hipError_t launch_gather_tasks(const GatherTask* tasks, int count, hipStream_t st);
// writes one solve record to device memory from a kernel argument (no host copy)
hipError_t launch_put_sinst(SInst<double>* dst, const SInst<double>& v, hipStream_t st);
// one workgroup per task (HSS x HSS product generators)
hipError_t launch_gemm_tasks(const GemmTask* tasks, int count, hipStream_t st);
// the convergence test of a lockstep refinement step on the device (k_refine_check)
hipError_t launch_refine_check(BatchVec* tab, char* sinst0, char* sinst1, int64_t sinst_size, int nrhs_off,
                               MvInst* mv, int cnt, int maxc, const double* norms, double tol, double* rel, int* out,
                               hipStream_t st);
hipError_t launch_resid_batch(const BatchVec* tab, int cnt, int maxc, double* norms, hipStream_t st);
hipError_t launch_axpy_batch(const BatchVec* tab, int cnt, int maxc, hipStream_t st);
hipError_t launch_residual(const double* B, int64_t ldb, double* R, int64_t ldr, int64_t n, int nrhs, hipStream_t st);
hipError_t launch_axpy(double* X, int64_t ldx, const double* D, int64_t ldd, int64_t n, int nrhs, hipStream_t st);
// unit columns j0 .. j0 + w - 1 of an identity into a zeroed rows x w block (v = 1), or cleared (v = 0)
hipError_t launch_unit_cols(double* E, int64_t ld, int64_t j0, int w, int64_t rows, double v, hipStream_t st);
hipError_t launch_colnorm2(const double* A, int64_t lda, int64_t n, int nrhs, double* out, hipStream_t st);
// Gram-Schmidt steps of the lockstep GMRES (basis block j at V + j * vs, n x nrhs, ld n):
// h[j + c * hld] = V_j(:, c)' W(:, c) for j < kc; W = (acc ? W : 0) + sg * sum_j h_j V_j per
// column; Y(:, c) = W(:, c) / sqrt(s[c]) (0 where s[c] == 0)
hipError_t launch_gs_dots(const double* V, int64_t vs, int kc, const double* W, int64_t n, int nrhs, double* h, int hld,
                          hipStream_t st);
hipError_t launch_gs_update(double* W, const double* V, int64_t vs, int kc, int64_t n, int nrhs, const double* h,
                            int hld, double sg, int acc, hipStream_t st);
hipError_t launch_col_scale(double* Y, const double* W, int64_t n, int nrhs, const double* s, hipStream_t st);
// tree cut (k_cut): one workgroup per (cut record, 16-column tile); max_k = the largest K
// of the records, at most cut_max_k().  cut_lds = its LDS bytes for element size w.
int cut_max_k();
size_t cut_lds(size_t w, int max_k);
template <typename TF, typename Inst, typename Rec>
hipError_t launch_cut(const Inst* insts, const Rec* work, int nwork, int max_nrhs, int max_k, hipStream_t st);
// a cut's build (nrhs columns): the identity seed of a stacked input (row k -> column
// col0 + k), and T (M x cols) gathered from stacked output rows
template <typename TF>
hipError_t launch_cut_seed(TF* W, const CutRow* map, int K, int64_t nrhs, int col0, hipStream_t st);
template <typename TF>
hipError_t launch_cut_gather(TF* T, int M, int cols, const CutRow* map, const TF* W, int64_t nrhs, hipStream_t st);

}  // namespace hss
