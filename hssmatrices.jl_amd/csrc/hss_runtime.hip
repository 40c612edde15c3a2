// hss_runtime.hip -- host runtime and C ABI of libhss_hip.so (include/hss_hip.h).
//
// Owns device memory, builds the level-major work lists from the node table,
// and sequences the kernel launches of hss_kernels.hip:
//   matvec  : up-sweep by height, down-sweep by depth, leaf products
//             (reference src/matmul.jl:18-62)
//   factor  : one launch per height (all nodes of that height, of every matrix
//             in a batch), then the root LU (src/ulvfactor.jl:60-96)
//   solve   : up-sweep by height, root solve, down-sweep by depth
//             (src/ulvfactor.jl:10-19, 98-107)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstddef>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/hss_hip.h"
#include "hss_host.h"
#include "hss_internal.h"
#include "hss_launch.h"

using namespace hss;

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------
static thread_local std::string g_err;
static thread_local double g_last_ms = -1.0;

static hss_status fail(hss_status st, const std::string& msg) {
    g_err = msg;
    return st;
}

#define HIP_TRY(expr)                                                                              \
    do {                                                                                           \
        hipError_t _e = (expr);                                                                    \
        if (_e != hipSuccess)                                                                      \
            return fail(_e == hipErrorOutOfMemory ? HSS_ERR_OOM : HSS_ERR_HIP,                     \
                        std::string(#expr) + ": " + hipGetErrorString(_e));                        \
    } while (0)

// ---------------------------------------------------------------------------
// optional per-launch timing (hss_prof_*).  Each timed scope records HIP events on the
// library stream around itself (e0, e1) and hands the kernel launcher a second pair (k0,
// k1) that its first launch records at the kernel's own start and end.  A scope with
// exactly one launch reports the kernel's duration; any other scope the time between its
// two events.  Resolved when the caller asks.
// ---------------------------------------------------------------------------
namespace hss {
ProfLaunch* g_prof_launch = nullptr;
}
struct ProfRec {
    const char* name;
    int level;
    int nwork;
    hipEvent_t e0, e1, k0, k1;
    int launches;
    float ms;
};
static bool g_prof = false;
// HSS_FACTOR_KERNEL=vector (or =mfma) puts every level on the per-reflector factor kernel
// (or every level that fits on the blocked MFMA kernel), for comparing the two; the results
// agree to rounding.  Unset: by precision (see use_mf in factor_batch).
enum class FactorKernel { Auto, Vector, Mfma };
static const FactorKernel g_factor_kernel = [] {
    const char* e = std::getenv("HSS_FACTOR_KERNEL");
    if (e && std::strcmp(e, "vector") == 0) return FactorKernel::Vector;
    if (e && std::strcmp(e, "mfma") == 0) return FactorKernel::Mfma;
    return FactorKernel::Auto;
}();
// HSS_CUT=0 runs every level of the matvec sweeps as a launch of its own (no tree cut), for
// comparing the two; the results agree to rounding.
static const bool g_cut = [] {
    const char* e = std::getenv("HSS_CUT");
    return !(e && std::strcmp(e, "0") == 0);
}();
static std::vector<ProfRec> g_prof_recs;
static std::vector<hipEvent_t> g_prof_pool;

static hipEvent_t prof_event() {
    if (!g_prof_pool.empty()) {
        hipEvent_t e = g_prof_pool.back();
        g_prof_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    (void)hipEventCreate(&e);
    return e;
}

struct ProfScope {
    int idx = -1;
    hipStream_t st;
    hss::ProfLaunch pl;
    hss::ProfLaunch* prev = nullptr;
    ProfScope(const char* name, int level, int nwork, hipStream_t s) : st(s) {
        if (!g_prof || nwork <= 0) return;
        ProfRec r{name, level, nwork, prof_event(), prof_event(), prof_event(), prof_event(), 0, -1.f};
        (void)hipEventRecord(r.e0, st);
        g_prof_recs.push_back(r);
        idx = (int)g_prof_recs.size() - 1;
        pl.k0 = r.k0;
        pl.k1 = r.k1;
        prev = hss::g_prof_launch;
        hss::g_prof_launch = &pl;
    }
    ~ProfScope() {
        if (idx < 0) return;
        hss::g_prof_launch = prev;
        (void)hipEventRecord(g_prof_recs[idx].e1, st);
        g_prof_recs[idx].launches = pl.launches;
    }
};

template <typename T>
static hipError_t dalloc(T** p, size_t count) {
    *p = nullptr;
    if (count == 0) count = 1;
    return hipMalloc(reinterpret_cast<void**>(p), count * sizeof(T));
}

// ---------------------------------------------------------------------------
// handles
// ---------------------------------------------------------------------------
struct Range {
    int off = 0, count = 0;
};

// One dense product of a tree cut (host copy; T and the row tables are device memory):
// y = T x, T M x K column-major.
struct CutInfo {
    const void* T = nullptr;
    int M = 0, K = 0;
    const CutRow* xrow = nullptr;
    const CutRow* yrow = nullptr;
};

// Device copies of per-call instance records.  A record is uploaded once and
// reused while the same (pointers, nrhs, ...) come back, so repeated calls
// enqueue no host-to-device copy.  Replacing a slot first waits for the
// stream, so no kernel still reads it, and then copies synchronously from
// host memory that stays owned by the cache.
struct SlotCache {
    static constexpr int SLOTS = 4;
    size_t rec = 0;
    char* d = nullptr;
    std::vector<char> h;
    int used = 0, next = 0;
    ~SlotCache() { (void)hipFree(d); }
    hipError_t get(const void* v, size_t bytes, hipStream_t st, const void** out) {
        if (!d) {
            rec = bytes;
            hipError_t e = hipMalloc(reinterpret_cast<void**>(&d), rec * SLOTS);
            if (e != hipSuccess) return e;
            h.assign(rec * SLOTS, 0);
        }
        if (bytes != rec) return hipErrorInvalidValue;
        for (int i = 0; i < used; ++i)
            if (std::memcmp(h.data() + i * rec, v, rec) == 0) {
                *out = d + i * rec;
                return hipSuccess;
            }
        int slot = used < SLOTS ? used++ : next;
        next = (slot + 1) % SLOTS;
        hipError_t e = hipStreamSynchronize(st);
        if (e != hipSuccess) return e;
        std::memcpy(h.data() + slot * rec, v, rec);
        e = hipMemcpy(d + slot * rec, v, rec, hipMemcpyHostToDevice);
        if (e != hipSuccess) return e;
        *out = d + slot * rec;
        return hipSuccess;
    }
};

// Level schedule of a set of solve requests: up-sweep by height, roots, down-sweep by
// depth; up_m = the largest nk, down_m = the largest n of a level's nodes; for the
// LDS-resident kernels, up_na / down_na = the largest operator (elements, rounded to 16)
// and up_k / down_k = the most input rows of a level's nodes; w = bytes per element.
// Launch choice for one sweep level: form 0 = LDS-resident (operator and input staged),
// 1 = operator streamed by the waves, 2 = chunked tile kernel; tw = columns per workgroup
// (16 or 32); nt = threads per workgroup (256, 512; 0 = by the level's width).
// form < 0: the built-in rules (pick_resident, pick_streamed).
struct LevelCfg {
    int form = -1, tw = 0, nt = 0;
};
struct Schedule {
    std::vector<WorkItem> work;
    std::vector<SweepRec> rec;  // the solve sweeps' records, same order as work
    std::vector<Range> up, down;
    std::vector<int> up_m, down_m, up_rows;
    std::vector<int64_t> up_na, down_na, up_k, down_k;
    size_t w = 8;
    Range roots;
    // per-level choices made by the autotuner (hss_plan_tune); empty = built-in rules
    std::vector<LevelCfg> up_cfg, down_cfg;
};
// Largest LDS of a workgroup (bytes); levels above it use the streamed kernels.
constexpr size_t kLdsResidentMax = 160 * 1024;
// A level that is not wide runs LDS-resident (operator and input staged) when the
// staging fits and either the operators are small (at most 72 KiB) or the level has at
// most one workgroup per CU: a level of many large operators keeps more bytes in flight
// streaming them.  tw =
// columns per workgroup, lds = bytes per workgroup; na = the level's largest operator
// (elements of w bytes, rounded to 16), k = its most input rows.
static bool pick_resident(size_t w, int64_t na, int64_t k, int count, int maxnrhs, int& tw, size_t& lds) {
    if (wide_level(count, maxnrhs)) return false;  // wide levels stream their operators
    for (int t : {resident_tw(count, maxnrhs), 16}) {
        lds = w * (size_t)(na + k * (t + 1));
        const int64_t wgs = (int64_t)count * ((maxnrhs + t - 1) / t);
        if (lds <= kLdsResidentMax && (w * (size_t)na <= 72 * 1024 || wgs <= 256)) {
            tw = t;
            return true;
        }
    }
    return false;
}
// Otherwise the operators are streamed by the waves with only the input staged (at most
// 80 KiB, two workgroups per CU); larger inputs go through the chunked tile kernels.
static bool pick_streamed(size_t w, int64_t k, int rows, int count, int maxnrhs, int& tw, size_t& lds) {
    (void)rows;
    for (int t : {resident_tw(count, maxnrhs), 16}) {
        lds = w * (size_t)(k * (t + 1));
        if (lds <= 80 * 1024) {
            tw = t;
            return true;
        }
    }
    return false;
}

// Matvec schedule: up-sweep steps by height, down-sweep steps by depth (1..), leaves;
// per step the largest operator height (m), operator size (na, elements rounded to 16)
// and input rows (k).
struct MvSched {
    std::vector<Range> up, down;
    Range leaf;
    std::vector<int> up_m, down_m;
    int leaf_m = 0;
    std::vector<int64_t> up_na, down_na, up_k, down_k;
    int64_t leaf_na = 0, leaf_k = 0;
    // the tree cuts of the instances that have one (k_cut records), run after the up-sweep;
    // cut_k = the largest Kc
    Range cut;
    int cut_k = 0;
};

struct PlanKey {
    int nrhs;
    const double* dB;
    int64_t ldb;
    double* dX;
    int64_t ldx;
    bool operator==(const PlanKey& o) const {
        return nrhs == o.nrhs && dB == o.dB && ldb == o.ldb && dX == o.dX && ldx == o.ldx;
    }
};
static void destroy_plan_cache(std::vector<std::pair<PlanKey, hss_plan_t>>& cache);
struct hss_fact_s;
// hold = the plan takes references on its factors and context (every plan but those a
// factor caches for itself)
static hss_status plan_create_impl(int32_t count, hss_fact_s* const* facts, const int32_t* nrhs,
                                   const double* const* dB, const int64_t* ldb, double* const* dX, const int64_t* ldx,
                                   int32_t with_refinement, bool hold, hss_plan_t* out);

// Handles hold references on what they use (a matrix on its context, a factor on
// its matrix, a plan on its factors and context, a server on its context), so
// destroying them in any order -- as garbage-collected bindings do -- is safe: a
// handle is released when it is destroyed and no other handle uses it.
struct hss_ctx_s {
    int refs = 1;
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
};

struct hss_mat_s {
    int refs = 1;
    hss_ctx_s* ctx = nullptr;
    int nn = 0;
    int64_t rows = 0, cols = 0;
    std::vector<hss_node_desc> desc;
    std::vector<DevNode> hn;
    double* d_blob = nullptr;
    int64_t blob_len = 0;
    DevNode* d_nodes = nullptr;
    // work lists (instance 0), concatenated in one device buffer
    std::vector<WorkItem> hwork;
    WorkItem* d_work = nullptr;
    std::vector<Range> by_height;  // non-root nodes of height h
    std::vector<Range> by_depth;   // nodes of depth d (d >= 1)
    Range leaves, root;
    int max_height = 0, max_depth = 0;
    int64_t zrows = 0, frows = 0;
    double* d_op = nullptr;  // matvec operators (DevNode::opZ/opF/opC)
    int64_t op_len = 0;
    MvSched mvs;             // matvec schedule of this matrix (instance 0)
    MvRec* d_mvwork = nullptr;
    // per-call scratch
    SlotCache mvinst;
    double* d_Z = nullptr;  // Z, then F at Z + zrows * nrhs
    size_t zf_cap = 0;      // nrhs capacity of Z/F
    double *d_io_b = nullptr, *d_io_c = nullptr;
    size_t io_cap = 0;  // elements
    int64_t dev_bytes = 0;
    // tree cut of the matvec (cut_depth 0: none): its product, T and the row tables
    int cut_depth = 0;
    std::vector<CutInfo> cut;
    double* d_cutT = nullptr;
    char* d_cut = nullptr;
};

struct hss_fact_s {
    int refs = 1;
    hss_mat_s* mat = nullptr;
    hss_prec prec = HSS_PREC_FP64;
    std::vector<FNode> hfn;
    FNode* d_fn = nullptr;
    void* d_fac = nullptr;
    int64_t fac_len = 0;
    int32_t* d_piv = nullptr;
    int32_t* d_status = nullptr;
    int64_t res_len = 0;               // resident part of the factor (G, Qt, root LU)
    int64_t yrows = 0, zrows = 0;      // solve workspace rows (times nrhs)
    size_t lds_groot = 0;              // LDS of the root solve kernel
    size_t lds_groot_res = 0;          // the same with the root's factor staged in LDS
    size_t lds_groot2 = 0;             // the inverse-and-refine root kernel (0: not available)
    size_t lds_groot2s = 0;            // the same with D and Mb staged too
    size_t lds_gform = 0, lds_qform = 0;
    int max_kx = 0, max_nq = 0;        // widest Gr / Qt (columns formed per node)
    void* d_wy = nullptr;  // y blocks, then the z blocks at wy + yrows * nrhs
    size_t ws_cap = 0;     // nrhs capacity
    SlotCache sinst;
    // single-request solve schedule (instance 0)
    SweepRec* d_swork = nullptr;
    Schedule sched;
    // refinement scratch (Float64): residual and correction
    double *d_r = nullptr, *d_d = nullptr, *d_nrm = nullptr;
    size_t rf_cap = 0;
    // GMRES-IR workspace (Krylov basis, the preconditioned vector, coefficients), gm_cap doubles
    double* d_gm = nullptr;
    size_t gm_cap = 0;
    double *d_io_b = nullptr, *d_io_x = nullptr;
    size_t io_cap = 0;
    int64_t dev_bytes = 0;
    // captured single-request plans keyed by their buffers (repeated solves replay a graph)
    std::vector<std::pair<PlanKey, hss_plan_t>> plan_cache;
};

// ---------------------------------------------------------------------------
// fixed device workspace of the allocation-free solve (hss_ulv_solve_dev_noalloc).  Each
// GPU the library is loaded on holds one copy until the process ends.  Host code never
// uses the array's name as an address: hss_ctx_create resolves the device address once
// per device (hipGetSymbolAddress, with that device current) and keeps it in
// g_scratch_ptr, and only that address is passed to kernels.
// ---------------------------------------------------------------------------
constexpr size_t kScratchBytes = size_t(64) << 20;
constexpr size_t kScratchRec = 256;  // the solve record at the start of the array
static __device__ __align__(256) double scratch[kScratchBytes / sizeof(double)];
constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[kMaxDevices] = {};

// once per device, with that device current
static hss_status scratch_init(int device) {
    if (device < 0 || device >= kMaxDevices) return fail(HSS_ERR_ARG, "device index beyond the library's table");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[device]) return HSS_OK;
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return fail(HSS_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[device] = p;
    return HSS_OK;
}

// ---------------------------------------------------------------------------
// context
// ---------------------------------------------------------------------------
extern "C" hss_status hss_ctx_create(int device, hss_ctx_t* out) {
    if (!out) return fail(HSS_ERR_ARG, "null output handle");
    int ndev = 0;
    HIP_TRY(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) return fail(HSS_ERR_ARG, "device index out of range");
    HIP_TRY(hipSetDevice(device));
    hss_status ss = scratch_init(device);
    if (ss != HSS_OK) return ss;
    auto* c = new (std::nothrow) hss_ctx_s;
    if (!c) return fail(HSS_ERR_OOM, "host allocation");
    c->device = device;
    HIP_TRY(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    HIP_TRY(hipEventCreate(&c->ev0));
    HIP_TRY(hipEventCreate(&c->ev1));
    *out = c;
    return HSS_OK;
}

static void ctx_release(hss_ctx_s* c) {
    if (!c || --c->refs > 0) return;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    (void)hipEventDestroy(c->ev0);
    (void)hipEventDestroy(c->ev1);
    (void)hipStreamDestroy(c->stream);
    delete c;
}
extern "C" hss_status hss_ctx_destroy(hss_ctx_t c) {
    ctx_release(c);
    return HSS_OK;
}

extern "C" void* hss_ctx_stream(hss_ctx_t c) { return c ? (void*)c->stream : nullptr; }

// ---------------------------------------------------------------------------
// matrix
// ---------------------------------------------------------------------------
static hss_status validate(const hss_node_desc* nd, int nn, int64_t blob_len) {
    if (nn <= 0) return fail(HSS_ERR_ARG, "empty node table");
    if (nd[0].parent != -1) return fail(HSS_ERR_ARG, "node 0 must be the root");
    auto in_blob = [&](int64_t off, int64_t count) { return count == 0 || (off >= 0 && off + count <= blob_len); };
    for (int i = 0; i < nn; ++i) {
        const hss_node_desc& x = nd[i];
        if (x.m < 0 || x.n < 0 || x.ku < 0 || x.kv < 0)
            return fail(HSS_ERR_ARG, "negative size in node " + std::to_string(i));
        const bool leaf = x.left < 0;
        if (leaf != (x.right < 0)) return fail(HSS_ERR_ARG, "node " + std::to_string(i) + " has one child");
        if (i > 0) {
            if (x.parent < 0 || x.parent >= i) return fail(HSS_ERR_ARG, "nodes must be breadth-first ordered");
            const hss_node_desc& p = nd[x.parent];
            if (p.left != i && p.right != i) return fail(HSS_ERR_ARG, "parent/child links disagree");
            if (!in_blob(x.off_R, (int64_t)x.ku * p.ku) || !in_blob(x.off_W, (int64_t)x.kv * p.kv))
                return fail(HSS_ERR_ARG, "translator out of the blob in node " + std::to_string(i));
        } else if (x.ku != 0 || x.kv != 0) {
            return fail(HSS_ERR_DIM, "root must have 0-column R/W (ku = kv = 0)");
        }
        if (leaf) {
            if (!in_blob(x.off_D, (int64_t)x.m * x.n) || !in_blob(x.off_U, (int64_t)x.m * x.ku) ||
                !in_blob(x.off_V, (int64_t)x.n * x.kv))
                return fail(HSS_ERR_ARG, "leaf generator out of the blob in node " + std::to_string(i));
            if ((int64_t)x.m * x.n > 0 && x.off_D < 0) return fail(HSS_ERR_ARG, "leaf without D");
        } else {
            if (x.left <= i || x.right <= i || x.left >= nn || x.right >= nn)
                return fail(HSS_ERR_ARG, "bad child index in node " + std::to_string(i));
            const hss_node_desc& l = nd[x.left];
            const hss_node_desc& r = nd[x.right];
            if (l.m + r.m != x.m || l.n + r.n != x.n)
                return fail(HSS_ERR_DIM, "children sizes do not add up in node " + std::to_string(i));
            if (l.row_off != x.row_off || r.row_off != x.row_off + l.m || l.col_off != x.col_off ||
                r.col_off != x.col_off + l.n)
                return fail(HSS_ERR_ARG, "row/column offsets inconsistent in node " + std::to_string(i));
            if (!in_blob(x.off_B12, (int64_t)l.ku * r.kv) || !in_blob(x.off_B21, (int64_t)r.ku * l.kv))
                return fail(HSS_ERR_ARG, "coupling out of the blob in node " + std::to_string(i));
        }
    }
    return HSS_OK;
}

static void free_mat(hss_mat_s* M) {
    if (!M) return;
    (void)hipFree(M->d_blob);
    (void)hipFree(M->d_op);
    (void)hipFree(M->d_mvwork);
    (void)hipFree(M->d_nodes);
    (void)hipFree(M->d_work);
    (void)hipFree(M->d_Z);
    (void)hipFree(M->d_io_b);
    (void)hipFree(M->d_io_c);
    (void)hipFree(M->d_cutT);
    (void)hipFree(M->d_cut);
    delete M;
}

// Operator shape of node i in a matvec step (kind 0 up, 1 down, 2 leaf): rows M, columns K.
static void mv_shape(const std::vector<DevNode>& hn, int i, int kind, int& rows, int& cols) {
    const DevNode& d = hn[i];
    if (kind == 0) {
        rows = d.kv;
        cols = d.is_leaf ? d.n : hn[d.left].kv + hn[d.right].kv;
    } else if (kind == 1) {
        const DevNode& p = hn[d.parent];
        const DevNode& sb = hn[d.is_left ? p.right : p.left];
        rows = d.ku;
        cols = sb.kv + p.ku;
    } else if (d.fold) {
        const DevNode& p = hn[d.parent];
        rows = d.m;
        cols = d.n + hn[d.is_left ? p.right : p.left].kv + p.ku;
    } else {
        rows = d.m;
        cols = d.n + d.ku;
    }
}

// The matvec record of node i (kind 0 up, 1 down, 2 leaf) of instance b, with the Z / F
// rows of the node table hn (a leaf's F rows are offset by zrows: F follows Z).
static MvRec mv_record(const std::vector<DevNode>& hn, int64_t zrows, int b, int i, int kind) {
    const DevNode& d = hn[i];
    MvRec q{};
    q.inst = b;
    int rows, cols;
    mv_shape(hn, i, kind, rows, cols);
    q.M = rows;
    q.K = cols;
    q.r1 = cols;
    if (kind == 0) {
        q.A = d.opZ;
        q.y = d.zrow;
        q.ldy = std::max(d.kv, 1);
        if (d.is_leaf) {
            q.from_b = 1;
            q.x0 = d.col_off;
            q.r0 = d.n;
            q.ld1 = 1;
        } else {
            const DevNode &L = hn[d.left], &R = hn[d.right];
            q.x0 = L.zrow;
            q.ld0 = std::max(L.kv, 1);
            q.r0 = L.kv;
            q.x1 = R.zrow;
            q.ld1 = std::max(R.kv, 1);
        }
    } else if (kind == 1) {
        const DevNode& p = hn[d.parent];
        const DevNode& sb = hn[d.is_left ? p.right : p.left];
        q.A = d.opF;
        q.x0 = sb.zrow;
        q.ld0 = std::max(sb.kv, 1);
        q.r0 = sb.kv;
        q.x1 = p.frow;
        q.ld1 = std::max(p.ku, 1);
        q.y = d.frow;
        q.ldy = std::max(d.ku, 1);
    } else {
        q.A = d.opC;
        q.from_b = 1;
        q.x0 = d.col_off;
        q.r0 = d.n;
        q.y = d.row_off;
        if (d.fold) {
            const DevNode& p = hn[d.parent];
            const DevNode& sb = hn[d.is_left ? p.right : p.left];
            q.x1 = sb.zrow;
            q.ld1 = std::max(sb.kv, 1);
            q.r1 = d.n + sb.kv;
            q.x2 = zrows + p.frow;
            q.ld2 = std::max(p.ku, 1);
        } else {
            q.x1 = zrows + d.frow;
            q.ld1 = std::max(d.ku, 1);
        }
    }
    return q;
}

// Appends the matvec work lists of `mats` (instance b = position in mats) to `work`.  A
// matrix with a tree cut at depth c runs no step of a node above depth c, nor the down step
// of a depth-c node: the cut product gives the depth-c nodes' F from their Z.
static void build_mv_schedule(const std::vector<hss_mat_s*>& mats, std::vector<MvRec>& work, MvSched& S) {
    int maxh = 0, maxd = 0;
    for (auto* M : mats) {
        maxh = std::max(maxh, M->max_height);
        maxd = std::max(maxd, M->max_depth);
    }
    auto level = [&](int kind, auto sel, Range& r, int& mmax, int64_t& na, int64_t& kk) {
        r.off = (int)work.size();
        for (int b = 0; b < (int)mats.size(); ++b) {
            hss_mat_s* M = mats[b];
            for (int i = 0; i < M->nn; ++i) {
                if (!sel(M, i)) continue;
                work.push_back(mv_record(M->hn, M->zrows, b, i, kind));
                int rows, cols;
                mv_shape(M->hn, i, kind, rows, cols);
                mmax = std::max(mmax, rows);
                na = std::max(na, ((int64_t)rows * cols + 15) / 16 * 16);
                kk = std::max(kk, (int64_t)cols);
            }
        }
        r.count = (int)work.size() - r.off;
    };
    S.up.assign(maxh, Range{});
    S.up_m.assign(maxh, 0);
    S.up_na.assign(maxh, 0);
    S.up_k.assign(maxh, 0);
    for (int h = 0; h < maxh; ++h)
        level(0, [&](const hss_mat_s* M, int i) { return i != 0 && M->hn[i].height == h && M->hn[i].depth >= M->cut_depth; },
              S.up[h], S.up_m[h], S.up_na[h], S.up_k[h]);
    S.down.assign(maxd + 1, Range{});
    S.down_m.assign(maxd + 1, 0);
    S.down_na.assign(maxd + 1, 0);
    S.down_k.assign(maxd + 1, 0);
    for (int d = 1; d <= maxd; ++d)
        level(1,
              [&](const hss_mat_s* M, int i) {
                  return i != 0 && M->hn[i].depth == d && d > M->cut_depth && !M->hn[i].fold;
              },
              S.down[d], S.down_m[d], S.down_na[d], S.down_k[d]);
    S.leaf_m = 0;
    S.leaf_na = S.leaf_k = 0;
    level(2, [&](const hss_mat_s* M, int i) { return M->hn[i].is_leaf != 0; }, S.leaf, S.leaf_m, S.leaf_na, S.leaf_k);
    // cut products: 16-row strips of T
    S.cut.off = (int)work.size();
    S.cut_k = 0;
    for (int b = 0; b < (int)mats.size(); ++b) {
        const hss_mat_s* M = mats[b];
        if (M->cut_depth <= 0) continue;
        const CutInfo& ci = M->cut[0];
        S.cut_k = std::max(S.cut_k, ci.K);
        for (int r0 = 0; r0 < ci.M; r0 += 16) {
            MvRec q{};
            q.inst = b;
            q.r0 = r0;
            q.M = std::min(16, ci.M - r0);
            q.K = ci.K;
            q.ld0 = ci.M;
            q.A = reinterpret_cast<int64_t>(ci.T);
            q.x0 = reinterpret_cast<int64_t>(ci.xrow);
            q.x1 = reinterpret_cast<int64_t>(ci.yrow);
            work.push_back(q);
        }
    }
    S.cut.count = (int)work.size() - S.cut.off;
}

// Each step: operator and input staged in LDS (pick_resident), else the operator
// streamed by the waves with the input staged (pick_streamed), else the chunked kernel.
static hss_status launch_mv_schedule(const MvInst* di, const MvRec* dw, const MvSched& S, int maxnrhs,
                                     hipStream_t st) {
    static const char* names[3][3] = {{"k_mv_up_s", "k_mv_up_x", "k_mv_up"},
                                      {"k_mv_down_s", "k_mv_down_x", "k_mv_down"},
                                      {"k_mv_leaf_s", "k_mv_leaf_x", "k_mv_leaf"}};
    auto step = [&](int kind, int level, const Range& r, int m, int64_t na, int64_t k) -> hss_status {
        if (r.count == 0) return HSS_OK;
        int tw = 32, form = 2;
        size_t lds = 0;
        if (pick_resident(sizeof(double), na, k, r.count, maxnrhs, tw, lds)) form = 0;
        else if (pick_streamed(sizeof(double), k, m, r.count, maxnrhs, tw, lds)) form = 1;
        else lds = 0;
        ProfScope ps(names[kind][form], level, r.count, st);
        HIP_TRY(launch_mv(kind, di, dw + r.off, r.count, m, maxnrhs, tw, lds, form == 0 ? 1 : 0, st));
        return HSS_OK;
    };
    for (size_t h = 0; h < S.up.size(); ++h) {
        hss_status s = step(0, (int)h, S.up[h], S.up_m[h], S.up_na[h], S.up_k[h]);
        if (s != HSS_OK) return s;
    }
    if (S.cut.count > 0) {
        ProfScope ps("k_cut", 0, S.cut.count, st);
        HIP_TRY((launch_cut<double, MvInst, MvRec>(di, dw + S.cut.off, S.cut.count, maxnrhs, S.cut_k, st)));
    }
    for (size_t d = 1; d < S.down.size(); ++d) {
        hss_status s = step(1, (int)d, S.down[d], S.down_m[d], S.down_na[d], S.down_k[d]);
        if (s != HSS_OK) return s;
    }
    return step(2, 0, S.leaf, S.leaf_m, S.leaf_na, S.leaf_k);
}

// ---------------------------------------------------------------------------
// tree cuts: the sweep steps above one depth collapsed into one dense product
// ---------------------------------------------------------------------------
// Depth of a tree cut.  sizes(c, k, m, t) gives, for a cut at depth c, the largest stacked
// input k and output m of its products and their total elements t.  Returns the deepest
// c >= 3 at or above every leaf (so at least the steps of depths 1 and 2 go) whose k and m
// are within cut_max_k() (a k_cut workgroup holds its stacked input in LDS) and whose
// products hold at most `budget` elements; 0 when there is none.  The budget (a 32nd of
// the operators the cut's own sweeps stream) keeps a cut small next to the work it saves
// launches on: a level of a batched launch serves many matrices at once and saves no
// launch by one matrix's cut, so there a cut only adds its own traffic.
template <typename Sizes>
static int choose_cut(const hss_mat_s* M, int64_t budget, Sizes sizes) {
    if (!g_cut) return 0;
    int lmin = M->max_depth;
    for (const DevNode& d : M->hn)
        if (d.is_leaf) lmin = std::min(lmin, d.depth);
    for (int c = lmin; c >= 3; --c) {
        int64_t k = 0, m = 0, t = 0;
        sizes(c, k, m, t);
        if (k >= 1 && m >= 1 && k <= cut_max_k() && m <= cut_max_k() && t <= budget) return c;
    }
    return 0;
}

// Device buffers of a build, freed (after the stream drains) when it goes out of scope.
struct TmpBufs {
    hipStream_t st;
    std::vector<void*> p;
    explicit TmpBufs(hipStream_t s) : st(s) {}
    ~TmpBufs() {
        (void)hipStreamSynchronize(st);
        for (void* q : p) (void)hipFree(q);
    }
    template <typename T>
    hipError_t alloc(T** q, size_t n) {
        hipError_t e = dalloc(q, n);
        if (e == hipSuccess) p.push_back(*q);
        return e;
    }
    template <typename T>
    hipError_t put(T** q, const std::vector<T>& v) {
        hipError_t e = alloc(q, v.size());
        if (e == hipSuccess && !v.empty()) e = hipMemcpy(*q, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice);
        return e;
    }
};

// One dense product of a cut: T (M x K) and its row tables (relative to the call's
// workspace).
struct CutSpec {
    const void* T;
    int M, K;
    std::vector<CutRow> xr, yr;
};
// The row tables of a cut's products in one device allocation (*out); info = the products
// with their device addresses.
static hipError_t upload_cuts(const std::vector<CutSpec>& cs, char** out, std::vector<CutInfo>& info, int64_t* bytes) {
    size_t tot = 0;
    for (const CutSpec& c : cs) tot += sizeof(CutRow) * (c.xr.size() + c.yr.size());
    char* d = nullptr;
    hipError_t e = dalloc(&d, tot);
    if (e != hipSuccess) return e;
    std::vector<char> h(std::max<size_t>(tot, 1), 0);
    info.clear();
    size_t o = 0;
    for (const CutSpec& c : cs) {
        const size_t ox = o, oy = ox + sizeof(CutRow) * c.xr.size();
        o = oy + sizeof(CutRow) * c.yr.size();
        CutInfo ci;
        ci.T = c.T;
        ci.M = c.M;
        ci.K = c.K;
        ci.xrow = reinterpret_cast<const CutRow*>(d + ox);
        ci.yrow = reinterpret_cast<const CutRow*>(d + oy);
        info.push_back(ci);
        std::memcpy(h.data() + ox, c.xr.data(), sizeof(CutRow) * c.xr.size());
        std::memcpy(h.data() + oy, c.yr.data(), sizeof(CutRow) * c.yr.size());
    }
    e = hipMemcpy(d, h.data(), tot, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        (void)hipFree(d);
        info.clear();
        return e;
    }
    *out = d;
    *bytes = (int64_t)tot;
    return hipSuccess;
}

// Tree cut of the matvec at depth c (mv_cut_depth chooses c, 0 = none): input = the depth-c
// nodes' Z (kv rows each), output = their F (ku rows each), F = T Z, one product.
// mv_cut_build builds T by running the sweep steps above the cut on the identity -- Z of the
// depth-c nodes = unit columns, Kc right-hand sides, on Z / F rows numbered for the nodes at
// depth <= c only -- and gathering the depth-c nodes' F; the matrix's schedule is then
// rebuilt without the steps the cut replaces.
static int mv_cut_depth(const hss_mat_s* M) {
    return choose_cut(M, M->op_len / 32, [&](int c, int64_t& k, int64_t& m, int64_t& t) {
        for (const DevNode& d : M->hn)
            if (d.depth == c) {
                k += d.kv;
                m += d.ku;
            }
        t = k * m;
    });
}
static hss_status mv_cut_build(hss_mat_s* M, int c) {
    if (c == 0) return HSS_OK;
    int Kc = 0, Mc = 0;
    for (const DevNode& d : M->hn)
        if (d.depth == c) {
            Kc += d.kv;
            Mc += d.ku;
        }
    hipStream_t st = M->ctx->stream;
    std::vector<DevNode> hn = M->hn;
    int64_t zt = 0, ft = 0;
    for (DevNode& d : hn) {
        if (d.depth > c) continue;
        d.zrow = zt;
        zt += d.kv;
        d.frow = ft;
        ft += d.ku;
    }
    // row tables of the build (compact rows) and of the matvec workspaces
    std::vector<CutRow> xb, yb, xr, yr;
    for (int i = 0; i < M->nn; ++i) {
        const DevNode &d = hn[i], &o = M->hn[i];
        if (d.depth != c) continue;
        for (int r = 0; r < d.kv; ++r) {
            xb.push_back(CutRow{d.zrow, r, d.kv});
            xr.push_back(CutRow{o.zrow, r, d.kv});
        }
        for (int r = 0; r < d.ku; ++r) {
            yb.push_back(CutRow{d.frow, r, d.ku});
            yr.push_back(CutRow{M->zrows + o.frow, r, d.ku});  // F follows Z in the workspace
        }
    }
    // the steps above the cut: up from depth c - 1 to 1, then down from depth 1 to c
    std::vector<MvRec> work;
    MvSched S;
    S.up.assign(c - 1, Range{});
    S.up_m.assign(c - 1, 0);
    S.up_na.assign(c - 1, 0);
    S.up_k.assign(c - 1, 0);
    S.down.assign(c + 1, Range{});
    S.down_m.assign(c + 1, 0);
    S.down_na.assign(c + 1, 0);
    S.down_k.assign(c + 1, 0);
    auto add = [&](int kind, int depth, Range& r, int& mm, int64_t& na, int64_t& kk) {
        r.off = (int)work.size();
        for (int i = 1; i < M->nn; ++i) {
            if (hn[i].depth != depth) continue;
            work.push_back(mv_record(hn, 0, 0, i, kind));
            int rows, cols;
            mv_shape(hn, i, kind, rows, cols);
            mm = std::max(mm, rows);
            na = std::max(na, ((int64_t)rows * cols + 15) / 16 * 16);
            kk = std::max(kk, (int64_t)cols);
        }
        r.count = (int)work.size() - r.off;
    };
    for (int h = 0; h < c - 1; ++h) add(0, c - 1 - h, S.up[h], S.up_m[h], S.up_na[h], S.up_k[h]);
    for (int d = 1; d <= c; ++d) add(1, d, S.down[d], S.down_m[d], S.down_na[d], S.down_k[d]);
    {
        TmpBufs tb(st);
        double *Zt = nullptr, *Ft = nullptr;
        MvRec* dw = nullptr;
        MvInst* di = nullptr;
        CutRow *dxb = nullptr, *dyb = nullptr;
        HIP_TRY(tb.alloc(&Zt, (size_t)zt * Kc));
        HIP_TRY(tb.alloc(&Ft, (size_t)std::max<int64_t>(ft, 1) * Kc));
        HIP_TRY(hipMemsetAsync(Zt, 0, sizeof(double) * (size_t)zt * Kc, st));
        HIP_TRY(hipMemsetAsync(Ft, 0, sizeof(double) * (size_t)std::max<int64_t>(ft, 1) * Kc, st));
        HIP_TRY(tb.put(&dw, work));
        HIP_TRY(tb.put(&dxb, xb));
        HIP_TRY(tb.put(&dyb, yb));
        std::vector<MvInst> I(1);
        std::memset(I.data(), 0, sizeof(MvInst));
        I[0].nodes = M->d_nodes;
        I[0].op = M->d_op;
        I[0].Z = Zt;
        I[0].F = Ft;
        I[0].alpha = 1.0;
        I[0].nrhs = Kc;
        HIP_TRY(tb.put(&di, I));
        HIP_TRY(launch_cut_seed<double>(Zt, dxb, Kc, Kc, 0, st));
        hss_status s = launch_mv_schedule(di, dw, S, Kc, st);
        if (s != HSS_OK) return s;
        HIP_TRY(dalloc(&M->d_cutT, (size_t)Mc * Kc));
        HIP_TRY(launch_cut_gather<double>(M->d_cutT, Mc, Kc, dyb, Ft, Kc, st));
        HIP_TRY(hipStreamSynchronize(st));
    }
    int64_t tab = 0;
    HIP_TRY(upload_cuts(std::vector<CutSpec>{CutSpec{M->d_cutT, Mc, Kc, xr, yr}}, &M->d_cut, M->cut, &tab));
    M->cut_depth = c;
    std::vector<MvRec> mw;
    build_mv_schedule(std::vector<hss_mat_s*>{M}, mw, M->mvs);
    MvRec* nw = nullptr;
    HIP_TRY(dalloc(&nw, mw.size()));
    hipError_t e = hipMemcpy(nw, mw.data(), sizeof(MvRec) * mw.size(), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        (void)hipFree(nw);
        return fail(HSS_ERR_HIP, std::string("matvec cut: ") + hipGetErrorString(e));
    }
    (void)hipFree(M->d_mvwork);
    M->d_mvwork = nw;
    M->dev_bytes += (int64_t)(sizeof(double) * (size_t)Mc * Kc) + tab;
    return HSS_OK;
}

// One product of a matvec operator: op[dst + r + c * m] = sum_l blob[a + r + l * m] *
// blob[b + l + c * k] for r < m, c < n, l < k (a folded leaf's U B_sibling and U R).
struct OpProduct {
    int64_t a, b, dst;
    int m, n, k;
};

// One piece of a matvec operator, copied from the generator blob:
//   op[dst + r + c * rows] = blob[src + r + c * ld]   (blob[src + c + r * ld] when trans)
// for r < rows, c < cols.
struct OpPiece {
    int64_t src, dst;
    int rows, cols, ld;
    bool trans;
};

// Layout of the matvec operators of every node, column-major, each starting on a
// 16-element boundary (gaps zero): sets DevNode::opZ/opF/opC, appends the pieces that fill
// it and returns its length.  Leaf opZ = V'; branch opZ = [W_l' W_r']; opF = [B_sibling R];
// leaf opC = [D U].
static int64_t plan_mv_operators(hss_mat_s* M, std::vector<OpPiece>& pc, std::vector<OpProduct>& pr) {
    int64_t len = 0;
    auto a16 = [&]() { len = (len + 15) & ~int64_t(15); };
    auto piece = [&](int64_t src, int rows, int cols, int ld, bool trans, int64_t dst) {
        if (rows > 0 && cols > 0) pc.push_back(OpPiece{src, dst, rows, cols, ld, trans});
    };
    for (int i = 0; i < M->nn; ++i) {
        DevNode& d = M->hn[i];
        d.opZ = d.opF = d.opC = -1;
        if (i != 0) {
            a16();
            d.opZ = len;
            const int kv = d.kv;
            if (d.is_leaf) {
                piece(d.V, kv, d.n, d.n, true, len);
                len += (int64_t)kv * d.n;
            } else {
                const DevNode& L = M->hn[d.left];
                const DevNode& R = M->hn[d.right];
                piece(L.W, kv, L.kv, L.kv, true, len);
                piece(R.W, kv, R.kv, R.kv, true, len + (int64_t)L.kv * kv);
                len += (int64_t)kv * (L.kv + R.kv);
            }
            const DevNode& p = M->hn[d.parent];
            const DevNode& sb = M->hn[d.is_left ? p.right : p.left];
            const int ku = d.ku;
            if (!d.fold) {
                a16();
                d.opF = len;
                piece(d.is_left ? p.B12 : p.B21, ku, sb.kv, ku, false, len);
                piece(d.R, ku, p.ku, ku, false, len + (int64_t)sb.kv * ku);
                len += (int64_t)ku * (sb.kv + p.ku);
            }
        }
        if (d.is_leaf) {
            a16();
            d.opC = len;
            piece(d.D, d.m, d.n, d.m, false, len);
            len += (int64_t)d.m * d.n;
            if (d.fold) {
                // [U B_sibling, U R]: U (m x ku) times B12 / B21 (ku x kv_sibling) and R (ku x ku_parent)
                const DevNode& p = M->hn[d.parent];
                const DevNode& sb = M->hn[d.is_left ? p.right : p.left];
                pr.push_back(OpProduct{d.U, d.is_left ? p.B12 : p.B21, len, d.m, sb.kv, d.ku});
                len += (int64_t)d.m * sb.kv;
                pr.push_back(OpProduct{d.U, d.R, len, d.m, p.ku, d.ku});
                len += (int64_t)d.m * p.ku;
            } else {
                piece(d.U, d.m, d.ku, d.m, false, len);
                len += (int64_t)d.m * d.ku;
            }
        }
    }
    a16();
    return len + 16;  // whole-vector reads past the last operator stay in bounds
}

// The matrix of a validated node table whose generators are either on the host (hblob,
// uploaded once) or already on the device (*dblob, an allocation of dalloc on ctx's device).
// On success the matrix owns *dblob and *dblob is set to null; on failure the caller still
// owns it.  The matvec operators are gathered from the device blob by copy tasks
// (k_gemm_tasks), so only the generators themselves cross PCIe, and a device blob not at all.
static hss_status mat_create_impl(hss_ctx_t ctx, const hss_node_desc* nodes, int32_t nn, const double* hblob,
                                  double** dblob, int64_t blob_len, hss_mat_t* out) {
    *out = nullptr;
    hss_status vs = validate(nodes, nn, blob_len);
    if (vs != HSS_OK) return vs;
    HIP_TRY(hipSetDevice(ctx->device));
    auto* M = new (std::nothrow) hss_mat_s;
    if (!M) return fail(HSS_ERR_OOM, "host allocation");
    M->ctx = ctx;
    M->nn = nn;
    M->rows = nodes[0].m;
    M->cols = nodes[0].n;
    M->desc.assign(nodes, nodes + nn);
    M->blob_len = blob_len;
    M->hn.resize(nn);
    int64_t zr = 0, fr = 0;
    for (int i = 0; i < nn; ++i) {
        const hss_node_desc& x = nodes[i];
        DevNode& d = M->hn[i];
        std::memset(&d, 0, sizeof(d));
        d.m = x.m; d.n = x.n; d.ku = x.ku; d.kv = x.kv;
        d.left = x.left; d.right = x.right; d.parent = x.parent;
        d.is_leaf = x.left < 0;
        d.is_left = (x.parent >= 0 && nodes[x.parent].left == i) ? 1 : 0;
        d.depth = x.depth; d.height = x.height;
        d.row_off = x.row_off; d.col_off = x.col_off;
        d.D = x.off_D; d.U = x.off_U; d.V = x.off_V; d.B12 = x.off_B12; d.B21 = x.off_B21;
        d.R = x.off_R; d.W = x.off_W;
        d.zrow = zr; zr += x.kv;
        d.frow = fr; fr += x.ku;
    }
    // recompute heights/depths (the caller's values are advisory)
    for (int i = nn - 1; i >= 0; --i) {
        DevNode& d = M->hn[i];
        d.height = d.is_leaf ? 0 : 1 + std::max(M->hn[d.left].height, M->hn[d.right].height);
    }
    for (int i = 0; i < nn; ++i) M->hn[i].depth = (i == 0) ? 0 : M->hn[M->hn[i].parent].depth + 1;
    M->zrows = zr;
    M->frows = fr;
    for (int i = 0; i < nn; ++i) {
        M->max_height = std::max(M->max_height, M->hn[i].height);
        M->max_depth = std::max(M->max_depth, M->hn[i].depth);
    }
    // work lists
    auto add_range = [&](std::vector<Range>& v, auto pred, int levels, auto level_of) {
        v.assign(levels, Range{});
        for (int L = 0; L < levels; ++L) {
            v[L].off = (int)M->hwork.size();
            for (int i = 0; i < nn; ++i)
                if (pred(i) && level_of(i) == L) M->hwork.push_back(WorkItem{0, i});
            v[L].count = (int)M->hwork.size() - v[L].off;
        }
    };
    add_range(M->by_height, [&](int i) { return i != 0; }, M->max_height, [&](int i) { return M->hn[i].height; });
    add_range(M->by_depth, [&](int i) { return i != 0; }, M->max_depth + 1, [&](int i) { return M->hn[i].depth; });
    M->leaves.off = (int)M->hwork.size();
    for (int i = 0; i < nn; ++i)
        if (M->hn[i].is_leaf) M->hwork.push_back(WorkItem{0, i});
    M->leaves.count = (int)M->hwork.size() - M->leaves.off;
    M->root.off = (int)M->hwork.size();
    M->hwork.push_back(WorkItem{0, 0});
    M->root.count = 1;
    // The matvec tree cut's depth comes first (from the unfolded operator size): the leaves
    // below it fold their down step into their leaf product.
    std::vector<OpPiece> pieces;
    std::vector<OpProduct> prods;
    M->op_len = plan_mv_operators(M, pieces, prods);
    const int cut = mv_cut_depth(M);
    for (DevNode& d : M->hn) d.fold = (d.is_leaf && d.parent >= 0 && d.depth > cut && d.ku > 0) ? 1 : 0;
    pieces.clear();
    prods.clear();
    M->op_len = plan_mv_operators(M, pieces, prods);
    std::vector<MvRec> mvwork;
    build_mv_schedule(std::vector<hss_mat_s*>{M}, mvwork, M->mvs);

    hipError_t e;
    if ((!dblob && (e = dalloc(&M->d_blob, (size_t)blob_len)) != hipSuccess) ||
        (e = dalloc(&M->d_op, (size_t)M->op_len)) != hipSuccess ||
        (e = dalloc(&M->d_mvwork, mvwork.size())) != hipSuccess ||
        (e = dalloc(&M->d_nodes, (size_t)nn)) != hipSuccess ||
        (e = dalloc(&M->d_work, M->hwork.size())) != hipSuccess) {
        free_mat(M);
        return fail(HSS_ERR_OOM, std::string("hss_mat_create: ") + hipGetErrorString(e));
    }
    // from here on a failure frees M and, for a device blob, leaves *dblob to the caller
    auto bail = [&](hipError_t err) {
        (void)hipStreamSynchronize(ctx->stream);
        if (dblob) M->d_blob = nullptr;
        free_mat(M);
        return fail(HSS_ERR_HIP, std::string("hss_mat_create: ") + hipGetErrorString(err));
    };
    hipStream_t st = ctx->stream;
    GemmTask* dT = nullptr;
    if (dblob) M->d_blob = *dblob;
    else if (blob_len > 0 &&
             (e = hipMemcpyAsync(M->d_blob, hblob, sizeof(double) * blob_len, hipMemcpyHostToDevice, st)) != hipSuccess)
        return bail(e);
    {
        std::vector<GemmTask> tasks;
        tasks.reserve(pieces.size());
        for (const OpPiece& q : pieces) {
            GemmTask t{};
            t.m = q.rows; t.n = q.cols; t.flags = GT_COPY | (q.trans ? GT_TA : 0);
            t.lda = std::max(q.ld, 1); t.ldb = 1; t.ldc = q.rows;
            t.a = M->d_blob + q.src; t.c = M->d_op + q.dst;
            tasks.push_back(t);
        }
        for (const OpProduct& q : prods) {
            if (q.m <= 0 || q.n <= 0 || q.k <= 0) continue;
            GemmTask t{};
            t.m = q.m; t.n = q.n; t.k = q.k; t.flags = 0;
            t.lda = q.m; t.ldb = q.k; t.ldc = q.m;
            t.a = M->d_blob + q.a; t.b = M->d_blob + q.b; t.c = M->d_op + q.dst;
            tasks.push_back(t);
        }
        if ((e = hipMemsetAsync(M->d_op, 0, sizeof(double) * (size_t)M->op_len, st)) != hipSuccess) return bail(e);
        if (!tasks.empty()) {
            if ((e = dalloc(&dT, tasks.size())) != hipSuccess ||
                (e = hipMemcpyAsync(dT, tasks.data(), sizeof(GemmTask) * tasks.size(), hipMemcpyHostToDevice, st)) !=
                    hipSuccess ||
                (e = launch_gemm_tasks(dT, (int)tasks.size(), st)) != hipSuccess) {
                (void)hipStreamSynchronize(st);
                (void)hipFree(dT);
                return bail(e);
            }
        }
    }
    if ((e = hipMemcpyAsync(M->d_mvwork, mvwork.data(), sizeof(MvRec) * mvwork.size(), hipMemcpyHostToDevice, st)) !=
            hipSuccess ||
        (e = hipMemcpyAsync(M->d_nodes, M->hn.data(), sizeof(DevNode) * nn, hipMemcpyHostToDevice, st)) !=
            hipSuccess ||
        (e = hipMemcpyAsync(M->d_work, M->hwork.data(), sizeof(WorkItem) * M->hwork.size(), hipMemcpyHostToDevice,
                            st)) != hipSuccess ||
        (e = hipStreamSynchronize(st)) != hipSuccess) {
        (void)hipStreamSynchronize(st);
        (void)hipFree(dT);
        return bail(e);
    }
    (void)hipFree(dT);
    M->dev_bytes = (int64_t)(sizeof(double) * (blob_len + M->op_len) + sizeof(DevNode) * nn +
                             sizeof(WorkItem) * M->hwork.size() + sizeof(MvRec) * mvwork.size());
    hss_status cs = mv_cut_build(M, cut);
    if (cs != HSS_OK) {
        (void)hipStreamSynchronize(st);
        if (dblob) M->d_blob = nullptr;
        free_mat(M);
        return cs;
    }
    if (dblob) *dblob = nullptr;
    ctx->refs++;
    *out = M;
    return HSS_OK;
}

namespace hss {
hss_status host_fail(hss_status st, const std::string& msg) { return fail(st, msg); }
hss_status host_validate_tree(const hss_node_desc* nodes, int nn) {
    if (nn <= 0 || !nodes) return fail(HSS_ERR_ARG, "empty node table");
    std::vector<hss_node_desc> t(nodes, nodes + nn);
    for (auto& x : t) {
        x.ku = x.kv = 0;
        x.off_D = x.off_U = x.off_V = x.off_B12 = x.off_B21 = x.off_R = x.off_W = -1;
    }
    for (int i = 0; i < nn; ++i)
        if (t[i].left < 0 && (int64_t)t[i].m * t[i].n > 0) t[i].off_D = 0;  // D is checked against a blob of m*n
    int64_t big = 0;
    for (auto& x : t) big = std::max<int64_t>(big, (int64_t)x.m * x.n);
    return validate(t.data(), nn, big);
}
hss_status host_mat_from_device(hss_ctx_t ctx, const hss_node_desc* nodes, int nn, double** dblob, int64_t blob_len,
                                hss_mat_t* out) {
    return mat_create_impl(ctx, nodes, nn, nullptr, dblob, blob_len, out);
}
int host_ctx_device(hss_ctx_t ctx) { return ctx ? ctx->device : -1; }
hss_ctx_t host_mat_ctx(hss_mat_t mat) { return mat ? mat->ctx : nullptr; }
const hss_node_desc* host_mat_nodes(hss_mat_t mat, int* nn) {
    if (nn) *nn = mat ? mat->nn : 0;
    return mat ? mat->desc.data() : nullptr;
}
const double* host_mat_blob(hss_mat_t mat) { return mat ? mat->d_blob : nullptr; }
int64_t host_fact_rows(hss_fact_t fact) { return fact && fact->mat ? fact->mat->rows : 0; }
}  // namespace hss

extern "C" hss_status hss_mat_create(hss_ctx_t ctx, const hss_node_desc* nodes, int32_t nn, const double* blob,
                                     int64_t blob_len, hss_mat_t* out) {
    if (!ctx || !nodes || !out || (blob_len > 0 && !blob)) return fail(HSS_ERR_ARG, "null argument");
    return mat_create_impl(ctx, nodes, nn, blob, nullptr, blob_len, out);
}

static void mat_release(hss_mat_s* M) {
    if (!M || --M->refs > 0) return;
    hss_ctx_s* c = M->ctx;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    free_mat(M);
    ctx_release(c);
}
// Adjoint of an HSS matrix as a new matrix on the same context (reference
// src/hssmatrix.jl:165-171): leaf D -> D', U <-> V; branch B12 -> B21', B21 -> B12';
// every node's R <-> W; rows and columns exchange sizes and offsets.  The generators are
// copied (and transposed) on the device into a new blob packed in the same node order.
extern "C" hss_status hss_mat_adjoint(hss_mat_t A, hss_mat_t* out) {
    if (!A || !out) return fail(HSS_ERR_ARG, "null argument");
    *out = nullptr;
    HIP_TRY(hipSetDevice(A->ctx->device));
    const int nn = A->nn;
    const std::vector<hss_node_desc>& d = A->desc;
    std::vector<hss_node_desc> t(d);
    std::vector<GemmTask> tasks;
    std::vector<int64_t> dst;  // each task's offset in the new blob
    int64_t len = 0;
    // a rows x cols block of A's blob, copied (or transposed into cols x rows) to the end
    // of the new blob; -1 for an absent generator
    auto put = [&](int64_t off, int64_t rows, int64_t cols, bool trans) -> int64_t {
        if (off < 0) return -1;
        const int64_t at = len;
        len += rows * cols;
        if (rows > 0 && cols > 0) {
            GemmTask g{};
            g.m = (int)(trans ? cols : rows);
            g.n = (int)(trans ? rows : cols);
            g.flags = GT_COPY | (trans ? GT_TA : 0);
            g.lda = (int)rows; g.ldb = 1; g.ldc = g.m;
            g.a = A->d_blob + off;
            tasks.push_back(g);
            dst.push_back(at);
        }
        return at;
    };
    for (int i = 0; i < nn; ++i) {
        const hss_node_desc& x = d[i];
        hss_node_desc& y = t[i];
        y.m = x.n;
        y.n = x.m;
        y.ku = x.kv;
        y.kv = x.ku;
        y.row_off = x.col_off;
        y.col_off = x.row_off;
        y.off_D = y.off_U = y.off_V = y.off_B12 = y.off_B21 = y.off_R = y.off_W = -1;
        if (x.left < 0) {
            y.off_D = put(x.off_D, x.m, x.n, true);
            y.off_U = put(x.off_V, x.n, x.kv, false);
            y.off_V = put(x.off_U, x.m, x.ku, false);
        } else {
            const hss_node_desc &l = d[x.left], &r = d[x.right];
            y.off_B12 = put(x.off_B21, r.ku, l.kv, true);
            y.off_B21 = put(x.off_B12, l.ku, r.kv, true);
        }
        if (x.parent >= 0) {
            const hss_node_desc& p = d[x.parent];
            y.off_R = put(x.off_W, x.kv, p.kv, false);
            y.off_W = put(x.off_R, x.ku, p.ku, false);
        }
    }
    hipStream_t st = A->ctx->stream;
    double* dB = nullptr;
    GemmTask* dT = nullptr;
    struct Free {
        double** b;
        GemmTask** t;
        ~Free() {
            (void)hipFree(*b);
            (void)hipFree(*t);
        }
    } guard{&dB, &dT};
    HIP_TRY(dalloc(&dB, (size_t)std::max<int64_t>(len, 1)));
    for (size_t k = 0; k < tasks.size(); ++k) tasks[k].c = dB + dst[k];
    HIP_TRY(hipMemsetAsync(dB, 0, sizeof(double) * (size_t)std::max<int64_t>(len, 1), st));
    if (!tasks.empty()) {
        HIP_TRY(dalloc(&dT, tasks.size()));
        HIP_TRY(hipMemcpyAsync(dT, tasks.data(), sizeof(GemmTask) * tasks.size(), hipMemcpyHostToDevice, st));
        HIP_TRY(launch_gemm_tasks(dT, (int)tasks.size(), st));
    }
    HIP_TRY(hipStreamSynchronize(st));
    // the new matrix takes dB over (the guard then frees null)
    return mat_create_impl(A->ctx, t.data(), nn, nullptr, &dB, len, out);
}

// ---------------------------------------------------------------------------
// generator interchange file: hss_gen_header | node table | Float64 blob
// ---------------------------------------------------------------------------
static_assert(sizeof(hss_gen_header) == 40, "hss_gen_header is 40 bytes");
static_assert(sizeof(hss_node_desc) == 112, "hss_node_desc is 112 bytes");

namespace {
struct OwnedFile {
    FILE* f = nullptr;
    explicit OwnedFile(FILE* p) : f(p) {}
    ~OwnedFile() {
        if (f) std::fclose(f);
    }
};
}  // namespace

extern "C" hss_status hss_gen_write(const char* path, const hss_node_desc* nodes, int32_t nn, const double* blob,
                                    int64_t blob_len) {
    if (!path || !nodes || (blob_len > 0 && !blob)) return fail(HSS_ERR_ARG, "null argument");
    if (blob_len < 0) return fail(HSS_ERR_ARG, "blob_len < 0");
    hss_status s = validate(nodes, nn, blob_len);
    if (s != HSS_OK) return s;
    hss_gen_header h{};
    std::memcpy(h.magic, "HSSG", 4);
    h.version = 1;
    h.elem_size = sizeof(double);
    h.num_nodes = nn;
    h.blob_len = blob_len;
    OwnedFile F(std::fopen(path, "wb"));
    if (!F.f) return fail(HSS_ERR_IO, std::string("cannot open ") + path + " for writing");
    bool ok = std::fwrite(&h, sizeof h, 1, F.f) == 1 &&
              std::fwrite(nodes, sizeof(hss_node_desc), (size_t)nn, F.f) == (size_t)nn &&
              (blob_len == 0 || std::fwrite(blob, sizeof(double), (size_t)blob_len, F.f) == (size_t)blob_len);
    ok = (std::fclose(F.f) == 0) && ok;
    F.f = nullptr;
    if (!ok) return fail(HSS_ERR_IO, std::string("writing ") + path + " failed");
    return HSS_OK;
}

extern "C" hss_status hss_gen_read(const char* path, hss_node_desc* nodes, int32_t nodes_cap, int32_t* num_nodes,
                                   double* blob, int64_t blob_cap, int64_t* blob_len) {
    if (!path) return fail(HSS_ERR_ARG, "null path");
    OwnedFile F(std::fopen(path, "rb"));
    if (!F.f) return fail(HSS_ERR_IO, std::string("cannot open ") + path);
    hss_gen_header h;
    if (std::fread(&h, sizeof h, 1, F.f) != 1) return fail(HSS_ERR_ARG, "generator file shorter than its header");
    if (std::memcmp(h.magic, "HSSG", 4) != 0) return fail(HSS_ERR_ARG, "not a generator file (magic)");
    if (h.version != 1) return fail(HSS_ERR_ARG, "unsupported generator file version " + std::to_string(h.version));
    if (h.elem_size != sizeof(double)) return fail(HSS_ERR_ARG, "generator file element size is not 8 (Float64)");
    if (h.num_nodes <= 0 || h.blob_len < 0 || h.blob_len > ((int64_t)1 << 56))
        return fail(HSS_ERR_ARG, "bad sizes in the generator file header");
    // the sizes must account for the whole file, before anything is allocated from them
    const int64_t want = (int64_t)sizeof h + (int64_t)h.num_nodes * (int64_t)sizeof(hss_node_desc) +
                         h.blob_len * (int64_t)sizeof(double);
    if (fseeko(F.f, 0, SEEK_END) != 0) return fail(HSS_ERR_IO, std::string("cannot seek in ") + path);
    const int64_t have = (int64_t)ftello(F.f);
    if (have != want)
        return fail(HSS_ERR_ARG, "generator file is " + std::to_string(have) + " bytes, its header says " +
                                     std::to_string(want));
    if (num_nodes) *num_nodes = h.num_nodes;
    if (blob_len) *blob_len = h.blob_len;
    if (!nodes || (!blob && h.blob_len > 0)) return HSS_OK;
    if (nodes_cap < h.num_nodes || blob_cap < h.blob_len)
        return fail(HSS_ERR_DIM, "buffers smaller than the generator file's node table or blob");
    if (fseeko(F.f, (off_t)sizeof h, SEEK_SET) != 0) return fail(HSS_ERR_IO, std::string("cannot seek in ") + path);
    if (std::fread(nodes, sizeof(hss_node_desc), (size_t)h.num_nodes, F.f) != (size_t)h.num_nodes ||
        (h.blob_len > 0 && std::fread(blob, sizeof(double), (size_t)h.blob_len, F.f) != (size_t)h.blob_len))
        return fail(HSS_ERR_IO, std::string("reading ") + path + " failed");
    return validate(nodes, h.num_nodes, h.blob_len);
}

extern "C" hss_status hss_mat_load(hss_ctx_t ctx, const char* path, hss_mat_t* out) {
    if (!ctx || !out) return fail(HSS_ERR_ARG, "null argument");
    *out = nullptr;
    int32_t nn = 0;
    int64_t bl = 0;
    hss_status s = hss_gen_read(path, nullptr, 0, &nn, nullptr, 0, &bl);
    if (s != HSS_OK) return s;
    std::vector<hss_node_desc> nodes((size_t)nn);
    std::vector<double> blob((size_t)std::max<int64_t>(bl, 1));
    s = hss_gen_read(path, nodes.data(), nn, &nn, blob.data(), bl, &bl);
    if (s != HSS_OK) return s;
    return hss_mat_create(ctx, nodes.data(), nn, blob.data(), bl, out);
}

extern "C" hss_status hss_mat_generators(hss_mat_t A, hss_node_desc* nodes, int32_t nodes_cap, int32_t* num_nodes,
                                         double* blob, int64_t blob_cap, int64_t* blob_len) {
    if (!A || !num_nodes || !blob_len) return fail(HSS_ERR_ARG, "null argument");
    *num_nodes = A->nn;
    *blob_len = A->blob_len;
    if ((nodes && nodes_cap < A->nn) || (blob && blob_cap < A->blob_len))
        return fail(HSS_ERR_DIM, "output buffers too small");
    if (nodes) std::memcpy(nodes, A->desc.data(), sizeof(hss_node_desc) * A->nn);
    if (blob && A->blob_len > 0) {
        HIP_TRY(hipSetDevice(A->ctx->device));
        HIP_TRY(hipStreamSynchronize(A->ctx->stream));
        HIP_TRY(hipMemcpy(blob, A->d_blob, sizeof(double) * A->blob_len, hipMemcpyDeviceToHost));
    }
    return HSS_OK;
}

extern "C" hss_status hss_mat_save(hss_mat_t A, const char* path) {
    if (!A || !path) return fail(HSS_ERR_ARG, "null argument");
    HIP_TRY(hipSetDevice(A->ctx->device));
    std::vector<double> blob((size_t)std::max<int64_t>(A->blob_len, 1));
    if (A->blob_len > 0)
        HIP_TRY(hipMemcpy(blob.data(), A->d_blob, sizeof(double) * (size_t)A->blob_len, hipMemcpyDeviceToHost));
    return hss_gen_write(path, A->desc.data(), A->nn, blob.data(), A->blob_len);
}

// ---------------------------------------------------------------------------
// HSS x HSS product (src/matmul.jl:65-118)
// ---------------------------------------------------------------------------
// C = A B for A and B on one tree (A's column clusters are B's row clusters).  C's ranks
// are the sums of A's and B's (ku_C = ku_A + ku_B, kv_C = kv_A + kv_B); like the
// reference, nothing is recompressed.  For a branch i with children l, r:
//   up:    Z_leaf = V_A' U_B,   Z_i = W_A,l' Z_l R_B,l + W_A,r' Z_r R_B,r
//   down:  F_l = B_A12 Z_r B_B21 + R_A,l F_i W_B,l',  F_r = B_A21 Z_l B_B12 + R_A,r F_i W_B,r'
//          (F_i = 0 at the root)
//   C:     B12 = [B_A12, R_A,l F_i W_B,r'; 0, B_B12],  B21 = [B_A21, R_A,r F_i W_B,l'; 0, B_B21]
//          R_l = [R_A,l, B_A12 Z_r R_B,r; 0, R_B,l],   W_l = [W_A,l, 0; B_B21' Z_r' W_A,r, W_B,l]
//          R_r = [R_A,r, B_A21 Z_l R_B,l; 0, R_B,r],   W_r = [W_A,r, 0; B_B12' Z_l' W_A,l, W_B,r]
//   leaf:  D = D_A D_B + U_A F V_B',  U = [U_A, D_A U_B],  V = [D_B' V_A, V_B]
// The products run on the device as ordered phases of small GEMM tasks (k_gemm_tasks)
// into a zeroed generator blob, which then becomes the matrix C without leaving the device.
extern "C" hss_status hss_mat_mul(hss_mat_t A, hss_mat_t B, hss_mat_t* out) {
    if (!A || !B || !out) return fail(HSS_ERR_ARG, "null argument");
    *out = nullptr;
    if (A->ctx != B->ctx) return fail(HSS_ERR_ARG, "the matrices live on different contexts");
    if (A->cols != B->rows)
        return fail(HSS_ERR_DIM, "second dimension of A (" + std::to_string(A->cols) +
                                     ") does not match first dimension of B (" + std::to_string(B->rows) + ")");
    const std::vector<hss_node_desc>& a = A->desc;
    const std::vector<hss_node_desc>& b = B->desc;
    const int nn = A->nn;
    if (B->nn != nn) return fail(HSS_ERR_DIM, "clusters don't match: the trees of A and B differ");
    for (int i = 0; i < nn; ++i)
        if (a[i].parent != b[i].parent || a[i].left != b[i].left || a[i].right != b[i].right || a[i].n != b[i].m)
            return fail(HSS_ERR_DIM, "clusters don't match at node " + std::to_string(i));
    // C's node table and generator layout
    std::vector<hss_node_desc> c(nn);
    int64_t len = 0;
    auto take = [](int64_t& cur, int64_t cnt) {
        const int64_t o = cur;
        cur += cnt;
        return o;
    };
    for (int i = 0; i < nn; ++i) {
        c[i] = a[i];
        c[i].n = b[i].n;
        c[i].ku = a[i].ku + b[i].ku;
        c[i].kv = a[i].kv + b[i].kv;
        c[i].col_off = b[i].col_off;
        c[i].off_D = c[i].off_U = c[i].off_V = c[i].off_B12 = c[i].off_B21 = c[i].off_R = c[i].off_W = -1;
    }
    for (int i = 0; i < nn; ++i) {
        hss_node_desc& x = c[i];
        if (x.left < 0) {
            x.off_D = take(len, (int64_t)x.m * x.n);
            x.off_U = take(len, (int64_t)x.m * x.ku);
            x.off_V = take(len, (int64_t)x.n * x.kv);
        } else {
            x.off_B12 = take(len, (int64_t)c[x.left].ku * c[x.right].kv);
            x.off_B21 = take(len, (int64_t)c[x.right].ku * c[x.left].kv);
        }
        if (x.parent >= 0) {
            x.off_R = take(len, (int64_t)x.ku * c[x.parent].ku);
            x.off_W = take(len, (int64_t)x.kv * c[x.parent].kv);
        }
    }
    // device workspace: per non-root node Z (kv_A x ku_B), F (ku_A x kv_B), TZ = Z R_B
    // (kv_A x ku_B of the parent), P (B_A12 Z_r or B_A21 Z_l: ku_A x ku_B of the sibling),
    // Y (Z_r B_B21 or Z_l B_B12: kv_A of the sibling x kv_B), Q = R_A F_parent (ku_A x
    // kv_B of the parent); per non-root leaf TL = U_A F (m x kv_B)
    std::vector<int64_t> oZ(nn, 0), oF(nn, 0), oTZ(nn, 0), oP(nn, 0), oY(nn, 0), oQ(nn, 0), oTL(nn, 0);
    int64_t wl = 0;
    for (int i = 1; i < nn; ++i) {
        const int p = a[i].parent, sib = (a[p].left == i) ? a[p].right : a[p].left;
        oZ[i] = take(wl, (int64_t)a[i].kv * b[i].ku);
        oF[i] = take(wl, (int64_t)a[i].ku * b[i].kv);
        oTZ[i] = take(wl, (int64_t)a[i].kv * b[p].ku);
        oP[i] = take(wl, (int64_t)a[i].ku * b[sib].ku);
        oY[i] = take(wl, (int64_t)a[sib].kv * b[i].kv);
        oQ[i] = take(wl, (int64_t)a[i].ku * b[p].kv);
        if (a[i].left < 0) oTL[i] = take(wl, (int64_t)a[i].m * b[i].kv);
    }
    HIP_TRY(hipSetDevice(A->ctx->device));
    hipStream_t st = A->ctx->stream;
    double* dC = nullptr;
    double* dW = nullptr;
    GemmTask* dT = nullptr;
    struct Free {
        double** c;
        double** w;
        GemmTask** t;
        ~Free() {
            (void)hipFree(*c);
            (void)hipFree(*w);
            (void)hipFree(*t);
        }
    } guard{&dC, &dW, &dT};
    HIP_TRY(dalloc(&dC, (size_t)std::max<int64_t>(len, 1)));
    HIP_TRY(dalloc(&dW, (size_t)std::max<int64_t>(wl, 1)));
    const double* pa = A->d_blob;
    const double* pb = B->d_blob;
    auto PA = [&](int64_t o) { return pa + o; };
    auto PB = [&](int64_t o) { return pb + o; };
    auto PC = [&](int64_t o) { return dC + o; };
    auto PW = [&](int64_t o) { return dW + o; };
    std::deque<std::vector<GemmTask>> ph;  // appending keeps references to earlier phases valid
    auto phase = [&]() -> std::vector<GemmTask>& {
        ph.emplace_back();
        return ph.back();
    };
    // c = op(x) op(y) (+ c); zero-sized or k = 0 products leave the zeroed c alone
    auto mm = [](std::vector<GemmTask>& v, int m, int n, int k, const double* x, int ldx, bool tx, const double* y,
                 int ldy, bool ty, double* z, int ldz, bool acc) {
        if (m <= 0 || n <= 0 || k <= 0) return;
        GemmTask t{};
        t.m = m; t.n = n; t.k = k;
        t.flags = (tx ? GT_TA : 0) | (ty ? GT_TB : 0) | (acc ? GT_ACC : 0);
        t.lda = std::max(ldx, 1); t.ldb = std::max(ldy, 1); t.ldc = std::max(ldz, 1);
        t.a = x; t.b = y; t.c = z;
        v.push_back(t);
    };
    auto cp = [](std::vector<GemmTask>& v, int m, int n, const double* x, int ldx, double* z, int ldz) {
        if (m <= 0 || n <= 0) return;
        GemmTask t{};
        t.m = m; t.n = n; t.flags = GT_COPY;
        t.lda = std::max(ldx, 1); t.ldb = 1; t.ldc = std::max(ldz, 1);
        t.a = x; t.c = z;
        v.push_back(t);
    };
    const std::vector<DevNode>& hn = A->hn;
    // up: Z by height
    for (int h = 0; h < A->max_height; ++h) {
        auto& u1 = phase();
        for (int i = 1; i < nn; ++i) {
            if (hn[i].height != h) continue;
            if (a[i].left < 0) {
                mm(u1, a[i].kv, b[i].ku, a[i].n, PA(a[i].off_V), a[i].n, true, PB(b[i].off_U), b[i].m, false,
                   PW(oZ[i]), a[i].kv, false);
            } else {
                for (int ch : {a[i].left, a[i].right})
                    mm(u1, a[ch].kv, b[i].ku, b[ch].ku, PW(oZ[ch]), a[ch].kv, false, PB(b[ch].off_R), b[ch].ku, false,
                       PW(oTZ[ch]), a[ch].kv, false);
            }
        }
        if (h == 0) continue;
        auto& u2 = phase();
        auto& u3 = phase();
        for (int i = 1; i < nn; ++i) {
            if (hn[i].height != h || a[i].left < 0) continue;
            const int l = a[i].left, r = a[i].right;
            mm(u2, a[i].kv, b[i].ku, a[l].kv, PA(a[l].off_W), a[l].kv, true, PW(oTZ[l]), a[l].kv, false, PW(oZ[i]),
               a[i].kv, false);
            mm(u3, a[i].kv, b[i].ku, a[r].kv, PA(a[r].off_W), a[r].kv, true, PW(oTZ[r]), a[r].kv, false, PW(oZ[i]),
               a[i].kv, true);
        }
    }
    // down: F, and C's couplings and translators, by depth
    for (int d = 0; d <= A->max_depth; ++d) {
        auto& d1 = phase();
        auto& d2 = phase();
        auto& d3 = phase();
        for (int i = 0; i < nn; ++i) {
            if (hn[i].depth != d || a[i].left < 0) continue;
            const int l = a[i].left, r = a[i].right;
            const bool root = i == 0;
            const int kuCl = c[l].ku, kuCr = c[r].ku;
            // D1: P, Y, Q and the blocks copied from A and B
            mm(d1, a[l].ku, b[r].ku, a[r].kv, PA(a[i].off_B12), a[l].ku, false, PW(oZ[r]), a[r].kv, false, PW(oP[l]),
               a[l].ku, false);
            mm(d1, a[r].ku, b[l].ku, a[l].kv, PA(a[i].off_B21), a[r].ku, false, PW(oZ[l]), a[l].kv, false, PW(oP[r]),
               a[r].ku, false);
            mm(d1, a[r].kv, b[l].kv, b[r].ku, PW(oZ[r]), a[r].kv, false, PB(b[i].off_B21), b[r].ku, false, PW(oY[l]),
               a[r].kv, false);
            mm(d1, a[l].kv, b[r].kv, b[l].ku, PW(oZ[l]), a[l].kv, false, PB(b[i].off_B12), b[l].ku, false, PW(oY[r]),
               a[l].kv, false);
            cp(d1, a[l].ku, a[r].kv, PA(a[i].off_B12), a[l].ku, PC(c[i].off_B12), kuCl);
            cp(d1, b[l].ku, b[r].kv, PB(b[i].off_B12), b[l].ku, PC(c[i].off_B12) + a[l].ku + (int64_t)a[r].kv * kuCl,
               kuCl);
            cp(d1, a[r].ku, a[l].kv, PA(a[i].off_B21), a[r].ku, PC(c[i].off_B21), kuCr);
            cp(d1, b[r].ku, b[l].kv, PB(b[i].off_B21), b[r].ku, PC(c[i].off_B21) + a[r].ku + (int64_t)a[l].kv * kuCr,
               kuCr);
            if (!root) {
                for (int ch : {l, r}) {
                    mm(d1, a[ch].ku, b[i].kv, a[i].ku, PA(a[ch].off_R), a[ch].ku, false, PW(oF[i]), a[i].ku, false,
                       PW(oQ[ch]), a[ch].ku, false);
                    cp(d1, a[ch].ku, a[i].ku, PA(a[ch].off_R), a[ch].ku, PC(c[ch].off_R), c[ch].ku);
                    cp(d1, b[ch].ku, b[i].ku, PB(b[ch].off_R), b[ch].ku,
                       PC(c[ch].off_R) + a[ch].ku + (int64_t)a[i].ku * c[ch].ku, c[ch].ku);
                    cp(d1, a[ch].kv, a[i].kv, PA(a[ch].off_W), a[ch].kv, PC(c[ch].off_W), c[ch].kv);
                    cp(d1, b[ch].kv, b[i].kv, PB(b[ch].off_W), b[ch].kv,
                       PC(c[ch].off_W) + a[ch].kv + (int64_t)a[i].kv * c[ch].kv, c[ch].kv);
                }
            }
            // D2: F = B_A Y, and the product blocks of C
            mm(d2, a[l].ku, b[l].kv, a[r].kv, PA(a[i].off_B12), a[l].ku, false, PW(oY[l]), a[r].kv, false, PW(oF[l]),
               a[l].ku, false);
            mm(d2, a[r].ku, b[r].kv, a[l].kv, PA(a[i].off_B21), a[r].ku, false, PW(oY[r]), a[l].kv, false, PW(oF[r]),
               a[r].ku, false);
            if (!root) {
                for (int ch : {l, r}) {
                    const int sb = ch == l ? r : l;
                    // R_ch[0:ku_A,ch, ku_A,i:] = P_ch R_B,sib;  W_ch[kv_A,ch:, 0:kv_A,i] = Y_ch' W_A,sib
                    mm(d2, a[ch].ku, b[i].ku, b[sb].ku, PW(oP[ch]), a[ch].ku, false, PB(b[sb].off_R), b[sb].ku, false,
                       PC(c[ch].off_R) + (int64_t)a[i].ku * c[ch].ku, c[ch].ku, false);
                    mm(d2, b[ch].kv, a[i].kv, a[sb].kv, PW(oY[ch]), a[sb].kv, true, PA(a[sb].off_W), a[sb].kv, false,
                       PC(c[ch].off_W) + a[ch].kv, c[ch].kv, false);
                }
                // B12[0:ku_A,l, kv_A,r:] = Q_l W_B,r';  B21[0:ku_A,r, kv_A,l:] = Q_r W_B,l'
                mm(d2, a[l].ku, b[r].kv, b[i].kv, PW(oQ[l]), a[l].ku, false, PB(b[r].off_W), b[r].kv, true,
                   PC(c[i].off_B12) + (int64_t)a[r].kv * kuCl, kuCl, false);
                mm(d2, a[r].ku, b[l].kv, b[i].kv, PW(oQ[r]), a[r].ku, false, PB(b[l].off_W), b[l].kv, true,
                   PC(c[i].off_B21) + (int64_t)a[l].kv * kuCr, kuCr, false);
                // D3: F += Q W_B'
                mm(d3, a[l].ku, b[l].kv, b[i].kv, PW(oQ[l]), a[l].ku, false, PB(b[l].off_W), b[l].kv, true, PW(oF[l]),
                   a[l].ku, true);
                mm(d3, a[r].ku, b[r].kv, b[i].kv, PW(oQ[r]), a[r].ku, false, PB(b[r].off_W), b[r].kv, true, PW(oF[r]),
                   a[r].ku, true);
            }
        }
    }
    // leaves: D = D_A D_B + (U_A F) V_B', U = [U_A, D_A U_B], V = [D_B' V_A, V_B]
    {
        auto& l1 = phase();
        auto& l2 = phase();
        for (int i = 0; i < nn; ++i) {
            if (a[i].left >= 0) continue;
            const int m = a[i].m, n = b[i].n, k = a[i].n;
            const bool root = i == 0;
            mm(l1, m, n, k, PA(a[i].off_D), m, false, PB(b[i].off_D), b[i].m, false, PC(c[i].off_D), m, false);
            cp(l1, m, a[i].ku, PA(a[i].off_U), m, PC(c[i].off_U), m);
            mm(l1, m, b[i].ku, k, PA(a[i].off_D), m, false, PB(b[i].off_U), b[i].m, false,
               PC(c[i].off_U) + (int64_t)a[i].ku * m, m, false);
            mm(l1, n, a[i].kv, b[i].m, PB(b[i].off_D), b[i].m, true, PA(a[i].off_V), a[i].n, false, PC(c[i].off_V), n,
               false);
            cp(l1, n, b[i].kv, PB(b[i].off_V), n, PC(c[i].off_V) + (int64_t)a[i].kv * n, n);
            if (!root) {
                mm(l1, m, b[i].kv, a[i].ku, PA(a[i].off_U), m, false, PW(oF[i]), a[i].ku, false, PW(oTL[i]), m, false);
                mm(l2, m, n, b[i].kv, PW(oTL[i]), m, false, PB(b[i].off_V), n, true, PC(c[i].off_D), m, true);
            }
        }
    }
    size_t ntask = 0;
    for (auto& v : ph) ntask += v.size();
    std::vector<GemmTask> all;
    all.reserve(ntask);
    for (auto& v : ph) all.insert(all.end(), v.begin(), v.end());
    HIP_TRY(dalloc(&dT, std::max<size_t>(ntask, 1)));
    HIP_TRY(hipMemsetAsync(dC, 0, sizeof(double) * (size_t)std::max<int64_t>(len, 1), st));
    HIP_TRY(hipMemsetAsync(dW, 0, sizeof(double) * (size_t)std::max<int64_t>(wl, 1), st));
    if (ntask) HIP_TRY(hipMemcpyAsync(dT, all.data(), sizeof(GemmTask) * ntask, hipMemcpyHostToDevice, st));
    size_t at = 0;
    for (auto& v : ph) {
        if (!v.empty()) {
            ProfScope ps("k_gemm_tasks", 0, (int)v.size(), st);
            HIP_TRY(launch_gemm_tasks(dT + at, (int)v.size(), st));
        }
        at += v.size();
    }
    HIP_TRY(hipStreamSynchronize(st));
    // C's generators stay on the device: the matrix takes dC over (the guard then frees null)
    return mat_create_impl(A->ctx, c.data(), nn, nullptr, &dC, len, out);
}

extern "C" hss_status hss_mat_destroy(hss_mat_t M) {
    mat_release(M);
    return HSS_OK;
}

extern "C" hss_status hss_mat_size(hss_mat_t M, int64_t* rows, int64_t* cols) {
    if (!M) return fail(HSS_ERR_ARG, "null matrix");
    if (rows) *rows = M->rows;
    if (cols) *cols = M->cols;
    return HSS_OK;
}

extern "C" int64_t hss_mat_device_bytes(hss_mat_t M) { return M ? M->dev_bytes : 0; }

// ---------------------------------------------------------------------------
// matvec
// ---------------------------------------------------------------------------
static hss_status ensure_zf(hss_mat_s* M, int nrhs) {
    if ((size_t)nrhs <= M->zf_cap) return HSS_OK;
    (void)hipFree(M->d_Z);
    M->d_Z = nullptr;
    M->zf_cap = 0;
    HIP_TRY(dalloc(&M->d_Z, (size_t)std::max<int64_t>(M->zrows + M->frows, 1) * nrhs));
    M->zf_cap = nrhs;
    return HSS_OK;
}

static hss_status matmul_enqueue(hss_mat_s* M, int nrhs, double alpha, const double* dB, int64_t ldb, double beta,
                                 double* dC, int64_t ldc) {
    if (nrhs < 0) return fail(HSS_ERR_ARG, "nrhs < 0");
    if (nrhs == 0 || M->rows == 0) return HSS_OK;
    if (ldb < std::max<int64_t>(M->cols, 1) || ldc < std::max<int64_t>(M->rows, 1))
        return fail(HSS_ERR_DIM, "leading dimension smaller than the matrix size");
    hss_status s = ensure_zf(M, nrhs);
    if (s != HSS_OK) return s;
    hipStream_t st = M->ctx->stream;
    MvInst I;
    std::memset(&I, 0, sizeof(I));
    I.nodes = M->d_nodes;
    I.op = M->d_op;
    I.B = dB;
    I.C = dC;
    I.ldb = ldb;
    I.ldc = ldc;
    I.Z = M->d_Z;
    I.F = I.Z + M->zrows * nrhs;
    I.alpha = alpha;
    I.beta = beta;
    I.nrhs = nrhs;
    const void* dv = nullptr;
    HIP_TRY(M->mvinst.get(&I, sizeof(I), st, &dv));
    return launch_mv_schedule(reinterpret_cast<const MvInst*>(dv), M->d_mvwork, M->mvs, nrhs, st);
}

extern "C" hss_status hss_matmul_dev(hss_mat_t M, int32_t nrhs, double alpha, const double* dB, int64_t ldb,
                                     double beta, double* dC, int64_t ldc) {
    if (!M) return fail(HSS_ERR_ARG, "null matrix");
    HIP_TRY(hipSetDevice(M->ctx->device));
    return matmul_enqueue(M, nrhs, alpha, dB, ldb, beta, dC, ldc);
}

static hss_status ensure_io(double** a, double** b, size_t* cap, size_t need) {
    if (need <= *cap) return HSS_OK;
    (void)hipFree(*a);
    (void)hipFree(*b);
    *a = *b = nullptr;
    *cap = 0;
    HIP_TRY(dalloc(a, need));
    HIP_TRY(dalloc(b, need));
    *cap = need;
    return HSS_OK;
}

extern "C" hss_status hss_matmul(hss_mat_t M, int32_t nrhs, double alpha, const double* B, int64_t ldb, double beta,
                                 double* C, int64_t ldc) {
    if (!M || (!B && nrhs > 0) || (!C && nrhs > 0)) return fail(HSS_ERR_ARG, "null argument");
    if (nrhs < 0) return fail(HSS_ERR_ARG, "nrhs < 0");
    if (nrhs == 0) return HSS_OK;
    if (ldb < std::max<int64_t>(M->cols, 1) || ldc < std::max<int64_t>(M->rows, 1))
        return fail(HSS_ERR_DIM, "leading dimension smaller than the matrix size");
    HIP_TRY(hipSetDevice(M->ctx->device));
    const int64_t nb = std::max<int64_t>(M->cols, 1), nc = std::max<int64_t>(M->rows, 1);
    hss_status s = ensure_io(&M->d_io_b, &M->d_io_c, &M->io_cap, (size_t)std::max(nb, nc) * nrhs);
    if (s != HSS_OK) return s;
    hipStream_t st = M->ctx->stream;
    HIP_TRY(hipMemcpy2DAsync(M->d_io_b, nb * sizeof(double), B, ldb * sizeof(double), M->cols * sizeof(double), nrhs,
                             hipMemcpyHostToDevice, st));
    if (beta != 0.0)
        HIP_TRY(hipMemcpy2DAsync(M->d_io_c, nc * sizeof(double), C, ldc * sizeof(double), M->rows * sizeof(double),
                                 nrhs, hipMemcpyHostToDevice, st));
    HIP_TRY(hipEventRecord(M->ctx->ev0, st));
    s = matmul_enqueue(M, nrhs, alpha, M->d_io_b, nb, beta, M->d_io_c, nc);
    if (s != HSS_OK) return s;
    HIP_TRY(hipEventRecord(M->ctx->ev1, st));
    HIP_TRY(hipMemcpy2DAsync(C, ldc * sizeof(double), M->d_io_c, nc * sizeof(double), M->rows * sizeof(double), nrhs,
                             hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, M->ctx->ev0, M->ctx->ev1) == hipSuccess) g_last_ms = ms;
    return HSS_OK;
}

// ---------------------------------------------------------------------------
// ULV factorization
// ---------------------------------------------------------------------------
static int pow2_at_least(int x) {
    int p = 1;
    while (p < x) p <<= 1;
    return p;
}

template <typename TF>
static size_t factor_lds(const FNode& f, const DevNode& d) {
    // Us (m x ku) | G, T (k x k each) | M (k x n) | two padded reflector rows | taus | scalars
    const size_t kk = (size_t)std::max(f.k, f.nk) + 1;
    const size_t vstride = (size_t)f.S * (FACT_CPT + 2) + 2;  // S chunks of CPT columns, padded
    return sizeof(TF) * ((size_t)f.m * d.ku + 2 * (size_t)f.k * f.k + ((size_t)f.k + 2) * vstride + kk + 8);
}

// Fill the factor metadata of one matrix; offsets are relative to this
// matrix's own factor / workspace buffers.  The factor buffer holds the resident
// solve operators first ([0, res_len): every G and Qt, the root's LU data and
// couplings) and the factorization-time data after them.
template <typename TF>
static hss_status plan_factor(hss_mat_s* M, hss_fact_s* F) {
    const int nn = M->nn;
    F->hfn.assign(nn, FNode{});
    auto take = [](int64_t& cur, int64_t count) {
        int64_t off = cur;
        cur += std::max<int64_t>(count, 0);
        return off;
    };
    // sizes, in post-order (reverse breadth-first: children have larger indices)
    for (int i = nn - 1; i >= 0; --i) {
        const DevNode& d = M->hn[i];
        FNode& f = F->hfn[i];
        std::memset(&f, 0, sizeof(f));
        f.Gz = f.Gr = f.Qt = -1;
        f.Yql = f.Tql = f.Ylq = f.taulq = f.Tlq = f.L1 = f.L1inv = f.L2a = f.Vt = f.P12 = f.P21 = f.LU = f.LUinv = f.piv = -1;
        f.rwD = f.rwU = f.rwV = f.rwM = -1;
        f.yrow = f.zrow = -1;
        f.Dr = f.Dinv = f.Mb = -1;
        if (d.is_leaf) {
            f.m = d.m;
            f.n = d.n;
        } else {
            const FNode& a = F->hfn[d.left];
            const FNode& b = F->hfn[d.right];
            f.m = a.mr + b.mr;
            f.n = a.nr + b.nr;
        }
        if (f.m != f.n)
            return fail(HSS_ERR_DIM, "ULV needs square diagonal blocks (node " + std::to_string(i) + ": " +
                                         std::to_string(f.m) + "x" + std::to_string(f.n) + ")");
        if (i == 0) {
            f.mr = f.m;
            f.nr = f.n;
            if (!d.is_leaf) {
                const FNode& a = F->hfn[d.left];
                const FNode& b = F->hfn[d.right];
                f.Kx = (a.My - a.nk) + (b.My - b.nk);
            }
            continue;
        }
        f.k = std::min(d.ku, f.m);
        f.pass = f.k >= f.m ? 1 : 0;
        f.nk = f.pass ? 0 : std::min(f.m - f.k, f.n);
        f.mr = f.pass ? f.m : f.k;
        f.nr = f.n - f.nk;
        f.nblk = (f.nk + NB - 1) / NB;
        f.My = f.nk + f.mr + d.kv;
        if (d.is_leaf) {
            f.Kx = f.m;
        } else {
            const FNode& a = F->hfn[d.left];
            const FNode& b = F->hfn[d.right];
            f.Kx = (a.My - a.nk) + (b.My - b.nk);
        }
        int S = pow2_at_least((f.n + FACT_CPT - 1) / FACT_CPT);
        if (S < 1) S = 1;
        f.S = S;
        if (S > 64 || (int64_t)(f.m + d.kv) * S > FACT_THREADS || f.m > 256 || f.n > 256)
            return fail(HSS_ERR_UNSUPPORTED, "node " + std::to_string(i) + " (" + std::to_string(f.m) + "x" +
                                                 std::to_string(f.n) + ", kv " + std::to_string(d.kv) +
                                                 ") exceeds the register-resident factor kernel");
        if (factor_lds<TF>(f, d) > 160 * 1024)
            return fail(HSS_ERR_UNSUPPORTED, "node " + std::to_string(i) + " exceeds the LDS of the factor kernel");
    }
    // resident solve operators; each node's [Gz Gr] and Qt start on a 16-element boundary
    // (whole-operator vector loads)
    int64_t fac = 0;
    auto align16 = [](int64_t& cur) { cur = (cur + 15) & ~int64_t(15); };
    for (int i = 1; i < nn; ++i) {
        FNode& f = F->hfn[i];
        align16(fac);
        if (f.nk > 0) f.Gz = take(fac, (int64_t)f.nk * f.Kx);
        f.Gr = take(fac, (int64_t)(f.My - f.nk) * (f.Kx + f.nk));
        align16(fac);
        if (!f.pass) f.Qt = take(fac, (int64_t)f.n * f.n);
    }
    align16(fac);
    FNode& fr = F->hfn[0];
    const DevNode& dr = M->hn[0];
    fr.LU = take(fac, (int64_t)fr.m * fr.m);
    fr.LUinv = take(fac, (int64_t)2 * ((fr.m + NB - 1) / NB) * NB * NB);  // inverted L/U diagonal blocks
    if (!dr.is_leaf) {
        fr.P12 = take(fac, (int64_t)F->hfn[dr.left].mr * M->hn[dr.right].kv);
        fr.P21 = take(fac, (int64_t)F->hfn[dr.right].mr * M->hn[dr.left].kv);
    }
    fr.piv = 0;
    fr.rlen = fac - fr.LU;
    if (!dr.is_leaf) {
        align16(fac);
        fr.Dinv = take(fac, (int64_t)fr.m * fr.m);
        fr.Dr = take(fac, (int64_t)fr.m * fr.m);
        fr.Mb = take(fac, (int64_t)fr.m * fr.Kx);
    }
    F->res_len = fac;
    // factorization-time data
    int64_t rw = 0, yrows = 0, zrows = 0;
    for (int i = nn - 1; i >= 1; --i) {
        const DevNode& d = M->hn[i];
        FNode& f = F->hfn[i];
        if (!d.is_leaf) {
            f.P12 = take(fac, (int64_t)F->hfn[d.left].mr * M->hn[d.right].kv);
            f.P21 = take(fac, (int64_t)F->hfn[d.right].mr * M->hn[d.left].kv);
        }
        if (!f.pass) {
            f.Yql = take(fac, (int64_t)f.m * f.k);
            f.Tql = take(fac, (int64_t)f.k * f.k);
            f.Ylq = take(fac, (int64_t)f.n * f.nk);
            f.taulq = take(fac, (int64_t)f.nk);
            f.Tlq = take(fac, (int64_t)f.nblk * NB * NB);
            f.L1 = take(fac, (int64_t)f.nk * f.nk);
            f.L1inv = take(fac, (int64_t)f.nblk * NB * NB);
            f.L2a = take(fac, (int64_t)f.k * f.nk);
            f.Vt = take(fac, (int64_t)d.kv * f.nk);
        }
        if (!d.is_leaf) f.rwM = take(rw, (int64_t)(f.m + d.kv) * f.n);
        f.rwD = take(rw, (int64_t)f.mr * f.nr);
        f.rwU = take(rw, (int64_t)f.mr * d.ku);
        f.rwV = take(rw, (int64_t)f.nr * d.kv);
    }
    F->fac_len = fac;
    // solve workspace: y blocks of non-root nodes, z blocks of branches and the root
    for (int i = 0; i < nn; ++i) {
        FNode& f = F->hfn[i];
        if (i > 0) f.yrow = take(yrows, f.My);
        if (!M->hn[i].is_leaf || i == 0) f.zrow = take(zrows, f.n);
    }
    F->yrows = yrows;
    F->zrows = zrows;
    // LDS (rows of NTP = RHS_TILE columns) and widths of the operator-forming kernels
    size_t gform = 0, qform = 0, groot = 0;
    int max_kx = 0, max_nq = 0;
    for (int i = 0; i < nn; ++i) {
        const FNode& f = F->hfn[i];
        const DevNode& d = M->hn[i];
        const int kvc = d.is_leaf ? 0 : M->hn[d.left].kv + M->hn[d.right].kv;
        if (i == 0) {
            groot = (size_t)f.m + kvc + NB;
        } else {
            const size_t tr = (size_t)std::max(std::max(f.k, NB), d.kv);
            gform = std::max(gform, (size_t)f.m + std::max(f.nk, 1) + kvc + 2 * tr);
            qform = std::max(qform, (size_t)f.n + 2 * NB);
            max_kx = std::max(max_kx, f.Kx + f.nk);
            if (f.Qt >= 0) max_nq = std::max(max_nq, f.n);
        }
    }
    const size_t tv = sizeof(TF) * RHS_TILE;
    F->lds_gform = gform * tv;
    F->lds_qform = qform * tv;
    F->lds_groot = groot * tv;
    F->lds_groot_res = groot * tv + sizeof(TF) * (size_t)((fr.rlen + 15) & ~int64_t(15));
    if (fr.Dinv >= 0) {
        const int64_t io = (int64_t)std::max(fr.Kx, 2 * fr.m) * 33 + (int64_t)fr.m * 33;
        F->lds_groot2 = sizeof(TF) * (size_t)(((int64_t)fr.m * fr.m + 15) / 16 * 16 + io);
        F->lds_groot2s = sizeof(TF) * (size_t)((2 * (int64_t)fr.m * fr.m + (int64_t)fr.m * fr.Kx + 15) / 16 * 16 + io);
    }
    F->max_kx = max_kx;
    F->max_nq = max_nq;
    if (F->lds_gform > 160 * 1024 || F->lds_groot > 160 * 1024)
        return fail(HSS_ERR_UNSUPPORTED, "a node exceeds the LDS of the solve-operator kernels");
    return HSS_OK;
}

static void free_fact(hss_fact_s* F) {
    if (!F) return;
    destroy_plan_cache(F->plan_cache);
    (void)hipFree(F->d_fn);
    (void)hipFree(F->d_fac);
    (void)hipFree(F->d_piv);
    (void)hipFree(F->d_status);
    (void)hipFree(F->d_wy);
    (void)hipFree(F->d_swork);
    (void)hipFree(F->d_r);
    (void)hipFree(F->d_d);
    (void)hipFree(F->d_nrm);
    (void)hipFree(F->d_gm);
    (void)hipFree(F->d_io_b);
    (void)hipFree(F->d_io_x);
    delete F;
}

static hss_status ensure_schedule(hss_fact_s* F);

// Factor a batch of matrices that share one context: one launch per height
// across all of them, then one root launch.
template <typename TF>
static hss_status factor_batch(const std::vector<hss_mat_s*>& mats, hss_prec prec, std::vector<hss_fact_s*>& out) {
    const int cnt = (int)mats.size();
    hss_ctx_s* ctx = mats[0]->ctx;
    hipStream_t st = ctx->stream;
    out.assign(cnt, nullptr);
    std::vector<int64_t> rw_off(cnt), rw_len(cnt);
    int64_t rw_total = 0;
    int maxh = 0;
    auto cleanup = [&]() {
        for (auto* f : out) free_fact(f);
        out.assign(cnt, nullptr);
    };
    for (int b = 0; b < cnt; ++b) {
        hss_mat_s* M = mats[b];
        if (M->ctx != ctx) { cleanup(); return fail(HSS_ERR_ARG, "batched matrices must share one context"); }
        auto* F = new (std::nothrow) hss_fact_s;
        if (!F) { cleanup(); return fail(HSS_ERR_OOM, "host allocation"); }
        out[b] = F;
        F->mat = M;
        F->prec = prec;
        hss_status s = plan_factor<TF>(M, F);
        if (s != HSS_OK) { cleanup(); return s; }
        int64_t rw = 0;
        for (const FNode& f : F->hfn) {
            if (f.rwD >= 0) rw = std::max(rw, f.rwD + (int64_t)f.mr * f.nr);
        }
        for (int i = 1; i < M->nn; ++i) {
            const FNode& f = F->hfn[i];
            if (f.rwM >= 0) rw = std::max(rw, f.rwM + (int64_t)(f.m + M->hn[i].kv) * f.n);
            rw = std::max(rw, f.rwV + (int64_t)f.nr * M->hn[i].kv);
            rw = std::max(rw, f.rwU + (int64_t)f.mr * M->hn[i].ku);
        }
        rw_off[b] = rw_total;
        rw_len[b] = rw;
        rw_total += rw;
        maxh = std::max(maxh, M->max_height);
        hipError_t e;
        if ((e = dalloc(&F->d_fn, F->hfn.size())) != hipSuccess ||
            (e = dalloc(reinterpret_cast<TF**>(&F->d_fac), (size_t)F->fac_len)) != hipSuccess ||
            (e = dalloc(&F->d_piv, (size_t)F->hfn[0].m)) != hipSuccess || (e = dalloc(&F->d_status, 1)) != hipSuccess) {
            cleanup();
            return fail(HSS_ERR_OOM, std::string("factor allocation: ") + hipGetErrorString(e));
        }
        HIP_TRY(hipMemcpyAsync(F->d_fn, F->hfn.data(), sizeof(FNode) * F->hfn.size(), hipMemcpyHostToDevice, st));
        HIP_TRY(hipMemsetAsync(F->d_status, 0, sizeof(int32_t), st));
        F->dev_bytes = (int64_t)(sizeof(TF) * F->fac_len + sizeof(FNode) * F->hfn.size() + 4 * (F->hfn[0].m + 1));
    }
    TF* d_rw = nullptr;
    FInst<TF>* d_inst = nullptr;
    WorkItem* d_work = nullptr;
    size_t lds_gform = 0, lds_qform = 0, lds_rinv = 0;
    int max_kx = 0, max_nq = 0, max_rcols = 0;
    for (auto* F : out) {
        lds_gform = std::max(lds_gform, F->lds_gform);
        lds_qform = std::max(lds_qform, F->lds_qform);
        max_kx = std::max(max_kx, F->max_kx);
        max_nq = std::max(max_nq, F->max_nq);
        const FNode& r = F->hfn[0];
        if (r.Dinv >= 0) {
            max_rcols = std::max(max_rcols, std::max(r.m, r.Kx));
            lds_rinv = std::max(lds_rinv, sizeof(TF) * (size_t)(r.m + NB) * RHS_TILE);
        }
    }
    // work lists across the batch; per level the launch shape of either factor kernel
    std::vector<WorkItem> work;
    std::vector<Range> lev(maxh);
    std::vector<size_t> lev_lds(maxh, 0), lev_lds_mf(maxh, 0);
    std::vector<int> lev_thr(maxh, 1), lev_rows(maxh, 0), lev_n(maxh, 0);
    for (int h = 0; h < maxh; ++h) {
        lev[h].off = (int)work.size();
        for (int b = 0; b < cnt; ++b) {
            hss_mat_s* M = mats[b];
            for (int i = 1; i < M->nn; ++i)
                if (M->hn[i].height == h) {
                    work.push_back(WorkItem{b, i});
                    const FNode& fi = out[b]->hfn[i];
                    lev_lds[h] = std::max(lev_lds[h], factor_lds<TF>(fi, M->hn[i]));
                    lev_thr[h] = std::max(lev_thr[h], (fi.m + M->hn[i].kv) * fi.S);
                    lev_rows[h] = std::max(lev_rows[h], fi.m + M->hn[i].kv);
                    lev_n[h] = std::max(lev_n[h], fi.n);
                    // Vs 2 x 16 x 130 | Ts 2 x 16 x 17 (rows padded) | taus 128 | Gram column 16
                    lev_lds_mf[h] = std::max(lev_lds_mf[h], sizeof(TF) * (size_t)(2 * 16 * 130 + 2 * 16 * 17 + 128 + 16));
                }
        }
        lev[h].count = (int)work.size() - lev[h].off;
        lev_lds_mf[h] = std::max(lev_lds_mf[h], lev_lds[h]);
        // a level of at most one workgroup per CU (256 CUs) gets all 160 KB of LDS: the
        // occupancy does not change, and the merge stages operands beyond the node's formula
        if (lev[h].count <= 256) {
            lev_lds[h] = std::max(lev_lds[h], (size_t)160 * 1024);
            lev_lds_mf[h] = std::max(lev_lds_mf[h], (size_t)160 * 1024);
        }
    }
    // Float64 levels run the blocked MFMA kernel when their nodes have n <= 128 and at most
    // factor_mf_max_rows rows of [D; V'] (others run the per-reflector kernel).  Float32
    // levels run the per-reflector kernel: measured on the n = 65536 log-kernel matrix
    // (leaf 128, ranks 11-35), the blocked kernel is 10 % faster per level in Float64 and
    // 0-20 % slower in Float32, whose per-reflector steps are cheaper.
    auto use_mf = [&](int h) {
        const bool want = g_factor_kernel == FactorKernel::Auto ? sizeof(TF) == 8 : g_factor_kernel == FactorKernel::Mfma;
        return want && lev_n[h] <= 128 && lev_rows[h] <= factor_mf_max_rows((int)sizeof(TF)) &&
               lev_lds_mf[h] <= 160 * 1024;
    };
    const int wy_count = (int)work.size();  // every non-root node of every matrix
    Range roots{(int)work.size(), cnt};
    for (int b = 0; b < cnt; ++b) work.push_back(WorkItem{b, 0});
    std::vector<FInst<TF>> hinst(cnt);
    hipError_t e;
    if ((e = dalloc(&d_rw, (size_t)rw_total)) != hipSuccess || (e = dalloc(&d_inst, (size_t)cnt)) != hipSuccess ||
        (e = dalloc(&d_work, work.size())) != hipSuccess) {
        (void)hipFree(d_rw); (void)hipFree(d_inst); (void)hipFree(d_work);
        cleanup();
        return fail(HSS_ERR_OOM, std::string("factor workspace: ") + hipGetErrorString(e));
    }
    for (int b = 0; b < cnt; ++b) {
        FInst<TF>& I = hinst[b];
        I.nodes = mats[b]->d_nodes;
        I.fn = out[b]->d_fn;
        I.blob = mats[b]->d_blob;
        I.fac = reinterpret_cast<TF*>(out[b]->d_fac);
        I.rw = d_rw + rw_off[b];
        I.ipiv = out[b]->d_piv;
        I.status = out[b]->d_status;
    }
    hss_status rs = HSS_OK;
    auto run = [&]() -> hss_status {
        HIP_TRY(hipMemcpyAsync(d_inst, hinst.data(), sizeof(FInst<TF>) * cnt, hipMemcpyHostToDevice, st));
        HIP_TRY(hipMemcpyAsync(d_work, work.data(), sizeof(WorkItem) * work.size(), hipMemcpyHostToDevice, st));
        HIP_TRY(hipEventRecord(ctx->ev0, st));
        for (int h = 0; h < maxh; ++h) {
            if (use_mf(h)) {
                ProfScope ps("k_ulv_factor_mf", h, lev[h].count, st);
                HIP_TRY((launch_ulv_factor_mf<TF>(d_inst, d_work + lev[h].off, lev[h].count, lev_lds_mf[h],
                                                  lev_rows[h], lev_n[h], st)));
            } else {
                ProfScope ps("k_ulv_factor", h, lev[h].count, st);
                HIP_TRY((launch_ulv_factor<TF>(d_inst, d_work + lev[h].off, lev[h].count, lev_lds[h], lev_thr[h], st)));
            }
        }
        {
            ProfScope ps("k_ulv_wy", 0, wy_count, st);
            HIP_TRY((launch_ulv_wy<TF>(d_inst, d_work, wy_count, st)));
        }
        {
            ProfScope ps("k_ulv_root", 0, roots.count, st);
            HIP_TRY((launch_ulv_root<TF>(d_inst, d_work + roots.off, roots.count, st)));
        }
        {
            ProfScope ps("k_ulv_rinv", 0, roots.count, st);
            HIP_TRY((launch_ulv_rinv<TF>(d_inst, d_work + roots.off, roots.count, max_rcols, lds_rinv, st)));
        }
        // solve operators G and Qt of every non-root node (one launch each)
        {
            ProfScope ps("k_ulv_gform", 0, wy_count, st);
            HIP_TRY((launch_ulv_gform<TF>(d_inst, d_work, wy_count, max_kx, lds_gform, st)));
        }
        {
            ProfScope ps("k_ulv_qform", 0, wy_count, st);
            HIP_TRY((launch_ulv_qform<TF>(d_inst, d_work, wy_count, max_nq, lds_qform, st)));
        }
        HIP_TRY(hipEventRecord(ctx->ev1, st));
        HIP_TRY(hipStreamSynchronize(st));
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1) == hipSuccess) g_last_ms = ms;
        return HSS_OK;
    };
    rs = run();
    (void)hipFree(d_rw);
    (void)hipFree(d_inst);
    (void)hipFree(d_work);
    if (rs != HSS_OK) { cleanup(); return rs; }
    for (int b = 0; b < cnt; ++b) {
        int32_t info = 0;
        HIP_TRY(hipMemcpy(&info, out[b]->d_status, sizeof(int32_t), hipMemcpyDeviceToHost));
        if (info != 0) {
            cleanup();
            return fail(HSS_ERR_SINGULAR, "SingularException(" + std::to_string(info) +
                                              "): exactly zero pivot in the root block of matrix " + std::to_string(b));
        }
    }
    // keep only the resident solve operators (the stream is idle here)
    for (int b = 0; b < cnt; ++b) {
        hss_fact_s* F = out[b];
        TF* res = nullptr;
        hipError_t e2 = dalloc(&res, (size_t)F->res_len + 16);
        if (e2 != hipSuccess) {
            cleanup();
            return fail(HSS_ERR_OOM, std::string("factor allocation: ") + hipGetErrorString(e2));
        }
        HIP_TRY(hipMemcpy(res, F->d_fac, sizeof(TF) * (size_t)F->res_len, hipMemcpyDeviceToDevice));
        (void)hipFree(F->d_fac);
        F->d_fac = res;
        F->fac_len = F->res_len;
        F->dev_bytes = (int64_t)(sizeof(TF) * F->fac_len + sizeof(FNode) * F->hfn.size() + 4 * (F->hfn[0].m + 1));
    }
    // the single-request solve schedule, built here so that solves (the allocation-free one
    // included) find it ready
    for (int b = 0; b < cnt; ++b) {
        hss_status s2 = ensure_schedule(out[b]);
        if (s2 != HSS_OK) { cleanup(); return s2; }
    }
    for (auto* F : out) F->mat->refs++;
    return HSS_OK;
}

extern "C" hss_status hss_ulv_factor_batch(int32_t count, const hss_mat_t* mats, hss_prec prec, hss_fact_t* out) {
    if (count <= 0 || !mats || !out) return fail(HSS_ERR_ARG, "bad batch arguments");
    std::vector<hss_mat_s*> v(mats, mats + count);
    for (auto* m : v)
        if (!m) return fail(HSS_ERR_ARG, "null matrix in batch");
    HIP_TRY(hipSetDevice(v[0]->ctx->device));
    std::vector<hss_fact_s*> f;
    hss_status s = (prec == HSS_PREC_FP32) ? factor_batch<float>(v, prec, f) : factor_batch<double>(v, prec, f);
    if (s != HSS_OK) return s;
    for (int i = 0; i < count; ++i) out[i] = f[i];
    return HSS_OK;
}

extern "C" hss_status hss_ulv_factor(hss_mat_t M, hss_prec prec, hss_fact_t* out) {
    if (!M || !out) return fail(HSS_ERR_ARG, "null argument");
    if (prec != HSS_PREC_FP64 && prec != HSS_PREC_FP32) return fail(HSS_ERR_ARG, "unknown precision");
    return hss_ulv_factor_batch(1, &M, prec, out);
}

static void fact_release(hss_fact_s* F) {
    if (!F || --F->refs > 0) return;
    hss_mat_s* M = F->mat;
    (void)hipSetDevice(M->ctx->device);
    (void)hipStreamSynchronize(M->ctx->stream);
    free_fact(F);
    mat_release(M);
}
extern "C" hss_status hss_fact_destroy(hss_fact_t F) {
    fact_release(F);
    return HSS_OK;
}

extern "C" int64_t hss_fact_device_bytes(hss_fact_t F) { return F ? F->dev_bytes : 0; }

// ---------------------------------------------------------------------------
// ULV solve
// ---------------------------------------------------------------------------
template <typename TF>
static hss_status ensure_solve_ws(hss_fact_s* F, int nrhs) {
    if ((size_t)nrhs <= F->ws_cap) return HSS_OK;
    (void)hipFree(F->d_wy);
    F->d_wy = nullptr;
    F->ws_cap = 0;
    HIP_TRY(dalloc(reinterpret_cast<TF**>(&F->d_wy), (size_t)std::max<int64_t>(F->yrows + F->zrows, 1) * nrhs));
    F->ws_cap = nrhs;
    return HSS_OK;
}

struct SolveReq {
    hss_fact_s* F;
    int nrhs;
    const double* dB;
    int64_t ldb;
    double* dX;
    int64_t ldx;
};

// Workspace elements of one solve of `nrhs` columns (y blocks, then z blocks).
static size_t solve_ws_elems(const hss_fact_s* F, int nrhs) { return (size_t)(F->yrows + F->zrows) * nrhs; }

template <typename TF>
static SInst<TF> make_sinst(const SolveReq& q, TF* ws = nullptr) {
    hss_fact_s* F = q.F;
    hss_mat_s* M = F->mat;
    SInst<TF> I;
    std::memset(&I, 0, sizeof(I));
    I.nodes = M->d_nodes;
    I.fn = F->d_fn;
    I.fac = reinterpret_cast<const TF*>(F->d_fac);
    I.ipiv = F->d_piv;
    I.B = q.dB;
    I.X = q.dX;
    I.ldb = q.ldb;
    I.ldx = q.ldx;
    if (ws) {
        I.wy = ws;
        I.wz = ws + F->yrows * q.nrhs;
    } else {
        I.wy = reinterpret_cast<TF*>(F->d_wy);
        I.wz = I.wy + F->yrows * q.nrhs;
    }
    I.nrhs = q.nrhs;
    return I;
}

// Appends the up-sweep record of `node` (instance b, solve rows of fn) to level h of S.
static void add_up_rec(Schedule& S, size_t h, const hss_mat_s* M, const std::vector<FNode>& fn, int b, int node) {
    auto r16 = [](int64_t x) { return (x + 15) & ~int64_t(15); };
    S.work.push_back(WorkItem{b, node});
    const FNode& f = fn[node];
    const DevNode& nd = M->hn[node];
    SweepRec q{};
    q.inst = b;
    q.node = node;
    q.leaf = nd.is_leaf;
    q.nk = f.nk;
    q.K = f.Kx;
    q.My = f.My;
    q.A = f.nk > 0 ? f.Gz : f.Gr;  // Gr follows Gz
    q.y = f.yrow;
    if (nd.is_leaf) {
        q.x0 = nd.row_off;
    } else {
        const FNode &fl = fn[nd.left], &fr = fn[nd.right];
        q.x0 = fl.yrow;
        q.add0 = fl.nk;
        q.l0 = fl.My;
        q.r0 = fl.My - fl.nk;
        q.x1 = fr.yrow;
        q.add1 = fr.nk;
        q.l1 = fr.My;
    }
    S.rec.push_back(q);
    const int64_t na = (int64_t)f.nk * f.Kx + (int64_t)(f.My - f.nk) * (f.Kx + f.nk);
    S.up_m[h] = std::max(S.up_m[h], f.nk);
    S.up_na[h] = std::max(S.up_na[h], r16(na));
    S.up_k[h] = std::max(S.up_k[h], (int64_t)(f.Kx + f.nk));
    S.up_rows[h] = std::max(S.up_rows[h], std::max(f.nk, f.My - f.nk));
}

// Appends the down-sweep record of `node` (instance b, solve rows of fn) to level d of S.
static void add_down_rec(Schedule& S, size_t d, const hss_mat_s* M, const std::vector<FNode>& fn, int b, int node) {
    auto r16 = [](int64_t x) { return (x + 15) & ~int64_t(15); };
    S.work.push_back(WorkItem{b, node});
    const FNode& f = fn[node];
    const DevNode& nd = M->hn[node];
    const FNode& fp = fn[nd.parent];
    SweepRec q{};
    q.inst = b;
    q.node = node;
    q.leaf = nd.is_leaf;
    q.nk = f.nk;
    q.K = f.n;
    q.My = f.My;
    q.A = f.Qt;
    q.x0 = f.yrow;
    q.x1 = fp.zrow;
    q.add1 = nd.is_left ? 0 : fn[M->hn[nd.parent].left].nr;
    q.l1 = fp.n;
    q.y = nd.is_leaf ? nd.col_off : f.zrow;
    S.rec.push_back(q);
    const int64_t na = f.Qt >= 0 ? (int64_t)f.n * f.n : 0;
    S.down_m[d] = std::max(S.down_m[d], f.n);
    S.down_na[d] = std::max(S.down_na[d], r16(na));
    S.down_k[d] = std::max(S.down_k[d], (int64_t)f.n);
}

static void schedule_levels(Schedule& S, int nup, int ndown) {
    S.work.clear();
    S.rec.clear();
    S.up.assign(nup, Range{});
    S.up_m.assign(nup, 0);
    S.up_na.assign(nup, 0);
    S.up_k.assign(nup, 0);
    S.up_rows.assign(nup, 0);
    S.down.assign(ndown, Range{});
    S.down_m.assign(ndown, 0);
    S.down_na.assign(ndown, 0);
    S.down_k.assign(ndown, 0);
}

static void build_schedule(const std::vector<hss_fact_s*>& facts, Schedule& S) {
    const int cnt = (int)facts.size();
    int maxh = 0, maxd = 0;
    for (auto* F : facts) {
        maxh = std::max(maxh, F->mat->max_height);
        maxd = std::max(maxd, F->mat->max_depth);
    }
    schedule_levels(S, maxh, maxd + 1);
    S.w = (!facts.empty() && facts[0]->prec == HSS_PREC_FP32) ? 4 : 8;
    for (int h = 0; h < maxh; ++h) {
        S.up[h].off = (int)S.work.size();
        for (int b = 0; b < cnt; ++b) {
            hss_mat_s* M = facts[b]->mat;
            if (h >= M->max_height) continue;
            const Range& r = M->by_height[h];
            for (int t = 0; t < r.count; ++t) add_up_rec(S, h, M, facts[b]->hfn, b, M->hwork[r.off + t].node);
        }
        S.up[h].count = (int)S.work.size() - S.up[h].off;
    }
    S.roots = Range{(int)S.work.size(), cnt};
    for (int b = 0; b < cnt; ++b) {
        S.work.push_back(WorkItem{b, 0});
        SweepRec q{};
        q.inst = b;
        S.rec.push_back(q);
    }
    for (int d = 1; d <= maxd; ++d) {
        S.down[d].off = (int)S.work.size();
        for (int b = 0; b < cnt; ++b) {
            hss_mat_s* M = facts[b]->mat;
            if (d > M->max_depth) continue;
            const Range& r = M->by_depth[d];
            for (int t = 0; t < r.count; ++t) add_down_rec(S, d, M, facts[b]->hfn, b, M->hwork[r.off + t].node);
        }
        S.down[d].count = (int)S.work.size() - S.down[d].off;
    }
}

// lgroot_res = LDS of the root kernel with the root factor staged (used when it fits)
// lgroot2 = LDS of the inverse-and-refine root kernel (0 when some root lacks it),
// lgroot2s = the same with every root operator staged
// LDS bytes of a sweep step in form 0 (operator na elements and k input rows staged) or
// form 1 (input only) with tw columns; false when they do not fit
static bool level_lds(size_t w, int form, int64_t na, int64_t k, int tw, size_t& lds) {
    lds = w * (size_t)((form == 0 ? na : 0) + k * (tw + 1));
    return lds <= kLdsResidentMax;
}
// The built-in choice for a level
static LevelCfg default_cfg(size_t w, int64_t na, int64_t k, int rows, int count, int maxnrhs) {
    int tw = 32;
    size_t lds = 0;
    if (pick_resident(w, na, k, count, maxnrhs, tw, lds)) return LevelCfg{0, tw, 0};
    if (pick_streamed(w, k, rows, count, maxnrhs, tw, lds)) return LevelCfg{1, tw, 0};
    return LevelCfg{2, 0, 0};
}

// One up-sweep level (height h) with choice c (c.form < 0: the built-in rules)
template <typename TF>
static hss_status launch_up_level(const SInst<TF>* di, const SweepRec* dw, const Schedule& S, size_t h, int maxnrhs,
                                  LevelCfg c, hipStream_t st) {
    const int cnt = S.up[h].count;
    if (cnt == 0) return HSS_OK;
    if (c.form < 0) c = default_cfg(S.w, S.up_na[h], S.up_k[h], S.up_rows[h], cnt, maxnrhs);
    if (c.form == 2) {
        ProfScope ps("k_ulv_gup", (int)h, cnt, st);
        HIP_TRY((launch_ulv_gup<TF>(di, dw + S.up[h].off, cnt, S.up_m[h], maxnrhs, st)));
        return HSS_OK;
    }
    size_t lds = 0;
    if (!level_lds(S.w, c.form, S.up_na[h], S.up_k[h], c.tw, lds))
        return fail(HSS_ERR_ARG, "level choice does not fit in LDS");
    ProfScope ps(c.form == 0 ? "k_ulv_gup_s" : "k_ulv_gup_x", (int)h, cnt, st);
    HIP_TRY((launch_ulv_gup_s<TF>(di, dw + S.up[h].off, cnt, maxnrhs, c.tw, lds, c.form == 0 ? 1 : 0, c.nt, st)));
    return HSS_OK;
}
// One down-sweep level (depth d) with choice c
template <typename TF>
static hss_status launch_down_level(const SInst<TF>* di, const SweepRec* dw, const Schedule& S, size_t d, int maxnrhs,
                                    LevelCfg c, hipStream_t st) {
    const int cnt = S.down[d].count;
    if (cnt == 0) return HSS_OK;
    if (c.form < 0) c = default_cfg(S.w, S.down_na[d], S.down_k[d], S.down_m[d], cnt, maxnrhs);
    if (c.form == 2) {
        ProfScope ps("k_ulv_gdown", (int)d, cnt, st);
        HIP_TRY((launch_ulv_gdown<TF>(di, dw + S.down[d].off, cnt, S.down_m[d], maxnrhs, st)));
        return HSS_OK;
    }
    size_t lds = 0;
    if (!level_lds(S.w, c.form, S.down_na[d], S.down_k[d], c.tw, lds))
        return fail(HSS_ERR_ARG, "level choice does not fit in LDS");
    ProfScope ps(c.form == 0 ? "k_ulv_gdown_s" : "k_ulv_gdown_x", (int)d, cnt, st);
    HIP_TRY((launch_ulv_gdown_s<TF>(di, dw + S.down[d].off, cnt, maxnrhs, c.tw, lds, c.form == 0 ? 1 : 0, c.nt,
                                    st)));
    return HSS_OK;
}

template <typename TF>
static hss_status launch_schedule(const SInst<TF>* di, const SweepRec* dw, const Schedule& S, int maxnrhs,
                                  size_t lgroot, size_t lgroot_res, size_t lgroot2, size_t lgroot2s, hipStream_t st) {
    for (size_t h = 0; h < S.up.size(); ++h) {
        hss_status s = launch_up_level<TF>(di, dw, S, h, maxnrhs, h < S.up_cfg.size() ? S.up_cfg[h] : LevelCfg{}, st);
        if (s != HSS_OK) return s;
    }
    if (S.roots.count > 0) {
        if (lgroot2 > 0 && lgroot2s > 0 && lgroot2s <= kLdsResidentMax) {
            ProfScope ps("k_ulv_groot2", 0, S.roots.count, st);
            HIP_TRY((launch_ulv_groot2<TF>(di, dw + S.roots.off, S.roots.count, maxnrhs, lgroot2s, 1, st)));
        } else if (lgroot2 > 0 && lgroot2 <= kLdsResidentMax) {
            ProfScope ps("k_ulv_groot2", 0, S.roots.count, st);
            HIP_TRY((launch_ulv_groot2<TF>(di, dw + S.roots.off, S.roots.count, maxnrhs, lgroot2, 0, st)));
        } else {
            ProfScope ps("k_ulv_groot", 0, S.roots.count, st);
            const bool res = lgroot_res <= kLdsResidentMax;
            HIP_TRY((launch_ulv_groot<TF>(di, dw + S.roots.off, S.roots.count, maxnrhs, res ? lgroot_res : lgroot,
                                          res ? 1 : 0, st)));
        }
    }
    for (size_t d = 1; d < S.down.size(); ++d) {
        hss_status s =
            launch_down_level<TF>(di, dw, S, d, maxnrhs, d < S.down_cfg.size() ? S.down_cfg[d] : LevelCfg{}, st);
        if (s != HSS_OK) return s;
    }
    return HSS_OK;
}

static hss_status ensure_schedule(hss_fact_s* F) {
    if (F->d_swork) return HSS_OK;
    build_schedule(std::vector<hss_fact_s*>{F}, F->sched);
    HIP_TRY(dalloc(&F->d_swork, F->sched.rec.size()));
    HIP_TRY(hipMemcpy(F->d_swork, F->sched.rec.data(), sizeof(SweepRec) * F->sched.rec.size(),
                      hipMemcpyHostToDevice));
    return HSS_OK;
}

template <typename TF>
static hss_status solve_one(hss_fact_s* F, int nrhs, const double* dB, int64_t ldb, double* dX, int64_t ldx,
                            hipStream_t st) {
    hss_status s = ensure_solve_ws<TF>(F, nrhs);
    if (s != HSS_OK) return s;
    s = ensure_schedule(F);
    if (s != HSS_OK) return s;
    SInst<TF> I = make_sinst<TF>(SolveReq{F, nrhs, dB, ldb, dX, ldx});
    const void* dv = nullptr;
    HIP_TRY(F->sinst.get(&I, sizeof(I), st, &dv));
    return launch_schedule<TF>(reinterpret_cast<const SInst<TF>*>(dv), F->d_swork, F->sched, nrhs, F->lds_groot,
                               F->lds_groot_res, F->lds_groot2, F->lds_groot2s, st);
}

static hss_status ensure_refine(hss_fact_s* F, int nrhs) {
    const size_t need = (size_t)std::max<int64_t>(F->mat->rows, 1) * nrhs;
    if (need <= F->rf_cap) return HSS_OK;
    (void)hipFree(F->d_r);
    (void)hipFree(F->d_d);
    (void)hipFree(F->d_nrm);
    F->d_r = F->d_d = F->d_nrm = nullptr;
    F->rf_cap = 0;
    HIP_TRY(dalloc(&F->d_r, need));
    HIP_TRY(dalloc(&F->d_d, need));
    HIP_TRY(dalloc(&F->d_nrm, (size_t)2 * nrhs));
    F->rf_cap = need;
    return HSS_OK;
}

// max_j ||r_j|| / ||b_j||
static hss_status rel_residual(hss_fact_s* F, int nrhs, const double* dB, int64_t ldb, const double* dR, int64_t ldr,
                               hipStream_t st, double* out) {
    const int64_t n = F->mat->rows;
    HIP_TRY(launch_colnorm2(dB, ldb, n, nrhs, F->d_nrm, st));
    HIP_TRY(launch_colnorm2(dR, ldr, n, nrhs, F->d_nrm + nrhs, st));
    std::vector<double> h(2 * (size_t)nrhs);
    HIP_TRY(hipMemcpyAsync(h.data(), F->d_nrm, sizeof(double) * 2 * nrhs, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    double worst = 0.0;
    for (int j = 0; j < nrhs; ++j) {
        const double nb = std::sqrt(h[j]), nr = std::sqrt(h[nrhs + j]);
        const double rel = nb > 0 ? nr / nb : nr;
        worst = std::max(worst, rel);
    }
    *out = worst;
    return HSS_OK;
}

// GMRES-IR correction (Carson and Higham, SIAM J. Sci. Comput. 40(2), 2018): d = A^{-1} r
// by GMRES in Float64, right-preconditioned by the resident factor M (Float32 or Float64):
// A M^{-1} u = r, d = M^{-1} u.  Each column runs its own Krylov space, all in lockstep.
// Step k: z = M^{-1} v_k, w = A z (Float64 HSS matvec), two passes of classical Gram-Schmidt
// against v_0 .. v_k (the second restores the orthogonality the first loses in rounding),
// v_{k+1} = w / ||w||.  The host keeps each column's Hessenberg matrix reduced by Givens
// rotations and retires a column once its GMRES residual is at most kGmresTol ||r_c|| (or
// the space stops growing); at most kGmresDim steps.  Then y = R^{-1} g per column and
// d = M^{-1} (V y).  Input r = F->d_r, output d = F->d_d (n x nrhs, ld n).
constexpr int kGmresDim = 20;
constexpr double kGmresTol = 1e-6;

static hss_status gmres_correction(hss_fact_s* F, int nrhs, hipStream_t st) {
    hss_mat_s* M = F->mat;
    const int64_t n = M->rows, blk = n * nrhs;
    const int m = kGmresDim, hld = m + 1;
    const bool fp32 = F->prec == HSS_PREC_FP32;
    // V (m + 1 blocks) | W | h1 | h2 | s (squared norms) | y
    const size_t need = (size_t)(m + 2) * blk + (size_t)3 * hld * nrhs + nrhs;
    if (need > F->gm_cap) {
        (void)hipFree(F->d_gm);
        F->d_gm = nullptr;
        F->gm_cap = 0;
        HIP_TRY(dalloc(&F->d_gm, need));
        F->gm_cap = need;
    }
    double* V = F->d_gm;
    double* W = V + (size_t)(m + 1) * blk;
    double* h1 = W + blk;
    double* h2 = h1 + (size_t)hld * nrhs;
    double* sq = h2 + (size_t)hld * nrhs;
    double* dy = sq + nrhs;
    auto precond = [&](const double* b, double* x) -> hss_status {
        return fp32 ? solve_one<float>(F, nrhs, b, n, x, n, st) : solve_one<double>(F, nrhs, b, n, x, n, st);
    };
    // v_0 = r / ||r|| per column
    HIP_TRY(launch_colnorm2(F->d_r, n, n, nrhs, sq, st));
    HIP_TRY(launch_col_scale(V, F->d_r, n, nrhs, sq, st));
    std::vector<double> hs((size_t)2 * hld * nrhs + nrhs);  // h1 | h2 | s, as on the device
    HIP_TRY(hipMemcpyAsync(hs.data(), sq, sizeof(double) * nrhs, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    // per column c: R column k at R[(c m + k) hld], rotations cs / sn, right-hand side g
    std::vector<double> beta(nrhs), R((size_t)hld * m * nrhs, 0.0), cs((size_t)m * nrhs), sn((size_t)m * nrhs),
        g((size_t)hld * nrhs, 0.0);
    std::vector<int> kc(nrhs, 0);
    std::vector<char> active(nrhs, 0);
    int nact = 0;
    for (int c = 0; c < nrhs; ++c) {
        beta[c] = std::sqrt(hs[c]);
        g[(size_t)c * hld] = beta[c];
        active[c] = beta[c] > 0.0;
        nact += active[c];
    }
    for (int k = 0; k < m && nact > 0; ++k) {
        hss_status s = precond(V + (size_t)k * blk, F->d_d);  // z = M^{-1} v_k
        if (s != HSS_OK) return s;
        s = matmul_enqueue(M, nrhs, 1.0, F->d_d, n, 0.0, W, n);  // w = A z
        if (s != HSS_OK) return s;
        HIP_TRY(launch_gs_dots(V, blk, k + 1, W, n, nrhs, h1, hld, st));
        HIP_TRY(launch_gs_update(W, V, blk, k + 1, n, nrhs, h1, hld, -1.0, 1, st));
        HIP_TRY(launch_gs_dots(V, blk, k + 1, W, n, nrhs, h2, hld, st));
        HIP_TRY(launch_gs_update(W, V, blk, k + 1, n, nrhs, h2, hld, -1.0, 1, st));
        HIP_TRY(launch_colnorm2(W, n, n, nrhs, sq, st));
        HIP_TRY(launch_col_scale(V + (size_t)(k + 1) * blk, W, n, nrhs, sq, st));
        HIP_TRY(hipMemcpyAsync(hs.data(), h1, sizeof(double) * hs.size(), hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        for (int c = 0; c < nrhs; ++c) {
            if (!active[c]) continue;
            double* r = &R[((size_t)c * m + k) * hld];
            for (int i = 0; i <= k; ++i) r[i] = hs[i + (size_t)c * hld] + hs[(size_t)hld * nrhs + i + (size_t)c * hld];
            const double hk1 = std::sqrt(hs[(size_t)2 * hld * nrhs + c]);
            double* csc = &cs[(size_t)c * m];
            double* snc = &sn[(size_t)c * m];
            for (int i = 0; i < k; ++i) {
                const double t = csc[i] * r[i] + snc[i] * r[i + 1];
                r[i + 1] = -snc[i] * r[i] + csc[i] * r[i + 1];
                r[i] = t;
            }
            const double den = std::hypot(r[k], hk1);
            if (!(den > 0.0)) {  // no new direction: the column keeps its first k steps
                active[c] = 0;
                --nact;
                continue;
            }
            double* gc = &g[(size_t)c * hld];
            csc[k] = r[k] / den;
            snc[k] = hk1 / den;
            r[k] = den;
            gc[k + 1] = -snc[k] * gc[k];
            gc[k] = csc[k] * gc[k];
            kc[c] = k + 1;
            if (std::fabs(gc[k + 1]) <= kGmresTol * beta[c] || hk1 == 0.0) {
                active[c] = 0;
                --nact;
            }
        }
    }
    // y = R^{-1} g over each column's kc steps (zero beyond), u = V y, d = M^{-1} u
    std::vector<double> y((size_t)hld * nrhs, 0.0);
    int kmax = 0;
    for (int c = 0; c < nrhs; ++c) {
        const int q = kc[c];
        kmax = std::max(kmax, q);
        double* yc = &y[(size_t)c * hld];
        const double* gc = &g[(size_t)c * hld];
        for (int i = q - 1; i >= 0; --i) {
            double t = gc[i];
            for (int l = i + 1; l < q; ++l) t -= R[((size_t)c * m + l) * hld + i] * yc[l];
            yc[i] = t / R[((size_t)c * m + i) * hld + i];
        }
    }
    HIP_TRY(hipMemcpyAsync(dy, y.data(), sizeof(double) * y.size(), hipMemcpyHostToDevice, st));
    HIP_TRY(launch_gs_update(W, V, blk, kmax, n, nrhs, dy, hld, 1.0, 0, st));
    hss_status s = precond(W, F->d_d);
    if (s != HSS_OK) return s;
    HIP_TRY(hipStreamSynchronize(st));  // y is read before it goes out of scope
    return HSS_OK;
}

static hss_status solve_dev_impl(hss_fact_s* F, int nrhs, const double* dB, int64_t ldb, double* dX, int64_t ldx,
                                 const hss_refine_opts* opts, hss_solve_report* rep) {
    hss_mat_s* M = F->mat;
    hipStream_t st = M->ctx->stream;
    if (nrhs < 0) return fail(HSS_ERR_ARG, "nrhs < 0");
    if (rep) { rep->iterations = 0; rep->rel_residual = -1.0; rep->seconds = 0.0; }
    if (nrhs == 0 || M->rows == 0) return HSS_OK;
    if (ldb < M->rows || ldx < M->rows) return fail(HSS_ERR_DIM, "leading dimension smaller than the matrix size");
    const bool fp32 = F->prec == HSS_PREC_FP32;
    hss_refine_opts o;
    if (opts) o = *opts;
    else { o.tol = fp32 ? 1e-10 : 0.0; o.max_iter = fp32 ? 10 : 0; o.mode = fp32 ? 1 : 0; }
    if (o.mode < 0 || o.mode > 2)
        return fail(HSS_ERR_ARG, "opts->mode must be 0 (none), 1 (iterative refinement) or 2 (GMRES-IR)");
    const bool refine = o.mode != 0 && o.max_iter >= 0;
    auto solve = [&](const double* b, int64_t lb, double* x, int64_t lx) -> hss_status {
        return fp32 ? solve_one<float>(F, nrhs, b, lb, x, lx, st) : solve_one<double>(F, nrhs, b, lb, x, lx, st);
    };
    if (!refine && !g_prof && !fp32) {
        // repeated unrefined solves on the same buffers replay a captured graph
        const PlanKey key{nrhs, dB, ldb, dX, ldx};
        hss_plan_t P = nullptr;
        for (auto& e : F->plan_cache)
            if (e.first == key) P = e.second;
        if (!P) {
            if (F->plan_cache.size() >= 4) {
                std::vector<std::pair<PlanKey, hss_plan_t>> old(F->plan_cache.begin(), F->plan_cache.begin() + 1);
                destroy_plan_cache(old);
                F->plan_cache.erase(F->plan_cache.begin());
            }
            hss_fact_t f1 = F;
            int32_t nr = nrhs;
            hss_status cs = plan_create_impl(1, &f1, &nr, &dB, &ldb, &dX, &ldx, 0, false, &P);
            if (cs != HSS_OK) return cs;
            F->plan_cache.push_back({key, P});
        }
        return hss_plan_execute(P, nullptr, nullptr);
    }
    HIP_TRY(hipEventRecord(M->ctx->ev0, st));
    hss_status s = solve(dB, ldb, dX, ldx);
    if (s != HSS_OK) return s;
    if (refine) {
        s = ensure_refine(F, nrhs);
        if (s != HSS_OK) return s;
        const int64_t n = M->rows;
        double rel = 0.0;
        int it = 0;
        for (;;) {
            // r = b - A x   (Float64 HSS matvec)
            s = matmul_enqueue(M, nrhs, 1.0, dX, ldx, 0.0, F->d_r, n);
            if (s != HSS_OK) return s;
            HIP_TRY(launch_residual(dB, ldb, F->d_r, n, n, nrhs, st));
            s = rel_residual(F, nrhs, dB, ldb, F->d_r, n, st, &rel);
            if (s != HSS_OK) return s;
            if (rel <= o.tol || it >= o.max_iter) break;
            // d = A \ r with the resident factor (mode 1) or by GMRES preconditioned with it
            // (mode 2), x += d
            s = o.mode == 2 ? gmres_correction(F, nrhs, st) : solve(F->d_r, n, F->d_d, n);
            if (s != HSS_OK) return s;
            HIP_TRY(launch_axpy(dX, ldx, F->d_d, n, n, nrhs, st));
            ++it;
        }
        if (rep) { rep->iterations = it; rep->rel_residual = rel; }
        HIP_TRY(hipEventRecord(M->ctx->ev1, st));
        HIP_TRY(hipStreamSynchronize(st));
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, M->ctx->ev0, M->ctx->ev1) == hipSuccess) {
            g_last_ms = ms;
            if (rep) rep->seconds = ms * 1e-3;
        }
        if (rel > o.tol) {
            char buf[160];
            std::snprintf(buf, sizeof buf, "refinement stopped at relative residual %.3e > tol %.3e after %d steps",
                          rel, o.tol, it);
            return fail(HSS_ERR_NOCONV, buf);
        }
        return HSS_OK;
    }
    HIP_TRY(hipEventRecord(M->ctx->ev1, st));
    return HSS_OK;
}

extern "C" hss_status hss_ulv_solve_dev(hss_fact_t F, int32_t nrhs, const double* dB, int64_t ldb, double* dX,
                                        int64_t ldx, const hss_refine_opts* opts, hss_solve_report* rep) {
    if (!F) return fail(HSS_ERR_ARG, "null factor");
    HIP_TRY(hipSetDevice(F->mat->ctx->device));
    return solve_dev_impl(F, nrhs, dB, ldb, dX, ldx, opts, rep);
}

// X = A \ B for a Float64 factor without refinement, allocating nothing: the y/z blocks of
// the solve and its instance record live in the fixed device array (scratch, 64 MiB per
// GPU).  The columns run in chunks that fit it; a matrix whose single column needs more
// workspace than the array holds is refused.  Calls on one GPU share the array: calls that
// overlap in time (other streams, other threads) race.  Nothing in the array has to start
// at zero: every sweep writes its outputs before a later sweep reads them.
extern "C" hss_status hss_ulv_solve_dev_noalloc(hss_fact_t F, int32_t nrhs, const double* dB, int64_t ldb,
                                                double* dX, int64_t ldx) {
    if (!F) return fail(HSS_ERR_ARG, "null factor");
    if (nrhs < 0) return fail(HSS_ERR_ARG, "nrhs < 0");
    hss_mat_s* M = F->mat;
    if (F->prec != HSS_PREC_FP64)
        return fail(HSS_ERR_UNSUPPORTED, "the allocation-free solve takes Float64 factors (a Float32 factor needs "
                                         "refinement buffers)");
    if (nrhs == 0 || M->rows == 0) return HSS_OK;
    if (!dB || !dX) return fail(HSS_ERR_ARG, "null argument");
    if (ldb < M->rows || ldx < M->rows) return fail(HSS_ERR_DIM, "leading dimension smaller than the matrix size");
    const int dev = M->ctx->device;
    HIP_TRY(hipSetDevice(dev));
    void* base = (dev >= 0 && dev < kMaxDevices) ? g_scratch_ptr[dev] : nullptr;
    if (!base) return fail(HSS_ERR_ARG, "device workspace not resolved (context not created on this device)");
    if (!F->d_swork) return fail(HSS_ERR_ARG, "factor has no solve schedule");
    static_assert(sizeof(SInst<double>) <= kScratchRec, "solve record larger than its slot");
    const size_t col_bytes = sizeof(double) * (size_t)(F->yrows + F->zrows);
    const size_t avail = kScratchBytes - kScratchRec;
    if (col_bytes > avail) {
        char buf[200];
        std::snprintf(buf, sizeof buf,
                      "one right-hand side needs %zu bytes of solve workspace, more than the %zu bytes of the "
                      "allocation-free solve's device array",
                      col_bytes, avail);
        return fail(HSS_ERR_UNSUPPORTED, buf);
    }
    const int chunk = (int)std::min<size_t>((size_t)nrhs, avail / std::max<size_t>(col_bytes, 1));
    hipStream_t st = M->ctx->stream;
    auto* rec = reinterpret_cast<SInst<double>*>(base);
    double* ws = reinterpret_cast<double*>(reinterpret_cast<char*>(base) + kScratchRec);
    for (int j0 = 0; j0 < nrhs; j0 += chunk) {
        const int c = std::min(chunk, nrhs - j0);
        // stream order: this record and workspace are overwritten only after the previous
        // chunk's sweeps, enqueued before, have run
        SInst<double> I = make_sinst<double>(SolveReq{F, c, dB + (int64_t)j0 * ldb, ldb, dX + (int64_t)j0 * ldx, ldx},
                                             ws);
        HIP_TRY(launch_put_sinst(rec, I, st));
        hss_status s = launch_schedule<double>(rec, F->d_swork, F->sched, c, F->lds_groot, F->lds_groot_res,
                                               F->lds_groot2, F->lds_groot2s, st);
        if (s != HSS_OK) return s;
    }
    return HSS_OK;
}

extern "C" hss_status hss_ulv_solve(hss_fact_t F, int32_t nrhs, const double* B, int64_t ldb, double* X, int64_t ldx,
                                    const hss_refine_opts* opts, hss_solve_report* rep) {
    if (!F || (nrhs > 0 && (!B || !X))) return fail(HSS_ERR_ARG, "null argument");
    if (nrhs < 0) return fail(HSS_ERR_ARG, "nrhs < 0");
    if (nrhs == 0) return HSS_OK;
    hss_mat_s* M = F->mat;
    if (ldb < M->rows || ldx < M->rows) return fail(HSS_ERR_DIM, "leading dimension smaller than the matrix size");
    HIP_TRY(hipSetDevice(M->ctx->device));
    const int64_t n = std::max<int64_t>(M->rows, 1);
    hss_status s = ensure_io(&F->d_io_b, &F->d_io_x, &F->io_cap, (size_t)n * nrhs);
    if (s != HSS_OK) return s;
    hipStream_t st = M->ctx->stream;
    HIP_TRY(hipMemcpy2DAsync(F->d_io_b, n * sizeof(double), B, ldb * sizeof(double), M->rows * sizeof(double), nrhs,
                             hipMemcpyHostToDevice, st));
    s = solve_dev_impl(F, nrhs, F->d_io_b, n, F->d_io_x, n, opts, rep);
    if (s != HSS_OK && s != HSS_ERR_NOCONV) return s;
    HIP_TRY(hipMemcpy2DAsync(X, ldx * sizeof(double), F->d_io_x, n * sizeof(double), M->rows * sizeof(double), nrhs,
                             hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, M->ctx->ev0, M->ctx->ev1) == hipSuccess) {
        g_last_ms = ms;
        if (rep) rep->seconds = ms * 1e-3;
    }
    return s;
}

extern "C" hss_status hss_ulvfactsolve(hss_mat_t M, int32_t nrhs, const double* B, int64_t ldb, double* X,
                                       int64_t ldx) {
    if (!M) return fail(HSS_ERR_ARG, "null matrix");
    if (M->rows != M->cols) return fail(HSS_ERR_DIM, "ulvfactsolve needs a square matrix");
    hss_fact_t F = nullptr;
    hss_status s = hss_ulv_factor(M, HSS_PREC_FP64, &F);
    if (s != HSS_OK) return s;
    hss_refine_opts none{0.0, 0, 0};
    s = hss_ulv_solve(F, nrhs, B, ldb, X, ldx, &none, nullptr);
    hss_fact_destroy(F);
    return s;
}

// ---------------------------------------------------------------------------
// batch plans: many solves (matrices of any size, repeated or not) in one
// launch per tree level; per-request workspaces; optional refinement run in
// lockstep with a batched Float64 matvec
// ---------------------------------------------------------------------------
struct hss_plan_s {
    hss_ctx_s* ctx = nullptr;
    bool holds_refs = false;  // false for the plans a factor caches for itself
    hss_prec prec = HSS_PREC_FP64;
    int cnt = 0, maxnrhs = 0;
    bool has_ir = false;
    std::vector<SolveReq> reqs;
    char* arena = nullptr;
    // solve tables: [0] solves b -> x, [1] solves r -> d (refinement)
    char* d_sinst[2] = {nullptr, nullptr};
    std::vector<char> h_sinst[2];
    size_t sinst_size = 0;
    MvInst* d_mv = nullptr;
    std::vector<MvInst> h_mv;
    BatchVec* d_vec = nullptr;
    std::vector<BatchVec> h_vec;
    MvRec* d_work = nullptr;     // matvec records
    SweepRec* d_rec = nullptr;   // solve sweep records
    Schedule S;
    MvSched mvs;
    size_t lgroot = 0, lgroot_res = 0, lgroot2 = 0, lgroot2s = 0;
    bool all_groot2 = true;
    double* d_norms = nullptr;
    std::vector<double> h_norms;
    // device-side convergence test of lockstep refinement: per-request residuals, the
    // active count of each step (a ring, copied to pinned host memory) and its events
    double* d_rel = nullptr;
    int* d_count = nullptr;
    int* h_count = nullptr;  // pinned
    std::vector<hipEvent_t> count_ev;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    // captured lockstep refinement: the check step for each of the two count slots, and the
    // correction step (solve + axpy); ir_tol = the tolerance the checks were captured with
    hipGraph_t ir_graph[3] = {nullptr, nullptr, nullptr};
    hipGraphExec_t ir_exec[3] = {nullptr, nullptr, nullptr};
    double ir_tol = -1.0;
    int64_t dev_bytes = 0;
    std::vector<char> uploaded_active;  // request activity the device tables currently hold
};

static void drop_ir_graphs(hss_plan_s* P) {
    for (int i = 0; i < 3; ++i) {
        if (P->ir_exec[i]) (void)hipGraphExecDestroy(P->ir_exec[i]);
        if (P->ir_graph[i]) (void)hipGraphDestroy(P->ir_graph[i]);
        P->ir_exec[i] = nullptr;
        P->ir_graph[i] = nullptr;
    }
    P->ir_tol = -1.0;
}

static void free_plan(hss_plan_s* P) {
    if (!P) return;
    drop_ir_graphs(P);
    if (P->gexec) (void)hipGraphExecDestroy(P->gexec);
    if (P->graph) (void)hipGraphDestroy(P->graph);
    (void)hipFree(P->arena);
    (void)hipFree(P->d_sinst[0]);
    (void)hipFree(P->d_sinst[1]);
    (void)hipFree(P->d_mv);
    (void)hipFree(P->d_vec);
    (void)hipFree(P->d_work);
    (void)hipFree(P->d_rec);
    (void)hipFree(P->d_norms);
    (void)hipFree(P->d_rel);
    (void)hipFree(P->d_count);
    if (P->h_count) (void)hipHostFree(P->h_count);
    for (hipEvent_t e : P->count_ev) (void)hipEventDestroy(e);
    delete P;
}

static size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

template <typename TF>
static hss_status plan_tables(hss_plan_s* P) {
    using TV = TF;
    const int cnt = P->cnt;
    // arena layout
    std::vector<size_t> off_ws(cnt), off_mv(cnt), off_r(cnt), off_d(cnt);
    size_t tot = 0;
    for (int i = 0; i < cnt; ++i) {
        const SolveReq& q = P->reqs[i];
        hss_fact_s* F = q.F;
        hss_mat_s* M = F->mat;
        off_ws[i] = tot;
        tot = align256(tot + sizeof(TV) * solve_ws_elems(F, q.nrhs));
        if (P->has_ir) {
            off_mv[i] = tot;
            tot = align256(tot + sizeof(double) * (size_t)(M->zrows + M->frows) * q.nrhs);
            off_r[i] = tot;
            tot = align256(tot + sizeof(double) * (size_t)M->rows * q.nrhs);
            off_d[i] = tot;
            tot = align256(tot + sizeof(double) * (size_t)M->rows * q.nrhs);
        }
    }
    HIP_TRY(hipMalloc(reinterpret_cast<void**>(&P->arena), std::max<size_t>(tot, 256)));
    P->dev_bytes = (int64_t)tot;
    P->sinst_size = sizeof(SInst<TF>);
    for (int t = 0; t < (P->has_ir ? 2 : 1); ++t) {
        std::vector<SInst<TF>> v(cnt);
        for (int i = 0; i < cnt; ++i) {
            const SolveReq& q = P->reqs[i];
            hss_fact_s* F = q.F;
            TV* ws = reinterpret_cast<TV*>(P->arena + off_ws[i]);
            SolveReq qq = q;
            if (t == 1) {
                qq.dB = reinterpret_cast<double*>(P->arena + off_r[i]);
                qq.ldb = std::max<int64_t>(F->mat->rows, 1);
                qq.dX = reinterpret_cast<double*>(P->arena + off_d[i]);
                qq.ldx = std::max<int64_t>(F->mat->rows, 1);
            }
            v[i] = make_sinst<TF>(qq, ws);
        }
        P->h_sinst[t].assign(reinterpret_cast<char*>(v.data()), reinterpret_cast<char*>(v.data() + cnt));
        HIP_TRY(hipMalloc(reinterpret_cast<void**>(&P->d_sinst[t]), P->h_sinst[t].size()));
        HIP_TRY(hipMemcpy(P->d_sinst[t], P->h_sinst[t].data(), P->h_sinst[t].size(), hipMemcpyHostToDevice));
    }
    if (P->has_ir) {
        P->h_mv.resize(cnt);
        P->h_vec.resize(cnt);
        for (int i = 0; i < cnt; ++i) {
            const SolveReq& q = P->reqs[i];
            hss_mat_s* M = q.F->mat;
            const int64_t n = std::max<int64_t>(M->rows, 1);
            MvInst& I = P->h_mv[i];
            std::memset(&I, 0, sizeof(I));
            I.nodes = M->d_nodes;
            I.op = M->d_op;
            I.B = q.dX;
            I.ldb = q.ldx;
            I.C = reinterpret_cast<double*>(P->arena + off_r[i]);
            I.ldc = n;
            I.Z = reinterpret_cast<double*>(P->arena + off_mv[i]);
            I.F = I.Z + M->zrows * q.nrhs;
            I.alpha = 1.0;
            I.beta = 0.0;
            I.nrhs = q.nrhs;
            BatchVec& v = P->h_vec[i];
            std::memset(&v, 0, sizeof(v));
            v.b = q.dB;
            v.ldb = q.ldb;
            v.x = q.dX;
            v.ldx = q.ldx;
            v.r = reinterpret_cast<double*>(P->arena + off_r[i]);
            v.d = reinterpret_cast<double*>(P->arena + off_d[i]);
            v.n = M->rows;
            v.nrhs = q.nrhs;
        }
        HIP_TRY(dalloc(&P->d_mv, (size_t)cnt));
        HIP_TRY(dalloc(&P->d_vec, (size_t)cnt));
        HIP_TRY(hipMemcpy(P->d_mv, P->h_mv.data(), sizeof(MvInst) * cnt, hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(P->d_vec, P->h_vec.data(), sizeof(BatchVec) * cnt, hipMemcpyHostToDevice));
        HIP_TRY(dalloc(&P->d_norms, (size_t)2 * cnt * P->maxnrhs));
        P->h_norms.assign((size_t)2 * cnt * P->maxnrhs, 0.0);
    }
    return HSS_OK;
}

static hss_status plan_create_impl(int32_t count, hss_fact_s* const* facts, const int32_t* nrhs,
                                   const double* const* dB, const int64_t* ldb, double* const* dX, const int64_t* ldx,
                                   int32_t with_refinement, bool hold, hss_plan_t* out) {
    if (count <= 0 || !facts || !nrhs || !dB || !ldb || !dX || !ldx || !out || !facts[0])
        return fail(HSS_ERR_ARG, "bad plan arguments");
    *out = nullptr;
    hss_ctx_s* ctx = facts[0]->mat->ctx;
    HIP_TRY(hipSetDevice(ctx->device));
    const hss_prec prec = facts[0]->prec;
    auto* P = new (std::nothrow) hss_plan_s;
    if (!P) return fail(HSS_ERR_OOM, "host allocation");
    P->ctx = ctx;
    P->prec = prec;
    P->has_ir = with_refinement != 0;
    for (int i = 0; i < count; ++i) {
        hss_fact_s* F = facts[i];
        if (!F || F->mat->ctx != ctx || F->prec != prec) {
            free_plan(P);
            return fail(HSS_ERR_ARG, "planned factors must share one context and one precision");
        }
        if (nrhs[i] <= 0 || ldb[i] < F->mat->rows || ldx[i] < F->mat->rows) {
            free_plan(P);
            return fail(HSS_ERR_DIM, "bad nrhs / leading dimension in request " + std::to_string(i));
        }
        P->reqs.push_back(SolveReq{F, nrhs[i], dB[i], ldb[i], dX[i], ldx[i]});
        P->maxnrhs = std::max(P->maxnrhs, (int)nrhs[i]);
        P->lgroot = std::max(P->lgroot, F->lds_groot);
        P->lgroot_res = std::max(P->lgroot_res, F->lds_groot_res);
        P->lgroot2 = std::max(P->lgroot2, F->lds_groot2);
        P->lgroot2s = std::max(P->lgroot2s, F->lds_groot2s);
        P->all_groot2 = P->all_groot2 && F->lds_groot2 > 0;
    }
    P->cnt = (int)P->reqs.size();
    P->uploaded_active.assign(P->cnt, 1);  // tables are built with every request active
    std::vector<hss_mat_s*> mats;
    std::vector<hss_fact_s*> pf;
    for (auto& q : P->reqs) {
        mats.push_back(q.F->mat);
        pf.push_back(q.F);
    }
    build_schedule(pf, P->S);
    // matvec records (refinement plans)
    std::vector<MvRec> work;
    if (P->has_ir) build_mv_schedule(mats, work, P->mvs);
    hipError_t e = hipSuccess;
    if (!work.empty()) {
        e = dalloc(&P->d_work, work.size());
        if (e == hipSuccess) e = hipMemcpy(P->d_work, work.data(), sizeof(MvRec) * work.size(), hipMemcpyHostToDevice);
    }
    if (e == hipSuccess) e = dalloc(&P->d_rec, P->S.rec.size());
    if (e == hipSuccess)
        e = hipMemcpy(P->d_rec, P->S.rec.data(), sizeof(SweepRec) * P->S.rec.size(), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        free_plan(P);
        return fail(HSS_ERR_OOM, std::string("plan tables: ") + hipGetErrorString(e));
    }
    hss_status s = (prec == HSS_PREC_FP32) ? plan_tables<float>(P) : plan_tables<double>(P);
    if (s != HSS_OK) {
        free_plan(P);
        return s;
    }
    if (hold) {
        for (auto& q : P->reqs) q.F->refs++;
        ctx->refs++;
        P->holds_refs = true;
    }
    *out = P;
    return HSS_OK;
}

extern "C" hss_status hss_plan_create(int32_t count, const hss_fact_t* facts, const int32_t* nrhs,
                                      const double* const* dB, const int64_t* ldb, double* const* dX,
                                      const int64_t* ldx, int32_t with_refinement, hss_plan_t* out) {
    return plan_create_impl(count, facts, nrhs, dB, ldb, dX, ldx, with_refinement, true, out);
}

template <typename TF>
static hss_status plan_solve(hss_plan_s* P, int t) {
    return launch_schedule<TF>(reinterpret_cast<const SInst<TF>*>(P->d_sinst[t]), P->d_rec, P->S, P->maxnrhs,
                               P->lgroot, P->lgroot_res, P->all_groot2 ? P->lgroot2 : 0, P->lgroot2s, P->ctx->stream);
}

// ---- autotuner: per-level launch choices of a plan's solve sweeps, by timing ----
// Candidates: LDS-resident or streamed at 16 or 32 columns and 256 or 512 threads (where
// the staging fits), and the chunked kernel.  A level's inputs are the previous levels'
// outputs, so one solve runs first; every sweep step is a pure function of its inputs,
// so launching one level repeatedly leaves the results as they were.  Choices are kept
// per level shape class in a process-wide cache (precision, direction, node count,
// columns, operator and input sizes), so identical levels are timed once.
namespace {
struct TuneKey {
    int64_t v[7];
    bool operator<(const TuneKey& o) const { return std::lexicographical_compare(v, v + 7, o.v, o.v + 7); }
};
std::mutex g_tune_mu;
std::map<TuneKey, LevelCfg> g_tune_cache;
}  // namespace

template <typename TF>
static hss_status tune_plan(hss_plan_s* P, int reps) {
    hipStream_t st = P->ctx->stream;
    const SInst<TF>* di = reinterpret_cast<const SInst<TF>*>(P->d_sinst[0]);
    Schedule& S = P->S;
    const int nrhs = P->maxnrhs;
    hss_status s = plan_solve<TF>(P, 0);  // every level's inputs in place
    if (s != HSS_OK) return s;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    struct Ev {
        hipEvent_t a, b;
        ~Ev() {
            (void)hipEventDestroy(a);
            (void)hipEventDestroy(b);
        }
    } ev{e0, e1};
    auto candidates = [&](int64_t na, int64_t k) {
        std::vector<LevelCfg> v;
        for (int form : {0, 1})
            for (int tw : {16, 32}) {
                if (tw == 32 && nrhs <= 16) continue;
                size_t lds = 0;
                if (!level_lds(S.w, form, na, k, tw, lds)) continue;
                for (int nt : {256, 512}) v.push_back(LevelCfg{form, tw, nt});
            }
        v.push_back(LevelCfg{2, 0, 0});
        return v;
    };
    // ms per launch of one level under choice c; < 0 if it cannot launch
    auto time_level = [&](auto&& launch, const LevelCfg& c) -> double {
        if (launch(c) != HSS_OK) return -1.0;
        if (hipEventRecord(e0, st) != hipSuccess) return -1.0;
        for (int r = 0; r < reps; ++r)
            if (launch(c) != HSS_OK) return -1.0;
        float ms = 0.f;
        if (hipEventRecord(e1, st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
            hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
            return -1.0;
        return ms / reps;
    };
    auto tune = [&](int dir, int count, int64_t na, int64_t k, int rows, auto&& launch) -> LevelCfg {
        const TuneKey key{{(int64_t)S.w, dir, count, nrhs, na, k, rows}};
        {
            std::lock_guard<std::mutex> g(g_tune_mu);
            auto it = g_tune_cache.find(key);
            if (it != g_tune_cache.end()) return it->second;
        }
        LevelCfg best = default_cfg(S.w, na, k, rows, count, nrhs);
        double tbest = time_level(launch, best);
        for (const LevelCfg& c : candidates(na, k)) {
            const double t = time_level(launch, c);
            if (t >= 0 && (tbest < 0 || t < tbest)) {
                tbest = t;
                best = c;
            }
        }
        std::lock_guard<std::mutex> g(g_tune_mu);
        g_tune_cache[key] = best;
        return best;
    };
    S.up_cfg.assign(S.up.size(), LevelCfg{});
    S.down_cfg.assign(S.down.size(), LevelCfg{});
    for (size_t h = 0; h < S.up.size(); ++h) {
        if (S.up[h].count == 0) continue;
        S.up_cfg[h] = tune(0, S.up[h].count, S.up_na[h], S.up_k[h], S.up_rows[h], [&](const LevelCfg& c) {
            return launch_up_level<TF>(di, P->d_rec, S, h, nrhs, c, st);
        });
    }
    for (size_t d = 1; d < S.down.size(); ++d) {
        if (S.down[d].count == 0) continue;
        S.down_cfg[d] = tune(1, S.down[d].count, S.down_na[d], S.down_k[d], S.down_m[d], [&](const LevelCfg& c) {
            return launch_down_level<TF>(di, P->d_rec, S, d, nrhs, c, st);
        });
    }
    // the results of the tuned launches were left as a solve leaves them; re-solve once so
    // the outputs are the tuned schedule's own
    return plan_solve<TF>(P, 0);
}

extern "C" hss_status hss_plan_tune(hss_plan_t P, int32_t reps, double* ms_before, double* ms_after) {
    if (!P) return fail(HSS_ERR_ARG, "null plan");
    if (reps < 1) return fail(HSS_ERR_ARG, "reps < 1");
    HIP_TRY(hipSetDevice(P->ctx->device));
    hipStream_t st = P->ctx->stream;
    const bool fp32 = P->prec == HSS_PREC_FP32;
    auto whole = [&](double* out) -> hss_status {
        if (!out) return HSS_OK;
        hss_status s = fp32 ? plan_solve<float>(P, 0) : plan_solve<double>(P, 0);  // warm
        if (s != HSS_OK) return s;
        HIP_TRY(hipEventRecord(P->ctx->ev0, st));
        for (int r = 0; r < reps; ++r) {
            s = fp32 ? plan_solve<float>(P, 0) : plan_solve<double>(P, 0);
            if (s != HSS_OK) return s;
        }
        HIP_TRY(hipEventRecord(P->ctx->ev1, st));
        HIP_TRY(hipEventSynchronize(P->ctx->ev1));
        float ms = 0.f;
        HIP_TRY(hipEventElapsedTime(&ms, P->ctx->ev0, P->ctx->ev1));
        *out = ms / reps;
        return HSS_OK;
    };
    hss_status s = whole(ms_before);
    if (s != HSS_OK) return s;
    s = fp32 ? tune_plan<float>(P, reps) : tune_plan<double>(P, reps);
    if (s != HSS_OK) return s;
    // a captured graph holds the old launches: capture again on the next execute
    if (P->gexec) (void)hipGraphExecDestroy(P->gexec);
    if (P->graph) (void)hipGraphDestroy(P->graph);
    P->gexec = nullptr;
    P->graph = nullptr;
    drop_ir_graphs(P);
    return whole(ms_after);
}

extern "C" hss_status hss_plan_tune_choice(hss_plan_t P, int32_t dir, int32_t level, int32_t* form, int32_t* tw,
                                           int32_t* nt) {
    if (!P || !form || !tw || !nt) return fail(HSS_ERR_ARG, "null argument");
    const std::vector<LevelCfg>& v = dir == 0 ? P->S.up_cfg : P->S.down_cfg;
    if (dir != 0 && dir != 1) return fail(HSS_ERR_ARG, "dir must be 0 (up) or 1 (down)");
    const int n = (int)(dir == 0 ? P->S.up.size() : P->S.down.size());
    if (level < 0 || level >= n) return fail(HSS_ERR_ARG, "level out of range");
    const LevelCfg c = level < (int)v.size() ? v[level] : LevelCfg{};
    *form = c.form;
    *tw = c.tw;
    *nt = c.nt;
    return HSS_OK;
}

static hss_status plan_matvec(hss_plan_s* P) {
    return launch_mv_schedule(P->d_mv, P->d_work, P->mvs, P->maxnrhs, P->ctx->stream);
}

// re-upload the tables whose nrhs changed (the stream is idle: callers synchronize first)
using SInstF = SInst<float>;
using SInstD = SInst<double>;

static hss_status plan_set_active(hss_plan_s* P, const std::vector<char>& active) {
    if (active == P->uploaded_active) return HSS_OK;  // the device tables already say so
    P->uploaded_active = active;
    const size_t nrhs_off = (P->prec == HSS_PREC_FP32) ? offsetof(SInstF, nrhs) : offsetof(SInstD, nrhs);
    for (int t = 0; t < 2; ++t) {
        if (!P->d_sinst[t]) continue;
        for (int i = 0; i < P->cnt; ++i) {
            int32_t* nr = reinterpret_cast<int32_t*>(P->h_sinst[t].data() + i * P->sinst_size + nrhs_off);
            *nr = active[i] ? P->reqs[i].nrhs : 0;
        }
        HIP_TRY(hipMemcpy(P->d_sinst[t], P->h_sinst[t].data(), P->h_sinst[t].size(), hipMemcpyHostToDevice));
    }
    for (int i = 0; i < P->cnt; ++i) {
        P->h_mv[i].nrhs = active[i] ? P->reqs[i].nrhs : 0;
        P->h_vec[i].nrhs = active[i] ? P->reqs[i].nrhs : 0;
    }
    HIP_TRY(hipMemcpy(P->d_mv, P->h_mv.data(), sizeof(MvInst) * P->cnt, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(P->d_vec, P->h_vec.data(), sizeof(BatchVec) * P->cnt, hipMemcpyHostToDevice));
    return HSS_OK;
}

extern "C" hss_status hss_plan_execute(hss_plan_t P, const hss_refine_opts* opts, hss_solve_report* rep) {
    if (!P) return fail(HSS_ERR_ARG, "null plan");
    HIP_TRY(hipSetDevice(P->ctx->device));
    hipStream_t st = P->ctx->stream;
    const bool fp32 = P->prec == HSS_PREC_FP32;
    if (rep) { rep->iterations = 0; rep->rel_residual = -1.0; rep->seconds = 0.0; }
    HIP_TRY(hipEventRecord(P->ctx->ev0, st));
    if (opts && opts->mode == 2)
        return fail(HSS_ERR_ARG, "GMRES-IR (opts->mode 2) runs on single solves (hss_ulv_solve, hss_ulv_solve_dev)");
    const bool refine = P->has_ir && opts && opts->mode == 1;
    if (fp32 && !refine)
        return fail(HSS_ERR_ARG, "a plan of Float32 factors executes with refinement only (with_refinement at "
                                 "creation and opts->mode 1): a Float32 solve alone is accurate to about 1e-7");
    if (!refine && !g_prof && !fp32) {
        // Float64 solve: replay one captured graph of all level launches
        if (!P->gexec) {
            HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            hss_status s = plan_solve<double>(P, 0);
            hipGraph_t g = nullptr;
            hipError_t ee = hipStreamEndCapture(st, &g);
            if (s != HSS_OK) { if (g) (void)hipGraphDestroy(g); return s; }
            HIP_TRY(ee);
            P->graph = g;
            HIP_TRY(hipGraphInstantiate(&P->gexec, P->graph, nullptr, nullptr, 0));
        }
        HIP_TRY(hipGraphLaunch(P->gexec, st));
        HIP_TRY(hipEventRecord(P->ctx->ev1, st));
        return HSS_OK;
    }
    hss_status s = fp32 ? plan_solve<float>(P, 0) : plan_solve<double>(P, 0);
    if (s != HSS_OK) return s;
    if (!refine) {
        HIP_TRY(hipEventRecord(P->ctx->ev1, st));
        return HSS_OK;
    }
    // Lockstep refinement with the convergence test on the device: each step is
    //   check k:  r = b - A x (batched Float64 matvec), norms, k_refine_check (retires the
    //             requests at tol in the device tables, counts the others)
    //   solve k:  d = A \ r for the active requests, x += d
    // The host reads check k's count (pinned memory, event) only after enqueuing check k + 1,
    // so the device never waits for the host; a step enqueued after the last request
    // retired finds every column count 0 and its launches do nothing.  Check k writes count
    // slot k % 2: the host has read slot k % 2 (check k - 2's) before it enqueues check k.
    // Outside profiling, the two check steps (one per slot) and the correction step are
    // captured once (per tolerance) and replayed as graphs; every kernel reads its tables
    // from device memory, which the checks and plan_set_active update in place.
    constexpr int kRing = 2;
    if (!P->d_rel) {
        HIP_TRY(dalloc(&P->d_rel, (size_t)P->cnt));
        HIP_TRY(hipMemsetAsync(P->d_rel, 0, sizeof(double) * P->cnt, st));
        HIP_TRY(dalloc(&P->d_count, (size_t)kRing));
        HIP_TRY(hipHostMalloc(reinterpret_cast<void**>(&P->h_count), sizeof(int) * kRing, hipHostMallocDefault));
        P->count_ev.resize(kRing);
        for (auto& e : P->count_ev) HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    }
    const size_t nrhs_off = fp32 ? offsetof(SInst<float>, nrhs) : offsetof(SInst<double>, nrhs);
    auto enqueue_check = [&](int slot) -> hss_status {
        hss_status c = plan_matvec(P);  // r = A x
        if (c != HSS_OK) return c;
        HIP_TRY(launch_resid_batch(P->d_vec, P->cnt, P->maxnrhs, P->d_norms, st));
        HIP_TRY(launch_refine_check(P->d_vec, P->d_sinst[0], P->d_sinst[1], (int64_t)P->sinst_size, (int)nrhs_off,
                                    P->d_mv, P->cnt, P->maxnrhs, P->d_norms, opts->tol, P->d_rel,
                                    P->d_count + slot, st));
        HIP_TRY(hipMemcpyAsync(P->h_count + slot, P->d_count + slot, sizeof(int), hipMemcpyDeviceToHost, st));
        return HSS_OK;
    };
    auto enqueue_solve = [&]() -> hss_status {
        hss_status c = fp32 ? plan_solve<float>(P, 1) : plan_solve<double>(P, 1);  // d = A \ r
        if (c != HSS_OK) return c;
        HIP_TRY(launch_axpy_batch(P->d_vec, P->cnt, P->maxnrhs, st));  // x += d
        return HSS_OK;
    };
    const bool graphs = !g_prof;
    if (graphs && P->ir_tol != opts->tol) drop_ir_graphs(P);
    auto capture = [&](int i, auto&& enqueue) -> hss_status {
        HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        hss_status c = enqueue();
        hipGraph_t g = nullptr;
        hipError_t ee = hipStreamEndCapture(st, &g);
        if (c != HSS_OK) {
            if (g) (void)hipGraphDestroy(g);
            return c;
        }
        HIP_TRY(ee);
        P->ir_graph[i] = g;
        HIP_TRY(hipGraphInstantiate(&P->ir_exec[i], g, nullptr, nullptr, 0));
        return HSS_OK;
    };
    if (graphs && !P->ir_exec[0]) {
        hss_status c = capture(0, [&] { return enqueue_check(0); });
        if (c == HSS_OK) c = capture(1, [&] { return enqueue_check(1); });
        if (c == HSS_OK) c = capture(2, enqueue_solve);
        if (c != HSS_OK) {
            drop_ir_graphs(P);
            return c;
        }
        P->ir_tol = opts->tol;
    }
    auto check = [&](int k) -> hss_status {
        const int slot = k % kRing;
        if (graphs) HIP_TRY(hipGraphLaunch(P->ir_exec[slot], st));
        else {
            hss_status c = enqueue_check(slot);
            if (c != HSS_OK) return c;
        }
        HIP_TRY(hipEventRecord(P->count_ev[slot], st));
        return HSS_OK;
    };
    auto count_of = [&](int k, int& c) -> hss_status {
        HIP_TRY(hipEventSynchronize(P->count_ev[k % kRing]));
        c = P->h_count[k % kRing];
        return HSS_OK;
    };
    int it = 0, remaining = -1;
    for (int k = 0;; ++k) {
        if ((s = check(k)) != HSS_OK) return s;
        if (k >= 1) {
            int c = 0;
            if ((s = count_of(k - 1, c)) != HSS_OK) return s;
            if (c == 0) {  // retired by check k - 1: solve k - 1 and check k did nothing
                it = k - 1;
                remaining = 0;
                break;
            }
        }
        if (k >= opts->max_iter) {
            if ((s = count_of(k, remaining)) != HSS_OK) return s;
            it = k;
            break;
        }
        if (graphs) HIP_TRY(hipGraphLaunch(P->ir_exec[2], st));
        else if ((s = enqueue_solve()) != HSS_OK) return s;
    }
    HIP_TRY(hipStreamSynchronize(st));
    std::vector<double> hrel(P->cnt);
    HIP_TRY(hipMemcpy(hrel.data(), P->d_rel, sizeof(double) * P->cnt, hipMemcpyDeviceToHost));
    double worst = 0.0;
    for (int i = 0; i < P->cnt; ++i)
        if (P->reqs[i].nrhs > 0) worst = hrel[i] > worst || hrel[i] != hrel[i] ? hrel[i] : worst;
    // the device tables retired requests on their own: upload the full activity again
    P->uploaded_active.assign(P->cnt, 2);
    // restore all requests as active for the next execution
    std::vector<char> all(P->cnt, 1);
    s = plan_set_active(P, all);
    if (s != HSS_OK) return s;
    HIP_TRY(hipEventRecord(P->ctx->ev1, st));
    HIP_TRY(hipStreamSynchronize(st));
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, P->ctx->ev0, P->ctx->ev1) == hipSuccess) g_last_ms = ms;
    if (rep) {
        rep->iterations = it;
        rep->rel_residual = worst;
        rep->seconds = ms * 1e-3;
    }
    if (remaining != 0 || !(worst <= opts->tol)) {
        char buf[160];
        std::snprintf(buf, sizeof buf, "batched refinement stopped at relative residual %.3e > tol %.3e after %d steps",
                      worst, opts->tol, it);
        return fail(HSS_ERR_NOCONV, buf);
    }
    return HSS_OK;
}

extern "C" hss_status hss_plan_destroy(hss_plan_t P) {
    if (!P) return HSS_OK;
    hss_ctx_s* c = P->ctx;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    std::vector<hss_fact_s*> held;
    if (P->holds_refs)
        for (auto& q : P->reqs) held.push_back(q.F);
    const bool hold = P->holds_refs;
    free_plan(P);
    for (auto* F : held) fact_release(F);
    if (hold) ctx_release(c);
    return HSS_OK;
}

extern "C" int64_t hss_plan_device_bytes(hss_plan_t P) { return P ? P->dev_bytes : 0; }

static void destroy_plan_cache(std::vector<std::pair<PlanKey, hss_plan_t>>& cache) {
    for (auto& e : cache) hss_plan_destroy(e.second);
    cache.clear();
}

// Batched solves without refinement: a one-shot plan.
extern "C" hss_status hss_ulv_solve_batch_dev(int32_t count, const hss_fact_t* facts, const int32_t* nrhs,
                                              const double* const* dB, const int64_t* ldb, double* const* dX,
                                              const int64_t* ldx) {
    if (count <= 0 || !facts || !nrhs || !dB || !ldb || !dX || !ldx || !facts[0])
        return fail(HSS_ERR_ARG, "bad batch arguments");
    std::vector<hss_fact_t> f;
    std::vector<int32_t> nr;
    std::vector<const double*> b;
    std::vector<double*> x;
    std::vector<int64_t> lb, lx;
    for (int i = 0; i < count; ++i) {
        if (nrhs[i] < 0) return fail(HSS_ERR_ARG, "nrhs < 0");
        if (nrhs[i] == 0) continue;
        f.push_back(facts[i]);
        nr.push_back(nrhs[i]);
        b.push_back(dB[i]);
        x.push_back(dX[i]);
        lb.push_back(ldb[i]);
        lx.push_back(ldx[i]);
    }
    if (f.empty()) return HSS_OK;
    for (hss_fact_t F : f)
        if (F && F->prec == HSS_PREC_FP32)
            return fail(HSS_ERR_ARG, "hss_ulv_solve_batch_dev solves Float64 factors only: a Float32 factor alone "
                                     "is accurate to about 1e-7; refine it (hss_plan_create with_refinement, "
                                     "hss_plan_execute with opts->mode 1)");
    hss_plan_t P = nullptr;
    hss_status s = hss_plan_create((int32_t)f.size(), f.data(), nr.data(), b.data(), lb.data(), x.data(), lx.data(), 0, &P);
    if (s != HSS_OK) return s;
    hss_refine_opts none{0.0, 0, 0};
    // without capture: a one-shot plan gains nothing from a graph
    hipStream_t st = P->ctx->stream;
    s = (P->prec == HSS_PREC_FP32) ? plan_solve<float>(P, 0) : plan_solve<double>(P, 0);
    (void)none;
    (void)hipStreamSynchronize(st);
    hss_plan_destroy(P);
    return s;
}

// ---------------------------------------------------------------------------
// serving: resident factorizations admitted by an HBM budget, and continuous
// batching of arriving solves.  Each round batches, for every request in
// flight, either its first solve (b -> x) or its next refinement correction
// (r -> d, x += d) into one launch per tree level, followed by one batched
// Float64 matvec + residual for all of them; requests leave as soon as they
// reach the tolerance and queued requests take their slots in the next round.
// ---------------------------------------------------------------------------
struct ServeReq {
    int64_t ticket = 0;
    int mid = 0;
    int nrhs = 0;
    const double* dB = nullptr;
    int64_t ldb = 0;
    double* dX = nullptr;
    int64_t ldx = 0;
    double* r = nullptr;  // n x nrhs residual / right-hand side of the correction
    double* d = nullptr;  // n x nrhs correction
    size_t rbytes = 0;
    int steps = 0;        // solves done so far (1 = first solve)
    double rel = -1.0;
    int status = 0;       // 0 pending/in flight, 1 converged, 2 stopped at max_iter
    std::chrono::steady_clock::time_point t_submit, t_done;
};

struct hss_server_s {
    hss_ctx_s* ctx = nullptr;
    hss_prec prec = HSS_PREC_FP32;
    hss_refine_opts opts{1e-10, 10, 1};
    int64_t budget = 0, used = 0;
    int max_inflight = 64;
    std::vector<hss_fact_s*> facts;  // by matrix id (nullptr after eviction)
    std::vector<int64_t> fact_bytes;
    std::vector<ServeReq*> pending, inflight;
    std::vector<ServeReq*> finished;  // kept until polled
    int64_t next_ticket = 1;
    // per-round device scratch (grown on demand)
    char* d_round = nullptr;
    size_t round_cap = 0;
    char* d_tab = nullptr;
    size_t tab_cap = 0;
    double* d_norms = nullptr;
    size_t norms_cap = 0;
    int64_t rounds = 0, solved = 0;
};

static hss_status grow(char** p, size_t* cap, size_t need) {
    if (need <= *cap) return HSS_OK;
    (void)hipFree(*p);
    *p = nullptr;
    *cap = 0;
    size_t c = std::max(need, (size_t)1 << 20);
    c = std::max(c, need + need / 2);
    HIP_TRY(hipMalloc(reinterpret_cast<void**>(p), c));
    *cap = c;
    return HSS_OK;
}

extern "C" hss_status hss_server_create(hss_ctx_t ctx, int64_t hbm_budget_bytes, hss_prec prec,
                                        const hss_refine_opts* opts, int32_t max_inflight, hss_server_t* out) {
    if (!ctx || !out || hbm_budget_bytes <= 0 || max_inflight <= 0) return fail(HSS_ERR_ARG, "bad server arguments");
    if (prec != HSS_PREC_FP64 && prec != HSS_PREC_FP32) return fail(HSS_ERR_ARG, "unknown precision");
    if (opts && opts->mode == 2)
        return fail(HSS_ERR_ARG, "GMRES-IR (opts->mode 2) runs on single solves (hss_ulv_solve, hss_ulv_solve_dev)");
    if (prec == HSS_PREC_FP32 && opts && opts->mode != 1)
        return fail(HSS_ERR_ARG, "a Float32 server needs refinement (opts->mode 1): a Float32 solve alone is "
                                 "accurate to about 1e-7");
    auto* S = new (std::nothrow) hss_server_s;
    if (!S) return fail(HSS_ERR_OOM, "host allocation");
    S->ctx = ctx;
    ctx->refs++;
    S->prec = prec;
    if (opts) S->opts = *opts;
    S->budget = hbm_budget_bytes;
    S->max_inflight = max_inflight;
    *out = S;
    return HSS_OK;
}

// Factor `mat` (server precision) and keep it resident, if the budget allows.
extern "C" hss_status hss_server_admit(hss_server_t S, hss_mat_t mat, int32_t* id) {
    if (!S || !mat || !id) return fail(HSS_ERR_ARG, "null argument");
    if (mat->ctx != S->ctx) return fail(HSS_ERR_ARG, "matrix belongs to another context");
    // admission estimate before factoring: the factor is about as large as the generators
    const int64_t estimate = mat->dev_bytes * (S->prec == HSS_PREC_FP32 ? 1 : 2);
    if (S->used + estimate > S->budget)
        return fail(HSS_ERR_BUDGET, "admission refused: " + std::to_string(S->used + estimate) + " bytes > budget " +
                                        std::to_string(S->budget));
    hss_fact_t F = nullptr;
    hss_status st = hss_ulv_factor(mat, S->prec, &F);
    if (st != HSS_OK) return st;
    const int64_t bytes = F->dev_bytes;
    if (S->used + bytes > S->budget) {
        hss_fact_destroy(F);
        return fail(HSS_ERR_BUDGET, "admission refused after factorization: over the HBM budget");
    }
    S->used += bytes;
    S->facts.push_back(F);
    S->fact_bytes.push_back(bytes);
    *id = (int32_t)S->facts.size() - 1;
    return HSS_OK;
}

extern "C" hss_status hss_server_evict(hss_server_t S, int32_t id) {
    if (!S || id < 0 || id >= (int32_t)S->facts.size() || !S->facts[id]) return fail(HSS_ERR_ARG, "bad matrix id");
    for (auto* q : S->pending)
        if (q->mid == id) return fail(HSS_ERR_ARG, "matrix has queued requests");
    for (auto* q : S->inflight)
        if (q->mid == id) return fail(HSS_ERR_ARG, "matrix has requests in flight");
    hss_fact_destroy(S->facts[id]);
    S->facts[id] = nullptr;
    S->used -= S->fact_bytes[id];
    return HSS_OK;
}

extern "C" int64_t hss_server_used_bytes(hss_server_t S) { return S ? S->used : 0; }

extern "C" hss_status hss_server_submit(hss_server_t S, int32_t id, int32_t nrhs, const double* dB, int64_t ldb,
                                        double* dX, int64_t ldx, int64_t* ticket) {
    if (!S || !ticket || id < 0 || id >= (int32_t)S->facts.size() || !S->facts[id])
        return fail(HSS_ERR_ARG, "bad submit arguments");
    hss_mat_s* M = S->facts[id]->mat;
    if (nrhs <= 0 || ldb < M->rows || ldx < M->rows) return fail(HSS_ERR_DIM, "bad nrhs / leading dimension");
    auto* q = new (std::nothrow) ServeReq;
    if (!q) return fail(HSS_ERR_OOM, "host allocation");
    q->ticket = S->next_ticket++;
    q->mid = id;
    q->nrhs = nrhs;
    q->dB = dB;
    q->ldb = ldb;
    q->dX = dX;
    q->ldx = ldx;
    q->t_submit = std::chrono::steady_clock::now();
    S->pending.push_back(q);
    *ticket = q->ticket;
    return HSS_OK;
}

template <typename TF>
static hss_status server_round(hss_server_s* S) {
    using TV = TF;
    hipStream_t st = S->ctx->stream;
    // 1. admit queued requests into free slots (their r/d vectors come out of the budget)
    while (!S->pending.empty() && (int)S->inflight.size() < S->max_inflight) {
        ServeReq* q = S->pending.front();
        hss_mat_s* M = S->facts[q->mid]->mat;
        const size_t bytes = sizeof(double) * 2 * (size_t)M->rows * q->nrhs;
        if (S->used + (int64_t)bytes > S->budget && !S->inflight.empty()) break;  // wait for memory to free up
        HIP_TRY(hipMalloc(reinterpret_cast<void**>(&q->r), bytes));
        q->d = q->r + (size_t)M->rows * q->nrhs;
        q->rbytes = bytes;
        S->used += (int64_t)bytes;
        S->pending.erase(S->pending.begin());
        S->inflight.push_back(q);
    }
    const int cnt = (int)S->inflight.size();
    if (cnt == 0) return HSS_OK;
    // 2. tables: solve (first solve b->x, or correction r->d), axpy, matvec x->r, residual
    size_t ws = 0;
    int maxnrhs = 0;
    size_t lgroot = 0, lgroot_res = 0, lgroot2 = 0, lgroot2s = 0;
    bool all_groot2 = true;
    std::vector<size_t> off_s(cnt), off_m(cnt);
    std::vector<hss_mat_s*> mats(cnt);
    for (int i = 0; i < cnt; ++i) {
        ServeReq* q = S->inflight[i];
        hss_fact_s* F = S->facts[q->mid];
        hss_mat_s* M = F->mat;
        mats[i] = M;
        off_s[i] = ws;
        ws = align256(ws + sizeof(TV) * solve_ws_elems(F, q->nrhs));
        off_m[i] = ws;
        ws = align256(ws + sizeof(double) * (size_t)(M->zrows + M->frows) * q->nrhs);
        maxnrhs = std::max(maxnrhs, q->nrhs);
        lgroot = std::max(lgroot, F->lds_groot);
        lgroot_res = std::max(lgroot_res, F->lds_groot_res);
        lgroot2 = std::max(lgroot2, F->lds_groot2);
        lgroot2s = std::max(lgroot2s, F->lds_groot2s);
        all_groot2 = all_groot2 && F->lds_groot2 > 0;
    }
    hss_status s = grow(&S->d_round, &S->round_cap, ws);
    if (s != HSS_OK) return s;
    std::vector<SInst<TF>> si(cnt);
    std::vector<MvInst> mv(cnt);
    std::vector<BatchVec> bv(cnt);
    for (int i = 0; i < cnt; ++i) {
        ServeReq* q = S->inflight[i];
        hss_fact_s* F = S->facts[q->mid];
        hss_mat_s* M = F->mat;
        const int64_t n = std::max<int64_t>(M->rows, 1);
        TV* ws = reinterpret_cast<TV*>(S->d_round + off_s[i]);
        SolveReq sr = (q->steps == 0) ? SolveReq{F, q->nrhs, q->dB, q->ldb, q->dX, q->ldx}
                                      : SolveReq{F, q->nrhs, q->r, n, q->d, n};
        si[i] = make_sinst<TF>(sr, ws);
        MvInst& I = mv[i];
        std::memset(&I, 0, sizeof(I));
        I.nodes = M->d_nodes;
        I.op = M->d_op;
        I.B = q->dX;
        I.ldb = q->ldx;
        I.C = q->r;
        I.ldc = n;
        I.Z = reinterpret_cast<double*>(S->d_round + off_m[i]);
        I.F = I.Z + M->zrows * q->nrhs;
        I.alpha = 1.0;
        I.beta = 0.0;
        I.nrhs = q->nrhs;
        BatchVec& v = bv[i];
        std::memset(&v, 0, sizeof(v));
        v.b = q->dB;
        v.ldb = q->ldb;
        v.x = q->dX;
        v.ldx = q->ldx;
        v.r = q->r;
        v.d = q->d;
        v.n = M->rows;
        v.nrhs = (q->steps == 0) ? 0 : q->nrhs;  // axpy only for corrections
    }
    Schedule sch;
    std::vector<hss_fact_s*> rf(cnt);
    for (int i = 0; i < cnt; ++i) rf[i] = S->facts[S->inflight[i]->mid];
    build_schedule(rf, sch);
    std::vector<MvRec> work;
    MvSched mvs;
    build_mv_schedule(mats, work, mvs);
    // one table upload per round (the stream is idle here)
    const size_t o_si = 0, o_mv = align256(sizeof(SInst<TF>) * cnt), o_bv = o_mv + align256(sizeof(MvInst) * cnt),
                 o_w = o_bv + align256(sizeof(BatchVec) * cnt), o_r = o_w + align256(sizeof(MvRec) * work.size()),
                 tab = o_r + sizeof(SweepRec) * sch.rec.size();
    s = grow(&S->d_tab, &S->tab_cap, tab);
    if (s != HSS_OK) return s;
    std::vector<char> h(tab);
    std::memcpy(h.data() + o_si, si.data(), sizeof(SInst<TF>) * cnt);
    std::memcpy(h.data() + o_mv, mv.data(), sizeof(MvInst) * cnt);
    std::memcpy(h.data() + o_bv, bv.data(), sizeof(BatchVec) * cnt);
    std::memcpy(h.data() + o_w, work.data(), sizeof(MvRec) * work.size());
    std::memcpy(h.data() + o_r, sch.rec.data(), sizeof(SweepRec) * sch.rec.size());
    HIP_TRY(hipMemcpy(S->d_tab, h.data(), tab, hipMemcpyHostToDevice));
    const size_t nb = sizeof(double) * 2 * (size_t)cnt * maxnrhs;
    s = grow(reinterpret_cast<char**>(&S->d_norms), &S->norms_cap, nb);
    if (s != HSS_OK) return s;
    auto* dsi = reinterpret_cast<const SInst<TF>*>(S->d_tab + o_si);
    auto* dmv = reinterpret_cast<const MvInst*>(S->d_tab + o_mv);
    auto* dbv = reinterpret_cast<BatchVec*>(S->d_tab + o_bv);
    auto* dw = reinterpret_cast<const MvRec*>(S->d_tab + o_w);
    auto* dr = reinterpret_cast<const SweepRec*>(S->d_tab + o_r);
    // 3. one batched solve, the corrections, one batched matvec + residual
    s = launch_schedule<TF>(dsi, dr, sch, maxnrhs, lgroot, lgroot_res, all_groot2 ? lgroot2 : 0, lgroot2s, st);
    if (s != HSS_OK) return s;
    HIP_TRY(launch_axpy_batch(dbv, cnt, maxnrhs, st));
    s = launch_mv_schedule(dmv, dw, mvs, maxnrhs, st);
    if (s != HSS_OK) return s;
    // the residual pass needs every request active: rewrite nrhs of the BatchVec rows
    for (int i = 0; i < cnt; ++i) bv[i].nrhs = S->inflight[i]->nrhs;
    HIP_TRY(hipMemcpyAsync(S->d_tab + o_bv, bv.data(), sizeof(BatchVec) * cnt, hipMemcpyHostToDevice, st));
    HIP_TRY(hipStreamSynchronize(st));  // bv must stay valid until the copy is done
    HIP_TRY(launch_resid_batch(dbv, cnt, maxnrhs, S->d_norms, st));
    std::vector<double> norms(2 * (size_t)cnt * maxnrhs);
    HIP_TRY(hipMemcpyAsync(norms.data(), S->d_norms, nb, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    // 4. retire converged requests
    const auto now = std::chrono::steady_clock::now();
    std::vector<ServeReq*> keep;
    for (int i = 0; i < cnt; ++i) {
        ServeReq* q = S->inflight[i];
        q->steps += 1;
        double rel = 0.0;
        for (int j = 0; j < q->nrhs; ++j) {
            const double b2 = norms[2 * ((size_t)i * maxnrhs + j)], r2 = norms[2 * ((size_t)i * maxnrhs + j) + 1];
            rel = std::max(rel, b2 > 0 ? std::sqrt(r2 / b2) : std::sqrt(r2));
        }
        q->rel = rel;
        const bool conv = rel <= S->opts.tol || S->opts.mode == 0;
        if (conv || q->steps > S->opts.max_iter) {
            q->status = conv ? 1 : 2;
            q->t_done = now;
            (void)hipFree(q->r);
            S->used -= (int64_t)q->rbytes;
            q->r = q->d = nullptr;
            S->finished.push_back(q);
            S->solved += 1;
        } else {
            keep.push_back(q);
        }
    }
    S->inflight.swap(keep);
    S->rounds += 1;
    return HSS_OK;
}

// One scheduling round; returns the number of requests still queued or in flight.
extern "C" hss_status hss_server_step(hss_server_t S, int32_t* outstanding) {
    if (!S) return fail(HSS_ERR_ARG, "null server");
    HIP_TRY(hipSetDevice(S->ctx->device));
    hss_status s = (S->prec == HSS_PREC_FP32) ? server_round<float>(S) : server_round<double>(S);
    if (outstanding) *outstanding = (int32_t)(S->pending.size() + S->inflight.size());
    return s;
}

// status: 0 = not finished, 1 = converged, 2 = stopped at max_iter.  A finished
// ticket is reported once and then forgotten.
extern "C" hss_status hss_server_poll(hss_server_t S, int64_t ticket, int32_t* status, double* latency_s,
                                      hss_solve_report* rep) {
    if (!S || !status) return fail(HSS_ERR_ARG, "null argument");
    *status = 0;
    for (size_t i = 0; i < S->finished.size(); ++i) {
        ServeReq* q = S->finished[i];
        if (q->ticket != ticket) continue;
        *status = q->status;
        if (latency_s) *latency_s = std::chrono::duration<double>(q->t_done - q->t_submit).count();
        if (rep) {
            rep->iterations = q->steps - 1;
            rep->rel_residual = q->rel;
            rep->seconds = 0.0;
        }
        S->finished.erase(S->finished.begin() + i);
        delete q;
        return HSS_OK;
    }
    return HSS_OK;
}

extern "C" hss_status hss_server_destroy(hss_server_t S) {
    if (!S) return HSS_OK;
    (void)hipSetDevice(S->ctx->device);
    (void)hipStreamSynchronize(S->ctx->stream);
    for (auto* q : S->pending) delete q;
    for (auto* q : S->inflight) {
        (void)hipFree(q->r);
        delete q;
    }
    for (auto* q : S->finished) delete q;
    for (auto* F : S->facts)
        if (F) hss_fact_destroy(F);
    (void)hipFree(S->d_round);
    (void)hipFree(S->d_tab);
    (void)hipFree(S->d_norms);
    hss_ctx_s* c = S->ctx;
    delete S;
    ctx_release(c);
    return HSS_OK;
}

// ---------------------------------------------------------------------------
// diagnostics
// ---------------------------------------------------------------------------
extern "C" const char* hss_last_error(void) { return g_err.c_str(); }

extern "C" void hss_prof_enable(int on) { g_prof = on != 0; }

extern "C" void hss_prof_reset(void) {
    for (auto& r : g_prof_recs)
        for (hipEvent_t e : {r.e0, r.e1, r.k0, r.k1}) g_prof_pool.push_back(e);
    g_prof_recs.clear();
}

extern "C" int32_t hss_prof_count(void) {
    // resolve pending timings (waits for the recorded events)
    for (auto& r : g_prof_recs)
        if (r.ms < 0.f) {
            (void)hipEventSynchronize(r.e1);
            float ms = 0.f;
            if (r.launches == 1 && hipEventElapsedTime(&ms, r.k0, r.k1) == hipSuccess) r.ms = ms;
            else if (hipEventElapsedTime(&ms, r.e0, r.e1) == hipSuccess) r.ms = ms;
        }
    return (int32_t)g_prof_recs.size();
}

extern "C" hss_status hss_prof_get(int32_t i, const char** name, int32_t* level, int32_t* nwork, double* ms) {
    if (i < 0 || i >= (int32_t)g_prof_recs.size()) return fail(HSS_ERR_ARG, "profile index out of range");
    const ProfRec& r = g_prof_recs[i];
    if (name) *name = r.name;
    if (level) *level = r.level;
    if (nwork) *nwork = r.nwork;
    if (ms) *ms = r.ms;
    return HSS_OK;
}

extern "C" hss_status hss_fact_node_info(hss_fact_t F, int32_t node, int32_t* info) {
    if (!F || node < 0 || node >= (int32_t)F->hfn.size() || !info) return fail(HSS_ERR_ARG, "bad node query");
    const FNode& f = F->hfn[node];
    info[0] = f.m; info[1] = f.n; info[2] = f.k; info[3] = f.nk;
    info[4] = f.mr; info[5] = f.nr; info[6] = f.pass; info[7] = f.nblk;
    info[8] = f.My; info[9] = f.Kx;
    return HSS_OK;
}
extern "C" hss_status hss_mat_matvec_info(hss_mat_t M, int32_t* cut_depth, int32_t* folded) {
    if (!M || !cut_depth || !folded) return fail(HSS_ERR_ARG, "null argument");
    *cut_depth = M->cut_depth;
    *folded = 0;
    for (const DevNode& d : M->hn) *folded += d.fold ? 1 : 0;
    return HSS_OK;
}
extern "C" double hss_last_device_ms(void) { return g_last_ms; }
