// hss_tasks.h -- host-side helpers of the device compressors (hss_compress.hip,
// hss_randcompress.hip): device buffers freed together, and lists of small GEMM / copy /
// gather tasks run as one launch each.  Not part of the public ABI.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "hss_internal.h"

namespace hss {

// device buffers of one compression, freed together after the stream has drained
struct Buffers {
    hipStream_t st = nullptr;
    std::vector<void*> ptrs;
    ~Buffers() {
        if (st) (void)hipStreamSynchronize(st);
        for (void* p : ptrs) (void)hipFree(p);
    }
    template <typename T>
    hipError_t get(T** p, size_t count) {
        void* q = nullptr;
        hipError_t e = hipMalloc(&q, std::max<size_t>(count, 1) * sizeof(T));
        if (e != hipSuccess) return e;
        ptrs.push_back(q);
        *p = reinterpret_cast<T*>(q);
        return hipSuccess;
    }
    void release(void* p) {
        auto it = std::find(ptrs.begin(), ptrs.end(), p);
        if (it == ptrs.end()) return;
        (void)hipStreamSynchronize(st);
        (void)hipFree(p);
        ptrs.erase(it);
    }
};

struct TaskList {
    std::vector<GemmTask> t;
    void gemm(int m, int n, int k, int flags, const double* a, int lda, const double* b, int ldb, double* c, int ldc) {
        if (m <= 0 || n <= 0) return;
        GemmTask g{};
        g.m = m; g.n = n; g.k = k; g.flags = flags;
        g.lda = std::max(lda, 1); g.ldb = std::max(ldb, 1); g.ldc = std::max(ldc, 1);
        g.a = a; g.b = b; g.c = c;
        t.push_back(g);
    }
    void copy(int m, int n, bool trans, const double* a, int lda, double* c, int ldc) {
        gemm(m, n, 0, GT_COPY | (trans ? GT_TA : 0), a, lda, nullptr, 1, c, ldc);
    }
};

struct GatherList {
    std::vector<GatherTask> t;
    // dst (m x n, ld ldd) = src[ri[i] + ci[j] * ld] (ri / ci null: i / j)
    void add(int m, int n, const double* src, int64_t ld, const int32_t* ri, const int32_t* ci, double* dst, int ldd) {
        if (m <= 0 || n <= 0) return;
        GatherTask g{};
        g.src = src; g.ld = ld; g.ri = ri; g.ci = ci; g.dst = dst;
        g.m = m; g.n = n; g.ldd = std::max(ldd, 1);
        t.push_back(g);
    }
};

}  // namespace hss
