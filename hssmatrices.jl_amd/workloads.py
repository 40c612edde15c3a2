"""Kernel-matrix workloads compressed on the device (``hip.compress_dev``).

The reference builds its test and benchmark matrices densely from a kernel and compresses
them with ``compress(A, rcl, ccl; atol, rtol)`` (``test/runtests.jl:19-20, 36``,
``test/benchmarks.jl:9-20``).  These builders do the same on one GPU: the dense kernel
matrix is formed in device memory by torch, compressed by ``hss_compress_dev`` and freed,
so only the generators stay.  The points come from numpy ``default_rng`` streams with the
seeds given, so every run builds the same inputs (the compressor's Gaussian samples are
fixed too).

* ``log_kernel``: log|x_i - x_j| with unit diagonal on n points sorted uniform on [-1, 1];
* ``cauchy_instances``: Cauchy 1/(x_i - x_j) with unit diagonal on sorted uniform points
  shifted by U(0, 0.5) per instance, n log-uniform in [1024, 16384], leaf 64;
* ``log_kernel_instances``: log-kernel matrices with n log-uniform in [4096, 65536], leaf 128.
"""
from __future__ import annotations

import numpy as np

from .compression import KernelMatrix
from .hssmatrix import bisection_cluster


def _compress(hip, K: KernelMatrix, leaf: int, atol: float, rtol: float, ctx=None):
    import torch
    n = K.shape[0]
    T = K.device_dense()
    cl = bisection_cluster(n, leaf)
    A = hip.compress_dev(T.data_ptr(), n, cl, cl, atol, rtol, ctx)
    del T
    torch.cuda.empty_cache()
    return A


def log_points(n: int, seed: int) -> np.ndarray:
    return np.sort(np.random.default_rng(seed).uniform(-1, 1, n))


def log_kernel(hip, n: int = 65536, leaf: int = 128, tol: float = 1e-9, seed: int = 0, ctx=None):
    """The n-point log-kernel matrix (unit diagonal), compressed at atol = rtol = tol."""
    x = log_points(n, seed)
    return _compress(hip, KernelMatrix(x, "log", 1.0), leaf, tol, tol, ctx)


def instance_sizes(count: int, lo: int, hi: int, rng) -> np.ndarray:
    """count sizes log-uniform in [lo, hi]."""
    return np.round(np.exp(rng.uniform(np.log(lo), np.log(hi), count))).astype(int)


def cauchy_instances(hip, count: int = 256, seed: int = 3, lo: int = 1024, hi: int = 16384, leaf: int = 64,
                     rtol: float = 1e-8, atol: float = 1e-9, ctx=None, progress=None):
    """count Cauchy-kernel matrices 1/(x_i - x_j), unit diagonal, x = sort(U(-1, 1)) + s,
    s ~ U(0, 0.5) per instance, n log-uniform in [lo, hi], compressed at rtol (atol: the
    reference's default 1e-9, src/HssMatrices.jl:62).  Returns (matrices, sizes, shifts,
    right-hand-side counts ~ U{1..32})."""
    rng = np.random.default_rng(seed)
    ns = instance_sizes(count, lo, hi, rng)
    shifts = rng.uniform(0.0, 0.5, count)
    rhs = rng.integers(1, 33, count)
    mats = []
    for i, (n, s) in enumerate(zip(ns, shifts)):
        x = np.sort(rng.uniform(-1, 1, int(n))) + s
        mats.append(_compress(hip, KernelMatrix(x, "cauchy", 1.0), leaf, atol, rtol, ctx))
        if progress is not None:
            progress(i + 1, count)
    return mats, ns, shifts, rhs


def log_kernel_instances(hip, count: int = 32, seed: int = 4, lo: int = 4096, hi: int = 65536, leaf: int = 128,
                         tol: float = 1e-9, ctx=None, progress=None):
    """count log-kernel matrices (unit diagonal) on sorted uniform points, n log-uniform in
    [lo, hi], compressed at atol = rtol = tol.  Returns (matrices, sizes)."""
    rng = np.random.default_rng(seed)
    ns = instance_sizes(count, lo, hi, rng)
    mats = []
    for i, n in enumerate(ns):
        x = np.sort(rng.uniform(-1, 1, int(n)))
        mats.append(_compress(hip, KernelMatrix(x, "log", 1.0), leaf, tol, tol, ctx))
        if progress is not None:
            progress(i + 1, count)
    return mats, ns
