"""Randomized adaptive HSS compression on the device (hss_randcompress_dev), the device
counterpart of randcompress_adaptive (reference src/compression.jl:312-430).

Like direct compression, it is pinned by tolerance: the reference draws its test matrices
from Julia's global RNG and checks ||A - full(hssB)|| / ||A|| <= 50 rtol
(test/runtests.jl:39-40).  The device compressor is held to that test, to the exact ranks of
a rank-3 matrix, to the adaptive loop growing the sample when the first one is too small for
the off-diagonal ranks, to rectangular and uneven trees, and its generators go through the
matvec and the ULV solve against the oracle like any others.
"""
import ctypes

import numpy as np
import pytest

from conftest import H
from test_gpu_compress import proportional_tree, ranks

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def dev(A):
    import torch
    T = torch.from_numpy(np.ascontiguousarray(np.asarray(A, dtype=np.float64).T)).to("cuda:0")
    torch.cuda.synchronize()
    return T


def test_randcompress_reference_tolerance(hip, oracle):
    """test/runtests.jl:19, 39-40 on the device: the reference's Cauchy test matrix, leaf 64,
    atol = rtol = 1e-6, the reference's default kest / noversampling / stepsize."""
    x = np.arange(-1, 1.0005, 0.001)[:2001]
    A = H.KernelMatrix(x, "cauchy_test").dense()
    cl = H.bisection_cluster(2001, 64)
    T = dev(A)
    D, smp = hip.randcompress_dev(T.data_ptr(), 2001, cl, cl, 1e-6, 1e-6, report=True)
    nodes, blob = D.generators()
    F = H.full(H.unpack(nodes, blob))
    err = np.linalg.norm(A - F)
    ku, kv = ranks(nodes)
    print(f"randcompress C1: rel err {err / np.linalg.norm(A):.3e}, samples {smp}, ranks {ku.max()} / {kv.max()}")
    assert err / np.linalg.norm(A) <= 50 * 1e-6 or err <= 50 * 1e-6
    assert smp >= 20 and max(ku.max(), kv.max()) <= smp
    X = np.random.default_rng(5).standard_normal((2001, 3))
    assert rel(D @ X, F @ X) < 1e-13
    B = np.random.default_rng(6).standard_normal((2001, 2))
    Z = D.ulvfactor("fp64").solve(B)
    assert rel(Z, oracle.ulvfactsolve(nodes, blob, B)) < 1e-12


def test_randcompress_adaptive_growth(hip, oracle):
    """log|x_i - x_j| (unit diagonal), n = 4096, leaf 128, 1e-9, from a sample of 2 + 2 columns
    growing by 5: the loop must enlarge the sample past the off-diagonal ranks (ranks are at
    most the sample width) and end within the tolerance."""
    rng = np.random.default_rng(7)
    x = np.sort(rng.uniform(-1, 1, 4096))
    K = H.KernelMatrix(x, "log", 1.0)
    cl = H.bisection_cluster(4096, 128)
    D, smp = hip.randcompress_dev(K.device_dense().data_ptr(), 4096, cl, cl, 1e-9, 1e-9, kest=2, noversampling=2,
                                  stepsize=5, report=True)
    nodes, blob = D.generators()
    ku, kv = ranks(nodes)
    A = K.dense()
    e = rel(H.full(H.unpack(nodes, blob)), A)
    print(f"randcompress log4096: rel err {e:.3e}, samples {smp}, ranks {ku.max()} / {kv.max()}")
    assert smp > 4 and (smp - 4) % 5 == 0
    assert max(ku.max(), kv.max()) <= smp
    assert e <= 50 * 1e-9
    Xv = rng.standard_normal((4096, 4))
    assert rel(D @ Xv, oracle.matmul(nodes, blob, Xv)) < 1e-14
    B = rng.standard_normal((4096, 3))
    Z = D.ulvfactor("fp64").solve(B)
    Zo = oracle.ulvfactsolve(nodes, blob, B)
    assert rel(Z, Zo) < 1e-10
    assert np.max(np.linalg.norm(A @ Z - B, axis=0) / np.linalg.norm(B, axis=0)) <= 50 * 1e-9 * np.linalg.cond(A)


def test_randcompress_exact_ranks(hip):
    """A rank-3 matrix U V' compresses to ranks 3 exactly (its samples have rank 3), and a
    block-diagonal one to its leaves' blocks with couplings at roundoff."""
    rng = np.random.default_rng(9)
    n = 1000
    cl = H.bisection_cluster(n, 64)
    U, V = rng.standard_normal((n, 3)), rng.standard_normal((n, 3))
    A = U @ V.T
    D = hip.randcompress_dev(dev(A).data_ptr(), n, cl, cl, 1e-9, 1e-9)
    nodes, blob = D.generators()
    ku, kv = ranks(nodes)
    assert ku.max() == 3 and kv.max() == 3 and ku.min() == 3 and kv.min() == 3
    assert rel(H.full(H.unpack(nodes, blob)), A) <= 1e-12
    Bd = np.zeros((n, n))
    for i in range(len(nodes)):
        if nodes["left"][i] < 0:
            r0, m = int(nodes["row_off"][i]), int(nodes["m"][i])
            Bd[r0:r0 + m, r0:r0 + m] = rng.standard_normal((m, m))
    D = hip.randcompress_dev(dev(Bd).data_ptr(), n, cl, cl, 1e-9, 1e-9)
    nodes, blob = D.generators()
    assert rel(H.full(H.unpack(nodes, blob)), Bd) <= 1e-12


def test_randcompress_uneven_rectangular(hip):
    """Leaves at two depths and a 1025 x 901 Cauchy matrix over compatible trees."""
    rng = np.random.default_rng(8)
    rcl = H.bisection_cluster(1025, 64)
    ccl = proportional_tree(rcl, 0, 901)
    xr = np.sort(rng.uniform(-1, 1, 1025))
    xc = np.sort(rng.uniform(-1, 1, 901)) + 2.5e-4
    A = 1.0 / (xr[:, None] - xc[None, :])
    D = hip.randcompress_dev(dev(A).data_ptr(), 1025, rcl, ccl, 1e-8, 1e-8)
    nodes, blob = D.generators()
    assert D.shape == (1025, 901)
    assert rel(H.full(H.unpack(nodes, blob)), A) <= 50 * 1e-8
    X = rng.standard_normal((901, 3))
    assert rel(D @ X, A @ X) <= 50 * 1e-8


def test_randcompress_errors(hip):
    cl = H.bisection_cluster(300, 64)
    T = dev(np.eye(300))
    lib = hip.load_library()
    nodes = H.cluster_nodes(cl)
    h = ctypes.c_void_p()
    ctx = hip.default_context()
    smp = ctypes.c_int32()
    args = lambda lda, atol, kest, r, bs: lib.hss_randcompress_dev(
        ctx.handle, T.data_ptr(), lda, nodes.ctypes.data, len(nodes), atol, 1e-9, kest, r, bs, ctypes.byref(h),
        ctypes.byref(smp))
    assert args(200, 1e-9, 10, 10, 20) == hip.HSS_ERR_DIM
    assert args(300, -1.0, 10, 10, 20) == hip.HSS_ERR_ARG
    assert args(300, 1e-9, 10, 10, 0) == hip.HSS_ERR_ARG
    assert args(300, 1e-9, -1, 10, 20) == hip.HSS_ERR_ARG
    assert not h.value
    # the identity is block diagonal: ranks may only hold roundoff, and the result is exact
    assert args(300, 1e-9, 10, 10, 20) == hip.HSS_OK
    lib.hss_mat_destroy(h)
