"""Direct HSS compression on the device (hss_compress_dev) against the reference's own
compression test and the host compressor.

Compression parity with the reference is pinned by tolerance only: its rank-revealing QR
(LowRankApprox.pqrfact) is a third-party routine, and its test checks
||A - full(hssA)|| / ||A|| <= 50 rtol (test/runtests.jl:36-38).  The device compressor
is held to the same test, to the exact answers a rank-3 and a block-diagonal matrix have
(ranks 3 and 0), to the ranks the host compressor finds (within a few: both truncate the
same block rows at the same tolerance), and its generators go through the matvec and ULV
solve against the oracle like any others.
"""
import numpy as np
import pytest

from conftest import H

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def compress_dense(hip, A, rcl, ccl=None, tol=1e-6):
    import torch
    T = torch.from_numpy(np.ascontiguousarray(np.asarray(A, dtype=np.float64).T)).to("cuda:0")
    return hip.compress_dev(T.data_ptr(), A.shape[0], rcl, ccl, tol, tol)


def proportional_tree(rcl, start, stop):
    """A column tree over [start, stop) compatible with rcl (same branching), its splits in
    the proportion of rcl's."""
    node = H.ClusterTree(start, stop)
    if rcl.isbranch():
        cut = start + int(round((stop - start) * rcl.left.size / rcl.size))
        cut = min(max(cut, start + 1), stop - 1)
        node.left = proportional_tree(rcl.left, start, cut)
        node.right = proportional_tree(rcl.right, cut, stop)
    return node


def ranks(nodes):
    return nodes["ku"][1:], nodes["kv"][1:]


def test_compress_reference_tolerance(hip):
    """test/runtests.jl:19, 36-38 on the device compressor."""
    x = np.arange(-1, 1.0005, 0.001)[:2001]
    A = H.KernelMatrix(x, "cauchy_test").dense()
    cl = H.bisection_cluster(2001, 64)
    D = compress_dense(hip, A, cl, tol=1e-6)
    nodes, blob = D.generators()
    F = H.full(H.unpack(nodes, blob))
    err = np.linalg.norm(A - F)
    assert err / np.linalg.norm(A) <= 50 * 1e-6 or err <= 50 * 1e-6
    # the device matrix is the one these generators describe
    X = np.random.default_rng(5).standard_normal((2001, 3))
    assert rel(D @ X, F @ X) < 1e-13


def test_compress_c1_against_host_compressor(hip, oracle):
    """C1 (Cauchy, n = 2001, leaf 64, 1e-6): accuracy, ranks next to the host compressor's,
    and the ULV solve of the device-compressed generators against the oracle."""
    x = -1 + 0.001 * np.arange(2001)
    K = H.KernelMatrix(x, "cauchy", 1.0)
    A = K.dense()
    cl = H.bisection_cluster(2001, 64)
    D = hip.compress_dev(K.device_dense().data_ptr(), 2001, cl, cl, 1e-6, 1e-6)
    nodes, blob = D.generators()
    assert rel(H.full(H.unpack(nodes, blob)), A) <= 50 * 1e-6
    hn, _ = H.pack(H.compress(K, cl, cl, 1e-6, 1e-6))
    for got, host in zip(ranks(nodes), ranks(hn)):
        print("device ranks", got.tolist(), "host", host.tolist())
        assert np.max(np.abs(got.astype(int) - host.astype(int))) <= 3
    B = np.random.default_rng(6).standard_normal((2001, 2))
    Z = D.ulvfactor("fp64").solve(B)
    assert rel(Z, oracle.ulvfactsolve(nodes, blob, B)) < 1e-12


def test_compress_log_kernel_tight_tolerance(hip, oracle):
    """log|x_i - x_j| (unit diagonal), n = 4096, leaf 128, rtol = atol = 1e-9."""
    rng = np.random.default_rng(7)
    x = np.sort(rng.uniform(-1, 1, 4096))
    K = H.KernelMatrix(x, "log", 1.0)
    cl = H.bisection_cluster(4096, 128)
    D = hip.compress_dev(K.device_dense().data_ptr(), 4096, cl, cl, 1e-9, 1e-9)
    nodes, blob = D.generators()
    A = K.dense()
    assert rel(H.full(H.unpack(nodes, blob)), A) <= 50 * 1e-9
    X = rng.standard_normal((4096, 4))
    assert rel(D @ X, oracle.matmul(nodes, blob, X)) < 1e-14
    B = rng.standard_normal((4096, 3))
    Z = D.ulvfactor("fp64").solve(B)
    Zo = oracle.ulvfactsolve(nodes, blob, B)
    assert np.max(np.linalg.norm(A @ Z - B, axis=0) / np.linalg.norm(B, axis=0)) <= 50 * 1e-9 * np.linalg.cond(A)
    assert rel(Z, Zo) < 1e-10


def test_compress_uneven_rectangular(hip):
    """Leaves at two depths (n = 1025, leaf 64) and a rectangular matrix (1025 x 901) over
    compatible row and column trees (binarytree.jl:68-74)."""
    rng = np.random.default_rng(8)
    rcl = H.bisection_cluster(1025, 64)
    ccl = proportional_tree(rcl, 0, 901)
    assert H.compatible(rcl, ccl)
    xr = np.sort(rng.uniform(-1, 1, 1025))
    xc = np.sort(rng.uniform(-1, 1, 901)) + 2.5e-4
    A = 1.0 / (xr[:, None] - xc[None, :])
    D = compress_dense(hip, A, rcl, ccl, tol=1e-8)
    nodes, blob = D.generators()
    assert D.shape == (1025, 901)
    assert len(set(nodes["depth"][nodes["left"] < 0].tolist())) == 2
    assert rel(H.full(H.unpack(nodes, blob)), A) <= 50 * 1e-8
    X = rng.standard_normal((901, 3))
    assert rel(D @ X, A @ X) <= 50 * 1e-8


def test_compress_exact_ranks(hip):
    """A rank-3 matrix U V' compresses to ranks 3 (runtests.jl:44-46 holds lowrank2hss to
    hssrank 3), and a block-diagonal one to ranks 0 with its leaves' blocks exactly."""
    rng = np.random.default_rng(9)
    n = 1000
    cl = H.bisection_cluster(n, 64)
    U, V = rng.standard_normal((n, 3)), rng.standard_normal((n, 3))
    A = U @ V.T
    D = compress_dense(hip, A, cl, tol=1e-9)
    nodes, blob = D.generators()
    ku, kv = ranks(nodes)
    assert ku.max() == 3 and kv.max() == 3 and ku.min() == 3 and kv.min() == 3
    assert rel(H.full(H.unpack(nodes, blob)), A) <= 1e-12
    Bd = np.zeros((n, n))
    leaves = [i for i in range(len(nodes)) if nodes["left"][i] < 0]
    for i in leaves:
        r0, m = int(nodes["row_off"][i]), int(nodes["m"][i])
        Bd[r0:r0 + m, r0:r0 + m] = rng.standard_normal((m, m))
    D = compress_dense(hip, Bd, cl, tol=1e-9)
    nodes, blob = D.generators()
    ku, kv = ranks(nodes)
    assert ku.max() == 0 and kv.max() == 0
    assert np.array_equal(H.full(H.unpack(nodes, blob)), Bd)


def test_compress_errors(hip):
    import ctypes
    cl = H.bisection_cluster(300, 64)
    A = np.eye(300)
    import torch
    T = torch.from_numpy(A).to("cuda:0")
    torch.cuda.synchronize()
    lib = hip.load_library()
    nodes = H.cluster_nodes(cl)
    h = ctypes.c_void_p()
    ctx = hip.default_context()
    st = lib.hss_compress_dev(ctx.handle, T.data_ptr(), 200, nodes.ctypes.data, len(nodes), 1e-9, 1e-9,
                              ctypes.byref(h))
    assert st == hip.HSS_ERR_DIM
    bad = nodes.copy()
    bad["m"][1] += 1  # children no longer add up to the parent
    st = lib.hss_compress_dev(ctx.handle, T.data_ptr(), 300, bad.ctypes.data, len(bad), 1e-9, 1e-9,
                              ctypes.byref(h))
    assert st in (hip.HSS_ERR_DIM, hip.HSS_ERR_ARG)
    st = lib.hss_compress_dev(ctx.handle, T.data_ptr(), 300, nodes.ctypes.data, len(nodes), -1.0, 1e-9,
                              ctypes.byref(h))
    assert st == hip.HSS_ERR_ARG
