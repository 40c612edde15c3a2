"""Recompression on the device (hss_mat_recompress), the device counterpart of
recompress!(hssA; atol, rtol) (reference src/compression.jl:148-260).

Pinned by tolerance and by exact answers: the reference's own test recompresses a
randomized compression and requires the same accuracy at no larger rank
(test/runtests.jl:39-43); an HSS product of rank-3 matrices, whose ranks add to 6, must come
back to rank 3; at a tolerance near roundoff the recompressed matrix is the same operator;
a leaf-only matrix has nothing to recompress.
"""
import numpy as np
import pytest

from conftest import H

pytestmark = pytest.mark.gpu
EPS = 2.220446049250313e-16


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def dev(A):
    import torch
    T = torch.from_numpy(np.ascontiguousarray(np.asarray(A, dtype=np.float64).T)).to("cuda:0")
    torch.cuda.synchronize()
    return T


def hssrank(nodes):
    return int(max(nodes["ku"][1:].max(), nodes["kv"][1:].max())) if len(nodes) > 1 else 0


def test_recompress_reference(hip, oracle):
    """test/runtests.jl:39-43: randcompress_adaptive, then recompress!, accuracy within 50 rtol
    and hssrank not above the randomized compression's."""
    x = np.arange(-1, 1.0005, 0.001)[:2001]
    A = H.KernelMatrix(x, "cauchy_test").dense()
    cl = H.bisection_cluster(2001, 64)
    T = dev(A)
    B = hip.randcompress_dev(T.data_ptr(), 2001, cl, cl, 1e-6, 1e-6)
    rk = hssrank(B.generators()[0])
    C = B.recompress(1e-6, 1e-6)
    nodes, blob = C.generators()
    F = H.full(H.unpack(nodes, blob))
    err = np.linalg.norm(A - F)
    print(f"recompress C1: rel err {err / np.linalg.norm(A):.3e}, hssrank {rk} -> {hssrank(nodes)}")
    assert err / np.linalg.norm(A) <= 50 * 1e-6 or err <= 50 * 1e-6
    assert hssrank(nodes) <= rk
    Xv = np.random.default_rng(3).standard_normal((2001, 3))
    assert rel(C @ Xv, F @ Xv) < 1e-13
    Bv = np.random.default_rng(4).standard_normal((2001, 2))
    Z = C.ulvfactor("fp64").solve(Bv)
    assert rel(Z, oracle.ulvfactsolve(nodes, blob, Bv)) < 1e-12


def test_recompress_product_exact_rank(hip):
    """U V' is rank 3; its HSS square (hss_mat_mul) has ranks 6 = 3 + 3 and is rank 3 again
    after recompression, the same operator to roundoff."""
    rng = np.random.default_rng(9)
    n = 1000
    cl = H.bisection_cluster(n, 64)
    Uu, Vv = rng.standard_normal((n, 3)), rng.standard_normal((n, 3)) / np.sqrt(n)
    A = Uu @ Vv.T + np.eye(n)
    D = hip.compress_dev(dev(A).data_ptr(), n, cl, cl, 1e-12, 1e-12)
    assert hssrank(D.generators()[0]) == 3
    P = D @ D
    assert hssrank(P.nodes) == 6
    C = P.recompress(1e-12, 1e-12)
    nodes, blob = C.generators()
    ku, kv = nodes["ku"][1:], nodes["kv"][1:]
    assert ku.max() == 3 and kv.max() == 3
    X = rng.standard_normal((n, 4))
    assert rel(C @ X, A @ (A @ X)) < 1e-12
    assert rel(C @ X, P @ X) < 1e-12


@pytest.mark.parametrize("name", ["cauchy_n120_leaf16", "uneven_n97_leaf10", "pruned_n128", "passthru_n60_leaf8",
                                  "blockdiag_rank0_n48", "leafonly_n20", "tiny_n12_leaf3"])
def test_recompress_golden_near_roundoff(hip, name):
    """At atol = rtol = 1e-15 the recompressed matrix is the same operator (uneven and pruned
    trees, pass-through nodes, rank 0, a leaf-only matrix), and no rank grows."""
    import pathlib
    G = np.load(pathlib.Path(__file__).with_name("golden") / "hss_golden.npz")
    nodes, blob = G[name + "/nodes"], G[name + "/blob"]
    A = hip.HipHssMatrix((nodes, blob))
    C = A.recompress(1e-15, 1e-15)
    cn, cb = C.generators()
    F0 = H.full(H.unpack(nodes, blob))
    F1 = H.full(H.unpack(cn, cb))
    assert rel(F1, F0) < 1e-13
    assert np.all(cn["ku"][1:] <= nodes["ku"][1:]) and np.all(cn["kv"][1:] <= nodes["kv"][1:])


def test_recompress_log_kernel_product(hip):
    """The product of an n = 4096 log-kernel matrix with itself (ranks doubled), recompressed at
    1e-9: the ranks drop, and the product's action is kept to within the truncation
    (relative to ||A A x||)."""
    rng = np.random.default_rng(12)
    x = np.sort(rng.uniform(-1, 1, 4096))
    K = H.KernelMatrix(x, "log", 1.0)
    cl = H.bisection_cluster(4096, 128)
    D = hip.compress_dev(K.device_dense().data_ptr(), 4096, cl, cl, 1e-9, 1e-9)
    r1 = hssrank(D.generators()[0])
    P = D @ D
    C = P.recompress(1e-9, 1e-9)
    nodes, _ = C.generators()
    r2 = hssrank(nodes)
    X = rng.standard_normal((4096, 3))
    e = rel(C @ X, P @ X)
    print(f"recompress log4096 A A: hssrank {hssrank(P.nodes)} -> {r2} (A alone {r1}), rel diff {e:.3e}")
    assert r2 < hssrank(P.nodes)
    assert e <= 1e-7
