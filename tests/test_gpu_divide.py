"""HSS right-hand sides on the device: hssA \\ hssB and hssA / hssB (hss_mat_ldiv,
hss_mat_rdiv; reference \\(hssA, hssB) and /(hssA, hssB), src/hssmatrix.jl:235-237, computed
there by the ULV divide of src/ulvdivide.jl).

The reference's test compares full(hssA \\ hssI) and full(hssI / hssA) with inv(A) at 50 rtol,
hssI an HSS identity (test/runtests.jl:69-77).  The same test runs here, with the divide's own
error also measured against inv(full(hssA)) (what the compression of A leaves aside), plus a
general HSS right-hand side and a tree mismatch.
"""
import numpy as np
import pytest

from conftest import H

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def dev(A):
    import torch
    T = torch.from_numpy(np.ascontiguousarray(np.asarray(A, dtype=np.float64).T)).to("cuda:0")
    torch.cuda.synchronize()
    return T


def full(M):
    nodes, blob = M.generators()
    return H.full(H.unpack(nodes, blob))


@pytest.fixture(scope="module")
def c1_pair(hip):
    x = np.arange(-1, 1.0005, 0.001)[:2001]
    A = H.KernelMatrix(x, "cauchy_test").dense()
    cl = H.bisection_cluster(2001, 64)
    hA = hip.compress_dev(dev(A).data_ptr(), 2001, cl, cl, 1e-6, 1e-6)
    hI = hip.compress_dev(dev(np.eye(2001)).data_ptr(), 2001, cl, cl, 1e-6, 1e-6)
    return A, hA, hI, cl


def test_ldiv_rdiv_identity_reference(c1_pair):
    """runtests.jl:72-74 compares full(hssA \\ hssI) and full(hssI / hssA) with inv(A) at 50 rtol.
    That comparison holds the compression of A as well as the divide: with cond(A) = 8.3e3 and
    A compressed at 1e-6 (relative error ~1.4e-6), inv(full(hssA)) is itself ~4e-5 from inv(A)
    (3.6e-5 with the host compressor), near the 5e-5 bar.  So the divide is held to 50 rtol
    against inv(full(hssA)), its own input, and the distance to inv(A) to its perturbation
    bound cond(A) ||A - full(hssA)|| / ||A|| plus that."""
    A, hA, hI, _ = c1_pair
    Ainv = np.linalg.inv(A)
    Fa = full(hA)
    Finv = np.linalg.inv(Fa)
    L = hA.hss_ldiv(hI, 1e-6, 1e-6)
    Rr = hI.hss_rdiv(hA, 1e-6, 1e-6)
    FL, FR = full(L), full(Rr)
    ca = rel(Fa, A)
    kappa = np.linalg.cond(A)
    print(f"A\\I vs inv(A) {rel(FL, Ainv):.3e}, vs inv(full(hssA)) {rel(FL, Finv):.3e}; "
          f"I/A vs inv(A) {rel(FR, Ainv):.3e}, vs inv(full(hssA)) {rel(FR, Finv):.3e}; "
          f"compression of A {ca:.3e}, cond {kappa:.3e}")
    assert rel(FL, Finv) <= 50 * 1e-6
    assert rel(FR, Finv) <= 50 * 1e-6
    assert rel(FL, Ainv) <= kappa * ca + 50 * 1e-6
    assert rel(FR, Ainv) <= kappa * ca + 50 * 1e-6
    X = np.random.default_rng(2).standard_normal((2001, 3))
    assert rel(L @ X, FL @ X) < 1e-13


def test_ldiv_general_rhs(hip, c1_pair):
    """hssA \\ hssB for a compressed shifted Cauchy matrix B, and hssB / hssA, against the
    dense solves of the compressed operators."""
    A, hA, _, cl = c1_pair
    x = np.arange(-1, 1.0005, 0.001)[:2001]
    Bd = 1.0 / (x[:, None] - x[None, :] + 0.0123)
    hB = hip.compress_dev(dev(Bd).data_ptr(), 2001, cl, cl, 1e-9, 1e-9)
    Fa, Fb = full(hA), full(hB)
    C = hA.hss_ldiv(hB, 1e-9, 1e-9)
    Cref = np.linalg.solve(Fa, Fb)
    D = hB.hss_rdiv(hA, 1e-9, 1e-9)
    Dref = np.linalg.solve(Fa.T, Fb.T).T
    eC, eD = rel(full(C), Cref), rel(full(D), Dref)
    print(f"A\\B {eC:.3e}, B/A {eD:.3e}")
    assert eC <= 50 * 1e-9 * np.linalg.cond(Fa) and eD <= 50 * 1e-9 * np.linalg.cond(Fa)
    assert eC <= 1e-6 and eD <= 1e-6


def test_ldiv_tree_mismatch(hip, c1_pair):
    _, hA, _, _ = c1_pair
    cl2 = H.bisection_cluster(2001, 128)
    hJ = hip.compress_dev(dev(np.eye(2001)).data_ptr(), 2001, cl2, cl2, 1e-6, 1e-6)
    from hssmatrices_jl_amd.hssmatrix import DimensionMismatch
    with pytest.raises(DimensionMismatch):
        hA.hss_ldiv(hJ)
