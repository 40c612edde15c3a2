"""Float32 at the boundary (hss_mat_create_f32, hss_matmul_f32, hss_ulvfactsolve_f32,
hss_ulv_solve_f32): the reference runs its hot path for T = Float32 too
(test/runtests.jl:16).  Float32 generators and data are widened to Float64 on the device
and the results rounded once to Float32, so each result equals the oracle's Float64 result
on the widened inputs to Float32 rounding (<= 2^-24 per entry), and the reference's own
Float32 tolerances (50 rtol on C1, runtests.jl:63-67) hold."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def test_f32_matvec_and_solves(hip, oracle, c1):
    nodes, blob = c1["nodes"], c1["blob"]
    b32 = blob.astype(np.float32)
    b64 = b32.astype(np.float64)  # the generators the device computes with
    A = hip.HipHssMatrix((nodes, b32))
    rng = np.random.default_rng(30)
    X = rng.standard_normal((2001, 5)).astype(np.float32)
    Y = A @ X
    assert Y.dtype == np.float32
    Yo = oracle.matmul(nodes, b64, X.astype(np.float64))
    assert rel(Y.astype(np.float64), Yo) < 2e-7
    y1 = A @ X[:, 0]
    assert y1.dtype == np.float32 and y1.shape == (2001,)
    B = rng.standard_normal((2001, 3)).astype(np.float32)
    Z = A.solve(B)
    Zo = oracle.ulvfactsolve(nodes, b64, B.astype(np.float64))
    assert Z.dtype == np.float32 and rel(Z.astype(np.float64), Zo) < 2e-7
    F = A.ulvfactor("fp64")
    Z2 = F.solve(B)
    assert Z2.dtype == np.float32 and np.array_equal(Z2, Z)
    # the reference's Float32 tolerances against the dense matrix
    Ad = c1["A"]
    assert rel((A @ X).astype(np.float64), Ad @ X.astype(np.float64)) <= 50 * 1e-6
    assert rel(Z.astype(np.float64), np.linalg.solve(Ad, B.astype(np.float64))) <= 50 * 1e-6


def test_f32_errors(hip, c1):
    from hssmatrices_jl_amd.hssmatrix import DimensionMismatch
    A = hip.HipHssMatrix((c1["nodes"], c1["blob"].astype(np.float32)))
    with pytest.raises(DimensionMismatch):
        A @ np.zeros((7, 2), dtype=np.float32)
    with pytest.raises(DimensionMismatch):
        A.solve(np.zeros((7, 1), dtype=np.float32))
