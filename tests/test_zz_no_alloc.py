"""The allocation-free solve (hss_ulv_solve_dev_noalloc) against hss_ulv_solve_dev.

Both run the same sweep kernels on the same Float64 factor; the allocation-free one keeps
its workspace in the library's fixed device array and splits the columns into chunks that
fit it.  Each column of the result must equal the allocating entry point's to 1e-13
relative (the kernels' column tiles may differ between the two calls, the per-column
arithmetic does not change the result beyond rounding), and the solve must match the
oracle as every other Float64 solve does.  Two calls in a row, the second with a smaller
input, because every call shares the one array and it keeps the previous call's contents.
"""
import ctypes

import numpy as np
import pytest

from conftest import H

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _dev(torch, A):
    """Column-major n x r host array -> device tensor (r, n) and a zero tensor alike."""
    t = torch.from_numpy(np.ascontiguousarray(A.T)).to("cuda:0")
    return t, torch.zeros_like(t)


def _solve_both(hip, torch, F, B):
    n, r = B.shape
    Bt, X1 = _dev(torch, B)
    X2 = torch.zeros_like(X1)
    X2.fill_(np.nan)  # the allocation-free result must overwrite every entry
    torch.cuda.synchronize()
    lib = hip.load_library()
    none = hip.RefineOpts(0.0, 0, 0)
    st = lib.hss_ulv_solve_dev(F.handle, r, Bt.data_ptr(), n, X1.data_ptr(), n, ctypes.byref(none), None)
    assert st == 0, lib.hss_last_error()
    st = lib.hss_ulv_solve_dev_noalloc(F.handle, r, Bt.data_ptr(), n, X2.data_ptr(), n)
    assert st == 0, lib.hss_last_error()
    torch.cuda.synchronize()
    return X1.cpu().numpy().T, X2.cpu().numpy().T


def test_noalloc_matches_allocating_solve(hip, oracle):
    import torch
    rng = np.random.default_rng(70)
    # first call: n = 32768, leaf 128, 200 columns (two chunks of the 64 MiB array);
    # second call: a smaller matrix (n = 3001) with fewer columns
    cases = [(32768, 128, 200), (3001, 64, 7)]
    for n, leaf, r in cases:
        print(f"noalloc case n={n} leaf={leaf} nrhs={r}", flush=True)
        h = H.random_hss(n, leaf, (10, 40), rng, diag=2.0)
        nodes, blob = H.pack(h)
        A = hip.HipHssMatrix((nodes, blob))
        F = A.ulvfactor("fp64")
        B = rng.standard_normal((n, r))
        X1, X2 = _solve_both(hip, torch, F, B)
        assert np.all(np.isfinite(X2))
        assert rel(X2, X1) < 1e-13
        for j in range(r):
            assert rel(X2[:, j], X1[:, j]) < 1e-13
        if n <= 4096:
            assert rel(X2, oracle.ulvfactsolve(nodes, blob, B)) < 1e-11
        R = (A @ X2) - B
        assert np.max(np.linalg.norm(R, axis=0) / np.linalg.norm(B, axis=0)) < 1e-12
        F.close()
        A.close()


def test_noalloc_refusals(hip):
    import torch
    rng = np.random.default_rng(71)
    h = H.random_hss(1500, 64, (10, 30), rng, diag=2.0)
    A = hip.HipHssMatrix(H.pack(h))
    F32 = A.ulvfactor("fp32")
    B = torch.zeros((2, 1500), dtype=torch.float64, device="cuda:0")
    X = torch.zeros_like(B)
    torch.cuda.synchronize()
    lib = hip.load_library()
    st = lib.hss_ulv_solve_dev_noalloc(F32.handle, 2, B.data_ptr(), 1500, X.data_ptr(), 1500)
    assert st == hip.HSS_ERR_UNSUPPORTED
    assert b"Float64" in lib.hss_last_error()
    F64 = A.ulvfactor("fp64")
    st = lib.hss_ulv_solve_dev_noalloc(F64.handle, 2, B.data_ptr(), 1000, X.data_ptr(), 1500)
    assert st == hip.HSS_ERR_DIM
    F32.close()
    F64.close()
    A.close()
