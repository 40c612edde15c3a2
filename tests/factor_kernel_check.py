"""Runs the ULV factor with the factor kernel named by HSS_FACTOR_KERNEL (vector or mfma,
read once when the library loads, hence a process of its own) on the golden fixtures, on
the n = 2001 Cauchy matrix and on an n = 4096 log-kernel matrix, in Float64 and Float32,
and checks the solves against the CPU oracle.  Called by test_gpu_factor_kernels.py;
prints one JSON line and exits non-zero on a failed check."""
import json
import pathlib
import sys

import numpy as np

ROOT = pathlib.Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "tests"))

from conftest import H, cauchy_c1, random_hss  # noqa: E402
import importlib  # noqa: E402

hip = importlib.import_module("hssmatrices_jl_amd.hip")
from oracle import oracle as O  # noqa: E402

EPS = 2.220446049250313e-16
GOLD = np.load(ROOT / "tests" / "golden" / "hss_golden.npz")


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def factor_kernels(A, prec):
    """the factor kernels that ran (from the launch records)"""
    hip.prof_reset()
    hip.prof_enable(True)
    F = A.ulvfactor(prec)
    recs = hip.prof_records()
    hip.prof_enable(False)
    hip.prof_reset()
    return F, sorted({r["kernel"] for r in recs if r["kernel"].startswith("k_ulv_factor")})


cases = [(str(nm), GOLD[str(nm) + "/nodes"], GOLD[str(nm) + "/blob"], float(GOLD[str(nm) + "/cond"]))
         for nm in GOLD["names"]]
_, h = cauchy_c1()
nodes, blob = H.pack(h)
cases.append(("c1", nodes, blob, None))
_, h = random_hss(np.random.default_rng(14), 4096, 128, 0, "log", 1e-9)
nodes, blob = H.pack(h)
# cond(A) = 6.66e5 (2-norm, from the SVD of the dense matrix)
cases.append(("log4096", nodes, blob, 6.66e5))

out = {"kernels": {}, "fp64_rel": {}, "fp32_residual": {}}
ok = True
rng = np.random.default_rng(5)
for name, nodes, blob, cond in cases:
    A = hip.HipHssMatrix((nodes, blob))
    n = A.shape[0]
    B = rng.standard_normal((n, 3))
    Xo = O.ulvfactsolve(nodes, blob, B)
    F, ks = factor_kernels(A, "fp64")
    out["kernels"][name] = ks
    X = F.solve(B)
    F.close()
    # two backward-stable solvers differ by at most ~cond * eps: the golden fixtures are held
    # to 50 cond eps as in test_gpu_parity, the log matrix to cond eps / 10 = 1.5e-11 (its
    # measured differences are about cond eps / 80), C1 (cond 3.1e3) to the 1e-12 parity bound
    if name == "c1":
        tol = 1e-12
    elif name == "log4096":
        tol = cond * EPS / 10
    else:
        tol = 50 * cond * EPS
    r = rel(X, Xo)
    out["fp64_rel"][name] = r
    ok &= r < tol
    F, ks32 = factor_kernels(A, "fp32")
    out["kernels"][name + "/fp32"] = ks32
    X, rep = F.solve(B, tol=1e-10, max_iter=30, report=True)
    F.close()
    R = O.matmul(nodes, blob, X) - B
    res = float(np.max(np.linalg.norm(R, axis=0) / np.linalg.norm(B, axis=0)))
    out["fp32_residual"][name] = res
    ok &= res <= 1e-10
out["ok"] = bool(ok)
print(json.dumps(out), flush=True)
sys.exit(0 if ok else 1)
