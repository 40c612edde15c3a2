"""Writes tests/golden/hss_golden.npz: small HSS matrices with expected
matvec and solve results from the closed form, in extended precision.

Expected values do NOT come from the oracle: the dense matrix is expanded from
the generators (reference src/hssmatrix.jl:284-305, ``_full!``) in numpy
longdouble, Y = A X is a longdouble product and Z = A \\ B a longdouble
Gaussian elimination with partial pivoting.  The reference's own tests hold no
golden vectors (test/runtests.jl is tolerance-only), so these fixtures pin
both the oracle and the GPU path to the mathematics.

Run:  python tests/golden/make_golden.py
"""
import pathlib
import sys

import numpy as np

ROOT = pathlib.Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
import __graft_entry__ as G  # noqa: E402

H = G.load_package()
LD = np.longdouble


def full_ld(h):
    """_full! (hssmatrix.jl:284-305) in longdouble."""
    def rec(h):
        # returns (A, U, V) of the subtree as longdouble arrays (U, V: tall generators)
        if h.isleaf():
            return h.D.astype(LD), h.U.astype(LD), h.V.astype(LD)
        A1, U1, V1 = rec(h.A11)
        A2, U2, V2 = rec(h.A22)
        B12, B21 = h.B12.astype(LD), h.B21.astype(LD)
        A = np.block([[A1, U1 @ B12 @ V2.T], [U2 @ B21 @ V1.T, A2]])
        U = np.vstack([U1 @ h.R1.astype(LD), U2 @ h.R2.astype(LD)])
        V = np.vstack([V1 @ h.W1.astype(LD), V2 @ h.W2.astype(LD)])
        return A, U, V
    return rec(h)[0]


def solve_ld(A, B):
    """Gaussian elimination with partial pivoting in longdouble."""
    A = A.astype(LD).copy()
    B = B.astype(LD).copy()
    n = A.shape[0]
    for i in range(n):
        p = i + int(np.argmax(np.abs(A[i:, i])))
        if A[p, i] == 0:
            raise ZeroDivisionError("singular")
        if p != i:
            A[[i, p]] = A[[p, i]]
            B[[i, p]] = B[[p, i]]
        f = A[i + 1:, i] / A[i, i]
        A[i + 1:, i:] -= np.outer(f, A[i, i:])
        B[i + 1:] -= np.outer(f, B[i])
    X = np.zeros_like(B)
    for i in range(n - 1, -1, -1):
        X[i] = (B[i] - A[i, i + 1:] @ X[i + 1:]) / A[i, i]
    return X


def kernel_hss(rng, n, leaf, kind, tol):
    x = np.sort(rng.uniform(-1, 1, n))
    K = H.KernelMatrix(x, kind, 1.0)
    cl = H.bisection_cluster(n, leaf)
    return H.compress(K, cl, cl, tol, tol)


def main():
    rng = np.random.default_rng(20240601)
    cases = {}
    cases["cauchy_n120_leaf16"] = kernel_hss(rng, 120, 16, "cauchy", 1e-10)
    cases["log_n200_leaf25"] = kernel_hss(rng, 200, 25, "log", 1e-9)
    cases["uneven_n97_leaf10"] = kernel_hss(rng, 97, 10, "cauchy", 1e-9)
    h = kernel_hss(rng, 128, 8, "log", 1e-9)
    h.A11 = H.prune_leaves(h.A11)  # unbalanced tree, as test/runtests.jl:75-76
    h.sz1 = h.A11.shape
    cases["pruned_n128"] = h
    cases["passthru_n60_leaf8"] = kernel_hss(rng, 60, 8, "cauchy", 1e-15)
    # rank-0 block diagonal: every coupling empty
    cl = H.bisection_cluster(48, 12)
    blocks = H.compress(rng.standard_normal((48, 48)) + 8 * np.eye(48), cl, cl, 1e-12, 1e-12)

    def zero_ranks(h):
        if h.isleaf():
            return H.HssMatrix(D=h.D, U=np.zeros((h.D.shape[0], 0)), V=np.zeros((h.D.shape[1], 0)),
                               rootnode=h.rootnode)
        a, b = zero_ranks(h.A11), zero_ranks(h.A22)
        z = np.zeros((0, 0))
        if h.rootnode:
            return H.HssMatrix(A11=a, A22=b, B12=z, B21=z, rootnode=True)
        return H.HssMatrix(A11=a, A22=b, B12=z, B21=z, R1=z, W1=z, R2=z, W2=z)

    cases["blockdiag_rank0_n48"] = zero_ranks(blocks)
    U = rng.standard_normal((64, 3))
    V = rng.standard_normal((64, 3))
    cl = H.bisection_cluster(64, 8)
    lr = H.lowrank2hss(U, V, cl, cl)

    def shift(h):
        if h.isleaf():
            h.D = h.D + 6 * np.eye(h.D.shape[0])
        else:
            shift(h.A11)
            shift(h.A22)

    shift(lr)
    cases["lowrank_shift_n64"] = lr
    cases["leafonly_n20"] = H.HssMatrix(D=rng.standard_normal((20, 20)) + 5 * np.eye(20), rootnode=True)
    cases["tiny_n12_leaf3"] = kernel_hss(rng, 12, 3, "cauchy", 1e-12)

    out = {}
    names = []
    for name, h in cases.items():
        nodes, blob = H.pack(h)
        n = int(nodes["m"][0])
        X = rng.standard_normal((n, 3))
        B = rng.standard_normal((n, 2))
        A = full_ld(h)
        Y = (A @ X.astype(LD)).astype(np.float64)
        Z = solve_ld(A, B).astype(np.float64)
        out[name + "/nodes"] = nodes
        out[name + "/blob"] = blob
        out[name + "/X"] = X
        out[name + "/Y"] = Y
        out[name + "/B"] = B
        out[name + "/Z"] = Z
        out[name + "/cond"] = np.array(np.linalg.cond(A.astype(np.float64)))
        names.append(name)
        print(f"{name}: n={n} {H.tree_summary(nodes)} cond={float(out[name + '/cond']):.2e}")
    out["names"] = np.array(names)
    np.savez_compressed(pathlib.Path(__file__).with_name("hss_golden.npz"), **out)


if __name__ == "__main__":
    main()
