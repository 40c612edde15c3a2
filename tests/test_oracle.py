"""Pins of the CPU oracle (oracle/hss_oracle.c) against the mathematics.

The reference ships no golden vectors (test/runtests.jl checks tolerances
against dense matrices only), so the oracle is pinned by
  * the extended-precision closed-form fixtures in tests/golden/ (Y = A X,
    Z = A \\ B from the dense expansion of the generators);
  * known answers: lowrank2hss (full = U V^T exactly, constructors.jl:10-20),
    leaf-only (dense GEMM / LU, matmul.jl:21-22, ulvfactor.jl:11-12), rank-0
    block diagonal, pass-through nodes (ulvfactor.jl:32-36);
  * invariants: linearity, mul! alpha/beta semantics (matmul.jl:18), the
    adjoint identity <y, A x> = <A^T y, x> (hssmatrix.jl:165-171);
  * the reference's own tolerance test (runtests.jl:63-67) on config C1;
  * planted mistakes (a transposed coupling, a dropped translator) that these
    checks must catch.
"""
import numpy as np
import pytest

from conftest import H, cauchy_c1, rect_hss

GOLD = np.load(__import__("pathlib").Path(__file__).with_name("golden") / "hss_golden.npz")
NAMES = [str(s) for s in GOLD["names"]]


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.mark.parametrize("name", NAMES)
def test_oracle_matvec_golden(oracle, name):
    nodes, blob = GOLD[name + "/nodes"], GOLD[name + "/blob"]
    Y = oracle.matmul(nodes, blob, GOLD[name + "/X"])
    assert rel(Y, GOLD[name + "/Y"]) < 1e-14


@pytest.mark.parametrize("name", NAMES)
def test_oracle_solve_golden(oracle, name):
    nodes, blob = GOLD[name + "/nodes"], GOLD[name + "/blob"]
    Z = oracle.ulvfactsolve(nodes, blob, GOLD[name + "/B"])
    cond = float(GOLD[name + "/cond"])
    # two backward-stable solves differ by at most ~cond * eps
    assert rel(Z, GOLD[name + "/Z"]) < 50 * cond * 2.2e-16


def test_oracle_c1_reference_tolerances(oracle, c1):
    """test/runtests.jl:63-67 on config C1: ||A x - hssA x|| / ||A x|| <= 50 rtol, same for the solve."""
    rng = np.random.default_rng(0)
    x = rng.standard_normal((2001, 5))
    A = c1["A"]
    y = oracle.matmul(c1["nodes"], c1["blob"], x)
    assert rel(y, A @ x) <= 50 * 1e-6
    b = rng.standard_normal((2001, 5))
    z = oracle.ulvfactsolve(c1["nodes"], c1["blob"], b)
    assert rel(z, np.linalg.solve(A, b)) <= 50 * 1e-6


def test_oracle_c1_closed_form(oracle, c1):
    rng = np.random.default_rng(1)
    x = rng.standard_normal((2001, 3))
    F = c1["F"]
    assert rel(oracle.matmul(c1["nodes"], c1["blob"], x), F @ x) < 1e-14
    z = oracle.ulvfactsolve(c1["nodes"], c1["blob"], x)
    assert rel(z, np.linalg.solve(F, x)) < 1e-12  # cond(F) ~ 3.1e3
    # backward error of a stable solve: ||F z - x|| / (||F|| ||z||) ~ eps
    assert np.linalg.norm(F @ z - x) / (np.linalg.norm(F, 2) * np.linalg.norm(z)) < 1e-14


def test_oracle_lowrank_known_answer(oracle):
    rng = np.random.default_rng(2)
    U = rng.standard_normal((150, 3))
    V = rng.standard_normal((150, 3))
    cl = H.bisection_cluster(150, 16)
    h = H.lowrank2hss(U, V, cl, cl)
    assert H.hssrank(h) == 3
    nodes, blob = H.pack(h)
    x = rng.standard_normal((150, 4))
    assert rel(oracle.matmul(nodes, blob, x), U @ (V.T @ x)) < 1e-14


def test_oracle_leaf_only(oracle):
    rng = np.random.default_rng(3)
    D = rng.standard_normal((30, 30)) + 4 * np.eye(30)
    nodes, blob = H.pack(H.HssMatrix(D=D, rootnode=True))
    x = rng.standard_normal((30, 2))
    assert rel(oracle.matmul(nodes, blob, x), D @ x) < 1e-15
    assert rel(oracle.ulvfactsolve(nodes, blob, x), np.linalg.solve(D, x)) < 1e-13


def test_oracle_alpha_beta_linearity(oracle, c1):
    rng = np.random.default_rng(4)
    n = 2001
    x1, x2 = rng.standard_normal((n, 2)), rng.standard_normal((n, 2))
    C0 = rng.standard_normal((n, 2))
    nodes, blob = c1["nodes"], c1["blob"]
    y1 = oracle.matmul(nodes, blob, x1)
    y2 = oracle.matmul(nodes, blob, x2)
    y12 = oracle.matmul(nodes, blob, 2.0 * x1 - 3.0 * x2)
    assert rel(y12, 2.0 * y1 - 3.0 * y2) < 1e-13
    C = oracle.matmul(nodes, blob, x1, alpha=-0.5, beta=2.0, C=C0)
    assert rel(C, -0.5 * y1 + 2.0 * C0) < 1e-14


def test_oracle_adjoint_identity(oracle):
    rng = np.random.default_rng(5)
    _, h = __import__("conftest").random_hss(rng, 300, 20, 0, "cauchy", 1e-9)
    nodes, blob = H.pack(h)
    nt, bt = H.pack(h.adjoint())
    x = rng.standard_normal(300)
    y = rng.standard_normal(300)
    lhs = y @ oracle.matmul(nodes, blob, x)
    rhs = oracle.matmul(nt, bt, y) @ x
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)


@pytest.mark.parametrize("name", NAMES)
def test_oracle_adjoint_closed_form(oracle, name):
    """The host adjoint (hssmatrix.jl:165-171: D', U <-> V, B12 <- B21', B21 <- B12', R <-> W)
    expands to the transpose (up to the rounding of the re-ordered products), and the oracle's
    matvec on it matches the dense transpose."""
    nodes, blob = GOLD[f"{name}/nodes"], GOLD[f"{name}/blob"]
    h = H.unpack(nodes, blob)
    F = H.full(h)
    assert rel(H.full(h.adjoint()), F.T) < 1e-14
    nt, bt = H.pack(h.adjoint())
    X = np.random.default_rng(8).standard_normal((F.shape[0], 3))
    assert rel(oracle.matmul(nt, bt, X), F.T @ X) < 1e-13


def test_oracle_rectangular(oracle):
    """Rectangular trees (m != n, ku != kv, hssmatrix.jl:11-68): the oracle's matvec of A and of
    its adjoint against the dense expansion."""
    rng = np.random.default_rng(3)
    h = rect_hss(rng, [7, 12, 9, 15, 5, 11, 8, 6], [10, 6, 14, 8, 9, 4, 13, 7], 3, 5)
    F = H.full(h)
    assert F.shape == (73, 71)
    nodes, blob = H.pack(h)
    X = rng.standard_normal((71, 4))
    assert rel(oracle.matmul(nodes, blob, X), F @ X) < 1e-14
    nt, bt = H.pack(h.adjoint())
    Y = rng.standard_normal((73, 4))
    assert rel(oracle.matmul(nt, bt, Y), F.T @ Y) < 1e-14


@pytest.mark.parametrize("name", NAMES)
def test_oracle_hss_product_closed_form(name):
    """The restated HSS x HSS product (matmul.jl:65-118) expands to the dense product, for
    A A and A A' on every golden tree, with ranks that add (no recompression)."""
    from oracle import hss_product as P
    h = H.unpack(GOLD[f"{name}/nodes"], GOLD[f"{name}/blob"])
    F = H.full(h)
    for b, Fb in ((h, F), (h.adjoint(), F.T)):
        c = P.product(h, b, H)
        assert rel(H.full(c), F @ Fb) < 1e-13
        if not h.leafnode:
            assert H.gensize(c.A11) == tuple(x + y for x, y in zip(H.gensize(h.A11), H.gensize(b.A11)))


def test_oracle_hss_product_rectangular():
    from oracle import hss_product as P
    rng = np.random.default_rng(9)
    ns = [10, 6, 14, 8, 9, 4, 13, 7]
    a = rect_hss(rng, [7, 12, 9, 15, 5, 11, 8, 6], ns, 3, 5)
    b = rect_hss(rng, ns, [4, 9, 6, 11, 7, 5, 10, 8], 4, 2)
    c = P.product(a, b, H)
    assert rel(H.full(c), H.full(a) @ H.full(b)) < 1e-13
    # a planted mistake (the cross term of W1 dropped) is caught
    bad = P.product(a, b, H)
    bad.A11.W1 = bad.A11.W1.copy()
    bad.A11.W1[a.A11.W1.shape[0]:, : a.A11.W1.shape[1]] = 0.0
    assert rel(H.full(bad), H.full(a) @ H.full(b)) > 1e-6


def singular_root_hss():
    """Two pass-through leaves (ku = m, ulvfactor.jl:32-36) whose merged root block is zero:
    only the dense root solve (ulvfactor.jl:83) sees the singularity."""
    I2 = np.eye(2)
    z = np.zeros((2, 2))
    a = H.HssMatrix(D=z, U=I2, V=I2)
    b = H.HssMatrix(D=z, U=I2, V=I2)
    return H.HssMatrix(A11=a, A22=b, B12=z, B21=z, rootnode=True)


def test_oracle_singular_root_raises(oracle):
    nodes, blob = H.pack(H.HssMatrix(D=np.zeros((5, 5)), rootnode=True))
    with pytest.raises(oracle.OracleSingular):
        oracle.ulvfactsolve(nodes, blob, np.ones(5))
    nodes, blob = H.pack(singular_root_hss())
    with pytest.raises(oracle.OracleSingular):
        oracle.ulvfactsolve(nodes, blob, np.ones(4))


def _perturbed(nodes, blob, what):
    """Plant a mistake in the oracle's input the way a plausible bug would act."""
    nodes = nodes.copy()
    blob = blob.copy()
    if what == "transpose_B12":
        # use B21's data as B12 (a swapped/transposed coupling)
        i = int(np.nonzero(nodes["left"] >= 0)[0][1])
        L, R = nodes[nodes[i]["left"]], nodes[nodes[i]["right"]]
        n12 = int(L["ku"]) * int(R["kv"])
        b = blob[nodes[i]["off_B12"]:nodes[i]["off_B12"] + n12].reshape(int(L["ku"]), int(R["kv"]), order="F")
        blob[nodes[i]["off_B12"]:nodes[i]["off_B12"] + n12] = b[::-1].ravel(order="F")
    elif what == "drop_translator":
        j = int(np.nonzero((nodes["parent"] > 0))[0][0])
        nr = int(nodes[j]["ku"]) * int(nodes[nodes[j]["parent"]]["ku"])
        blob[nodes[j]["off_R"]:nodes[j]["off_R"] + nr] = 0.0
    elif what == "sign_V":
        k = int(np.nonzero(nodes["left"] < 0)[0][0])
        nv = int(nodes[k]["n"]) * int(nodes[k]["kv"])
        blob[nodes[k]["off_V"]:nodes[k]["off_V"] + nv] *= -1
    return nodes, blob


@pytest.mark.parametrize("what", ["transpose_B12", "drop_translator", "sign_V"])
def test_planted_mistakes_are_caught(oracle, what):
    """A matvec that used a wrong operand must fail the closed-form check."""
    name = "cauchy_n120_leaf16"
    nodes, blob = GOLD[name + "/nodes"], GOLD[name + "/blob"]
    pn, pb = _perturbed(nodes, blob, what)
    Y = oracle.matmul(pn, pb, GOLD[name + "/X"])
    assert rel(Y, GOLD[name + "/Y"]) > 1e-6
    Z = oracle.ulvfactsolve(pn, pb, GOLD[name + "/B"])
    assert rel(Z, GOLD[name + "/Z"]) > 1e-6
