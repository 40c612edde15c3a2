"""GPU parity: libhss_hip.so (through its C ABI) against the CPU oracle.

Tolerances (Float64):
  * matvec: relative difference <= 1e-14 (a sum of the same products in a
    different order);
  * ULV solve: relative difference <= 1e-12 on config C1 (n = 2001 Cauchy
    kernel, leaf 64; cond(A) ~ 3.1e3) and <= 50 * cond(A) * eps on the golden fixtures
    (two backward-stable solvers differ by at most ~cond * eps);
  * mixed precision: refinement must reach the requested residual 1e-10 and
    the result is compared with the Float64 oracle.
"""
import pathlib

import numpy as np
import pytest

from conftest import H, cauchy_c1, random_hss, rect_hss

pytestmark = pytest.mark.gpu

GOLD = np.load(pathlib.Path(__file__).with_name("golden") / "hss_golden.npz")
NAMES = [str(s) for s in GOLD["names"]]
EPS = 2.220446049250313e-16
ROOT = pathlib.Path(__file__).resolve().parents[1]


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.fixture(scope="module")
def c1_dev(hip, c1):
    return hip.HipHssMatrix((c1["nodes"], c1["blob"]))


@pytest.mark.parametrize("name", NAMES)
def test_matvec_golden(hip, oracle, name):
    nodes, blob = GOLD[name + "/nodes"], GOLD[name + "/blob"]
    A = hip.HipHssMatrix((nodes, blob))
    X = GOLD[name + "/X"]
    Y = A @ X
    assert rel(Y, GOLD[name + "/Y"]) < 1e-14
    assert rel(Y, oracle.matmul(nodes, blob, X)) < 1e-14


@pytest.mark.parametrize("name", NAMES)
def test_solve_golden(hip, oracle, name):
    nodes, blob = GOLD[name + "/nodes"], GOLD[name + "/blob"]
    A = hip.HipHssMatrix((nodes, blob))
    B = GOLD[name + "/B"]
    Z = A.solve(B)
    cond = float(GOLD[name + "/cond"])
    assert rel(Z, GOLD[name + "/Z"]) < 50 * cond * EPS
    assert rel(Z, oracle.ulvfactsolve(nodes, blob, B)) < 50 * cond * EPS


@pytest.mark.parametrize("nrhs", [1, 5, 31, 32, 33, 64, 65])
def test_c1_matvec_parity(c1_dev, oracle, c1, nrhs):
    rng = np.random.default_rng(nrhs)
    X = rng.standard_normal((2001, nrhs))
    assert rel(c1_dev @ X, oracle.matmul(c1["nodes"], c1["blob"], X)) < 1e-14


def test_c1_matvec_vector_and_alpha_beta(hip, c1_dev, oracle, c1):
    rng = np.random.default_rng(7)
    x = rng.standard_normal(2001)
    y = c1_dev @ x
    assert y.shape == (2001,)
    assert rel(y, oracle.matmul(c1["nodes"], c1["blob"], x)) < 1e-14
    B = rng.standard_normal((2001, 3))
    C0 = np.asfortranarray(rng.standard_normal((2001, 3)))
    C = C0.copy(order="F")
    hip.mul(C, c1_dev, B, -0.75, 1.5)
    ref = oracle.matmul(c1["nodes"], c1["blob"], B, alpha=-0.75, beta=1.5, C=C0)
    assert rel(C, ref) < 1e-14


def test_matvec_tree_cut_at_leaf_depth(hip, oracle):
    """A tree whose only depth of leaves is the cut: 8 rectangular leaves (row / column ranks
    3 / 2) make the cut product hold every coupling above the leaves, and no leaf folds its
    down step.  Against the oracle for 1, 20 and 40 columns, and with alpha / beta."""
    rng = np.random.default_rng(31)
    ms, ns = [64, 60, 70, 58, 66, 61, 72, 63], [50, 71, 66, 59, 64, 69, 57, 62]
    h = rect_hss(rng, ms, ns, 3, 2)
    nodes, blob = H.pack(h)
    A = hip.HipHssMatrix((nodes, blob))
    assert A.matvec_info() == {"cut_depth": 3, "folded_leaves": 0}
    F = H.full(h)
    for r in (1, 20, 40):
        X = rng.standard_normal((sum(ns), r))
        assert rel(A @ X, oracle.matmul(nodes, blob, X)) < 1e-14
        assert rel(A @ X, F @ X) < 1e-13
    B = rng.standard_normal((sum(ns), 3))
    C0 = np.asfortranarray(rng.standard_normal((sum(ms), 3)))
    C = C0.copy(order="F")
    hip.mul(C, A, B, 2.5, -0.5)
    assert rel(C, oracle.matmul(nodes, blob, B, alpha=2.5, beta=-0.5, C=C0)) < 1e-14


@pytest.mark.parametrize("nrhs", [1, 5, 33, 64])
def test_c1_solve_parity(c1_dev, oracle, c1, nrhs):
    """North star: the FP64 GPU path matches the reference's hssA\\x to 1e-12 relative."""
    rng = np.random.default_rng(100 + nrhs)
    B = rng.standard_normal((2001, nrhs))
    Z = c1_dev.solve(B)
    assert rel(Z, oracle.ulvfactsolve(c1["nodes"], c1["blob"], B)) <= 1e-12


def test_c1_reference_tolerances(c1_dev, c1):
    """test/runtests.jl:63-67 against the dense matrix, on the GPU path."""
    rng = np.random.default_rng(11)
    A = c1["A"]
    x = rng.standard_normal((2001, 5))
    assert rel(c1_dev @ x, A @ x) <= 50 * 1e-6
    rhs = rng.standard_normal((2001, 5))
    assert rel(c1_dev.solve(rhs), np.linalg.solve(A, rhs)) <= 50 * 1e-6


def test_resident_factor_reuse(c1_dev, oracle, c1):
    F = c1_dev.ulvfactor("fp64")
    rng = np.random.default_rng(12)
    for nrhs in (1, 40, 3, 64, 17, 64):
        B = rng.standard_normal((2001, nrhs))
        assert rel(F.solve(B), oracle.ulvfactsolve(c1["nodes"], c1["blob"], B)) <= 1e-12


@pytest.mark.parametrize("name", NAMES)
def test_resident_factor_tiles_reuse(hip, oracle, name):
    """Repeated solves of one resident factor, with the number of column tiles changing
    between them (workspace regrown, per-level tile widths re-chosen), match the oracle and
    repeat bit for bit."""
    nodes, blob = GOLD[name + "/nodes"], GOLD[name + "/blob"]
    A = hip.HipHssMatrix((nodes, blob))
    F = A.ulvfactor("fp64")
    cond = float(GOLD[name + "/cond"])
    n = A.shape[0]
    rng = np.random.default_rng(40)
    for nrhs in (3, 33, 3, 70):
        B = rng.standard_normal((n, nrhs))
        Z = F.solve(B)
        assert rel(Z, oracle.ulvfactsolve(nodes, blob, B)) < 50 * cond * EPS
        assert np.array_equal(F.solve(B), Z)


def test_solve_is_deterministic(c1_dev):
    rng = np.random.default_rng(13)
    B = rng.standard_normal((2001, 8))
    F = c1_dev.ulvfactor("fp64")
    a = F.solve(B)
    b = F.solve(B)
    assert np.array_equal(a, b)


@pytest.mark.parametrize("kind,n,leaf", [("cauchy", 2001, 64), ("log", 4096, 128)])
def test_mixed_precision_refinement(hip, oracle, kind, n, leaf):
    rng = np.random.default_rng(14)
    if kind == "cauchy":
        _, h = cauchy_c1()
    else:
        _, h = random_hss(rng, n, leaf, 0, "log", 1e-9)
    nodes, blob = H.pack(h)
    A = hip.HipHssMatrix((nodes, blob))
    F = A.ulvfactor("fp32")
    B = rng.standard_normal((n, 7))
    X, rep = F.solve(B, tol=1e-10, max_iter=20, report=True)
    assert rep["rel_residual"] <= 1e-10
    R = oracle.matmul(nodes, blob, X) - B
    assert np.max(np.linalg.norm(R, axis=0) / np.linalg.norm(B, axis=0)) <= 1e-10
    Xo = oracle.ulvfactsolve(nodes, blob, B)
    # error vs the FP64 oracle is bounded by cond * residual
    assert rel(X, Xo) < 1e-6


def _shift_diagonal(h, sigma):
    """h - sigma I in place: the leaves' diagonal blocks carry the whole diagonal."""
    if h.leafnode:
        h.D = np.asfortranarray(h.D - sigma * np.eye(*h.D.shape))
    else:
        _shift_diagonal(h.A11, sigma)
        _shift_diagonal(h.A22, sigma)


def test_gmres_ir_converges_where_refinement_stalls(hip, oracle):
    """GMRES-IR (mode 2) against plain refinement with the same Float32 factor, on a log-kernel
    matrix shifted next to one of its eigenvalues (cond ~1e9, so cond * eps_fp32 ~ 60): plain
    refinement contracts by about 0.4 per step and misses 1e-9 after 10 steps; GMRES
    preconditioned with the factor reaches it in a few outer steps.  Residuals are checked
    with the oracle's matvec."""
    rng = np.random.default_rng(21)
    n = 2048
    _, h = random_hss(rng, n, 128, 0, "log", 1e-9)
    A = H.full(h)
    ev = np.linalg.eigvalsh((A + A.T) / 2)
    sigma = ev[n // 3] + 1e-7 * abs(ev[-1])
    _shift_diagonal(h, sigma)
    nodes, blob = H.pack(h)
    assert 3e8 < np.linalg.cond(H.full(h)) < 3e9
    F = hip.HipHssMatrix((nodes, blob)).ulvfactor("fp32")
    B = rng.standard_normal((n, 3))
    with pytest.raises(hip.NotConverged):
        F.solve(B, tol=1e-9, max_iter=10)
    X, rep = F.solve(B, tol=1e-9, max_iter=10, report=True, method="gmres")
    assert 1 <= rep["iterations"] <= 5 and rep["rel_residual"] <= 1e-9
    R = oracle.matmul(nodes, blob, X) - B
    assert np.max(np.linalg.norm(R, axis=0) / np.linalg.norm(B, axis=0)) <= 1e-9
    # a Float64 factor as preconditioner (its first solve may already meet the tolerance)
    X64, rep64 = hip.HipHssMatrix((nodes, blob)).ulvfactor("fp64").solve(B, tol=1e-9, max_iter=5, report=True,
                                                                          method="gmres")
    R = oracle.matmul(nodes, blob, X64) - B
    assert rep64["rel_residual"] <= 1e-9
    assert np.max(np.linalg.norm(R, axis=0) / np.linalg.norm(B, axis=0)) <= 1e-9


def test_gmres_ir_well_conditioned(hip, oracle):
    """On a well-conditioned log-kernel matrix (n = 4096), GMRES-IR with the Float32 factor
    reaches 1e-12 in no more outer steps than plain refinement, and its answer agrees with
    the Float64 oracle's to cond * 1e-12."""
    prec = "fp32"
    rng = np.random.default_rng(14)
    n = 4096
    _, h = random_hss(rng, n, 128, 0, "log", 1e-9)
    nodes, blob = H.pack(h)
    F = hip.HipHssMatrix((nodes, blob)).ulvfactor(prec)
    B = rng.standard_normal((n, 5))
    B[:, 2] = 0.0  # a zero column has nothing to correct
    Xi, ri = F.solve(B, tol=1e-12, max_iter=20, report=True)
    Xg, rg = F.solve(B, tol=1e-12, max_iter=20, report=True, method="gmres")
    assert rg["rel_residual"] <= 1e-12 and rg["iterations"] <= ri["iterations"]
    assert not np.any(Xg[:, 2])
    R = oracle.matmul(nodes, blob, Xg) - B
    nb = np.linalg.norm(B, axis=0)
    assert np.max(np.linalg.norm(R, axis=0)[nb > 0] / nb[nb > 0]) <= 1e-12
    Xo = oracle.ulvfactsolve(nodes, blob, B)
    assert rel(Xg, Xo) < 1e-7 and rel(Xi, Xo) < 1e-7


def test_refinement_mode_checks(hip):
    """GMRES-IR (refinement mode 2) runs on single solves only: a plan or a server given it
    returns HSS_ERR_ARG instead of refining another way, and a single solve refuses an
    unknown mode; the plan still refines with mode 1 afterwards."""
    import ctypes

    import torch
    rng = np.random.default_rng(23)
    _, h = cauchy_c1()
    nodes, blob = H.pack(h)
    F = hip.HipHssMatrix((nodes, blob)).ulvfactor("fp32")
    lib = hip.load_library()
    n = 2001
    B = np.asfortranarray(rng.standard_normal((n, 2)))
    X = np.zeros((n, 2), order="F")
    rep = hip.SolveReport()
    bad = hip.RefineOpts(1e-10, 5, 3)
    assert lib.hss_ulv_solve(F._h, 2, B.ctypes.data, n, X.ctypes.data, n, ctypes.byref(bad),
                             ctypes.byref(rep)) == hip.HSS_ERR_ARG
    gm = hip.RefineOpts(1e-10, 10, 2)
    srv = ctypes.c_void_p()
    assert lib.hss_server_create(hip.default_context().handle, 1 << 30, hip.PREC_FP32, ctypes.byref(gm), 8,
                                 ctypes.byref(srv)) == hip.HSS_ERR_ARG
    dev = torch.device("cuda:0")
    Bt = torch.from_numpy(np.ascontiguousarray(B.T)).to(dev)
    Xt = torch.zeros_like(Bt)
    torch.cuda.synchronize()
    plan = hip.Plan([F], [2], [Bt.data_ptr()], [n], [Xt.data_ptr()], [n], refinement=True)
    assert lib.hss_plan_execute(plan._h, ctypes.byref(gm), ctypes.byref(rep)) == hip.HSS_ERR_ARG
    assert plan.execute(tol=1e-10, max_iter=20)["rel_residual"] <= 1e-10
    plan.close()


def test_batched_factor_and_solve_match_single(hip, oracle):
    """Uneven instances in one launch per level give the same answers as one at a time."""
    import torch

    rng = np.random.default_rng(15)
    mats, hs = [], []
    for n, leaf in ((700, 64), (1500, 64), (333, 32), (1024, 64)):
        _, h = random_hss(rng, n, leaf, 0, "cauchy", 1e-8)
        nodes, blob = H.pack(h)
        hs.append((nodes, blob))
        mats.append(hip.HipHssMatrix((nodes, blob)))
    facts = hip.ulvfactor_batch(mats, "fp64")
    dev = torch.device("cuda:0")
    Bs, Xs, nrhs = [], [], []
    for i, (nodes, blob) in enumerate(hs):
        n = int(nodes["n"][0])
        r = 1 + 7 * i
        B = rng.standard_normal((n, r))
        Bs.append(torch.from_numpy(np.ascontiguousarray(B.T)).to(dev))
        Xs.append(torch.zeros((r, n), dtype=torch.float64, device=dev))
        nrhs.append(r)
    torch.cuda.synchronize()  # torch fills on its own stream; the library uses its context stream
    hip.solve_batch_dev(facts, nrhs, [b.data_ptr() for b in Bs], [int(b.shape[1]) for b in Bs],
                        [x.data_ptr() for x in Xs], [int(x.shape[1]) for x in Xs])
    torch.cuda.synchronize()
    for i, (nodes, blob) in enumerate(hs):
        B = Bs[i].cpu().numpy().T
        Xb = Xs[i].cpu().numpy().T
        Xs1 = mats[i].ulvfactor("fp64").solve(B)
        assert np.array_equal(Xb, Xs1)
        assert rel(Xb, oracle.ulvfactsolve(nodes, blob, B)) < 1e-11


def test_handles_destroyed_in_any_order(hip, oracle, c1):
    """Destroying a context, matrix or factor before the handles that use it keeps it alive
    until they are destroyed too (garbage-collected bindings finalize in any order)."""
    import torch

    ctx = hip.Context(0)
    A = hip.HipHssMatrix((c1["nodes"], c1["blob"]), ctx)
    F = A.ulvfactor("fp64")
    rng = np.random.default_rng(21)
    B = rng.standard_normal((2001, 3))
    Bt = torch.from_numpy(np.ascontiguousarray(B.T)).to("cuda:0")
    Xt = torch.zeros_like(Bt)
    torch.cuda.synchronize()
    plan = hip.Plan([F], [3], [Bt.data_ptr()], [2001], [Xt.data_ptr()], [2001])
    ctx.close()
    A.close()
    X1 = F.solve(B)  # the factor still holds its matrix and context
    F.close()
    plan.execute()  # the plan still holds the factor
    torch.cuda.synchronize()
    plan.close()
    ref = oracle.ulvfactsolve(c1["nodes"], c1["blob"], B)
    assert rel(X1, ref) < 1e-12
    assert rel(Xt.cpu().numpy().T, ref) < 1e-12


def test_errors_map_to_reference_exceptions(hip, c1_dev):
    from hssmatrices_jl_amd.hssmatrix import DimensionMismatch, SingularException
    with pytest.raises(DimensionMismatch):
        c1_dev @ np.zeros((2000, 2))
    with pytest.raises(DimensionMismatch):
        c1_dev.solve(np.zeros((7, 1)))
    from test_oracle import singular_root_hss
    S = hip.HipHssMatrix(H.pack(singular_root_hss()))
    with pytest.raises(SingularException):
        S.solve(np.ones(4))
    Z = hip.HipHssMatrix(H.pack(H.HssMatrix(D=np.zeros((6, 6)), rootnode=True)))
    with pytest.raises(SingularException):
        Z.solve(np.ones(6))
    # zero right-hand sides is a no-op
    assert (c1_dev @ np.zeros((2001, 0))).shape == (2001, 0)


def test_refusals(hip):
    """What the factor and the batched entry points refuse, with the documented status:
    * a root block with a NaN column: pivoted in range and reported singular (LAPACK's
      idamax never returns an index outside the column either);
    * nodes beyond the factor's size limits (leaves of 250 points: n = 250 needs
      m + kv <= 128): HSS_ERR_UNSUPPORTED naming the node, before any device work;
    * Float32 factors where nothing refines them (hss_ulv_solve_batch_dev, a plan executed
      without refinement, a server without refinement): HSS_ERR_ARG, not a 1e-7 answer."""
    import ctypes

    import torch
    from hssmatrices_jl_amd.hssmatrix import SingularException
    D = np.eye(6)
    D[:, 2] = np.nan
    Nn = hip.HipHssMatrix(H.pack(H.HssMatrix(D=D, rootnode=True)))
    with pytest.raises(SingularException):
        Nn.solve(np.ones(6))

    _, h = random_hss(np.random.default_rng(3), 1000, 256, 0, "cauchy", 1e-8)
    nodes, blob = H.pack(h)
    assert nodes["n"][nodes["left"] < 0].max() == 250
    big = hip.HipHssMatrix((nodes, blob))
    for call in (lambda: big.ulvfactor("fp64"), lambda: big.ulvfactor("fp32"), lambda: big.solve(np.ones((1000, 1)))):
        with pytest.raises(hip.HssError) as ei:
            call()
        assert ei.value.status == hip.HSS_ERR_UNSUPPORTED and "exceeds" in str(ei.value)

    _, h = random_hss(np.random.default_rng(4), 600, 64, 0, "cauchy", 1e-8)
    A = hip.HipHssMatrix(H.pack(h))
    F32 = A.ulvfactor("fp32")
    dev = torch.device("cuda:0")
    B = torch.ones((2, 600), dtype=torch.float64, device=dev)
    X = torch.zeros((2, 600), dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    with pytest.raises(ValueError, match="Float64 factors only"):
        hip.solve_batch_dev([F32], [2], [B.data_ptr()], [600], [X.data_ptr()], [600])
    for refinement in (False, True):
        plan = hip.Plan([F32], [2], [B.data_ptr()], [600], [X.data_ptr()], [600], refinement=refinement)
        with pytest.raises(ValueError, match="refinement only"):
            plan.execute()
    lib = hip.load_library()
    ctx = hip.default_context()
    hs = ctypes.c_void_p()
    st = lib.hss_server_create(ctx.handle, 1 << 30, hip.PREC_FP32, ctypes.byref(hip.RefineOpts(1e-10, 10, 0)), 4,
                               ctypes.byref(hs))
    assert st == hip.HSS_ERR_ARG and not hs.value


@pytest.mark.parametrize("name", NAMES)
def test_adjoint_golden(hip, oracle, name):
    """hss_mat_adjoint (hssmatrix.jl:165-171): A' X, X A (matmul.jl:14) and X / A
    (hssmatrix.jl:236) against the oracle on the host adjoint; A'' X against A X."""
    nodes, blob = GOLD[name + "/nodes"], GOLD[name + "/blob"]
    A = hip.HipHssMatrix((nodes, blob))
    At = A.adjoint()
    assert At.shape == (A.shape[1], A.shape[0])
    nt, bt = H.pack(H.unpack(nodes, blob).adjoint())
    X = GOLD[name + "/X"]
    Yt = oracle.matmul(nt, bt, X)
    assert rel(At @ X, Yt) < 1e-14
    assert rel(X.T @ A, Yt.T) < 1e-14
    assert rel(At.adjoint() @ X, A @ X) < 1e-15
    B = GOLD[name + "/B"]
    cond = float(GOLD[name + "/cond"])
    Zo = oracle.ulvfactsolve(nt, bt, B)
    assert rel(B.T / A, Zo.T) < 50 * cond * EPS
    assert rel(A.rdiv(B.T), Zo.T) < 50 * cond * EPS
    # (X / A) A = X
    assert rel((B.T / A) @ A, B.T) < 50 * cond * EPS


@pytest.mark.parametrize("name", NAMES)
def test_generator_file_load_and_save(hip, oracle, tmp_path, name):
    """hss_mat_load / hss_mat_save: a matrix loaded from a generator file gives the products of
    the one created from memory bit for bit; saving writes back the same table and blob; an
    adjoint saved to a file reloads as the transpose."""
    nodes, blob = GOLD[name + "/nodes"], GOLD[name + "/blob"]
    p = tmp_path / "a.hssg"
    H.save_generators(p, nodes, blob)
    A = hip.HipHssMatrix((nodes, blob))
    L = hip.HipHssMatrix.load(p)
    assert L.shape == A.shape
    X = GOLD[name + "/X"]
    assert np.array_equal(L @ X, A @ X)
    q = tmp_path / "b.hssg"
    L.save(q)
    n2, b2 = H.load_generators(q)
    assert n2.tobytes() == np.ascontiguousarray(nodes).tobytes()
    assert np.array_equal(b2, blob)
    t = tmp_path / "t.hssg"
    A.adjoint().save(t)
    nt, bt = H.load_generators(t)
    F = H.full(H.unpack(nodes, blob))
    assert rel(oracle.matmul(nt, bt, X), F.T @ X) < 1e-13


def _tree_rel(t1, t2):
    """Largest relative difference between the generators of two trees of one shape."""
    worst = 0.0
    gens = ("D", "U", "V") if t1.leafnode else ("B12", "B21", "R1", "W1", "R2", "W2")
    for g in gens:
        x, y = getattr(t1, g), getattr(t2, g)
        assert x.shape == y.shape, (g, x.shape, y.shape)
        if y.size:
            worst = max(worst, float(np.abs(x - y).max() / max(np.abs(y).max(), 1e-300)))
    if not t1.leafnode:
        worst = max(worst, _tree_rel(t1.A11, t2.A11), _tree_rel(t1.A22, t2.A22))
    return worst


def _device_tree(M, tmp_path):
    p = tmp_path / "m.hssg"
    M.save(p)
    return H.unpack(*H.load_generators(p))


@pytest.mark.parametrize("name", NAMES)
def test_hss_product_golden(hip, tmp_path, name):
    """hss_mat_mul (matmul.jl:65-118): A A and A A' formed on the device, generator by
    generator against the oracle's restatement, against the dense product, and through the
    product's own matvec against A (B X)."""
    from oracle import hss_product as P
    nodes, blob = GOLD[name + "/nodes"], GOLD[name + "/blob"]
    h = H.unpack(nodes, blob)
    F = H.full(h)
    A = hip.HipHssMatrix((nodes, blob))
    X = GOLD[name + "/X"]
    for B, hb, Fb in ((A, h, F), (A.adjoint(), h.adjoint(), F.T)):
        C = A @ B
        assert C.shape == (F.shape[0], Fb.shape[1])
        tc = _device_tree(C, tmp_path)
        assert _tree_rel(tc, P.product(h, hb, H)) < 1e-12
        assert rel(H.full(tc), F @ Fb) < 1e-12
        assert rel(C @ X, A @ (B @ X)) < 1e-12


def test_hss_product_rectangular_and_mismatch(hip, tmp_path):
    from oracle import hss_product as P
    rng = np.random.default_rng(9)
    ns = [10, 6, 14, 8, 9, 4, 13, 7]
    a = rect_hss(rng, [7, 12, 9, 15, 5, 11, 8, 6], ns, 3, 5)
    b = rect_hss(rng, ns, [4, 9, 6, 11, 7, 5, 10, 8], 4, 2)
    A, B = hip.HipHssMatrix(a), hip.HipHssMatrix(b)
    C = A @ B
    assert C.shape == (73, 60)
    tc = _device_tree(C, tmp_path)
    assert _tree_rel(tc, P.product(a, b, H)) < 1e-12
    assert rel(H.full(tc), H.full(a) @ H.full(b)) < 1e-12
    X = rng.standard_normal((60, 3))
    assert rel(C @ X, A @ (B @ X)) < 1e-12
    # clusters that do not match (A A: A's columns are 71, its rows 73)
    with pytest.raises(H.DimensionMismatch):
        A @ A


def test_hss_product_large_leaves(hip, tmp_path):
    """Leaf blocks of 64-130 rows and columns: the leaf products D_A D_B take the task
    kernel's LDS-tiled path (64 x 64 output tiles, ragged edges); generators against the
    oracle's product, the dense product and the product's matvec."""
    from oracle import hss_product as P
    rng = np.random.default_rng(17)
    ns = [80, 66, 96, 72]
    a = rect_hss(rng, [70, 90, 65, 100], ns, 3, 5)
    b = rect_hss(rng, ns, [75, 64, 130, 68], 4, 2)
    A, B = hip.HipHssMatrix(a), hip.HipHssMatrix(b)
    C = A @ B
    assert C.shape == (325, 337)
    tc = _device_tree(C, tmp_path)
    assert _tree_rel(tc, P.product(a, b, H)) < 1e-12
    assert rel(H.full(tc), H.full(a) @ H.full(b)) < 1e-12
    X = rng.standard_normal((337, 5))
    assert rel(C @ X, A @ (B @ X)) < 1e-12


def test_adjoint_rectangular(hip, oracle):
    """A rectangular tree (73 x 71, ku = 3, kv = 5): the adjoint exchanges rows and columns,
    sizes and offsets; A' X and X A match the oracle and the dense transpose."""
    rng = np.random.default_rng(31)
    h = rect_hss(rng, [7, 12, 9, 15, 5, 11, 8, 6], [10, 6, 14, 8, 9, 4, 13, 7], 3, 5)
    F = H.full(h)
    nodes, blob = H.pack(h)
    nt, bt = H.pack(h.adjoint())
    A = hip.HipHssMatrix((nodes, blob))
    At = A.adjoint()
    assert A.shape == (73, 71) and At.shape == (71, 73)
    X = rng.standard_normal((71, 5))
    Y = rng.standard_normal((73, 6))
    assert rel(A @ X, oracle.matmul(nodes, blob, X)) < 1e-14
    assert rel(A @ X, F @ X) < 1e-13
    assert rel(At @ Y, oracle.matmul(nt, bt, Y)) < 1e-14
    assert rel(At @ Y, F.T @ Y) < 1e-13
    assert rel(Y.T @ A, (F.T @ Y).T) < 1e-13
    with pytest.raises(H.DimensionMismatch):
        X.T @ A
    with pytest.raises(H.DimensionMismatch):
        Y.T / A


def test_c1_left_multiply_and_rdiv(c1_dev, oracle, c1):
    """Config C1 through the adjoint: X A and X / A against the oracle (1e-12 for the solve)."""
    rng = np.random.default_rng(12)
    X = rng.standard_normal((7, 2001))
    nt, bt = H.pack(c1["h"].adjoint())
    assert rel(X @ c1_dev, oracle.matmul(nt, bt, X.T).T) < 1e-14
    assert rel(X @ c1_dev, X @ c1["F"]) < 1e-13
    Z = X / c1_dev
    assert rel(Z, oracle.ulvfactsolve(nt, bt, X.T).T) < 1e-12
    assert rel(Z @ c1["F"], X) < 1e-12


def _batch_inputs(hip, rng, specs, prec):
    import torch
    mats, packs = [], []
    for n, leaf in specs:
        _, h = random_hss(rng, n, leaf, 0, "cauchy", 1e-8)
        packs.append(H.pack(h))
        mats.append(hip.HipHssMatrix(packs[-1]))
    facts = hip.ulvfactor_batch(mats, prec)
    return mats, packs, facts


def test_plan_repeated_factors_and_graph_replay(hip, oracle):
    """A plan may name one factor several times; graph replays equal the eager solve."""
    import torch
    rng = np.random.default_rng(21)
    mats, packs, facts = _batch_inputs(hip, rng, [(900, 64), (400, 32), (1300, 64)], "fp64")
    dev = torch.device("cuda:0")
    order = [0, 1, 0, 2, 1]  # factors 0 and 1 appear twice
    Bs, Xs, nrhs = [], [], []
    for t, i in enumerate(order):
        n = int(packs[i][0]["n"][0])
        r = 2 + 5 * t
        Bs.append(torch.from_numpy(np.ascontiguousarray(rng.standard_normal((r, n)))).to(dev))
        Xs.append(torch.zeros((r, n), dtype=torch.float64, device=dev))
        nrhs.append(r)
    torch.cuda.synchronize()
    plan = hip.Plan([facts[i] for i in order], nrhs, [b.data_ptr() for b in Bs], [int(b.shape[1]) for b in Bs],
                    [x.data_ptr() for x in Xs], [int(x.shape[1]) for x in Xs])
    plan.execute()
    torch.cuda.synchronize()
    first = [x.cpu().numpy().copy() for x in Xs]
    for rep in range(3):  # replays of the captured graph, with host work between them
        for x in Xs:
            x.zero_()
        torch.cuda.synchronize()
        plan.execute()
        torch.cuda.synchronize()
        for x, f0 in zip(Xs, first):
            assert np.array_equal(x.cpu().numpy(), f0)
    for t, i in enumerate(order):
        nodes, blob = packs[i]
        B = Bs[t].cpu().numpy().T
        X = first[t].T
        assert np.array_equal(X, mats[i].ulvfactor("fp64").solve(B))
        assert rel(X, oracle.ulvfactsolve(nodes, blob, B)) < 1e-11
    plan.close()


def test_plan_mixed_precision_lockstep_refinement(hip, oracle):
    import torch
    rng = np.random.default_rng(22)
    mats, packs, facts = _batch_inputs(hip, rng, [(1024, 64), (700, 64), (2000, 64)], "fp32")
    dev = torch.device("cuda:0")
    Bs, Xs, nrhs = [], [], []
    for i, (nodes, blob) in enumerate(packs):
        n = int(nodes["n"][0])
        r = 1 + 3 * i
        Bs.append(torch.from_numpy(np.ascontiguousarray(rng.standard_normal((r, n)))).to(dev))
        Xs.append(torch.zeros((r, n), dtype=torch.float64, device=dev))
        nrhs.append(r)
    torch.cuda.synchronize()
    plan = hip.Plan(facts, nrhs, [b.data_ptr() for b in Bs], [int(b.shape[1]) for b in Bs],
                    [x.data_ptr() for x in Xs], [int(x.shape[1]) for x in Xs], refinement=True)
    rep = plan.execute(tol=1e-10, max_iter=20)
    assert rep["rel_residual"] <= 1e-10
    for i, (nodes, blob) in enumerate(packs):
        B = Bs[i].cpu().numpy().T
        X = Xs[i].cpu().numpy().T
        R = oracle.matmul(nodes, blob, X) - B
        assert np.max(np.linalg.norm(R, axis=0) / np.linalg.norm(B, axis=0)) <= 1e-10
    # a second execution starts again from the right-hand sides
    rep2 = plan.execute(tol=1e-10, max_iter=20)
    assert rep2["rel_residual"] <= 1e-10
    plan.close()


def test_plan_refinement_graph_replays_match_eager(hip, oracle):
    """Lockstep refinement replays captured graphs (a check step per count slot, the
    correction step): every replay, with the program's own eager work and synchronizations
    in between, gives the launch-by-launch result (profiling on runs the steps eagerly) bit
    for bit; a new tolerance captures again and meets it against the oracle's matvec."""
    import torch
    rng = np.random.default_rng(23)
    mats, packs, facts = _batch_inputs(hip, rng, [(1500, 64), (900, 64)], "fp32")
    Bs, Xs = _plan_inputs(rng, packs, [4, 2])
    plan = hip.Plan(facts, [4, 2], [b.data_ptr() for b in Bs], [int(b.shape[1]) for b in Bs],
                    [x.data_ptr() for x in Xs], [int(x.shape[1]) for x in Xs], refinement=True)
    hip.prof_enable(True)
    try:
        eager = plan.execute(tol=1e-10, max_iter=20)
    finally:
        hip.prof_enable(False)
        hip.prof_reset()
    X0 = [x.cpu().numpy().copy() for x in Xs]
    assert eager["rel_residual"] <= 1e-10
    for rep in range(3):
        junk = torch.randn(1 << 20, dtype=torch.float64, device="cuda:0").sum().item()  # eager work between replays
        assert np.isfinite(junk)
        got = plan.execute(tol=1e-10, max_iter=20)
        torch.cuda.synchronize()
        assert got["iterations"] == eager["iterations"] and got["rel_residual"] == eager["rel_residual"]
        for x, x0 in zip(Xs, X0):
            assert np.array_equal(x.cpu().numpy(), x0), f"replay {rep} differs from the eager run"
    got = plan.execute(tol=1e-11, max_iter=30)
    assert got["rel_residual"] <= 1e-11
    for i, (nodes, blob) in enumerate(packs):
        B = Bs[i].cpu().numpy().T
        X = Xs[i].cpu().numpy().T
        R = oracle.matmul(nodes, blob, X) - B
        assert np.max(np.linalg.norm(R, axis=0) / np.linalg.norm(B, axis=0)) <= 1e-11
    plan.close()


def _plan_inputs(rng, packs, rs):
    import torch
    dev = torch.device("cuda:0")
    Bs, Xs = [], []
    for (nodes, _), r in zip(packs, rs):
        n = int(nodes["n"][0])
        Bs.append(torch.from_numpy(np.ascontiguousarray(rng.standard_normal((r, n)))).to(dev))
        Xs.append(torch.zeros((r, n), dtype=torch.float64, device=dev))
    torch.cuda.synchronize()
    return Bs, Xs


def test_plan_autotune(hip, oracle):
    """hss_plan_tune: every level gets a valid launch choice; the tuned plan's solves equal
    the untuned plan's up to rounding and match the oracle; its graph replays repeat bit for
    bit; an FP32 lockstep plan still refines to its tolerance after tuning."""
    import torch
    rng = np.random.default_rng(33)
    specs = [(900, 64), (400, 32), (1300, 64), (700, 32)]
    rs = [3, 17, 32, 1]
    mats, packs, facts = _batch_inputs(hip, rng, specs, "fp64")
    Bs, Xs = _plan_inputs(rng, packs, rs)
    plan = hip.Plan(facts, rs, [b.data_ptr() for b in Bs], [int(b.shape[1]) for b in Bs],
                    [x.data_ptr() for x in Xs], [int(x.shape[1]) for x in Xs])
    plan.execute()
    torch.cuda.synchronize()
    before = [x.cpu().numpy().copy() for x in Xs]
    rep = plan.tune(reps=2, levels=12)
    assert rep["ms_before"] > 0 and rep["ms_after"] > 0
    assert rep["up"] and rep["down"]
    for f, tw, nt in rep["up"] + rep["down"]:
        assert f in (-1, 0, 1, 2)
        if f in (0, 1):
            assert tw in (16, 32) and nt in (0, 256, 512)
    for x in Xs:
        x.zero_()
    torch.cuda.synchronize()
    plan.execute()
    torch.cuda.synchronize()
    after = [x.cpu().numpy().copy() for x in Xs]
    for a, b0 in zip(after, before):
        assert rel(a, b0) < 1e-12
    for _ in range(2):  # replays of the graph captured with the tuned launches
        for x in Xs:
            x.zero_()
        torch.cuda.synchronize()
        plan.execute()
        torch.cuda.synchronize()
        for x, a in zip(Xs, after):
            assert np.array_equal(x.cpu().numpy(), a)
    for i, (nodes, blob) in enumerate(packs):
        B = Bs[i].cpu().numpy().T
        assert rel(after[i].T, oracle.ulvfactsolve(nodes, blob, B)) < 1e-11
    plan.close()
    # mixed precision: FP32 factors, lockstep refinement after tuning
    facts32 = hip.ulvfactor_batch(mats, "fp32")
    Bs, Xs = _plan_inputs(rng, packs, rs)
    plan = hip.Plan(facts32, rs, [b.data_ptr() for b in Bs], [int(b.shape[1]) for b in Bs],
                    [x.data_ptr() for x in Xs], [int(x.shape[1]) for x in Xs], refinement=True)
    plan.tune(reps=2)
    r = plan.execute(tol=1e-10, max_iter=20)
    assert r["rel_residual"] <= 1e-10
    for i, (nodes, blob) in enumerate(packs):
        B = Bs[i].cpu().numpy().T
        R = oracle.matmul(nodes, blob, Xs[i].cpu().numpy().T) - B
        assert np.max(np.linalg.norm(R, axis=0) / np.linalg.norm(B, axis=0)) <= 1e-10
    plan.close()


def test_wide_batch_levels(hip, oracle):
    """Levels of at least 2048 (node, 32-column) tiles run 256-thread workgroups with the
    operators streamed; every instance matches its own single solve and the oracle, and an
    FP32 lockstep plan over the same levels (wide matvec too) reaches its residual."""
    import torch
    rng = np.random.default_rng(31)
    _, h = random_hss(rng, 4096, 32, 0, "cauchy", 1e-8)
    nodes, blob = H.pack(h)
    copies = 20  # 20 x 128 leaves
    mats = [hip.HipHssMatrix((nodes, blob)) for _ in range(copies)]
    n = int(nodes["n"][0])
    dev = torch.device("cuda:0")
    Bs, Xs, nrhs = [], [], []
    for i in range(copies):
        r = 1 + 7 * (i % 5)
        Bs.append(torch.from_numpy(np.ascontiguousarray(rng.standard_normal((r, n)))).to(dev))
        Xs.append(torch.zeros((r, n), dtype=torch.float64, device=dev))
        nrhs.append(r)
    torch.cuda.synchronize()
    facts = hip.ulvfactor_batch(mats, "fp64")
    hip.solve_batch_dev(facts, nrhs, [b.data_ptr() for b in Bs], [n] * copies, [x.data_ptr() for x in Xs],
                        [n] * copies)
    torch.cuda.synchronize()
    single = mats[0].ulvfactor("fp64")
    for i in range(copies):
        B = Bs[i].cpu().numpy().T
        Xb = Xs[i].cpu().numpy().T
        assert rel(Xb, single.solve(B)) < 1e-13
    def resid(X, B):
        R = oracle.matmul(nodes, blob, X) - B
        return np.max(np.linalg.norm(R, axis=0) / np.linalg.norm(B, axis=0))

    for i in (0, 3, copies - 1):
        # as accurate as the reference: residual within 10x the oracle's own
        B = Bs[i].cpu().numpy().T
        X = Xs[i].cpu().numpy().T
        Xo = oracle.ulvfactsolve(nodes, blob, B)
        assert resid(X, B) <= 10 * resid(Xo, B) + 1e-14
        assert rel(X, Xo) < 1e-10
    f32 = hip.ulvfactor_batch(mats, "fp32")
    for x in Xs:
        x.zero_()
    torch.cuda.synchronize()
    plan = hip.Plan(f32, nrhs, [b.data_ptr() for b in Bs], [n] * copies, [x.data_ptr() for x in Xs], [n] * copies,
                    refinement=True)
    rep = plan.execute(tol=1e-10, max_iter=20)
    torch.cuda.synchronize()
    assert rep["rel_residual"] <= 1e-10
    for i in (0, 7, copies - 1):
        assert resid(Xs[i].cpu().numpy().T, Bs[i].cpu().numpy().T) <= 1e-10
    plan.close()


def test_server_alone_equals_mixed(hip, oracle):
    """Continuous batching: each result is the same alone and in a mix with other requests."""
    import torch
    rng = np.random.default_rng(23)
    packs, mats = [], []
    for n, leaf in ((1024, 64), (1500, 64), (700, 32)):
        _, h = random_hss(rng, n, leaf, 0, "cauchy", 1e-8)
        packs.append(H.pack(h))
        mats.append(hip.HipHssMatrix(packs[-1]))
    srv = hip.Server(budget_bytes=1 << 30, prec="fp32", tol=1e-10, max_iter=20, max_inflight=8)
    ids = [srv.admit(m) for m in mats]
    dev = torch.device("cuda:0")

    def make(i, r):
        n = int(packs[i][0]["n"][0])
        B = torch.from_numpy(np.ascontiguousarray(rng.standard_normal((r, n)))).to(dev)
        X = torch.zeros_like(B)
        torch.cuda.synchronize()
        return B, X

    # alone
    B0, X0 = make(0, 3)
    t = srv.submit(ids[0], 3, B0.data_ptr(), B0.shape[1], X0.data_ptr(), X0.shape[1])
    while srv.step() > 0:
        pass
    res = srv.poll(t)
    assert res is not None and res["status"] == 1 and res["rel_residual"] <= 1e-10
    alone = X0.cpu().numpy().copy()
    # the same request in a mix, arriving between rounds of other requests
    X0.zero_()
    torch.cuda.synchronize()
    others = [make(i % 3, 1 + i) for i in range(6)]
    tickets = []
    for k, (B, X) in enumerate(others[:3]):
        tickets.append(srv.submit(ids[k % 3], B.shape[0], B.data_ptr(), B.shape[1], X.data_ptr(), X.shape[1]))
    srv.step()
    t0 = srv.submit(ids[0], 3, B0.data_ptr(), B0.shape[1], X0.data_ptr(), X0.shape[1])
    for k, (B, X) in enumerate(others[3:]):
        tickets.append(srv.submit(ids[k % 3], B.shape[0], B.data_ptr(), B.shape[1], X.data_ptr(), X.shape[1]))
    while srv.step() > 0:
        pass
    r0 = srv.poll(t0)
    assert r0["status"] == 1
    mixed = X0.cpu().numpy()
    assert rel(mixed, alone) <= 1e-12
    for k, ((B, X), tk) in enumerate(zip(others, tickets)):
        r = srv.poll(tk)
        assert r["status"] == 1
        nodes, blob = packs[k % 3]
        Bn, Xn = B.cpu().numpy().T, X.cpu().numpy().T
        R = oracle.matmul(nodes, blob, Xn) - Bn
        assert np.max(np.linalg.norm(R, axis=0) / np.linalg.norm(Bn, axis=0)) <= 1e-10
    # admission is refused beyond the budget
    small = hip.Server(budget_bytes=1024, prec="fp32")
    with pytest.raises(hip.HssError):
        small.admit(mats[0])
    srv.close()


def test_launch_profiling_records(hip, c1_dev):
    """hss_prof_*: a profiled solve gives the unprofiled solve's result bit for bit; every
    sweep level is one record with a positive kernel time; the records of one solve add up
    to less than the solve's wall time (each is a kernel's own duration)."""
    import time
    F = c1_dev.ulvfactor("fp64")
    rng = np.random.default_rng(21)
    B = rng.standard_normal((2001, 32))
    Z = F.solve(B)
    hip.prof_reset()
    hip.prof_enable(True)
    try:
        t0 = time.perf_counter()
        Zp = F.solve(B)
        wall_ms = (time.perf_counter() - t0) * 1e3
        recs = hip.prof_records()
    finally:
        hip.prof_enable(False)
        hip.prof_reset()
    assert np.array_equal(Zp, Z)
    assert recs, "no launch recorded"
    assert all(r["ms"] > 0 and r["nwork"] > 0 for r in recs)
    up = [r for r in recs if r["kernel"].startswith("k_ulv_gup")]
    down = [r for r in recs if r["kernel"].startswith("k_ulv_gdown")]
    roots = [r for r in recs if r["kernel"].startswith("k_ulv_groot")]
    assert len(roots) == 1 and len(up) >= 1 and len(down) >= 1
    # one record per up-sweep height (0, 1, ...) and per down-sweep depth
    assert sorted(r["level"] for r in up) == list(range(len(up)))
    assert len({r["level"] for r in down}) == len(down)
    assert sum(r["ms"] for r in recs) < wall_ms
