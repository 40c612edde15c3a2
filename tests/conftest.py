"""Shared fixtures.  GPU tests are marked ``gpu`` and run on an MI355X;
everything else runs on the CPU (oracle, host logic, ABI exports)."""
import pathlib
import sys

import numpy as np
import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

import __graft_entry__ as _entry  # noqa: E402

H = _entry.load_package()


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libhss_hip.so)")


@pytest.fixture(scope="session")
def pkg():
    return H


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O
    O.build()
    return O


@pytest.fixture(scope="session")
def hip():
    import importlib
    return importlib.import_module("hssmatrices_jl_amd.hip")


def cauchy_c1(leafsize=64, tol=1e-6):
    """Config C1: Cauchy 1/(x_i-x_j) with unit diagonal on x = -1:0.001:1 (n = 2001)."""
    x = -1 + 0.001 * np.arange(2001)
    K = H.KernelMatrix(x, "cauchy", 1.0)
    rcl = H.bisection_cluster(2001, leafsize)
    return K, H.compress(K, rcl, rcl, tol, tol)


def random_hss(rng, n, leafsize, k, kind="cauchy", tol=1e-10):
    x = np.sort(rng.uniform(-1, 1, n))
    K = H.KernelMatrix(x, kind, 1.0 if kind != "log" else 1.0)
    rcl = H.bisection_cluster(n, leafsize)
    return K, H.compress(K, rcl, rcl, tol, tol)


@pytest.fixture(scope="session")
def c1():
    K, h = cauchy_c1()
    nodes, blob = H.pack(h)
    return {"K": K, "h": h, "nodes": nodes, "blob": blob, "A": K.dense(), "F": H.full(h)}


def rect_hss(rng, ms, ns, ku, kv, root=True):
    """A random rectangular HSS tree over leaves of ms[i] x ns[i] (len a power of two), with
    generator ranks ku (rows) and kv (columns) on every non-root node (hssmatrix.jl:11-68)."""
    g = lambda r, c: rng.standard_normal((r, c))
    if len(ms) == 1:
        return H.HssMatrix(D=g(ms[0], ns[0]), U=g(ms[0], ku), V=g(ns[0], kv), rootnode=root)
    h = len(ms) // 2
    a = rect_hss(rng, ms[:h], ns[:h], ku, kv, False)
    b = rect_hss(rng, ms[h:], ns[h:], ku, kv, False)
    if root:
        return H.HssMatrix(A11=a, A22=b, B12=g(ku, kv), B21=g(ku, kv), rootnode=True)
    return H.HssMatrix(A11=a, A22=b, B12=g(ku, kv), B21=g(ku, kv), R1=g(ku, ku), W1=g(kv, kv),
                       R2=g(ku, ku), W2=g(kv, kv))
