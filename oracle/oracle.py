"""Python driver of the CPU oracle (TEST INFRASTRUCTURE ONLY).

Loads ``oracle/build/libhss_oracle.so`` (built by ``oracle/Makefile``) and
hands it the addresses of the LAPACK/BLAS routines the reference calls
(dgemm, dgeqlf, dormql, dgelqf, dormlq, dtrsm, dgetrf, dgetrs), taken from
scipy's ``cython_lapack``/``cython_blas`` capsules (OpenBLAS).

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline
leg may import this module; the product path never does.

Functions:
  matmul(nodes, blob, B, alpha=1, beta=0, C=None)  -- reference matmul.jl:18-62
  ulvfactsolve(nodes, blob, b)                      -- reference ulvfactor.jl:10-107
"""
from __future__ import annotations

import ctypes
import os
import pathlib
import subprocess

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
LIB = HERE / "build" / "libhss_oracle.so"

HSSO_OK, HSSO_DIM, HSSO_SINGULAR = 0, 1, 3


class OracleError(RuntimeError):
    pass


class OracleSingular(OracleError):
    pass


def build(force: bool = False) -> pathlib.Path:
    if force or not LIB.exists() or LIB.stat().st_mtime < (HERE / "hss_oracle.c").stat().st_mtime:
        subprocess.run(["make", "-s", "-C", str(HERE), "-B" if force else "build/libhss_oracle.so"],
                       check=True)
    return LIB


def _capsule_ptr(mod, name):
    cap = mod.__pyx_capi__[name]
    gn = ctypes.pythonapi.PyCapsule_GetName
    gn.restype = ctypes.c_char_p
    gn.argtypes = [ctypes.py_object]
    gp = ctypes.pythonapi.PyCapsule_GetPointer
    gp.restype = ctypes.c_void_p
    gp.argtypes = [ctypes.py_object, ctypes.c_char_p]
    return gp(cap, gn(cap))


class _Lapack(ctypes.Structure):
    _fields_ = [(n, ctypes.c_void_p) for n in
                ("dgemm", "dgeqlf", "dormql", "dgelqf", "dormlq", "dtrsm", "dgetrf", "dgetrs")]


_lib = None
_lapack = None


def _load():
    global _lib, _lapack
    if _lib is None:
        build()
        from scipy.linalg import cython_blas, cython_lapack
        lib = ctypes.CDLL(str(LIB))
        lp = _Lapack()
        lp.dgemm = _capsule_ptr(cython_blas, "dgemm")
        lp.dtrsm = _capsule_ptr(cython_blas, "dtrsm")
        for n in ("dgeqlf", "dormql", "dgelqf", "dormlq", "dgetrf", "dgetrs"):
            setattr(lp, n, _capsule_ptr(cython_lapack, n))
        P = ctypes.c_void_p
        lib.hsso_matmul.restype = ctypes.c_int
        lib.hsso_matmul.argtypes = [P, P, ctypes.c_int, P, ctypes.c_int, ctypes.c_double, P, ctypes.c_int,
                                    ctypes.c_double, P, ctypes.c_int]
        lib.hsso_ulvfactsolve.restype = ctypes.c_int
        lib.hsso_ulvfactsolve.argtypes = [P, P, ctypes.c_int, P, ctypes.c_int, P, ctypes.c_int, P, ctypes.c_int]
        _lib, _lapack = lib, lp
    return _lib, _lapack


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def matmul(nodes, blob, B, alpha: float = 1.0, beta: float = 0.0, C=None):
    """C = alpha*A*B + beta*C (reference mul!, matmul.jl:18-28)."""
    lib, lp = _load()
    B = np.asfortranarray(B, dtype=np.float64)
    vec = B.ndim == 1
    if vec:
        B = B.reshape(-1, 1, order="F")
    m, n = int(nodes["m"][0]), int(nodes["n"][0])
    if B.shape[0] != n:
        raise OracleError("dimension mismatch")
    nrhs = B.shape[1]
    if C is None:
        C = np.zeros((m, nrhs), order="F")
    else:
        C = np.asfortranarray(np.array(C, dtype=np.float64, copy=True)).reshape(m, nrhs, order="F")
    nodes = np.ascontiguousarray(nodes)
    blob = np.ascontiguousarray(blob, dtype=np.float64)
    st = lib.hsso_matmul(ctypes.byref(lp), _ptr(nodes), len(nodes), _ptr(blob), nrhs, alpha, _ptr(B),
                         max(n, 1), beta, _ptr(C), max(m, 1))
    if st != HSSO_OK:
        raise OracleError(f"hsso_matmul status {st}")
    return C[:, 0] if vec else C


def ulvfactsolve(nodes, blob, b):
    """z = A \\ b through the reference's fused ULV factor-solve (ulvfactor.jl:10-107)."""
    lib, lp = _load()
    b = np.asfortranarray(b, dtype=np.float64)
    vec = b.ndim == 1
    if vec:
        b = b.reshape(-1, 1, order="F")
    n = int(nodes["n"][0])
    nrhs = b.shape[1]
    z = np.zeros((n, nrhs), order="F")
    nodes = np.ascontiguousarray(nodes)
    blob = np.ascontiguousarray(blob, dtype=np.float64)
    st = lib.hsso_ulvfactsolve(ctypes.byref(lp), _ptr(nodes), len(nodes), _ptr(blob), nrhs, _ptr(b), max(n, 1),
                               _ptr(z), max(n, 1))
    if st == HSSO_SINGULAR:
        raise OracleSingular("singular root block")
    if st != HSSO_OK:
        raise OracleError(f"hsso_ulvfactsolve status {st}")
    return z[:, 0] if vec else z
