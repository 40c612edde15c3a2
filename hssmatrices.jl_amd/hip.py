"""ctypes binding of libhss_hip.so and the device-side HSS objects.

This is the Python mirror of the reference's operator interface for the hot
path (the Julia mirror is ``julia/HssHip.jl``):

=====================================  =============================================
reference (bonevbs/HssMatrices.jl)     here
=====================================  =============================================
``hssA * x`` / ``*(hssA, B)``          ``hssA @ x``  (``HipHssMatrix.__matmul__``)
  ``src/matmul.jl:13,15``
``mul!(C, hssA, B, a, b)``             ``mul(C, hssA, B, a, b)``
  ``src/matmul.jl:18``
``hssA \\ B`` (``hssmatrix.jl:234``)   ``ldiv(hssA, B)`` / ``hssA.solve(B)``
``ulvfactsolve(hssA, b)``              ``ulvfactsolve(hssA, b)``
  ``src/ulvfactor.jl:10``
(no counterpart: refactors each call)  ``hssA.ulvfactor(prec)`` -> resident ``HipUlvFactor``
``hssA'`` (``hssmatrix.jl:165-171``)   ``hssA.adjoint()`` / ``hssA.H``
``A * hssB`` (``matmul.jl:14``)        ``A @ hssB``  (``HipHssMatrix.__rmatmul__``)
``A / hssB`` (``hssmatrix.jl:236``)    ``A / hssB`` / ``hssB.rdiv(A)``
``hssA * hssB`` (``matmul.jl:65-118``) ``hssA @ hssB`` / ``hssA.hss_mul(hssB)``
(in memory only, hssmatrix.jl:11-68)   ``HipHssMatrix.load(path)`` / ``.save(path)``,
                                       ``gen_read`` / ``gen_write`` (generator files)
=====================================  =============================================

Errors follow the reference: ``DimensionMismatch`` (``matmul.jl:19-20``),
``ValueError`` for ``ArgumentError``, ``SingularException`` for a zero pivot of
the dense root solve (``ulvfactor.jl:83``).

The shared library must have been built (``__graft_entry__.build()`` or
``make -C hssmatrices.jl_amd/csrc``); importing this module without it raises.
"""
from __future__ import annotations

import ctypes
import pathlib

import numpy as np

from .hssmatrix import NODE_DTYPE, DimensionMismatch, HssMatrix, SingularException, cluster_nodes, pack

LIB_PATH = pathlib.Path(__file__).resolve().parent / "libhss_hip.so"

HSS_OK, HSS_ERR_DIM, HSS_ERR_ARG, HSS_ERR_SINGULAR, HSS_ERR_NOCONV = 0, 1, 2, 3, 4
HSS_ERR_OOM, HSS_ERR_BUDGET, HSS_ERR_HIP, HSS_ERR_UNSUPPORTED, HSS_ERR_IO = 5, 6, 7, 8, 9
PREC_FP64, PREC_FP32 = 0, 1

# every symbol include/hss_hip.h declares
EXPORTS = [
    "hss_ctx_create", "hss_ctx_destroy", "hss_ctx_stream",
    "hss_mat_create", "hss_mat_destroy", "hss_mat_adjoint", "hss_mat_mul", "hss_mat_size", "hss_mat_device_bytes",
    "hss_gen_write", "hss_gen_read", "hss_mat_load", "hss_mat_save", "hss_compress_dev", "hss_randcompress_dev", "hss_mat_recompress", "hss_mat_ldiv", "hss_mat_rdiv", "hss_mat_generators",
    "hss_mat_create_f32", "hss_matmul_f32", "hss_ulvfactsolve_f32", "hss_ulv_solve_f32",
    "hss_matmul", "hss_matmul_dev",
    "hss_ulvfactsolve", "hss_ulv_factor", "hss_fact_destroy", "hss_fact_device_bytes",
    "hss_ulv_solve", "hss_ulv_solve_dev", "hss_ulv_solve_dev_noalloc", "hss_ulv_solve_batch_dev",
    "hss_ulv_factor_batch",
    "hss_last_error", "hss_last_device_ms",
    "hss_prof_enable", "hss_prof_reset", "hss_prof_count", "hss_prof_get", "hss_fact_node_info",
    "hss_mat_matvec_info",
    "hss_plan_create", "hss_plan_execute", "hss_plan_destroy", "hss_plan_device_bytes",
    "hss_plan_tune", "hss_plan_tune_choice",
    "hss_server_create", "hss_server_admit", "hss_server_evict", "hss_server_used_bytes", "hss_server_submit",
    "hss_server_step", "hss_server_poll", "hss_server_destroy",
]


class HssError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"[hss status {status}] {msg}")
        self.status = status


class NotConverged(HssError):
    pass


class RefineOpts(ctypes.Structure):
    _fields_ = [("tol", ctypes.c_double), ("max_iter", ctypes.c_int), ("mode", ctypes.c_int)]


class SolveReport(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int), ("rel_residual", ctypes.c_double), ("seconds", ctypes.c_double)]


_lib = None


def load_library(path: pathlib.Path | str | None = None):
    """Load libhss_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    p = pathlib.Path(path) if path else LIB_PATH
    if not p.exists():
        raise ImportError(f"{p} is missing: build it with __graft_entry__.build() "
                          f"or `make -C hssmatrices.jl_amd/csrc`")
    lib = ctypes.CDLL(str(p))
    V, I32, I64, D = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_double
    PV = ctypes.POINTER(ctypes.c_void_p)
    sig = {
        "hss_ctx_create": ([ctypes.c_int, PV], ctypes.c_int),
        "hss_ctx_destroy": ([V], ctypes.c_int),
        "hss_ctx_stream": ([V], V),
        "hss_mat_create": ([V, V, I32, V, I64, PV], ctypes.c_int),
        "hss_mat_destroy": ([V], ctypes.c_int),
        "hss_mat_adjoint": ([V, PV], ctypes.c_int),
        "hss_mat_mul": ([V, V, PV], ctypes.c_int),
        "hss_mat_size": ([V, ctypes.POINTER(I64), ctypes.POINTER(I64)], ctypes.c_int),
        "hss_mat_device_bytes": ([V], I64),
        "hss_gen_write": ([ctypes.c_char_p, V, I32, V, I64], ctypes.c_int),
        "hss_gen_read": ([ctypes.c_char_p, V, I32, ctypes.POINTER(I32), V, I64, ctypes.POINTER(I64)], ctypes.c_int),
        "hss_mat_load": ([V, ctypes.c_char_p, PV], ctypes.c_int),
        "hss_mat_save": ([V, ctypes.c_char_p], ctypes.c_int),
        "hss_compress_dev": ([V, V, I64, V, I32, D, D, PV], ctypes.c_int),
        "hss_randcompress_dev": ([V, V, I64, V, I32, D, D, I32, I32, I32, PV, V], ctypes.c_int),
        "hss_mat_recompress": ([V, D, D, PV], ctypes.c_int),
        "hss_mat_ldiv": ([V, V, D, D, PV], ctypes.c_int),
        "hss_mat_create_f32": ([V, V, I32, V, I64, PV], ctypes.c_int),
        "hss_matmul_f32": ([V, I32, ctypes.c_float, V, I64, ctypes.c_float, V, I64], ctypes.c_int),
        "hss_ulvfactsolve_f32": ([V, I32, V, I64, V, I64], ctypes.c_int),
        "hss_ulv_solve_f32": ([V, I32, V, I64, V, I64], ctypes.c_int),
        "hss_mat_rdiv": ([V, V, D, D, PV], ctypes.c_int),
        "hss_mat_generators": ([V, V, I32, ctypes.POINTER(I32), V, I64, ctypes.POINTER(I64)], ctypes.c_int),
        "hss_matmul": ([V, I32, D, V, I64, D, V, I64], ctypes.c_int),
        "hss_matmul_dev": ([V, I32, D, V, I64, D, V, I64], ctypes.c_int),
        "hss_ulvfactsolve": ([V, I32, V, I64, V, I64], ctypes.c_int),
        "hss_ulv_factor": ([V, ctypes.c_int, PV], ctypes.c_int),
        "hss_fact_destroy": ([V], ctypes.c_int),
        "hss_fact_device_bytes": ([V], I64),
        "hss_ulv_solve": ([V, I32, V, I64, V, I64, ctypes.POINTER(RefineOpts), ctypes.POINTER(SolveReport)],
                          ctypes.c_int),
        "hss_ulv_solve_dev": ([V, I32, V, I64, V, I64, ctypes.POINTER(RefineOpts), ctypes.POINTER(SolveReport)],
                              ctypes.c_int),
        "hss_ulv_solve_dev_noalloc": ([V, I32, V, I64, V, I64], ctypes.c_int),
        "hss_ulv_solve_batch_dev": ([I32, V, V, V, V, V, V], ctypes.c_int),
        "hss_ulv_factor_batch": ([I32, V, ctypes.c_int, V], ctypes.c_int),
        "hss_last_error": ([], ctypes.c_char_p),
        "hss_last_device_ms": ([], D),
        "hss_prof_enable": ([ctypes.c_int], None),
        "hss_prof_reset": ([], None),
        "hss_prof_count": ([], I32),
        "hss_prof_get": ([I32, ctypes.POINTER(ctypes.c_char_p), ctypes.POINTER(I32), ctypes.POINTER(I32),
                          ctypes.POINTER(D)], ctypes.c_int),
        "hss_fact_node_info": ([V, I32, ctypes.POINTER(I32)], ctypes.c_int),
        "hss_mat_matvec_info": ([V, ctypes.POINTER(I32), ctypes.POINTER(I32)], ctypes.c_int),
        "hss_plan_create": ([I32, V, V, V, V, V, V, I32, PV], ctypes.c_int),
        "hss_plan_execute": ([V, ctypes.POINTER(RefineOpts), ctypes.POINTER(SolveReport)], ctypes.c_int),
        "hss_plan_destroy": ([V], ctypes.c_int),
        "hss_plan_tune": ([V, I32, ctypes.POINTER(D), ctypes.POINTER(D)], ctypes.c_int),
        "hss_plan_tune_choice": ([V, I32, I32, ctypes.POINTER(I32), ctypes.POINTER(I32), ctypes.POINTER(I32)],
                                 ctypes.c_int),
        "hss_plan_device_bytes": ([V], I64),
        "hss_server_create": ([V, I64, ctypes.c_int, ctypes.POINTER(RefineOpts), I32, PV], ctypes.c_int),
        "hss_server_admit": ([V, V, ctypes.POINTER(I32)], ctypes.c_int),
        "hss_server_evict": ([V, I32], ctypes.c_int),
        "hss_server_used_bytes": ([V], I64),
        "hss_server_submit": ([V, I32, I32, V, I64, V, I64, ctypes.POINTER(I64)], ctypes.c_int),
        "hss_server_step": ([V, ctypes.POINTER(I32)], ctypes.c_int),
        "hss_server_poll": ([V, I64, ctypes.POINTER(I32), ctypes.POINTER(D), ctypes.POINTER(SolveReport)],
                            ctypes.c_int),
        "hss_server_destroy": ([V], ctypes.c_int),
    }
    for name, (args, res) in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    _lib = lib
    return lib


def _check(status: int):
    if status == HSS_OK:
        return
    msg = _lib.hss_last_error().decode(errors="replace")
    if status == HSS_ERR_DIM:
        raise DimensionMismatch(msg)
    if status == HSS_ERR_ARG:
        raise ValueError(msg)
    if status == HSS_ERR_SINGULAR:
        raise SingularException(msg)
    if status == HSS_ERR_NOCONV:
        raise NotConverged(status, msg)
    if status == HSS_ERR_IO:
        raise OSError(msg)
    raise HssError(status, msg)


def _refine_opts(tol, max_iter, method):
    """hss_refine_opts of a single solve: none unless tol or max_iter is given; mode 1
    (iterative refinement) or 2 (GMRES-IR) when tol > 0."""
    if method not in ("ir", "gmres"):
        raise ValueError(f"method must be 'ir' or 'gmres', not {method!r}")
    if tol is None and max_iter is None:
        return None
    mode = (2 if method == "gmres" else 1) if (tol is not None and tol > 0) else 0
    return RefineOpts(tol if tol is not None else 0.0, max_iter if max_iter is not None else 10, mode)


def last_device_ms() -> float:
    return float(load_library().hss_last_device_ms())


def prof_enable(on: bool = True):
    """Time every library kernel launch (``hss_prof_enable``): one record per launch, timed by
    the kernel's own start and end events."""
    load_library().hss_prof_enable(1 if on else 0)


def prof_reset():
    load_library().hss_prof_reset()


def prof_records():
    """[{kernel, level, nwork, ms}] of the launches recorded since the last reset (waits for them)."""
    lib = load_library()
    out = []
    for i in range(lib.hss_prof_count()):
        name = ctypes.c_char_p()
        lev = ctypes.c_int32()
        nw = ctypes.c_int32()
        ms = ctypes.c_double()
        _check(lib.hss_prof_get(i, ctypes.byref(name), ctypes.byref(lev), ctypes.byref(nw), ctypes.byref(ms)))
        out.append({"kernel": name.value.decode(), "level": lev.value, "nwork": nw.value, "ms": ms.value})
    return out


class Context:
    """One device and its HIP stream (``hss_ctx_create``)."""

    def __init__(self, device: int = 0):
        lib = load_library()
        h = ctypes.c_void_p()
        _check(lib.hss_ctx_create(device, ctypes.byref(h)))
        self._h = h
        self.device = device

    @property
    def handle(self):
        return self._h

    @property
    def stream(self) -> int:
        return int(_lib.hss_ctx_stream(self._h) or 0)

    def close(self):
        if self._h:
            _lib.hss_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default_ctx = None


def default_context() -> Context:
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx


def _as_f64_fortran(B):
    B = np.asarray(B, dtype=np.float64)
    vec = B.ndim == 1
    if vec:
        B = B.reshape(-1, 1)
    return np.asfortranarray(B), vec


class _Adopt:
    """A matrix handle created by the library, with its host node table."""

    def __init__(self, handle, nodes):
        self.handle = handle
        self.nodes = nodes


class HipHssMatrix:
    """Device-resident HSS matrix (``hss_mat_create``); the drop-in for ``HssMatrix{Float64}``."""

    # numpy operands hand ``X @ A`` and ``X / A`` to __rmatmul__ / __rtruediv__
    __array_ufunc__ = None

    def __init__(self, hss, ctx: Context | None = None):
        lib = load_library()
        self.ctx = ctx or default_context()
        if isinstance(hss, _Adopt):  # a handle made by the library (hss_mat_adjoint)
            self.nodes = hss.nodes
            self._h = hss.handle
            self.shape = (int(self.nodes["m"][0]), int(self.nodes["n"][0]))
            return
        if isinstance(hss, HssMatrix):
            nodes, blob = pack(hss)
        else:
            nodes, blob = hss
        self.nodes = np.ascontiguousarray(nodes, dtype=NODE_DTYPE)
        h = ctypes.c_void_p()
        if isinstance(blob, np.ndarray) and blob.dtype == np.float32:
            # HssMatrix{Float32} generators (hss_mat_create_f32: widened to Float64 on creation)
            blob = np.ascontiguousarray(blob)
            _check(lib.hss_mat_create_f32(self.ctx.handle, self.nodes.ctypes.data, len(self.nodes),
                                          blob.ctypes.data, blob.size, ctypes.byref(h)))
        else:
            blob = np.ascontiguousarray(blob, dtype=np.float64)
            _check(lib.hss_mat_create(self.ctx.handle, self.nodes.ctypes.data, len(self.nodes), blob.ctypes.data,
                                      blob.size, ctypes.byref(h)))
        self._h = h
        self.shape = (int(self.nodes["m"][0]), int(self.nodes["n"][0]))

    @property
    def handle(self):
        return self._h

    # ---- generator files (hss_gen_header layout) ----
    @classmethod
    def load(cls, path, ctx: Context | None = None) -> "HipHssMatrix":
        """A device matrix from a generator file (``hss_mat_load``)."""
        lib = load_library()
        ctx = ctx or default_context()
        nodes, _ = gen_read(path)
        h = ctypes.c_void_p()
        _check(lib.hss_mat_load(ctx.handle, str(path).encode(), ctypes.byref(h)))
        return cls(_Adopt(h, nodes), ctx)

    def generators(self):
        """The matrix's ``(nodes, blob)`` in host memory (``hss_mat_generators``): the layout
        ``pack`` produces and ``unpack`` turns back into an ``HssMatrix``."""
        nn, bl = ctypes.c_int32(), ctypes.c_int64()
        _check(_lib.hss_mat_generators(self._h, None, 0, ctypes.byref(nn), None, 0, ctypes.byref(bl)))
        nodes = np.zeros(nn.value, dtype=NODE_DTYPE)
        blob = np.zeros(max(bl.value, 1), dtype=np.float64)
        _check(_lib.hss_mat_generators(self._h, nodes.ctypes.data, nn.value, ctypes.byref(nn), blob.ctypes.data,
                                       blob.size, ctypes.byref(bl)))
        return nodes, blob[:bl.value]

    def save(self, path) -> None:
        """The matrix's generators, read back from HBM, to a generator file (``hss_mat_save``)."""
        _check(_lib.hss_mat_save(self._h, str(path).encode()))

    def device_bytes(self) -> int:
        return int(_lib.hss_mat_device_bytes(self._h))

    def matvec_info(self) -> dict:
        """The matvec's tree cut depth (0: none) and folded leaves (``hss_mat_matvec_info``)."""
        c, f = ctypes.c_int32(), ctypes.c_int32()
        _check(_lib.hss_mat_matvec_info(self._h, ctypes.byref(c), ctypes.byref(f)))
        return {"cut_depth": c.value, "folded_leaves": f.value}

    def size(self, dim=None):
        return self.shape if dim is None else self.shape[dim - 1]

    def close(self):
        if getattr(self, "_h", None):
            _lib.hss_mat_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- matvec ----
    def __matmul__(self, B):
        """``hssA * B`` (reference ``matmul.jl:13,15``); with B a device HSS matrix, the HSS
        product ``hssA * hssB`` (``matmul.jl:65-118``, ``hss_mat_mul``) as a new one."""
        if isinstance(B, HipHssMatrix):
            return self.hss_mul(B)
        if isinstance(B, np.ndarray) and B.dtype == np.float32:
            return _f32_call(self, B, "matmul")
        Bf, vec = _as_f64_fortran(B)
        if Bf.shape[0] != self.shape[1]:
            raise DimensionMismatch(
                f"First dimension of B does not match second dimension of A. Expected {self.shape[1]}, "
                f"got {Bf.shape[0]}")
        C = np.zeros((self.shape[0], Bf.shape[1]), order="F")
        mul(C, self, Bf, 1.0, 0.0)
        return C[:, 0] if vec else C

    def hss_mul(self, B: "HipHssMatrix") -> "HipHssMatrix":
        """``hssA * hssB`` (``matmul.jl:65-118``): ranks add, nothing is recompressed."""
        h = ctypes.c_void_p()
        _check(_lib.hss_mat_mul(self._h, B._h, ctypes.byref(h)))
        nodes = np.zeros(len(self.nodes), dtype=NODE_DTYPE)
        # the product's node table, as hss_mat_mul lays it out (sizes and offsets of rows
        # from A, of columns from B); only its shape fields are read on the host
        nodes[:] = self.nodes
        nodes["n"], nodes["col_off"] = B.nodes["n"], B.nodes["col_off"]
        nodes["ku"] = self.nodes["ku"] + B.nodes["ku"]
        nodes["kv"] = self.nodes["kv"] + B.nodes["kv"]
        return HipHssMatrix(_Adopt(h, nodes), self.ctx)

    def recompress(self, atol: float = 1e-9, rtol: float = 1e-9) -> "HipHssMatrix":
        """``recompress!(hssA; atol, rtol)`` (``compression.jl:148-260``) as a new device matrix
        (``hss_mat_recompress``): orthonormal bases, then every coupling and basis truncated
        top-down; the ranks never grow."""
        h = ctypes.c_void_p()
        _check(_lib.hss_mat_recompress(self._h, float(atol), float(rtol), ctypes.byref(h)))
        nodes = np.zeros(len(self.nodes), dtype=NODE_DTYPE)
        nn, bl = ctypes.c_int32(), ctypes.c_int64()
        st = _lib.hss_mat_generators(h, nodes.ctypes.data, len(nodes), ctypes.byref(nn), None, 0, ctypes.byref(bl))
        if st != HSS_OK:
            _lib.hss_mat_destroy(h)
            _check(st)
        return HipHssMatrix(_Adopt(h, nodes), self.ctx)

    def _adopted(self, h) -> "HipHssMatrix":
        nodes = np.zeros(len(self.nodes), dtype=NODE_DTYPE)
        nn, bl = ctypes.c_int32(), ctypes.c_int64()
        st = _lib.hss_mat_generators(h, nodes.ctypes.data, len(nodes), ctypes.byref(nn), None, 0, ctypes.byref(bl))
        if st != HSS_OK:
            _lib.hss_mat_destroy(h)
            _check(st)
        return HipHssMatrix(_Adopt(h, nodes), self.ctx)

    def hss_ldiv(self, B: "HipHssMatrix", atol: float = 1e-9, rtol: float = 1e-9) -> "HipHssMatrix":
        """``hssA \\ hssB`` (``hssmatrix.jl:235``, ``ulvdivide.jl``): A^-1 B as an HSS matrix
        (``hss_mat_ldiv``)."""
        h = ctypes.c_void_p()
        _check(_lib.hss_mat_ldiv(self._h, B._h, float(atol), float(rtol), ctypes.byref(h)))
        return self._adopted(h)

    def hss_rdiv(self, B: "HipHssMatrix", atol: float = 1e-9, rtol: float = 1e-9) -> "HipHssMatrix":
        """``hssA / hssB`` (``hssmatrix.jl:237``): A B^-1 as an HSS matrix (``hss_mat_rdiv``)."""
        h = ctypes.c_void_p()
        _check(_lib.hss_mat_rdiv(self._h, B._h, float(atol), float(rtol), ctypes.byref(h)))
        return self._adopted(h)

    def matmul_dev(self, nrhs: int, alpha: float, dB: int, ldb: int, beta: float, dC: int, ldc: int):
        """``hss_matmul_dev`` on device pointers (e.g. ``torch.Tensor.data_ptr()``)."""
        _check(_lib.hss_matmul_dev(self._h, nrhs, alpha, dB, ldb, beta, dC, ldc))

    # ---- adjoint (reference hssmatrix.jl:165-171) ----
    def adjoint(self) -> "HipHssMatrix":
        """``hssA'`` as a new device matrix (``hss_mat_adjoint``): D', U <-> V, B12 <- B21',
        B21 <- B12', R <-> W."""
        h = ctypes.c_void_p()
        _check(_lib.hss_mat_adjoint(self._h, ctypes.byref(h)))
        t = self.nodes.copy()
        t["m"], t["n"] = self.nodes["n"], self.nodes["m"]
        t["ku"], t["kv"] = self.nodes["kv"], self.nodes["ku"]
        t["row_off"], t["col_off"] = self.nodes["col_off"], self.nodes["row_off"]
        return HipHssMatrix(_Adopt(h, t), self.ctx)

    @property
    def H(self) -> "HipHssMatrix":
        return self.adjoint()

    def __rmatmul__(self, A):
        """``A * hssB`` = ``(hssB' * A')'`` (reference ``matmul.jl:14``)."""
        Af = np.asarray(A, dtype=np.float64)
        if Af.ndim != 2 or Af.shape[1] != self.shape[0]:
            raise DimensionMismatch(f"Second dimension of A does not match first dimension of B. "
                                    f"Expected {self.shape[0]}, got {Af.shape[-1]}")
        At = self.adjoint()
        try:
            return (At @ np.asfortranarray(Af.T)).T
        finally:
            At.close()

    def rdiv(self, A):
        """``A / hssB`` = ``ulvfactsolve(hssB', A')'`` (reference ``hssmatrix.jl:236``)."""
        Af = np.asarray(A, dtype=np.float64)
        if Af.ndim != 2 or Af.shape[1] != self.shape[1]:
            raise DimensionMismatch(f"Second dimension of A does not match second dimension of B. "
                                    f"Expected {self.shape[1]}, got {Af.shape[-1]}")
        At = self.adjoint()
        try:
            return ulvfactsolve(At, np.asfortranarray(Af.T)).T
        finally:
            At.close()

    __rtruediv__ = rdiv

    # ---- ULV ----
    def ulvfactor(self, prec: str = "fp64") -> "HipUlvFactor":
        return HipUlvFactor(self, prec)

    def solve(self, B):
        """``hssA \\ B`` (reference ``hssmatrix.jl:234`` -> ``ulvfactsolve``)."""
        if isinstance(B, np.ndarray) and B.dtype == np.float32:
            return _f32_call(self, B, "solve")
        return ulvfactsolve(self, B)


def _f32_call(A, B, what: str, fact=None):
    """Float32 data at the boundary (``hss_matmul_f32`` / ``hss_ulvfactsolve_f32`` /
    ``hss_ulv_solve_f32``): the device computes in Float64, the result is Float32."""
    vec = B.ndim == 1
    Bf = np.asfortranarray(B.reshape(-1, 1) if vec else B, dtype=np.float32)
    rows = A.shape[0]
    if what == "matmul":
        if Bf.shape[0] != A.shape[1]:
            raise DimensionMismatch(f"First dimension of B does not match second dimension of A. Expected "
                                    f"{A.shape[1]}, got {Bf.shape[0]}")
        C = np.zeros((rows, Bf.shape[1]), dtype=np.float32, order="F")
        _check(_lib.hss_matmul_f32(A.handle, Bf.shape[1], 1.0, Bf.ctypes.data, max(Bf.shape[0], 1), 0.0,
                                   C.ctypes.data, max(rows, 1)))
    else:
        if Bf.shape[0] != rows:
            raise DimensionMismatch(f"matrix is {A.shape}, right-hand side has {Bf.shape[0]} rows")
        C = np.zeros((A.shape[1], Bf.shape[1]), dtype=np.float32, order="F")
        if fact is None:
            _check(_lib.hss_ulvfactsolve_f32(A.handle, Bf.shape[1], Bf.ctypes.data, max(rows, 1), C.ctypes.data,
                                             max(A.shape[1], 1)))
        else:
            _check(_lib.hss_ulv_solve_f32(fact, Bf.shape[1], Bf.ctypes.data, max(rows, 1), C.ctypes.data,
                                          max(A.shape[1], 1)))
    return C[:, 0] if vec else C


def gen_write(path, nodes, blob) -> None:
    """Write a node table and blob as a generator file with the library (``hss_gen_write``;
    host only)."""
    lib = load_library()
    nodes = np.ascontiguousarray(nodes, dtype=NODE_DTYPE)
    blob = np.ascontiguousarray(blob, dtype=np.float64)
    _check(lib.hss_gen_write(str(path).encode(), nodes.ctypes.data, len(nodes), blob.ctypes.data, blob.size))


def gen_read(path):
    """(nodes, blob) of a generator file read by the library (``hss_gen_read``; host only)."""
    lib = load_library()
    nn, bl = ctypes.c_int32(), ctypes.c_int64()
    p = str(path).encode()
    _check(lib.hss_gen_read(p, None, 0, ctypes.byref(nn), None, 0, ctypes.byref(bl)))
    nodes = np.zeros(nn.value, dtype=NODE_DTYPE)
    blob = np.zeros(max(bl.value, 1), dtype=np.float64)
    _check(lib.hss_gen_read(p, nodes.ctypes.data, nn.value, ctypes.byref(nn), blob.ctypes.data, bl.value,
                            ctypes.byref(bl)))
    return nodes, blob[:bl.value]


def mul(C, hssA: HipHssMatrix, B, alpha: float = 1.0, beta: float = 0.0):
    """``mul!(C, hssA, B, alpha, beta)`` (reference ``matmul.jl:18-28``); C is updated in place."""
    if not isinstance(C, np.ndarray) or C.dtype != np.float64 or not C.flags.f_contiguous:
        raise ValueError("C must be a Fortran-ordered float64 array")
    Bf, _ = _as_f64_fortran(B)
    C2 = C.reshape(C.shape[0], -1, order="F")
    if Bf.shape[0] != hssA.shape[1]:
        raise DimensionMismatch(
            f"First dimension of B does not match second dimension of A. Expected {hssA.shape[1]}, got {Bf.shape[0]}")
    if C2.shape != (hssA.shape[0], Bf.shape[1]):
        raise DimensionMismatch("Dimensions of C don't match up with A and B.")
    nrhs = Bf.shape[1]
    _check(_lib.hss_matmul(hssA.handle, nrhs, float(alpha), Bf.ctypes.data, max(Bf.shape[0], 1), float(beta),
                           C2.ctypes.data, max(C2.shape[0], 1)))
    return C


def compress_dev(dA: int, lda: int, rcl, ccl=None, atol: float = 1e-9, rtol: float = 1e-9,
                 ctx: Context | None = None) -> HipHssMatrix:
    """Direct HSS compression of a dense device matrix (``hss_compress_dev``), the device
    counterpart of ``compress(A, rcl, ccl; atol, rtol)`` (reference ``compression.jl:47-63``).

    ``dA`` is the address of the column-major matrix (``lda`` >= its rows; a torch tensor
    ``T`` of shape (cols, rows) in row-major order is such a matrix with ``lda = rows``).
    ``rcl`` / ``ccl`` are ``ClusterTree``s (``ccl`` defaults to ``rcl``).  The device work is
    enqueued on the context's stream after a device-wide synchronization, so buffers that
    torch filled on its own stream are complete."""
    lib = load_library()
    ctx = ctx or default_context()
    ccl = rcl if ccl is None else ccl
    nodes = cluster_nodes(rcl, ccl)
    try:
        import torch
        if torch.cuda.is_available():
            torch.cuda.synchronize()
    except ImportError:  # pragma: no cover
        pass
    h = ctypes.c_void_p()
    _check(lib.hss_compress_dev(ctx.handle, dA, lda, nodes.ctypes.data, len(nodes), float(atol), float(rtol),
                                ctypes.byref(h)))
    # the node table with the ranks and offsets found; the blob stays on the device
    found = np.zeros(len(nodes), dtype=NODE_DTYPE)
    nn, bl = ctypes.c_int32(), ctypes.c_int64()
    st = lib.hss_mat_generators(h, found.ctypes.data, len(found), ctypes.byref(nn), None, 0, ctypes.byref(bl))
    if st != HSS_OK:
        lib.hss_mat_destroy(h)
        _check(st)
    return HipHssMatrix(_Adopt(h, found), ctx)


def randcompress_dev(dA: int, lda: int, rcl, ccl=None, atol: float = 1e-9, rtol: float = 1e-9, kest: int = 10,
                     noversampling: int = 10, stepsize: int = 20, ctx: Context | None = None,
                     report: bool = False):
    """Randomized adaptive HSS compression of a dense device matrix (``hss_randcompress_dev``),
    the device counterpart of ``randcompress_adaptive(A, rcl, ccl; kest, atol, rtol,
    noversampling, stepsize)`` (reference ``compression.jl:312-430``; defaults as the
    reference's ``HssOptions``).  Arguments as ``compress_dev``.  With ``report`` returns
    (matrix, the sample width it was built with)."""
    lib = load_library()
    ctx = ctx or default_context()
    ccl = rcl if ccl is None else ccl
    nodes = cluster_nodes(rcl, ccl)
    try:
        import torch
        if torch.cuda.is_available():
            torch.cuda.synchronize()
    except ImportError:  # pragma: no cover
        pass
    h = ctypes.c_void_p()
    smp = ctypes.c_int32()
    _check(lib.hss_randcompress_dev(ctx.handle, dA, lda, nodes.ctypes.data, len(nodes), float(atol), float(rtol),
                                    int(kest), int(noversampling), int(stepsize), ctypes.byref(h), ctypes.byref(smp)))
    found = np.zeros(len(nodes), dtype=NODE_DTYPE)
    nn, bl = ctypes.c_int32(), ctypes.c_int64()
    st = lib.hss_mat_generators(h, found.ctypes.data, len(found), ctypes.byref(nn), None, 0, ctypes.byref(bl))
    if st != HSS_OK:
        lib.hss_mat_destroy(h)
        _check(st)
    A = HipHssMatrix(_Adopt(h, found), ctx)
    return (A, int(smp.value)) if report else A


def ulvfactsolve(hssA: HipHssMatrix, b):
    """``ulvfactsolve(hssA, b)`` (reference ``ulvfactor.jl:10-19``): factor + solve in one call."""
    bf, vec = _as_f64_fortran(b)
    n = hssA.shape[1]
    if hssA.shape[0] != hssA.shape[1] or bf.shape[0] != hssA.shape[0]:
        raise DimensionMismatch(f"matrix is {hssA.shape}, right-hand side has {bf.shape[0]} rows")
    X = np.zeros((n, bf.shape[1]), order="F")
    _check(_lib.hss_ulvfactsolve(hssA.handle, bf.shape[1], bf.ctypes.data, max(n, 1), X.ctypes.data, max(n, 1)))
    return X[:, 0] if vec else X


ldiv = ulvfactsolve


class HipUlvFactor:
    """Resident ULV factorization (``hss_ulv_factor``)."""

    def __init__(self, hssA: HipHssMatrix, prec: str = "fp64"):
        self.mat = hssA
        self.prec = prec
        p = PREC_FP32 if prec in ("fp32", "mixed", PREC_FP32) else PREC_FP64
        h = ctypes.c_void_p()
        _check(_lib.hss_ulv_factor(hssA.handle, p, ctypes.byref(h)))
        self._h = h
        self.factor_ms = last_device_ms()

    @classmethod
    def _wrap(cls, hssA, prec, handle):
        obj = cls.__new__(cls)
        obj.mat = hssA
        obj.prec = prec
        obj._h = handle
        obj.factor_ms = last_device_ms()
        return obj

    @property
    def handle(self):
        return self._h

    def device_bytes(self) -> int:
        return int(_lib.hss_fact_device_bytes(self._h))

    def node_info(self, node: int) -> dict:
        """Factor sizes of one node (``hss_fact_node_info``)."""
        buf = (ctypes.c_int32 * 10)()
        _check(_lib.hss_fact_node_info(self._h, node, buf))
        keys = ("m", "n", "k", "nk", "mr", "nr", "pass", "nblk", "My", "Kx")
        return dict(zip(keys, list(buf)))

    def close(self):
        if getattr(self, "_h", None):
            _lib.hss_fact_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def solve(self, b, tol: float | None = None, max_iter: int | None = None, report: bool = False,
              method: str = "ir"):
        """X = A \\ b with the resident factor; FP32 factors refine to ``tol`` (default 1e-10).
        ``method``: "ir" (each correction is one solve with the factor) or "gmres" (GMRES-IR:
        each correction by Float64 GMRES preconditioned with the factor).
        Float32 data (a numpy float32 array) go through ``hss_ulv_solve_f32``."""
        if isinstance(b, np.ndarray) and b.dtype == np.float32 and tol is None and max_iter is None and not report:
            return _f32_call(self.mat, b, "solve", fact=self._h)
        bf, vec = _as_f64_fortran(b)
        n = self.mat.shape[0]
        if bf.shape[0] != n:
            raise DimensionMismatch(f"matrix is {self.mat.shape}, right-hand side has {bf.shape[0]} rows")
        X = np.zeros((n, bf.shape[1]), order="F")
        opts = _refine_opts(tol, max_iter, method)
        rep = SolveReport()
        _check(_lib.hss_ulv_solve(self._h, bf.shape[1], bf.ctypes.data, max(n, 1), X.ctypes.data, max(n, 1),
                                  ctypes.byref(opts) if opts is not None else None, ctypes.byref(rep)))
        out = X[:, 0] if vec else X
        if report:
            return out, {"iterations": rep.iterations, "rel_residual": rep.rel_residual, "seconds": rep.seconds}
        return out

    def solve_dev(self, nrhs: int, dB: int, ldb: int, dX: int, ldx: int, tol: float | None = None,
                  max_iter: int | None = None, method: str = "ir"):
        opts = _refine_opts(tol, max_iter, method)
        rep = SolveReport()
        _check(_lib.hss_ulv_solve_dev(self._h, nrhs, dB, ldb, dX, ldx,
                                      ctypes.byref(opts) if opts is not None else None, ctypes.byref(rep)))
        return {"iterations": rep.iterations, "rel_residual": rep.rel_residual, "seconds": rep.seconds}


def ulvfactor_batch(mats, prec: str = "fp64"):
    """Factor many matrices with one launch per tree level (``hss_ulv_factor_batch``)."""
    lib = load_library()
    p = PREC_FP32 if prec in ("fp32", "mixed") else PREC_FP64
    hs = (ctypes.c_void_p * len(mats))(*[m.handle.value for m in mats])
    out = (ctypes.c_void_p * len(mats))()
    _check(lib.hss_ulv_factor_batch(len(mats), ctypes.cast(hs, ctypes.c_void_p), p, ctypes.cast(out, ctypes.c_void_p)))
    return [HipUlvFactor._wrap(m, prec, ctypes.c_void_p(out[i])) for i, m in enumerate(mats)]


def solve_batch_dev(facts, nrhs, dB, ldb, dX, ldx):
    """Batched solves on device pointers, one launch per tree level (``hss_ulv_solve_batch_dev``)."""
    lib = load_library()
    cnt = len(facts)
    hs = (ctypes.c_void_p * cnt)(*[f.handle.value for f in facts])
    nr = (ctypes.c_int32 * cnt)(*nrhs)
    b = (ctypes.c_void_p * cnt)(*dB)
    lb = (ctypes.c_int64 * cnt)(*ldb)
    x = (ctypes.c_void_p * cnt)(*dX)
    lx = (ctypes.c_int64 * cnt)(*ldx)
    _check(lib.hss_ulv_solve_batch_dev(cnt, ctypes.cast(hs, ctypes.c_void_p), ctypes.cast(nr, ctypes.c_void_p),
                                       ctypes.cast(b, ctypes.c_void_p), ctypes.cast(lb, ctypes.c_void_p),
                                       ctypes.cast(x, ctypes.c_void_p), ctypes.cast(lx, ctypes.c_void_p)))


class Plan:
    """Batched solve plan (``hss_plan_create``): one launch per tree level across all requests.

    ``facts`` may repeat and differ in size; ``dB``/``dX`` are device pointers
    (column-major, leading dimensions ``ldb``/``ldx``) fixed for the plan's life.
    """

    def __init__(self, facts, nrhs, dB, ldb, dX, ldx, refinement: bool = False):
        lib = load_library()
        cnt = len(facts)
        self.facts = list(facts)  # keep the factors alive
        hs = (ctypes.c_void_p * cnt)(*[f.handle.value for f in facts])
        nr = (ctypes.c_int32 * cnt)(*nrhs)
        b = (ctypes.c_void_p * cnt)(*dB)
        lb = (ctypes.c_int64 * cnt)(*ldb)
        x = (ctypes.c_void_p * cnt)(*dX)
        lx = (ctypes.c_int64 * cnt)(*ldx)
        h = ctypes.c_void_p()
        _check(lib.hss_plan_create(cnt, ctypes.cast(hs, ctypes.c_void_p), ctypes.cast(nr, ctypes.c_void_p),
                                   ctypes.cast(b, ctypes.c_void_p), ctypes.cast(lb, ctypes.c_void_p),
                                   ctypes.cast(x, ctypes.c_void_p), ctypes.cast(lx, ctypes.c_void_p),
                                   1 if refinement else 0, ctypes.byref(h)))
        self._h = h

    def tune(self, reps: int = 5, levels: int = 0):
        """Autotune the per-level launch choices (``hss_plan_tune``): returns the plan's solve
        time before and after (ms) and, for the first ``levels`` up/down levels, the choices
        as (form, tw, nt) with form 0 resident, 1 streamed, 2 chunked."""
        b, a = ctypes.c_double(), ctypes.c_double()
        _check(_lib.hss_plan_tune(self._h, reps, ctypes.byref(b), ctypes.byref(a)))
        out = {"ms_before": b.value, "ms_after": a.value, "up": [], "down": []}
        for d, key in ((0, "up"), (1, "down")):
            for lv in range(levels):
                f, t, n = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
                if _lib.hss_plan_tune_choice(self._h, d, lv, ctypes.byref(f), ctypes.byref(t), ctypes.byref(n)):
                    break
                out[key].append((f.value, t.value, n.value))
        return out

    def execute(self, tol: float | None = None, max_iter: int = 10):
        """Enqueue the batched solve; with ``tol`` refine in lockstep (synchronous)."""
        opts = RefineOpts(tol, max_iter, 1) if tol is not None else None
        rep = SolveReport()
        _check(_lib.hss_plan_execute(self._h, ctypes.byref(opts) if opts is not None else None, ctypes.byref(rep)))
        return {"iterations": rep.iterations, "rel_residual": rep.rel_residual, "seconds": rep.seconds}

    def device_bytes(self) -> int:
        return int(_lib.hss_plan_device_bytes(self._h))

    def close(self):
        if getattr(self, "_h", None):
            _lib.hss_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Serving engine (``hss_server_*``): resident factors under an HBM budget and
    continuous batching of solves (first solves and refinement steps of all
    requests in flight share one launch per tree level)."""

    def __init__(self, budget_bytes: int, prec: str = "fp32", tol: float = 1e-10, max_iter: int = 10,
                 max_inflight: int = 64, ctx: Context | None = None):
        lib = load_library()
        self.ctx = ctx or default_context()
        p = PREC_FP32 if prec in ("fp32", "mixed") else PREC_FP64
        opts = RefineOpts(tol, max_iter, 1)
        h = ctypes.c_void_p()
        _check(lib.hss_server_create(self.ctx.handle, int(budget_bytes), p, ctypes.byref(opts), max_inflight,
                                     ctypes.byref(h)))
        self._h = h
        self.mats = {}

    def admit(self, mat: HipHssMatrix) -> int:
        mid = ctypes.c_int32()
        _check(_lib.hss_server_admit(self._h, mat.handle, ctypes.byref(mid)))
        self.mats[mid.value] = mat
        return mid.value

    def evict(self, mid: int):
        _check(_lib.hss_server_evict(self._h, mid))
        self.mats.pop(mid, None)

    def used_bytes(self) -> int:
        return int(_lib.hss_server_used_bytes(self._h))

    def submit(self, mid: int, nrhs: int, dB: int, ldb: int, dX: int, ldx: int) -> int:
        t = ctypes.c_int64()
        _check(_lib.hss_server_submit(self._h, mid, nrhs, dB, ldb, dX, ldx, ctypes.byref(t)))
        return t.value

    def step(self) -> int:
        out = ctypes.c_int32()
        _check(_lib.hss_server_step(self._h, ctypes.byref(out)))
        return out.value

    def poll(self, ticket: int):
        """None while running; else {status, latency_s, iterations, rel_residual}."""
        st = ctypes.c_int32()
        lat = ctypes.c_double()
        rep = SolveReport()
        _check(_lib.hss_server_poll(self._h, ticket, ctypes.byref(st), ctypes.byref(lat), ctypes.byref(rep)))
        if st.value == 0:
            return None
        return {"status": st.value, "latency_s": lat.value, "iterations": rep.iterations,
                "rel_residual": rep.rel_residual}

    def close(self):
        if getattr(self, "_h", None):
            _lib.hss_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
