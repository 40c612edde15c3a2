"""Direct HSS compression on the host (input producer for the GPU path).

Restates the reference's direct compression, ``src/compression.jl:47-133``
(Xia, Chandrasekaran, Gu, Li 2010): a post-order recursion that carries the
not-yet-compressed HSS block row ``Brow`` and block column ``Bcol`` of every
node, compresses them with a rank-revealing QR at each non-root node, and
reads the couplings ``B12``/``B21`` off the compressed blocks of the siblings.

Two deliberate differences from the reference, both outside the GPU path:

* The matrix may be given lazily as a kernel (``KernelMatrix``): the
  uncompressed pieces of ``Brow``/``Bcol`` are only evaluated at the leaves, so
  an n = 65536 kernel matrix never exists densely in host memory.  The
  arithmetic is the same as with the dense matrix.
* The rank-revealing step.  The reference calls ``LowRankApprox.pqrfact``
  (``compression.jl:22-30``), a third-party package that is not available here;
  its truncation is restated as column-pivoted Householder QR (LAPACK
  ``dgeqp3``) truncated at the first diagonal entry of R with
  ``|R[k,k]| <= max(atol, rtol*|R[1,1]|)``.  For a wide block (more columns
  than twice the rows) the pivoted QR runs on ``R1^T`` where ``A^T = Q1 R1`` is
  an unpivoted QR: the column space and singular values are those of A, and
  the cost drops from O(m*N*m) pivoting passes to one blocked QR.  For very
  wide blocks (N > 4m) the range is taken from a fixed-seed Gaussian sketch
  Y = A*Omega with m + 32 columns (rank(A) <= m, so range(Y) = range(A)), and
  the truncation is read off the pivoted QR of Y/sqrt(m+32).  Compression
  parity with the reference is therefore unpinned (its generators are not
  reproducible without LowRankApprox); every parity test of the GPU path feeds
  identical generators to the GPU and to the oracle instead.
"""
from __future__ import annotations

import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import scipy.linalg as sla

try:
    from threadpoolctl import ThreadpoolController as _TPC
except Exception:  # pragma: no cover
    _TPC = None
_TP = [None]


def _single_thread_blas():
    if _TPC is None:
        return None
    if _TP[0] is None:
        _TP[0] = _TPC()
    return _TP[0].limit(limits=1, user_api="blas")

try:  # torch's CPU LAPACK/BLAS (MKL, multi-threaded) is several times faster
    import torch as _torch  # than numpy's for the tall QR and the kernel blocks
except Exception:  # pragma: no cover - torch is part of this image
    _torch = None

from .hssmatrix import ClusterTree, HssMatrix, compatible

__all__ = ["KernelMatrix", "compress", "compress_block"]


class KernelMatrix:
    """Lazily evaluated kernel matrix A[i, j] = kernel(x[i], x[j]).

    ``kind`` selects the synthetic kernels of the benchmark workloads:

    * ``"cauchy"``:  1/(x_i - x_j) off the diagonal, ``diag`` on it
      (reference ``test/benchmarks.jl:9`` uses diag 10000; the workloads use 1);
    * ``"cauchy_test"``: ``(x-y) > 0 ? 0.001/(x-y) : 2`` (reference
      ``test/runtests.jl:19``);
    * ``"log"``:  log|x_i - x_j| off the diagonal, ``diag`` on it.
    """

    def __init__(self, x, kind: str = "cauchy", diag: float = 1.0):
        self.x = np.ascontiguousarray(np.asarray(x, dtype=np.float64))
        self.kind = kind
        self.diag = float(diag)
        self.shape = (self.x.size, self.x.size)

    def block(self, r0: int, r1: int, c0: int, c1: int) -> np.ndarray:
        """A[r0:r1, c0:c1] as a Fortran-ordered array (evaluated in column chunks on host threads)."""
        out = np.empty((r1 - r0, c1 - c0), order="F")
        nc = c1 - c0
        chunk = 2048
        if (r1 - r0) * nc <= 4 * 65536:
            self._fill(out, r0, r1, c0, c1)
        else:
            spans = [(a, min(a + chunk, nc)) for a in range(0, nc, chunk)]
            list(_pool().map(lambda ab: self._fill(out[:, ab[0]:ab[1]], r0, r1, c0 + ab[0], c0 + ab[1]), spans))
        return out

    def _fill(self, out, r0, r1, c0, c1):
        d = np.subtract.outer(self.x[r0:r1], self.x[c0:c1])
        with np.errstate(divide="ignore", invalid="ignore"):
            if self.kind == "cauchy":
                np.divide(1.0, d, out=out)
            elif self.kind == "cauchy_test":
                out[...] = np.where(d > 0, 0.001 / d, 2.0)
            elif self.kind == "log":
                np.abs(d, out=d)
                np.log(d, out=out)
            else:
                raise ValueError(f"unknown kernel {self.kind}")
        if self.kind != "cauchy_test":
            lo, hi = max(r0, c0), min(r1, c1)
            if lo < hi:
                idx = np.arange(lo, hi)
                out[idx - r0, idx - c0] = self.diag

    def dense(self) -> np.ndarray:
        n = self.shape[0]
        return self.block(0, n, 0, n)

    def device_dense(self, device="cuda:0"):
        """The dense matrix in device memory, built there by torch: a tensor T of shape
        (n, n) with T[j, i] = A[i, j], i.e. A column-major with leading dimension n (the
        input of ``hip.compress_dev``).  The same entries as ``dense`` up to the rounding of
        the device's division and logarithm."""
        import torch
        x = torch.from_numpy(self.x).to(device)
        d = x[None, :] - x[:, None]  # d[j, i] = x_i - x_j
        if self.kind == "cauchy":
            T = torch.reciprocal(d)
        elif self.kind == "cauchy_test":
            T = torch.where(d > 0, 0.001 / d, torch.full_like(d, 2.0))
        elif self.kind == "log":
            T = torch.log(torch.abs(d))
        else:
            raise ValueError(f"unknown kernel {self.kind}")
        del d
        if self.kind != "cauchy_test":
            T.fill_diagonal_(self.diag)
        return T


_POOL = None


def _pool():
    global _POOL
    if _POOL is None:
        _POOL = ThreadPoolExecutor(max_workers=min(16, os.cpu_count() or 1))
    return _POOL


class _Dense:
    """Dense matrix with the same ``block`` accessor as ``KernelMatrix``."""

    def __init__(self, A):
        self.A = np.asfortranarray(np.asarray(A, dtype=np.float64))
        self.shape = self.A.shape

    def block(self, r0, r1, c0, c1):
        return self.A[r0:r1, c0:c1]


def compress_block(A: np.ndarray, atol: float, rtol: float):
    """Rank-revealing split A ~= Q R with orthonormal Q (m x k), R (k x N).

    Counterpart of the reference's ``_compress_block!`` (``compression.jl:22-30``),
    which returns ``F.Q, F.R[:, invperm(F.p)]``; here ``R = Q^T A`` directly,
    which equals the un-permuted truncated R factor of the pivoted QR.
    """
    m, N = A.shape
    if m == 0 or N == 0:
        return np.zeros((m, 0), order="F"), np.zeros((0, N), order="F")
    if N > 4 * m and _torch is not None and m + 32 <= _SKETCH_COLS:
        # wide block: Y = A Omega spans range(A) (rank(A) <= m < s almost surely);
        # the pivoted QR of Y orders that range, and R = Q^T A below re-projects A
        s_cols = m + 32
        Om = _sketch(N, s_cols)
        B = np.asfortranarray((_torch.from_numpy(A) @ Om).numpy()) / np.sqrt(s_cols)
    elif N > 2 * m:
        if _torch is not None:
            # A is column-major, so A.T viewed row-major is (N x m) without a copy
            at = _torch.from_numpy(np.ascontiguousarray(A.T))
            r1 = _torch.linalg.qr(at, mode="r")[1].numpy()
        else:
            r1 = np.linalg.qr(np.ascontiguousarray(A.T), mode="r")  # (m x m)
        B = np.asfortranarray(r1.T)
    else:
        B = A
    lim = _single_thread_blas()  # small problem: BLAS threading only adds overhead
    try:
        Q, R, _ = sla.qr(B, mode="economic", pivoting=True)
    finally:
        if lim is not None:
            lim.restore_original_limits()
    d = np.abs(np.diag(R))
    if d.size == 0 or d[0] == 0.0:
        k = 0
    else:
        tol = max(atol, rtol * d[0])
        k = int(np.count_nonzero(d > tol))
        # pivoted QR has non-increasing |R[i,i]| in exact arithmetic; truncate at
        # the first entry below the tolerance
        below = np.nonzero(d <= tol)[0]
        if below.size:
            k = int(below[0])
    Qk = np.asfortranarray(Q[:, :k])
    if _torch is not None and k * N >= 65536:
        Rk = np.asfortranarray((_torch.from_numpy(np.ascontiguousarray(Qk.T)) @ _torch.from_numpy(A)).numpy())
    else:
        Rk = np.asfortranarray(Qk.T @ A)
    return Qk, Rk


_SKETCH = [None]
_SKETCH_COLS = 512


def _sketch(N: int, s: int):
    """Fixed-seed Gaussian test matrix: the leading N x s block of one cached matrix."""
    cur = _SKETCH[0]
    if cur is None or cur.shape[0] < N:
        rows = max(N, 65536 if cur is None else 2 * cur.shape[0])
        g = _torch.Generator().manual_seed(1234567)
        cur = _torch.randn(rows, _SKETCH_COLS, generator=g, dtype=_torch.float64)
        _SKETCH[0] = cur
    return cur[:N, :s]


# --- lazily evaluated block rows / columns ----------------------------------
# A block row is a list of pieces concatenated horizontally; a piece is either
# ("k", r0, r1, c0, c1) meaning A[r0:r1, c0:c1] or a dense ndarray.
# Block columns are lists of pieces stacked vertically.

def _rows(pieces, a, b):
    out = []
    for p in pieces:
        if isinstance(p, tuple):
            _, r0, r1, c0, c1 = p
            out.append(("k", r0 + a, r0 + b, c0, c1))
        else:
            out.append(p[a:b, :])
    return out


def _cols(pieces, a, b):
    out = []
    for p in pieces:
        if isinstance(p, tuple):
            _, r0, r1, c0, c1 = p
            out.append(("k", r0, r1, c0 + a, c0 + b))
        else:
            out.append(p[:, a:b])
    return out


def _hcat(Am, pieces, nrows):
    mats = [Am.block(p[1], p[2], p[3], p[4]) if isinstance(p, tuple) else p for p in pieces]
    mats = [m for m in mats if m.shape[1] > 0]
    if not mats:
        return np.zeros((nrows, 0), order="F")
    return np.asfortranarray(np.hstack(mats))


def _vcat(Am, pieces, ncols):
    mats = [Am.block(p[1], p[2], p[3], p[4]) if isinstance(p, tuple) else p for p in pieces]
    mats = [m for m in mats if m.shape[0] > 0]
    if not mats:
        return np.zeros((0, ncols), order="F")
    return np.asfortranarray(np.vstack(mats))


class _Progress:
    """Counts compressed leaves; calls ``cb(done, total)`` every ``every`` leaves and at the last."""

    def __init__(self, cb, total: int, every: int = 32):
        self.cb, self.total, self.every, self.done = cb, total, every, 0

    def leaf(self):
        self.done += 1
        if self.done % self.every == 0 or self.done == self.total:
            self.cb(self.done, self.total)


def _nleaves(c: ClusterTree) -> int:
    return 1 if c.isleaf() else _nleaves(c.left) + _nleaves(c.right)


def compress(A, rcl: ClusterTree, ccl: ClusterTree, atol: float = 1e-9, rtol: float = 1e-9,
             progress=None) -> HssMatrix:
    """Direct HSS compression, reference ``compression.jl:47-63``.

    ``A`` is a dense array or a ``KernelMatrix``.  Defaults atol = rtol = 1e-9
    follow ``src/HssMatrices.jl:62-63``.  ``progress(done, total)``, if given, is
    called every 32 compressed leaves and after the last (long compressions report
    as they go).
    """
    Am = A if hasattr(A, "block") else _Dense(A)
    if atol < 0:
        raise ValueError("atol")
    if rtol < 0:
        raise ValueError("rtol")
    if not compatible(rcl, ccl):
        raise ValueError("row and column clusters are not compatible")
    m, n = rcl.size, ccl.size
    if tuple(Am.shape) != (m, n):
        raise ValueError("size of row- and column-cluster-trees must match")
    if rcl.isleaf():
        return HssMatrix(D=Am.block(rcl.start, rcl.stop, ccl.start, ccl.stop), rootnode=True)
    prog = _Progress(progress, _nleaves(rcl)) if progress is not None else None
    h, _, _ = _compress(Am, [], [], m, n, rcl, ccl, atol, rtol, True, prog)
    return h


def _compress(Am, Brow, Bcol, nbrow_rows, nbcol_cols, rcl, ccl, atol, rtol, rootnode, prog=None):
    """Post-order compression; ``Brow``/``Bcol`` are piece lists.

    Returns (node, Brow, Bcol) with the compressed block row (k x W) and
    block column (H x k) as dense arrays, as ``compression.jl:66-133``.
    """
    if rcl.isleaf():
        # leaf, compression.jl:66-74
        D = Am.block(rcl.start, rcl.stop, ccl.start, ccl.stop)
        m, n = D.shape
        brow = _hcat(Am, Brow, m)
        bcol = _vcat(Am, Bcol, n)
        U, brow_c = compress_block(brow, atol, rtol)
        V, bcolT_c = compress_block(np.asfortranarray(bcol.T), atol, rtol)
        if prog is not None:
            prog.leaf()
        return HssMatrix(D=D, U=U, V=V), brow_c, np.asfortranarray(bcolT_c.T)

    # branch, compression.jl:77-133
    m1, m2 = rcl.left.size, rcl.right.size
    n1, n2 = ccl.left.size, ccl.right.size
    rows1, rows2 = (rcl.left.start, rcl.left.stop), (rcl.right.start, rcl.right.stop)
    cols1, cols2 = (ccl.left.start, ccl.left.stop), (ccl.right.start, ccl.right.stop)

    Brow1 = [("k", rows1[0], rows1[1], cols2[0], cols2[1])] + _rows(Brow, 0, m1)
    Bcol1 = [("k", rows2[0], rows2[1], cols1[0], cols1[1])] + _cols(Bcol, 0, n1)
    A11, Brow1c, Bcol1c = _compress(Am, Brow1, Bcol1, m1, n1, rcl.left, ccl.left, atol, rtol, False, prog)

    Brow2 = [Bcol1c[0:m2, :]] + _rows(Brow, m1, m1 + m2)
    Bcol2 = [Brow1c[:, 0:n2]] + _cols(Bcol, n1, n1 + n2)
    A22, Brow2c, Bcol2c = _compress(Am, Brow2, Bcol2, m2, n2, rcl.right, ccl.right, atol, rtol, False, prog)

    rm1 = Brow1c.shape[0]; rn1 = Bcol1c.shape[1]
    B12 = Bcol2c[0:rm1, :]
    B21 = Brow2c[:, 0:rn1]

    Brow_new = np.vstack([Brow1c[:, n2:], Brow2c[:, rn1:]])
    Bcol_new = np.hstack([Bcol1c[m2:, :], Bcol2c[rm1:, :]])
    if not rootnode:
        R, Brow_c = compress_block(np.asfortranarray(Brow_new), atol, rtol)
        W, BcolT_c = compress_block(np.asfortranarray(Bcol_new.T), atol, rtol)
        h = HssMatrix(A11=A11, A22=A22, B12=B12, B21=B21, R1=R[:rm1], W1=W[:rn1], R2=R[rm1:], W2=W[rn1:])
        return h, Brow_c, np.asfortranarray(BcolT_c.T)
    h = HssMatrix(A11=A11, A22=A22, B12=B12, B21=B21, rootnode=True)
    return h, Brow_new, Bcol_new
