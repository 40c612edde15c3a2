"""Host-side HSS data model: cluster trees, the recursive HssMatrix node, dense
expansion, and the level-major packing that the C-ABI library consumes.

This module mirrors the reference's data structures so that a test or a user
script reads like the reference's own code:

* ``HssMatrix``             -- reference ``src/hssmatrix.jl:11-68`` (leaf: D, U, V;
                               branch: A11, A22, B12, B21, R1, W1, R2, W2, sz1, sz2)
* ``bisection_cluster``     -- reference ``src/clustertree.jl:14-35``
* ``full`` / ``_full``      -- reference ``src/hssmatrix.jl:270-305``
* ``gensize``, ``hssrank``  -- reference ``src/hssmatrix.jl:251-262``
* ``prune_leaves``          -- reference ``src/hssmatrix.jl:325-346``
* ``lowrank2hss``           -- reference ``src/constructors.jl:5-20``

Indices are 0-based and ranges half-open (Python convention); the reference
uses 1-based inclusive ranges.  Generators are Fortran-ordered float64 arrays.

The packed format (``pack``/``unpack``) is the input of ``hss_mat_create`` in
``include/hss_hip.h``: one node record per tree node in breadth-first
(level-major) order, root first, and one float64 blob holding every generator
column-major, level by level.  The translators R1/W1 (R2/W2) of a branch are
stored with its left (right) child, since they map that child's basis into the
parent's.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

__all__ = [
    "ClusterTree", "bisection_cluster", "HssMatrix", "full", "hssrank", "gensize",
    "prune_leaves", "lowrank2hss", "NODE_DTYPE", "pack", "unpack", "tree_summary",
    "DimensionMismatch", "SingularException", "GEN_HEADER_DTYPE", "save_generators", "load_generators",
]


class DimensionMismatch(ValueError):
    """Raised where the reference throws ``DimensionMismatch``."""


class SingularException(np.linalg.LinAlgError):
    """Raised where the reference's dense ``\\`` throws ``SingularException``."""


# ----------------------------------------------------------------------------
# cluster trees (reference src/clustertree.jl, src/binarytree.jl)
# ----------------------------------------------------------------------------
@dataclass
class ClusterTree:
    """Binary cluster tree over the index range ``[start, stop)``.

    Reference: ``const ClusterTree = BinaryNode{UnitRange{Int}}``
    (``src/clustertree.jl:10``).
    """
    start: int
    stop: int
    left: Optional["ClusterTree"] = None
    right: Optional["ClusterTree"] = None

    @property
    def size(self) -> int:
        return self.stop - self.start

    def isleaf(self) -> bool:
        return self.left is None and self.right is None

    def isbranch(self) -> bool:
        return self.left is not None and self.right is not None

    def depth(self) -> int:
        """Number of levels (a single leaf has depth 1), ``binarytree.jl:27-40``."""
        if self.isleaf():
            return 1
        return 1 + max(self.left.depth(), self.right.depth())

    def nleaves(self) -> int:
        if self.isleaf():
            return 1
        return self.left.nleaves() + self.right.nleaves()


def bisection_cluster(n, leafsize: int = 64) -> ClusterTree:
    """Bisection cluster tree, reference ``src/clustertree.jl:14-35``.

    A range of length ``len > leafsize`` is split after its first
    ``ceil(len/2)`` indices (``clustertree.jl:29-32``).  ``n`` is either an int
    (range ``[0, n)``) or a ``(start, stop)`` pair.
    """
    if leafsize < 1:
        raise ValueError("leafsize")
    if isinstance(n, tuple):
        start, stop = n
    else:
        start, stop = 0, int(n)
    if stop - start <= 0:
        raise ValueError("Index range must be larger or equal to 0")
    return _bisection_cluster(start, stop, leafsize)


def _bisection_cluster(start: int, stop: int, leafsize: int) -> ClusterTree:
    node = ClusterTree(start, stop)
    length = stop - start
    if length > leafsize:
        h = (length + 1) // 2  # ceil(length/2)
        node.left = _bisection_cluster(start, start + h, leafsize)
        node.right = _bisection_cluster(start + h, stop, leafsize)
    return node


def compatible(a: ClusterTree, b: ClusterTree) -> bool:
    """Same branching structure, reference ``binarytree.jl:68-74``."""
    if a.isleaf() and b.isleaf():
        return True
    if a.isbranch() and b.isbranch():
        return compatible(a.left, b.left) and compatible(a.right, b.right)
    return False


# ----------------------------------------------------------------------------
# HSS node
# ----------------------------------------------------------------------------
def _f(a) -> np.ndarray:
    return np.asfortranarray(np.asarray(a, dtype=np.float64))


class HssMatrix:
    """Recursive HSS node, reference ``src/hssmatrix.jl:11-68``.

    Leaf: ``D`` (m x n), ``U`` (m x ku), ``V`` (n x kv).
    Branch: children ``A11``/``A22``, couplings ``B12`` (ku1 x kv2) and ``B21``
    (ku2 x kv1), translators ``R1`` (ku1 x kr), ``R2`` (ku2 x kr), ``W1``
    (kv1 x kw), ``W2`` (kv2 x kw).  A root branch has 0-column R/W
    (``hssmatrix.jl:46-55``).
    """

    def __init__(self, *, D=None, U=None, V=None, A11=None, A22=None, B12=None, B21=None,
                 R1=None, W1=None, R2=None, W2=None, rootnode: bool = False):
        self.rootnode = bool(rootnode)
        if A11 is None:
            self.leafnode = True
            self.D = _f(D)
            m, n = self.D.shape
            self.U = _f(U) if U is not None else np.zeros((m, 0), order="F")
            self.V = _f(V) if V is not None else np.zeros((n, 0), order="F")
            if self.U.shape[0] != m:
                raise ValueError("D and U must have same number of rows")
            if self.V.shape[0] != n:
                raise ValueError("D and V must have same number of columns")
        else:
            self.leafnode = False
            self.A11, self.A22 = A11, A22
            self.B12, self.B21 = _f(B12), _f(B21)
            self.sz1, self.sz2 = A11.shape, A22.shape
            kr1, kw1 = gensize(A11)
            kr2, kw2 = gensize(A22)
            if R1 is None:  # root constructor, hssmatrix.jl:46-55
                R1 = np.zeros((kr1, 0)); R2 = np.zeros((kr2, 0))
                W1 = np.zeros((kw1, 0)); W2 = np.zeros((kw2, 0))
            self.R1, self.R2, self.W1, self.W2 = _f(R1), _f(R2), _f(W1), _f(W2)
            if self.R1.shape[1] != self.R2.shape[1]:
                raise DimensionMismatch("R1 and R2 must have same number of columns")
            if self.W1.shape[1] != self.W2.shape[1]:
                raise DimensionMismatch("W1 and W2 must have same number of rows")

    # reference hssmatrix.jl:88-95
    def isleaf(self) -> bool:
        return self.leafnode

    def isbranch(self) -> bool:
        return not self.leafnode

    def isroot(self) -> bool:
        return self.rootnode

    @property
    def shape(self):
        if self.leafnode:
            return self.D.shape
        return (self.sz1[0] + self.sz2[0], self.sz1[1] + self.sz2[1])

    def size(self, dim=None):
        s = self.shape
        return s if dim is None else s[dim - 1]

    def __repr__(self):
        m, n = self.shape
        return f"{m}x{n} HssMatrix{{Float64}}"

    def copy(self) -> "HssMatrix":
        if self.leafnode:
            return HssMatrix(D=self.D.copy(), U=self.U.copy(), V=self.V.copy(), rootnode=self.rootnode)
        return HssMatrix(A11=self.A11.copy(), A22=self.A22.copy(), B12=self.B12.copy(), B21=self.B21.copy(),
                         R1=self.R1.copy(), W1=self.W1.copy(), R2=self.R2.copy(), W2=self.W2.copy(),
                         rootnode=self.rootnode)

    def adjoint(self) -> "HssMatrix":
        """Reference ``hssmatrix.jl:165-171`` (real case: transpose)."""
        if self.leafnode:
            return HssMatrix(D=self.D.T.copy(), U=self.V.copy(), V=self.U.copy(), rootnode=self.rootnode)
        return HssMatrix(A11=self.A11.adjoint(), A22=self.A22.adjoint(), B12=self.B21.T.copy(),
                         B21=self.B12.T.copy(), R1=self.W1.copy(), W1=self.R1.copy(), R2=self.W2.copy(),
                         W2=self.R2.copy(), rootnode=self.rootnode)

    @property
    def T(self):
        return self.adjoint()

    def full(self) -> np.ndarray:
        return full(self)


def gensize(h: HssMatrix):
    """Reference ``hssmatrix.jl:254-262``."""
    if h.isleaf():
        return h.U.shape[1], h.V.shape[1]
    kr = h.R1.shape[1]
    kw = h.W1.shape[1]
    if kr != h.R2.shape[1]:
        raise DimensionMismatch("dimensions of column-translators do not match")
    if kw != h.W2.shape[1]:
        raise DimensionMismatch("dimensions of row-translators do not match")
    return kr, kw


def hssrank(h: HssMatrix) -> int:
    """Reference ``hssmatrix.jl:251``."""
    if h.isleaf():
        return 0
    return max(hssrank(h.A11), hssrank(h.A22), *h.B12.shape, *h.B21.shape)


def full(h: HssMatrix) -> np.ndarray:
    """Dense expansion, reference ``hssmatrix.jl:270-305``."""
    if h.isleaf():
        return h.D.copy()
    m, n = h.shape
    k = max(hssrank(h), 1)
    kmax = max(k, _max_gen_cols(h))
    A = np.empty((m, n), order="F")
    U = np.zeros((m, kmax), order="F")
    V = np.zeros((n, kmax), order="F")
    _full(h, A, U, V, 0, 0, rootnode=True)
    return A


def _max_gen_cols(h: HssMatrix) -> int:
    if h.isleaf():
        return max(h.U.shape[1], h.V.shape[1])
    return max(_max_gen_cols(h.A11), _max_gen_cols(h.A22), h.R1.shape[1], h.W1.shape[1])


def _full(h: HssMatrix, A, U, V, ro: int, co: int, rootnode: bool = False):
    if h.isleaf():
        m, n = h.shape
        A[ro:ro + m, co:co + n] = h.D
        U[ro:ro + m, :h.U.shape[1]] = h.U
        V[co:co + n, :h.V.shape[1]] = h.V
        return
    m1, n1 = h.sz1
    m2, n2 = h.sz2
    ru1, rv1 = gensize(h.A11)
    ru2, rv2 = gensize(h.A22)
    _full(h.A11, A, U, V, ro, co)
    _full(h.A22, A, U, V, ro + m1, co + n1)
    A[ro:ro + m1, co + n1:co + n1 + n2] = U[ro:ro + m1, :ru1] @ h.B12 @ V[co + n1:co + n1 + n2, :rv2].T
    A[ro + m1:ro + m1 + m2, co:co + n1] = U[ro + m1:ro + m1 + m2, :ru2] @ h.B21 @ V[co:co + n1, :rv1].T
    if not rootnode:
        kr, kw = h.R1.shape[1], h.W1.shape[1]
        U1 = U[ro:ro + m1, :ru1] @ h.R1
        U2 = U[ro + m1:ro + m1 + m2, :ru2] @ h.R2
        V1 = V[co:co + n1, :rv1] @ h.W1
        V2 = V[co + n1:co + n1 + n2, :rv2] @ h.W2
        U[ro:ro + m1, :kr] = U1
        U[ro + m1:ro + m1 + m2, :kr] = U2
        V[co:co + n1, :kw] = V1
        V[co + n1:co + n1 + n2, :kw] = V2


def _hssleaf(h: HssMatrix):
    """Reference ``hssmatrix.jl:338-346``: (D, U, V) of the subtree as one leaf."""
    if h.isleaf():
        return h.D, h.U, h.V
    A11, U1, V1 = _hssleaf(h.A11)
    A22, U2, V2 = _hssleaf(h.A22)
    D = np.block([[A11, U1 @ h.B12 @ V2.T], [U2 @ h.B21 @ V1.T, A22]])
    return D, np.vstack([U1 @ h.R1, U2 @ h.R2]), np.vstack([V1 @ h.W1, V2 @ h.W2])


def prune_leaves(h: HssMatrix) -> HssMatrix:
    """Merge sibling leaves on the bottom level, reference ``hssmatrix.jl:325-335``."""
    if h.isleaf():
        return h
    if h.A11.isleaf() and h.A22.isleaf():
        D, U, V = _hssleaf(h)
        return HssMatrix(D=D, U=U, V=V, rootnode=h.rootnode)
    h.A11 = prune_leaves(h.A11)
    h.A22 = prune_leaves(h.A22)
    h.sz1, h.sz2 = h.A11.shape, h.A22.shape
    return h


def cluster(h: HssMatrix):
    """Row/column cluster trees of an HSS matrix, reference ``hssmatrix.jl:349-359``."""
    return _cluster(h, 0, 0)


def _cluster(h: HssMatrix, ro: int, co: int):
    if h.isleaf():
        m, n = h.shape
        return ClusterTree(ro, ro + m), ClusterTree(co, co + n)
    r1, c1 = _cluster(h.A11, ro, co)
    r2, c2 = _cluster(h.A22, r1.stop, c1.stop)
    return ClusterTree(ro, r2.stop, r1, r2), ClusterTree(co, c2.stop, c1, c2)


def lowrank2hss(U, V, rcl: ClusterTree, ccl: ClusterTree) -> HssMatrix:
    """HSS form of U V^T with identity translators, reference ``constructors.jl:5-20``."""
    U = _f(U); V = _f(V)
    k = U.shape[1]
    if V.shape[1] != k:
        raise ValueError("second dimension of U and V must agree")
    return _lowrank2hss(U, V, rcl, ccl, k, True)


def _lowrank2hss(U, V, rcl, ccl, k, rootnode=False):
    if rcl.isleaf() and ccl.isleaf():
        Ul = U[rcl.start:rcl.stop]; Vl = V[ccl.start:ccl.stop]
        return HssMatrix(D=Ul @ Vl.T, U=Ul, V=Vl, rootnode=rootnode)
    if rcl.isbranch() and ccl.isbranch():
        A11 = _lowrank2hss(U, V, rcl.left, ccl.left, k)
        A22 = _lowrank2hss(U, V, rcl.right, ccl.right, k)
        I = np.eye(k)
        if rootnode:
            return HssMatrix(A11=A11, A22=A22, B12=I, B21=I, rootnode=True)
        return HssMatrix(A11=A11, A22=A22, B12=I, B21=I, R1=I, W1=I, R2=I, W2=I, rootnode=False)
    raise ValueError("row and column clusters are not compatible")


# ----------------------------------------------------------------------------
# packing: level-major node table + generator blob (input of hss_mat_create)
# ----------------------------------------------------------------------------
NODE_DTYPE = np.dtype([
    ("parent", "<i4"), ("left", "<i4"), ("right", "<i4"),
    ("depth", "<i4"), ("height", "<i4"),
    ("m", "<i4"), ("n", "<i4"), ("ku", "<i4"), ("kv", "<i4"), ("pad", "<i4"),
    ("row_off", "<i8"), ("col_off", "<i8"),
    ("off_D", "<i8"), ("off_U", "<i8"), ("off_V", "<i8"),
    ("off_B12", "<i8"), ("off_B21", "<i8"),
    ("off_R", "<i8"), ("off_W", "<i8"),
])
assert NODE_DTYPE.itemsize == 112


def cluster_nodes(rcl: ClusterTree, ccl: ClusterTree | None = None) -> np.ndarray:
    """The node table of a pair of compatible cluster trees (row tree ``rcl``, column tree
    ``ccl``, reference ``binarytree.jl:68-74``): ``NODE_DTYPE`` records in the breadth-first
    order ``pack`` uses, with the links, depths, heights, sizes and offsets filled, ranks 0
    and no generators.  The tree argument of ``hss_compress_dev``."""
    ccl = rcl if ccl is None else ccl
    if not compatible(rcl, ccl):
        raise ValueError("row and column clusters are not compatible")
    order = [(rcl, ccl, -1, 0)]
    i = 0
    while i < len(order):
        r, c, _, d = order[i]
        if r.isbranch():
            order.append((r.left, c.left, i, d + 1))
            order.append((r.right, c.right, i, d + 1))
        i += 1
    nodes = np.zeros(len(order), dtype=NODE_DTYPE)
    for f in ("off_D", "off_U", "off_V", "off_B12", "off_B21", "off_R", "off_W", "left", "right"):
        nodes[f] = -1
    for i, (r, c, parent, d) in enumerate(order):
        rec = nodes[i]
        rec["parent"], rec["depth"] = parent, d
        rec["m"], rec["n"] = r.size, c.size
        rec["row_off"], rec["col_off"] = r.start - rcl.start, c.start - ccl.start
        if parent >= 0:
            p = nodes[parent]
            if p["left"] < 0:
                p["left"] = i
            else:
                p["right"] = i
    for i in range(len(order) - 1, -1, -1):
        l, r = nodes["left"][i], nodes["right"][i]
        if l >= 0:
            nodes["height"][i] = 1 + max(nodes["height"][l], nodes["height"][r])
    return nodes


def pack(h: HssMatrix):
    """Pack a (root) HSS matrix into ``(nodes, blob)``.

    ``nodes`` is a structured array of ``NODE_DTYPE`` in breadth-first order
    (root = 0, then depth 1 left to right, ...); ``blob`` holds the generators
    column-major, level by level.  The layout is the one ``hss_mat_create``
    expects (``include/hss_hip.h``).
    """
    # breadth-first walk: (hss node, parent index, R, W, depth, row_off, col_off)
    order = [(h, -1, None, None, 0, 0, 0)]
    i = 0
    while i < len(order):
        node, _, _, _, d, ro, co = order[i]
        if node.isbranch():
            m1, n1 = node.sz1
            order.append((node.A11, i, node.R1, node.W1, d + 1, ro, co))
            order.append((node.A22, i, node.R2, node.W2, d + 1, ro + m1, co + n1))
        i += 1
    nn = len(order)
    nodes = np.zeros(nn, dtype=NODE_DTYPE)
    for f in ("off_D", "off_U", "off_V", "off_B12", "off_B21", "off_R", "off_W"):
        nodes[f] = -1
    nodes["left"] = -1
    nodes["right"] = -1
    chunks = []
    pos = 0

    def put(a):
        nonlocal pos
        a = np.asarray(a, dtype=np.float64)
        off = pos
        if a.size:
            chunks.append(np.ravel(a, order="F"))
            pos += a.size
        return off

    for i, (node, parent, R, W, d, ro, co) in enumerate(order):
        rec = nodes[i]
        rec["parent"] = parent
        rec["depth"] = d
        m, n = node.shape
        rec["m"], rec["n"] = m, n
        rec["row_off"], rec["col_off"] = ro, co
        ku, kv = gensize(node)
        if parent < 0:
            ku = kv = 0
        rec["ku"], rec["kv"] = ku, kv
        if parent >= 0:
            p = nodes[parent]
            if p["left"] < 0:
                p["left"] = i
            else:
                p["right"] = i
        if node.isleaf():
            rec["off_D"] = put(node.D)
            uu = node.U if parent >= 0 else node.U[:, :0]
            vv = node.V if parent >= 0 else node.V[:, :0]
            rec["off_U"] = put(uu)
            rec["off_V"] = put(vv)
        else:
            rec["off_B12"] = put(node.B12)
            rec["off_B21"] = put(node.B21)
        if parent >= 0:
            rec["off_R"] = put(R)
            rec["off_W"] = put(W)
    # heights (leaf = 0)
    heights = np.zeros(nn, dtype=np.int32)
    for i in range(nn - 1, -1, -1):
        l, r = nodes["left"][i], nodes["right"][i]
        if l >= 0:
            heights[i] = 1 + max(heights[l], heights[r])
    nodes["height"] = heights
    blob = np.concatenate(chunks) if chunks else np.zeros(0)
    return nodes, np.ascontiguousarray(blob, dtype=np.float64)


def unpack(nodes, blob) -> HssMatrix:
    """Inverse of ``pack``."""
    def mat(off, r, c):
        if r * c == 0:
            return np.zeros((r, c), order="F")
        return np.asfortranarray(blob[off:off + r * c].reshape((r, c), order="F"))

    def build(i):
        rec = nodes[i]
        root = rec["parent"] < 0
        if rec["left"] < 0:
            m, n, ku, kv = int(rec["m"]), int(rec["n"]), int(rec["ku"]), int(rec["kv"])
            return HssMatrix(D=mat(rec["off_D"], m, n), U=mat(rec["off_U"], m, ku), V=mat(rec["off_V"], n, kv),
                             rootnode=root)
        l, r = int(rec["left"]), int(rec["right"])
        A11, A22 = build(l), build(r)
        L, Rr = nodes[l], nodes[r]
        B12 = mat(rec["off_B12"], int(L["ku"]), int(Rr["kv"]))
        B21 = mat(rec["off_B21"], int(Rr["ku"]), int(L["kv"]))
        ku, kv = int(rec["ku"]), int(rec["kv"])
        R1 = mat(L["off_R"], int(L["ku"]), ku); W1 = mat(L["off_W"], int(L["kv"]), kv)
        R2 = mat(Rr["off_R"], int(Rr["ku"]), ku); W2 = mat(Rr["off_W"], int(Rr["kv"]), kv)
        return HssMatrix(A11=A11, A22=A22, B12=B12, B21=B21, R1=R1, W1=W1, R2=R2, W2=W2, rootnode=root)

    return build(0)


# ----------------------------------------------------------------------------
# generator interchange file (the layout of include/hss_hip.h's hss_gen_header):
# 40-byte header ("HSSG", version 1, element size 8, node count, blob length, two
# reserved int64), the node table (NODE_DTYPE records), the Float64 blob; little-endian.
# The reference keeps HssMatrix{T} in memory only (src/hssmatrix.jl:11-68).
# ----------------------------------------------------------------------------
GEN_HEADER_DTYPE = np.dtype([("magic", "S4"), ("version", "<u4"), ("elem_size", "<u4"),
                             ("num_nodes", "<i4"), ("blob_len", "<i8"), ("reserved", "<i8", (2,))])


def save_generators(path, nodes, blob) -> None:
    """Write a node table and blob (``pack``'s output) as a generator file."""
    nodes = np.ascontiguousarray(nodes, dtype=NODE_DTYPE)
    blob = np.ascontiguousarray(blob, dtype="<f8")
    hdr = np.zeros((), GEN_HEADER_DTYPE)
    hdr["magic"], hdr["version"], hdr["elem_size"] = b"HSSG", 1, 8
    hdr["num_nodes"], hdr["blob_len"] = len(nodes), blob.size
    with open(path, "wb") as f:
        f.write(hdr.tobytes())
        f.write(nodes.tobytes())
        f.write(blob.tobytes())


def load_generators(path):
    """(nodes, blob) of a generator file; ValueError for a malformed one."""
    raw = np.fromfile(path, dtype=np.uint8)
    if raw.size < GEN_HEADER_DTYPE.itemsize:
        raise ValueError(f"{path}: shorter than a generator file header")
    hdr = raw[:GEN_HEADER_DTYPE.itemsize].view(GEN_HEADER_DTYPE)[0]
    if hdr["magic"] != b"HSSG" or hdr["version"] != 1 or hdr["elem_size"] != 8:
        raise ValueError(f"{path}: not a version-1 Float64 generator file")
    nn, bl = int(hdr["num_nodes"]), int(hdr["blob_len"])
    if nn <= 0 or bl < 0:
        raise ValueError(f"{path}: bad sizes in the header")
    o1 = GEN_HEADER_DTYPE.itemsize + nn * NODE_DTYPE.itemsize
    if raw.size != o1 + 8 * bl:
        raise ValueError(f"{path}: {raw.size} bytes, the header says {o1 + 8 * bl}")
    nodes = raw[GEN_HEADER_DTYPE.itemsize:o1].view(NODE_DTYPE).copy()
    blob = raw[o1:].view("<f8").copy()
    return nodes, blob


def tree_summary(nodes) -> dict:
    """Shape statistics used in reports: leaves, depth, rank range, generator count G."""
    leaf = nodes["left"] < 0
    nonroot = nodes["parent"] >= 0
    G = 0
    for rec in nodes:
        if rec["left"] < 0:
            G += int(rec["m"]) * int(rec["n"]) + int(rec["m"]) * int(rec["ku"]) + int(rec["n"]) * int(rec["kv"])
    for rec in nodes:
        if rec["left"] >= 0:
            L, R = nodes[rec["left"]], nodes[rec["right"]]
            G += int(L["ku"]) * int(R["kv"]) + int(R["ku"]) * int(L["kv"])
            G += int(L["ku"]) * int(rec["ku"]) + int(R["ku"]) * int(rec["ku"])
            G += int(L["kv"]) * int(rec["kv"]) + int(R["kv"]) * int(rec["kv"])
    ks = nodes["ku"][nonroot]
    return {
        "n": int(nodes["m"][0]),
        "nodes": int(len(nodes)),
        "leaves": int(leaf.sum()),
        "max_depth": int(nodes["depth"].max()),
        "rank_min": int(ks.min()) if ks.size else 0,
        "rank_max": int(ks.max()) if ks.size else 0,
        "G": int(G),
    }


def random_hss(n: int, leafsize: int, rank_range=(10, 40), rng=None, diag: float = 4.0,
               coupling: float = 1.0) -> HssMatrix:
    """Synthetic HSS matrix with the shapes of a compressed kernel matrix.

    Bisection tree (``bisection_cluster``), node ranks drawn uniformly from
    ``rank_range`` (capped by what the children can carry), orthonormal nested
    bases (leaf U/V and stacked translators [R1; R2], [W1; W2] with orthonormal
    columns), couplings N(0, coupling^2 / k), and leaf blocks
    N(0, 1/m) + ``diag`` * I.  Used for batch and serving benchmarks, where
    compressing hundreds of kernel matrices on the host would dominate.
    """
    rng = np.random.default_rng(rng)
    lo, hi = rank_range

    def orth(r, c):
        if r == 0 or c == 0:
            return np.zeros((r, c), order="F")
        q, _ = np.linalg.qr(rng.standard_normal((r, c)))
        return np.asfortranarray(q)

    def build(cl: ClusterTree, root: bool):
        if cl.isleaf():
            m = cl.size
            D = rng.standard_normal((m, m)) / np.sqrt(m) + diag * np.eye(m)
            if root:
                return HssMatrix(D=D, rootnode=True), 0
            k = int(min(rng.integers(lo, hi + 1), m))
            return HssMatrix(D=D, U=orth(m, k), V=orth(m, k)), k
        a, ka = build(cl.left, False)
        b, kb = build(cl.right, False)
        B12 = rng.standard_normal((ka, kb)) * coupling / np.sqrt(max(ka, 1))
        B21 = rng.standard_normal((kb, ka)) * coupling / np.sqrt(max(kb, 1))
        if root:
            return HssMatrix(A11=a, A22=b, B12=B12, B21=B21, rootnode=True), 0
        k = int(min(rng.integers(lo, hi + 1), ka + kb))
        R = orth(ka + kb, k)
        W = orth(ka + kb, k)
        return HssMatrix(A11=a, A22=b, B12=B12, B21=B21, R1=R[:ka], W1=W[:ka], R2=R[ka:], W2=W[ka:]), k

    h, _ = build(bisection_cluster(n, leafsize), True)
    return h
