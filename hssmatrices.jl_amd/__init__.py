"""hssmatrices.jl_amd -- MI355X-native HSS matvec and ULV solve.

Host mirror of the reference's data model (``hssmatrix``), its CPU input
producers (``compression``), and the device path (``hip``) that drives the
hand-written gfx950 kernels in ``libhss_hip.so`` through its C ABI.
"""
from .hssmatrix import (ClusterTree, bisection_cluster, HssMatrix, full, hssrank, gensize,
                        prune_leaves, lowrank2hss, pack, unpack, tree_summary, NODE_DTYPE,
                        DimensionMismatch, SingularException, cluster, random_hss,
                        save_generators, load_generators, cluster_nodes, compatible)
from .compression import KernelMatrix, compress, compress_block
