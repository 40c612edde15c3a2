# HssHip.jl -- Julia binding of libhss_hip.so (include/hss_hip.h) for HssMatrices.jl.
#
# The reference (bonevbs/HssMatrices.jl 0.1.6) has no FFI: its boundary is
# dispatch on HssMatrix{T}.  This module is the binding a maintainer would add:
# it packs an HssMatrix{Float64} level-major (breadth-first node table +
# column-major generator blob, the layout hss_mat_create expects), keeps the
# generators in HBM, and provides the hot-path methods on a wrapper type
# (no type piracy on HssMatrix itself):
#
#   *(A::HipHss, B)            <- *(hssA, B)            src/matmul.jl:13,15
#   mul!(C, A::HipHss, B, a, b) <- mul!(C, hssA, B, a, b) src/matmul.jl:18
#   \(A::HipHss, B)            <- \(hssA, B::Matrix)     src/hssmatrix.jl:234
#   ulvfactsolve(A::HipHss, b) <- ulvfactsolve           src/ulvfactor.jl:10
#   ulvfactor(A::HipHss)       -> resident factor (no reference counterpart)
#   adjoint(A::HipHss), A'     <- adjoint(hssA)         src/hssmatrix.jl:165-171
#   *(X, A::HipHss)            <- *(A, hssB)            src/matmul.jl:14
#   /(X, A::HipHss)            <- /(A, hssB)            src/hssmatrix.jl:236
#   *(A::HipHss, B::HipHss)    <- *(hssA, hssB)         src/matmul.jl:65-118
#   HssHip.save / HssHip.load  generator files (include/hss_hip.h, hss_gen_header):
#                              HssMatrix{Float64} <-> file in Julia, HipHss <-> file
#                              through hss_mat_save / hss_mat_load
#
# Status codes map to the reference's exceptions: DimensionMismatch,
# ArgumentError, SingularException.  Julia is not available in the build
# container, so this file is not executed by the test suite; the same ABI is
# exercised by the Python binding (hssmatrices.jl_amd/hip.py) and the tests.
module HssHip

using LinearAlgebra
using HssMatrices
import Base: *, \, /, adjoint, size
import LinearAlgebra: mul!
# ulvfactsolve is HssMatrices' own generic function (exported at src/HssMatrices.jl:44):
# the methods below extend it for HipHss, so `using HssMatrices, HssHip` sees one function
import HssMatrices: ulvfactsolve

const libhss = joinpath(@__DIR__, "..", "libhss_hip.so")

# hss_node_desc (112 bytes), field order of include/hss_hip.h
struct NodeDesc
    parent::Int32; left::Int32; right::Int32; depth::Int32; height::Int32
    m::Int32; n::Int32; ku::Int32; kv::Int32; pad::Int32
    row_off::Int64; col_off::Int64
    off_D::Int64; off_U::Int64; off_V::Int64
    off_B12::Int64; off_B21::Int64
    off_R::Int64; off_W::Int64
end

struct HssHipError <: Exception
    status::Cint
    msg::String
end

function check(st::Cint)
    st == 0 && return nothing
    msg = unsafe_string(ccall((:hss_last_error, libhss), Cstring, ()))
    st == 1 && throw(DimensionMismatch(msg))
    st == 2 && throw(ArgumentError(msg))
    st == 3 && throw(SingularException(0))
    st == 9 && throw(Base.IOError(msg, 0))
    throw(HssHipError(st, msg))
end

mutable struct Context
    h::Ptr{Cvoid}
end
function Context(device::Integer=0)
    r = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:hss_ctx_create, libhss), Cint, (Cint, Ptr{Ptr{Cvoid}}), device, r))
    c = Context(r[])
    finalizer(c -> ccall((:hss_ctx_destroy, libhss), Cint, (Ptr{Cvoid},), c.h), c)
    return c
end
const DEFAULT_CTX = Ref{Union{Nothing,Context}}(nothing)
default_context() = something(DEFAULT_CTX[], (DEFAULT_CTX[] = Context(0)))

# Breadth-first packing; R1/W1 (R2/W2) are stored with the left (right) child.
function pack(hssA::HssMatrix{Float64})
    order = Any[(hssA, -1, nothing, nothing, 0, 0, 0)]
    i = 1
    while i <= length(order)
        node, _, _, _, d, ro, co = order[i]
        if !HssMatrices.isleaf(node)
            m1, n1 = node.sz1
            push!(order, (node.A11, i - 1, node.R1, node.W1, d + 1, ro, co))
            push!(order, (node.A22, i - 1, node.R2, node.W2, d + 1, ro + m1, co + n1))
        end
        i += 1
    end
    blob = Float64[]
    put(A) = (off = length(blob); append!(blob, vec(A)); Int64(off))
    recs = Vector{Vector{Int64}}()
    left = fill(Int32(-1), length(order)); right = fill(Int32(-1), length(order))
    for (k, (node, parent, R, W, d, ro, co)) in enumerate(order)
        if parent >= 0
            left[parent + 1] < 0 ? (left[parent + 1] = k - 1) : (right[parent + 1] = k - 1)
        end
    end
    nodes = NodeDesc[]
    for (k, (node, parent, R, W, d, ro, co)) in enumerate(order)
        m, n = size(node)
        ku, kv = parent < 0 ? (0, 0) : HssMatrices.gensize(node)
        oD = oU = oV = oB12 = oB21 = oR = oW = Int64(-1)
        if HssMatrices.isleaf(node)
            oD = put(node.D)
            oU = put(parent < 0 ? zeros(m, 0) : node.U)
            oV = put(parent < 0 ? zeros(n, 0) : node.V)
        else
            oB12 = put(node.B12); oB21 = put(node.B21)
        end
        if parent >= 0
            oR = put(R); oW = put(W)
        end
        push!(nodes, NodeDesc(parent, left[k], right[k], d, 0, m, n, ku, kv, 0, ro, co,
                              oD, oU, oV, oB12, oB21, oR, oW))
    end
    return nodes, blob   # heights are recomputed by the library
end

# Rebuilds an HssMatrix{Float64} from a node table and blob (the inverse of pack).
function unpack(nodes::Vector{NodeDesc}, blob::Vector{Float64})
    mat(off, r, c) = off < 0 ? zeros(r, c) : reshape(blob[off+1:off+r*c], r, c)
    function build(i)
        x = nodes[i]
        root = x.parent < 0
        if x.left < 0
            return HssMatrix(mat(x.off_D, x.m, x.n), mat(x.off_U, x.m, x.ku), mat(x.off_V, x.n, x.kv), root)
        end
        l, r = nodes[x.left + 1], nodes[x.right + 1]
        A11, A22 = build(x.left + 1), build(x.right + 1)
        return HssMatrix(A11, A22, mat(x.off_B12, l.ku, r.kv), mat(x.off_B21, r.ku, l.kv),
                         mat(l.off_R, l.ku, x.ku), mat(l.off_W, l.kv, x.kv),
                         mat(r.off_R, r.ku, x.ku), mat(r.off_W, r.kv, x.kv), root)
    end
    return build(1)
end

# Generator file: 40-byte header ("HSSG", version 1, element size 8, node count, blob
# length, two reserved Int64), the node table, the Float64 blob; little-endian.
const GEN_MAGIC = UInt8['H', 'S', 'S', 'G']
function save(path::AbstractString, hssA::HssMatrix{Float64})
    nodes, blob = pack(hssA)
    open(path, "w") do io
        write(io, GEN_MAGIC, UInt32(1), UInt32(8), Int32(length(nodes)), Int64(length(blob)), Int64(0), Int64(0))
        write(io, nodes)
        write(io, blob)
    end
    return path
end
function load(path::AbstractString)
    open(path, "r") do io
        read(io, 4) == GEN_MAGIC || throw(ArgumentError("$path: not a generator file"))
        (read(io, UInt32) == 1 && read(io, UInt32) == 8) ||
            throw(ArgumentError("$path: not a version-1 Float64 generator file"))
        nn, bl = read(io, Int32), read(io, Int64)
        read(io, Int64); read(io, Int64)
        (nn > 0 && bl >= 0 && filesize(path) == 40 + 112 * Int64(nn) + 8 * bl) ||
            throw(ArgumentError("$path: sizes in the header do not match the file"))
        nodes = read!(io, Vector{NodeDesc}(undef, nn))
        blob = read!(io, Vector{Float64}(undef, bl))
        return unpack(nodes, blob)
    end
end

mutable struct HipHss
    h::Ptr{Cvoid}
    m::Int
    n::Int
    ctx::Context
end

function HipHss(hssA::HssMatrix{Float64}, ctx::Context=default_context())
    nodes, blob = pack(hssA)
    r = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:hss_mat_create, libhss), Cint,
                (Ptr{Cvoid}, Ptr{NodeDesc}, Int32, Ptr{Float64}, Int64, Ptr{Ptr{Cvoid}}),
                ctx.h, nodes, length(nodes), blob, length(blob), r))
    A = HipHss(r[], size(hssA, 1), size(hssA, 2), ctx)
    finalizer(A -> ccall((:hss_mat_destroy, libhss), Cint, (Ptr{Cvoid},), A.h), A)
    return A
end

# a device matrix from a generator file, and back
function HipHss(path::AbstractString, ctx::Context=default_context())
    r = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:hss_mat_load, libhss), Cint, (Ptr{Cvoid}, Cstring, Ptr{Ptr{Cvoid}}), ctx.h, path, r))
    m, n = Ref{Int64}(0), Ref{Int64}(0)
    check(ccall((:hss_mat_size, libhss), Cint, (Ptr{Cvoid}, Ref{Int64}, Ref{Int64}), r[], m, n))
    A = HipHss(r[], m[], n[], ctx)
    finalizer(A -> ccall((:hss_mat_destroy, libhss), Cint, (Ptr{Cvoid},), A.h), A)
    return A
end
function save(path::AbstractString, A::HipHss)
    check(ccall((:hss_mat_save, libhss), Cint, (Ptr{Cvoid}, Cstring), A.h, path))
    return path
end

size(A::HipHss) = (A.m, A.n)
size(A::HipHss, d::Integer) = size(A)[d]

function mul!(C::StridedMatrix{Float64}, A::HipHss, B::StridedMatrix{Float64}, α::Number, β::Number)
    size(B, 1) == A.n || throw(DimensionMismatch("First dimension of B does not match second dimension of A. Expected $(A.n), got $(size(B, 1))"))
    size(C) == (A.m, size(B, 2)) || throw(DimensionMismatch("Dimensions of C don't match up with A and B."))
    check(ccall((:hss_matmul, libhss), Cint,
                (Ptr{Cvoid}, Int32, Float64, Ptr{Float64}, Int64, Float64, Ptr{Float64}, Int64),
                A.h, size(B, 2), α, B, stride(B, 2), β, C, stride(C, 2)))
    return C
end
*(A::HipHss, B::AbstractMatrix) = mul!(zeros(A.m, size(B, 2)), A, Matrix{Float64}(B), 1.0, 0.0)
*(A::HipHss, x::AbstractVector) = reshape(A * reshape(x, length(x), 1), length(x))

# a method of HssMatrices.ulvfactsolve (src/ulvfactor.jl:10), not a new function
function ulvfactsolve(A::HipHss, b::Matrix{Float64})
    A.m == A.n == size(b, 1) || throw(DimensionMismatch("matrix is $(size(A)), right-hand side has $(size(b, 1)) rows"))
    X = zeros(A.n, size(b, 2))
    check(ccall((:hss_ulvfactsolve, libhss), Cint,
                (Ptr{Cvoid}, Int32, Ptr{Float64}, Int64, Ptr{Float64}, Int64),
                A.h, size(b, 2), b, stride(b, 2), X, stride(X, 2)))
    return X
end
\(A::HipHss, B::Matrix{Float64}) = ulvfactsolve(A, B)
# the vector method the reference lacks (its Vector RHS falls back to a dense solve)
\(A::HipHss, x::Vector{Float64}) = vec(ulvfactsolve(A, reshape(x, length(x), 1)))

# hssA' as a new device matrix: D', U <-> V, B12 <- B21', B21 <- B12', R <-> W
function adjoint(A::HipHss)
    r = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:hss_mat_adjoint, libhss), Cint, (Ptr{Cvoid}, Ptr{Ptr{Cvoid}}), A.h, r))
    At = HipHss(r[], A.n, A.m, A.ctx)
    finalizer(At -> ccall((:hss_mat_destroy, libhss), Cint, (Ptr{Cvoid},), At.h), At)
    return At
end
*(X::AbstractMatrix, A::HipHss) = copy((A' * copy(X'))')
# HSS x HSS product (ranks add; nothing is recompressed, as in the reference)
function *(A::HipHss, B::HipHss)
    r = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:hss_mat_mul, libhss), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Ptr{Cvoid}}), A.h, B.h, r))
    C = HipHss(r[], A.m, B.n, A.ctx)
    finalizer(C -> ccall((:hss_mat_destroy, libhss), Cint, (Ptr{Cvoid},), C.h), C)
    return C
end
/(X::Matrix{Float64}, A::HipHss) = copy(ulvfactsolve(A', collect(X'))')

# a device call that returns a new matrix handle of size m x n
function adopt(A::HipHss, r::Ref{Ptr{Cvoid}}, m::Integer, n::Integer)
    C = HipHss(r[], m, n, A.ctx)
    finalizer(C -> ccall((:hss_mat_destroy, libhss), Cint, (Ptr{Cvoid},), C.h), C)
    return C
end
# hssA \ hssB and hssA / hssB (src/hssmatrix.jl:235-237), at the tolerances of HssOptions
function \(A::HipHss, B::HipHss; atol::Float64=HssOptions(Float64).atol, rtol::Float64=HssOptions(Float64).rtol)
    r = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:hss_mat_ldiv, libhss), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Float64, Float64, Ptr{Ptr{Cvoid}}),
                A.h, B.h, atol, rtol, r))
    return adopt(A, r, A.n, B.n)
end
function /(A::HipHss, B::HipHss; atol::Float64=HssOptions(Float64).atol, rtol::Float64=HssOptions(Float64).rtol)
    r = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:hss_mat_rdiv, libhss), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Float64, Float64, Ptr{Ptr{Cvoid}}),
                A.h, B.h, atol, rtol, r))
    return adopt(A, r, A.m, B.m)
end
# recompress!(hssA; atol, rtol) (src/compression.jl:148-203) as a new device matrix
function recompress(A::HipHss; atol::Float64=HssOptions(Float64).atol, rtol::Float64=HssOptions(Float64).rtol)
    r = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:hss_mat_recompress, libhss), Cint, (Ptr{Cvoid}, Float64, Float64, Ptr{Ptr{Cvoid}}),
                A.h, atol, rtol, r))
    return adopt(A, r, A.m, A.n)
end

mutable struct HipUlv
    h::Ptr{Cvoid}
    A::HipHss
end
function ulvfactor(A::HipHss; prec::Symbol=:fp64)
    r = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:hss_ulv_factor, libhss), Cint, (Ptr{Cvoid}, Cint, Ptr{Ptr{Cvoid}}),
                A.h, prec === :fp32 ? 1 : 0, r))
    F = HipUlv(r[], A)
    finalizer(F -> ccall((:hss_fact_destroy, libhss), Cint, (Ptr{Cvoid},), F.h), F)
    return F
end
function \(F::HipUlv, B::Matrix{Float64})
    X = zeros(F.A.n, size(B, 2))
    check(ccall((:hss_ulv_solve, libhss), Cint,
                (Ptr{Cvoid}, Int32, Ptr{Float64}, Int64, Ptr{Float64}, Int64, Ptr{Cvoid}, Ptr{Cvoid}),
                F.h, size(B, 2), B, stride(B, 2), X, stride(X, 2), C_NULL, C_NULL))
    return X
end

# hss_refine_opts: mode 1 = iterative refinement, 2 = GMRES-IR (Float64 GMRES preconditioned
# with the factor); the solve stops at max_j ||b_j - A x_j|| / ||b_j|| <= tol
struct RefineOpts
    tol::Float64
    max_iter::Cint
    mode::Cint
end
function solve(F::HipUlv, B::Matrix{Float64}; tol::Float64=1e-10, maxiter::Integer=10, method::Symbol=:ir)
    method in (:ir, :gmres) || throw(ArgumentError("method must be :ir or :gmres"))
    X = zeros(F.A.n, size(B, 2))
    opts = Ref(RefineOpts(tol, Cint(maxiter), method === :gmres ? Cint(2) : Cint(1)))
    check(ccall((:hss_ulv_solve, libhss), Cint,
                (Ptr{Cvoid}, Int32, Ptr{Float64}, Int64, Ptr{Float64}, Int64, Ref{RefineOpts}, Ptr{Cvoid}),
                F.h, size(B, 2), B, stride(B, 2), X, stride(X, 2), opts, C_NULL))
    return X
end

# ulvfactsolve is not exported again: HssMatrices exports it, and the methods above extend it
export HipHss, ulvfactor, recompress

end # module
