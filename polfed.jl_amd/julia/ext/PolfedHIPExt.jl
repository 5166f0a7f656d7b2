# PolfedHIPExt -- the MI355X backend of Polfed.jl through libpolfed_hip.
#
# A package extension of the reference (Polfed.jl), loaded when AMDGPU.jl is
# loaded beside it: the HIP counterpart of ext/PolfedCUDAExt.jl:1-53.  Add to
# the reference's Project.toml (see ../Project.toml.fragment):
#
#     [weakdeps]   AMDGPU = "21141c5a-9bdb-4563-92ae-f87d6854732e"
#     [extensions] PolfedHIPExt = "AMDGPU"
#
# and point POLFED_HIP_LIB at libpolfed_hip.so (default: the in-tree build).
#
# What it plugs in (reference file:line):
#   * the GPU backend registry _set_cuda_backend! (src/Polfed.jl:34-72), as
#     PolfedCUDAExt.__init__ does (ext/PolfedCUDAExt.jl:19-40);
#   * is_gpu_array for ROCArray (src/Polfed.jl:82);
#   * HIPOperator: a matrix-free or packed operator callable as f!(Y, X)
#     (MappingConfig.jl:28-33), the first argument of polfed(f!, x0, ...);
#   * HIPRescaled / HIPRecurrence / HIPFinalSum: the MappingConfig hooks
#     f!_rescaled, clenshaw_recurrence (six-argument form, kept unwrapped by
#     normalize_clenshaw_recurrence, DefineTransformation.jl:61-82) and
#     clenshaw_finalsum (MappingConfig.jl:82-103);
#   * clenshaw(::Clenshaw, W, V, b, ::GPU, ::NoParallel)
#     (ParallelizationClenshaw.jl:41-51): one pf_filter call per filter
#     application when the transform's recurrence is an HIPRecurrence;
#   * _dos_moments_impl(f!, N, R, D, E, ::GPU) (Moments.jl:58-73): one
#     pf_kpm_moments call when f! is an HIPRescaled;
#   * reorthogonalization!(B, W, ::ProcessingUnit) (reorthogonalization.jl:
#     77-84): two CGS passes on the FP64 matrix cores (pf_reorth).
#
# Julia is not installed where this extension was written, so it has not been
# run; every entry point it calls is exercised through the same C ABI by the
# Python tests.
module PolfedHIPExt

using AMDGPU
using LinearAlgebra

import Polfed
const PC = Polfed.PolfedCore
const LZ = Polfed.Lanczos

const lib = get(ENV, "POLFED_HIP_LIB",
                joinpath(@__DIR__, "..", "..", "lib", "libpolfed_hip.so"))

const PF_F64 = Int32(0)
const PF_F32 = Int32(1)
const PF_PROBE_GAUSSIAN = Int32(0)

check(st::Cint) = st == 0 ? nothing :
    error("libpolfed_hip: ", unsafe_string(ccall((:pf_last_error, lib), Cstring, ())))

# hipStream_t of the task-local AMDGPU stream
stream() = AMDGPU.stream().stream

# ----------------------------------------------------------------- handle
mutable struct Handle
    ptr::Ptr{Cvoid}
end

function Handle(device::Integer)
    r = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:pf_create, lib), Cint, (Ref{Ptr{Cvoid}}, Cint), r, device))
    h = Handle(r[])
    finalizer(x -> ccall((:pf_destroy, lib), Cint, (Ptr{Cvoid},), x.ptr), h)
    return h
end

const HANDLE = Ref{Handle}()
handle() = isassigned(HANDLE) ? HANDLE[] :
    (HANDLE[] = Handle(AMDGPU.device_id(AMDGPU.device()) - 1))

# ---------------------------------------------------------------- operator
"""Operator on the device: `op(Y, X)` writes H X (unscaled), as the f! the
reference's `polfed(f!, x0, nev, target)` takes."""
mutable struct HIPOperator <: Function
    op::Ptr{Cvoid}
    D::Int
end

function _wrap(r::Ref{Ptr{Cvoid}})
    D = ccall((:pf_op_dim, lib), Int64, (Ptr{Cvoid},), r[])
    o = HIPOperator(r[], D)
    finalizer(x -> ccall((:pf_op_destroy, lib), Cint, (Ptr{Cvoid},), x.op), o)
    return o
end

"""Spin-1/2 chain in the sector with `Lup` up spins: bonds (i, j) 1-based with
couplings J and anisotropies Δ, z-fields per site (xxz_hamiltonian /
j1j2_hamiltonian, src/Models/SpinHalfChains.jl:74-220)."""
function HIPOperator(L::Integer, Lup::Integer, bonds::AbstractVector{<:Tuple{Integer,Integer}},
                     J::AbstractVector, Δ::AbstractVector, fields::AbstractVector)
    r = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:pf_op_create_spin_chain, lib), Cint,
        (Ptr{Cvoid}, Int32, Int32, Int32, Ptr{Int32}, Ptr{Int32}, Ptr{Float64}, Ptr{Float64},
         Ptr{Float64}, Ref{Ptr{Cvoid}}),
        handle().ptr, L, Lup, length(bonds), Int32[first.(bonds)...], Int32[last.(bonds)...],
        Float64.(J), Float64.(Δ), Float64.(fields), r))
    return _wrap(r)
end

"""Quantum Sun, spin 1/2, full space (build_hamiltonian, src/QSun/QuantumSun.jl:1-102):
`H_grain` the scaled dg x dg grain matrix, `neighbors` 1-based grain sites,
`couplings` α^u_j, `fields` the outer z-fields."""
function HIPOperator(L_grain::Integer, L_loc::Integer, g0::Real, H_grain::AbstractMatrix,
                     neighbors::AbstractVector, couplings::AbstractVector,
                     fields::AbstractVector)
    r = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:pf_op_create_qsun, lib), Cint,
        (Ptr{Cvoid}, Int32, Int32, Float64, Ptr{Float64}, Ptr{Int32}, Ptr{Float64},
         Ptr{Float64}, Ref{Ptr{Cvoid}}),
        handle().ptr, L_grain, L_loc, g0, Matrix{Float64}(H_grain), Int32.(neighbors .- 1),
        Float64.(couplings), Float64.(fields), r))
    return _wrap(r)
end

"""Any real symmetric matrix, packed by value with the reference's own
get_diags_and_offdiagonals_by_value (src/Optimization/optimization.jl:134-182)
and handed over as 0-based columns with absolute start offsets."""
function HIPOperator(mat::AbstractMatrix{Float64})
    diag, buckets = PC.get_diags_and_offdiagonals_by_value(PC.materialize_host_matrix(mat))
    D = length(diag)
    nv = length(buckets)
    flat = Int32[]
    starts = zeros(Int64, D + 1, nv)
    vals = Float64[]
    for (v, bucket) in enumerate(buckets)
        value, f, st = bucket
        base = length(flat)
        starts[1:D, v] .= base .+ (st[1:D] .- 1)
        starts[D + 1, v] = base + length(f)
        append!(flat, Int32.(f .- 1))
        push!(vals, value)
    end
    r = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:pf_op_create_packed, lib), Cint,
        (Ptr{Cvoid}, Int64, Ptr{Float64}, Int32, Ptr{Float64}, Ptr{Int64}, Ptr{Int32}, Int64,
         Ref{Ptr{Cvoid}}),
        handle().ptr, D, Float64.(diag), nv, vals, starts, flat, length(flat), r))
    return _wrap(r)
end

ld(A::ROCMatrix) = stride(A, 2)
ld(v::ROCVector) = length(v)
ncols(A::ROCMatrix) = size(A, 2)
ncols(::ROCVector) = 1

function _apply(o::HIPOperator, rescaled::Bool, dtype::Int32, Y, X)
    check(ccall((:pf_apply, lib), Cint,
        (Ptr{Cvoid}, Int32, Int32, Ptr{Float64}, Int64, Ptr{Float64}, Int64, Int32, Ptr{Cvoid}),
        o.op, dtype, rescaled, X, ld(X), Y, ld(Y), ncols(X), stream()))
    return nothing
end

(o::HIPOperator)(Y::ROCVecOrMat{Float64}, X::ROCVecOrMat{Float64}) = _apply(o, false, PF_F64, Y, X)

set_rescale!(o::HIPOperator, Emin::Real, Emax::Real) =
    check(ccall((:pf_op_set_rescale, lib), Cint, (Ptr{Cvoid}, Float64, Float64), o.op, Emin, Emax))

# ------------------------------------------------------------ mapping hooks
"""f!_rescaled: Y = (H - b) X / a with the bounds of `set_rescale!`."""
struct HIPRescaled <: Function
    o::HIPOperator
end
(r::HIPRescaled)(Y, X) = _apply(r.o, true, PF_F64, Y, X)

"""clenshaw_recurrence(b1, b2, b3, c, k, X): b1 = c X + 2 H~ b2 - b3 (Chebyshev);
`dtype` PF_F32 runs the whole filter of `clenshaw` below in FP32."""
struct HIPRecurrence <: Function
    o::HIPOperator
    dtype::Int32
end
HIPRecurrence(o::HIPOperator) = HIPRecurrence(o, PF_F64)
function (r::HIPRecurrence)(b1, b2, b3, c, k, X)
    check(ccall((:pf_clenshaw_recurrence, lib), Cint,
        (Ptr{Cvoid}, Int32, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Float64, Float64, Float64,
         Ptr{Float64}, Int64, Int32, Ptr{Cvoid}),
        r.o.op, PF_F64, b1, b2, b3, c, 2.0, -1.0, X, ld(X), ncols(X), stream()))
    return nothing
end

"""clenshaw_finalsum(b2, b3, c0, Y, X): Y = c0 X + H~ b2 - b3 (Chebyshev)."""
struct HIPFinalSum <: Function
    o::HIPOperator
end
function (f::HIPFinalSum)(b2, b3, c0, Y, X)
    check(ccall((:pf_clenshaw_finalsum, lib), Cint,
        (Ptr{Cvoid}, Int32, Ptr{Float64}, Ptr{Float64}, Float64, Float64, Float64, Float64,
         Ptr{Float64}, Ptr{Float64}, Int64, Int32, Ptr{Cvoid}),
        f.o.op, PF_F64, b2, b3, c0, -1.0, 1.0, 1.0, Y, X, ld(X), ncols(X), stream()))
    return nothing
end

"""The MappingConfig keywords that route a polfed call through the library:
`polfed(op, x0, nev, target; mapping = MappingConfig(; hip_mapping(op, Emin, Emax)...))`."""
function hip_mapping(o::HIPOperator, Emin::Real, Emax::Real; dtype::Int32 = PF_F64)
    set_rescale!(o, Emin, Emax)
    return (; var"f!_rescaled" = HIPRescaled(o), clenshaw_recurrence = HIPRecurrence(o, dtype),
            clenshaw_finalsum = HIPFinalSum(o), Emin = Emin, Emax = Emax)
end

# ------------------------------------------------------- whole-filter call
const ROCBlock = Union{ROCMatrix{Float64},ROCVector{Float64}}

function PC.clenshaw(t::PC.ClenshawType, W::ROCBlock, V::ROCBlock, b::AbstractVector,
                     pu::PC.GPU, ps::PC.NoParallel)
    rec = t.recurrence!
    rec isa HIPRecurrence || return t(W, V, b)   # user hooks: the generic Clenshaw loop
    t.type === :Chebyshev || return t(W, V, b)
    K = t.order
    c = Float64[t.coefficients(n) for n in 0:K]
    check(ccall((:pf_filter, lib), Cint,
        (Ptr{Cvoid}, Int32, Int32, Ptr{Float64}, Int32, Ptr{Float64}, Int64, Ptr{Float64},
         Int64, Int32, Ptr{Cvoid}),
        rec.o.op, rec.dtype, 0, c, K, V, ld(V), W, ld(W), ncols(V), stream()))
    return nothing
end

# ------------------------------------------------------------ KPM moments
function PC._dos_moments_impl(f!::HIPRescaled, N::Int, R::Int, D::Int, E::Type{Float64},
                              pu::PC.GPU)
    P = AMDGPU.zeros(Float64, D, R)
    check(ccall((:pf_fill_probes, lib), Cint,
        (Int32, UInt64, Int64, Int32, Ptr{Float64}, Int64, Ptr{Cvoid}),
        PF_PROBE_GAUSSIAN, rand(UInt64), D, R, P, D, stream()))
    μ = zeros(Float64, N)
    check(ccall((:pf_kpm_moments, lib), Cint,
        (Ptr{Cvoid}, Int32, Int32, Int32, Ptr{Float64}, Int64, Ptr{Float64}, Ptr{Cvoid}),
        f!.o.op, PF_F64, N, R, P, D, μ, stream()))
    return μ
end

# --------------------------------------------- full reorthogonalization
const ROCBasis = Union{ROCMatrix{Float64},SubArray{Float64,2,<:ROCMatrix{Float64}}}

function LZ.reorthogonalization!(B::ROCBasis, W::ROCMatrix{Float64}, ::LZ.ProcessingUnit)
    size(B, 2) == 0 && return nothing
    check(ccall((:pf_reorth, lib), Cint,
        (Ptr{Cvoid}, Ptr{Float64}, Int64, Int32, Ptr{Float64}, Int64, Int32, Int64, Int32,
         Ptr{Cvoid}),
        handle().ptr, pointer(B), stride(B, 2), size(B, 2), W, stride(W, 2), size(W, 2),
        size(W, 1), 2, stream()))
    return nothing
end

# ------------------------------------------------------- backend registry
_memory_info() = AMDGPU.Mem.info()              # (free, total) bytes of the device
_available_memory() = first(AMDGPU.Mem.info())
_total_memory() = last(AMDGPU.Mem.info())

function __init__()
    Polfed._set_cuda_backend!(
        array = AMDGPU.ROCArray,
        matrix = AMDGPU.ROCMatrix,
        vector = AMDGPU.ROCVector,
        matrix_type = AMDGPU.ROCMatrix,
        vector_type = AMDGPU.ROCVector,
        zeros = AMDGPU.zeros,
        ones = AMDGPU.ones,
        rand = AMDGPU.rand,
        randn = AMDGPU.randn,
        memory_info = _memory_info,
        available_memory = _available_memory,
        total_memory = _total_memory,
        functional = AMDGPU.functional,
        device_count = () -> length(AMDGPU.devices()),
        synchronize = AMDGPU.synchronize,
        reclaim = () -> GC.gc(),
        allowscalar = (f -> AMDGPU.@allowscalar f()),
    )
    return nothing
end

Polfed.is_gpu_array(::AMDGPU.ROCArray) = true

end # module
