"""
    JBlasAMD

Julia front end of libjblas_hip.so, the MI355X (gfx950) GEMM engine behind
jBLAS.jl's GEMM entry point.  Drop-ins:

  * `jmul!(D, A, X)`  -- D = A*X with jBLAS's per-element fma recurrence
    (src/gemm.jl:244-348 of jBLAS.jl), bit-identical, fringe included;
    accepts any strided `Matrix{Float64}` (the reference takes `MMatrix`).
  * `gemm!(tA, tB, α, A, B, β, C; method)` -- LinearAlgebra.BLAS.gemm!
    argument order; Float64 methods `:native` (FP64 MFMA), `:recurrence`,
    `:ozaki` (split-BF16, keyword `nslices`, default 6 slices of 9 bits), `:ozaki_int8` (INT8 slices of
    8 bits, default 7); Float32 `:native`, `:bf16x3`.
  * `fastmul!(D, A, X)` -- the same recurrence as one small-GEMM launch
    (src/kernels.jl:202-208), Float64 or Float32, dense matrices.
  * `kernel!(D, A, X)` / `initkernel!(D, A, X)` -- D += A*X (each chain
    starting from D) / D = A*X on strided sub-blocks (src/kernels.jl:212-275).
  * `gemm(tA, tB, α, A, B)` -- non-mutating form with a ChainRulesCore
    `rrule` (package extension) backed by the library's pullback entry point.
  * `GroupedPlan` -- many GEMMs in one persistent, hipGraph-captured launch.

All calls go through `ccall` into the C ABI declared in include/jblas_hip.h.
Host arrays use the synchronous host-pointer entry points.  Package
extensions add the reference's own signatures and device arrays:
  * StaticArrays: `jmul!(D::MMatrix, A::MMatrix, X::MMatrix, ::Val...)` and
    `fastmul!(::MMatrix...)` on `pointer()` (src/gemm.jl:244-246,
    src/kernels.jl:202-208);
  * jBLAS: methods of `jBLAS.jmul!`, `jBLAS.fastmul!`, `jBLAS.kernel!` and
    `jBLAS.initkernel!` for Float64 (and Float32 `jmul!`), so that an unchanged
    jBLAS call runs here;
  * AMDGPU: `jmul!`, `gemm!`, `fastmul!`, `kernel!`, `initkernel!` on
    `ROCMatrix` through the `*_dev` entry points, on the array's HIP stream.
"""
module JBlasAMD

using LinearAlgebra

export jmul!, fastmul!, gemm!, gemm, kernel!, initkernel!, Kernel, GroupedPlan, launch!, destroy!, nonfinite_check,
       gemm_grouped!, autotune!, tune_save, tune_load, tune_reset, tune_lookup, grouped_choices

const libjblas = get(ENV, "JB_LIBRARY", normpath(joinpath(@__DIR__, "..", "..", "..", "lib", "libjblas_hip.so")))

const JB_F64_NATIVE = Cint(0)
const JB_F64_RECURRENCE = Cint(1)
const JB_F32_NATIVE = Cint(2)
const JB_F32_BF16X3 = Cint(3)

struct JBError <: Exception
    func::Symbol
    code::Cint
    msg::String
end
Base.showerror(io::IO, e::JBError) =
    print(io, "JBError in ", e.func, ": ", e.code < 0 ? "argument $(-e.code) invalid" : "code $(e.code)", ": ", e.msg)

last_error() = unsafe_string(ccall((:jb_last_error, libjblas), Cstring, ()))
check(func::Symbol, rc::Cint) = rc == 0 ? nothing : throw(JBError(func, rc, last_error()))

colmajor(A::StridedMatrix) = stride(A, 1) == 1
ld(A::StridedMatrix) = max(1, stride(A, 2), size(A, 1))

# The reference's jmul! takes `Val` prefetch knobs after the operands
# (src/gemm.jl:244-246); they tune CPU software prefetching and have no
# counterpart on the GPU, so they are accepted and ignored.
function jmul!(D::StridedMatrix{Float64}, A::StridedMatrix{Float64}, X::StridedMatrix{Float64},
               ::Val = Val(7), ::Val = Val(7), ::Val = Val(3), ::Val = Val(3), ::Val = Val(3))
    M, N = size(A)
    N2, P = size(X)
    (N == N2 && size(D) == (M, P)) || throw(DimensionMismatch("D $(size(D)), A $(size(A)), X $(size(X))"))
    if colmajor(D) && colmajor(A) && colmajor(X) && ld(D) == max(1, M) && ld(A) == max(1, M) && ld(X) == max(1, N)
        check(:jb_jmul, ccall((:jb_jmul, libjblas), Cint,
                              (Int64, Int64, Int64, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}), M, N, P, D, A, X))
        return D
    end
    gemm!('N', 'N', 1.0, A, X, 0.0, D; method = :recurrence)
end

"Float32 jmul! (the reference's is generic in T): the binary32 recurrence, dense matrices."
function jmul!(D::StridedMatrix{Float32}, A::StridedMatrix{Float32}, X::StridedMatrix{Float32},
               ::Val = Val(7), ::Val = Val(7), ::Val = Val(3), ::Val = Val(3), ::Val = Val(3))
    M, N, P = _fastmul_dims(D, A, X)
    check(:jb_jmul_f32, ccall((:jb_jmul_f32, libjblas), Cint,
                              (Int64, Int64, Int64, Ptr{Float32}, Ptr{Float32}, Ptr{Float32}), M, N, P, D, A, X))
    D
end

"""
    jmul_ptr!(pD, pA, pX, M, N, P)

jmul! on raw host pointers to dense column-major storage (leading dimensions
M, N, M), the form the reference's own generated code works on (`pointer(D)`,
src/gemm.jl:309-311).  Used by the StaticArrays extension for `MMatrix`.
"""
jmul_ptr!(pD::Ptr{Float64}, pA::Ptr{Float64}, pX::Ptr{Float64}, M::Integer, N::Integer, P::Integer) =
    check(:jb_jmul, ccall((:jb_jmul, libjblas), Cint,
                          (Int64, Int64, Int64, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}), M, N, P, pD, pA, pX))
jmul_ptr!(pD::Ptr{Float32}, pA::Ptr{Float32}, pX::Ptr{Float32}, M::Integer, N::Integer, P::Integer) =
    check(:jb_jmul_f32, ccall((:jb_jmul_f32, libjblas), Cint,
                              (Int64, Int64, Int64, Ptr{Float32}, Ptr{Float32}, Ptr{Float32}), M, N, P, pD, pA, pX))
fastmul_ptr!(pD::Ptr{Float64}, pA::Ptr{Float64}, pX::Ptr{Float64}, M::Integer, N::Integer, P::Integer) =
    check(:jb_fastmul, ccall((:jb_fastmul, libjblas), Cint,
                             (Int64, Int64, Int64, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}), M, N, P, pD, pA, pX))
fastmul_ptr!(pD::Ptr{Float32}, pA::Ptr{Float32}, pX::Ptr{Float32}, M::Integer, N::Integer, P::Integer) =
    check(:jb_fastmul_f32, ccall((:jb_fastmul_f32, libjblas), Cint,
                                 (Int64, Int64, Int64, Ptr{Float32}, Ptr{Float32}, Ptr{Float32}), M, N, P, pD, pA, pX))

function _fastmul_dims(D, A, X)
    M, N = size(A)
    N2, P = size(X)
    (N == N2 && size(D) == (M, P)) || throw(DimensionMismatch("D $(size(D)), A $(size(A)), X $(size(X))"))
    (colmajor(D) && colmajor(A) && colmajor(X) && ld(D) == max(1, M) && ld(A) == max(1, M) &&
     ld(X) == max(1, N)) || throw(ArgumentError("fastmul! takes dense column-major matrices"))
    M, N, P
end

function fastmul!(D::StridedMatrix{Float64}, A::StridedMatrix{Float64}, X::StridedMatrix{Float64})
    M, N, P = _fastmul_dims(D, A, X)
    check(:jb_fastmul, ccall((:jb_fastmul, libjblas), Cint,
                             (Int64, Int64, Int64, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}), M, N, P, D, A, X))
    D
end

function fastmul!(D::StridedMatrix{Float32}, A::StridedMatrix{Float32}, X::StridedMatrix{Float32})
    M, N, P = _fastmul_dims(D, A, X)
    check(:jb_fastmul_f32, ccall((:jb_fastmul_f32, libjblas), Cint,
                                 (Int64, Int64, Int64, Ptr{Float32}, Ptr{Float32}, Ptr{Float32}), M, N, P, D, A, X))
    D
end

function _dims(tA, tB, A, B, C)
    m, k = tA == 'N' ? size(A) : reverse(size(A))
    k2, n = tB == 'N' ? size(B) : reverse(size(B))
    (k == k2 && size(C) == (m, n)) || throw(DimensionMismatch("op(A) $m×$k, op(B) $k2×$n, C $(size(C))"))
    m, n, k
end

function gemm!(tA::AbstractChar, tB::AbstractChar, α::Real, A::StridedMatrix{Float64}, B::StridedMatrix{Float64},
               β::Real, C::StridedMatrix{Float64}; method::Symbol = :native,
               nslices::Integer = method === :ozaki_int8 ? 7 : 6)
    (colmajor(A) && colmajor(B) && colmajor(C)) || throw(ArgumentError("unit-stride columns required"))
    m, n, k = _dims(tA, tB, A, B, C)
    if method === :native
        check(:jb_dgemm, ccall((:jb_dgemm, libjblas), Cint,
            (Cchar, Cchar, Int64, Int64, Int64, Float64, Ptr{Float64}, Int64, Ptr{Float64}, Int64, Float64,
             Ptr{Float64}, Int64), tA, tB, m, n, k, α, A, ld(A), B, ld(B), β, C, ld(C)))
    elseif method === :recurrence
        check(:jb_dgemm_recurrence, ccall((:jb_dgemm_recurrence, libjblas), Cint,
            (Cchar, Cchar, Int64, Int64, Int64, Float64, Ptr{Float64}, Int64, Ptr{Float64}, Int64, Float64,
             Ptr{Float64}, Int64), tA, tB, m, n, k, α, A, ld(A), B, ld(B), β, C, ld(C)))
    elseif method === :ozaki
        check(:jb_dgemm_ozaki, ccall((:jb_dgemm_ozaki, libjblas), Cint,
            (Cchar, Cchar, Int64, Int64, Int64, Float64, Ptr{Float64}, Int64, Ptr{Float64}, Int64, Float64,
             Ptr{Float64}, Int64, Cint), tA, tB, m, n, k, α, A, ld(A), B, ld(B), β, C, ld(C), nslices))
    elseif method === :ozaki_int8
        check(:jb_dgemm_ozaki_i8, ccall((:jb_dgemm_ozaki_i8, libjblas), Cint,
            (Cchar, Cchar, Int64, Int64, Int64, Float64, Ptr{Float64}, Int64, Ptr{Float64}, Int64, Float64,
             Ptr{Float64}, Int64, Cint), tA, tB, m, n, k, α, A, ld(A), B, ld(B), β, C, ld(C), nslices))
    else
        throw(ArgumentError("Float64 method must be :native, :recurrence, :ozaki or :ozaki_int8"))
    end
    C
end

function gemm!(tA::AbstractChar, tB::AbstractChar, α::Real, A::StridedMatrix{Float32}, B::StridedMatrix{Float32},
               β::Real, C::StridedMatrix{Float32}; method::Symbol = :native, nprod::Integer = 6)
    (colmajor(A) && colmajor(B) && colmajor(C)) || throw(ArgumentError("unit-stride columns required"))
    m, n, k = _dims(tA, tB, A, B, C)
    if method === :native
        check(:jb_sgemm, ccall((:jb_sgemm, libjblas), Cint,
            (Cchar, Cchar, Int64, Int64, Int64, Float32, Ptr{Float32}, Int64, Ptr{Float32}, Int64, Float32,
             Ptr{Float32}, Int64), tA, tB, m, n, k, α, A, ld(A), B, ld(B), β, C, ld(C)))
    elseif method === :bf16x3
        check(:jb_sgemm_bf16x3, ccall((:jb_sgemm_bf16x3, libjblas), Cint,
            (Cchar, Cchar, Int64, Int64, Int64, Float32, Ptr{Float32}, Int64, Ptr{Float32}, Int64, Float32,
             Ptr{Float32}, Int64, Cint), tA, tB, m, n, k, α, A, ld(A), B, ld(B), β, C, ld(C), nprod))
    else
        throw(ArgumentError("Float32 method must be :native or :bf16x3"))
    end
    C
end

function _kernel_dims(D, A, X)
    M, N = size(A)
    N2, P = size(X)
    (N == N2 && size(D) == (M, P)) || throw(DimensionMismatch("D $(size(D)), A $(size(A)), X $(size(X))"))
    (colmajor(D) && colmajor(A) && colmajor(X)) || throw(ArgumentError("unit-stride columns required"))
    M, N, P
end

"""D += A*X on strided sub-blocks (kernel!, src/kernels.jl:212-241): each
element's fma chain starts from D."""
function kernel!(D::StridedMatrix{Float64}, A::StridedMatrix{Float64}, X::StridedMatrix{Float64})
    M, N, P = _kernel_dims(D, A, X)
    check(:jb_kernel, ccall((:jb_kernel, libjblas), Cint,
        (Int64, Int64, Int64, Ptr{Float64}, Int64, Ptr{Float64}, Int64, Ptr{Float64}, Int64),
        M, N, P, D, ld(D), A, ld(A), X, ld(X)))
    D
end

"D = A*X on strided sub-blocks (initkernel!, src/kernels.jl:242-275), the jmul! recurrence."
function initkernel!(D::StridedMatrix{Float64}, A::StridedMatrix{Float64}, X::StridedMatrix{Float64})
    M, N, P = _kernel_dims(D, A, X)
    check(:jb_initkernel, ccall((:jb_initkernel, libjblas), Cint,
        (Int64, Int64, Int64, Ptr{Float64}, Int64, Ptr{Float64}, Int64, Ptr{Float64}, Int64),
        M, N, P, D, ld(D), A, ld(A), X, ld(X)))
    D
end

"""
    Kernel{Mₖ,Pₖ,stride_AD,stride_X,N}
    Kernel(Mₖ, N, Pₖ, stride_AD, stride_X)

The reference's type-level microkernel descriptor (src/kernel_structure.jl:8-9):
an Mₖ×Pₖ tile of D = A·X with K = N, A and D sharing the leading dimension
stride_AD, X with leading dimension stride_X.  Same parameters, same
constructor argument order.
"""
struct Kernel{Mₖ,Pₖ,stride_AD,stride_X,N} end
Kernel(Mₖ, N, Pₖ, stride_AD, stride_X) = Kernel{Mₖ,Pₖ,stride_AD,stride_X,N}()

"""
    kernel!(pD::Ptr{Float64}, pA, pX, ::Kernel{Mₖ,Pₖ,stride_AD,stride_X,N}[, prefetch])

The reference's pointer microkernel (src/kernels.jl:239-241, prefetch variants
:277-537): D += A·X on the Mₖ×Pₖ tile at pD, each element's fma chain starting
from D, with ldd = lda = stride_AD and ldx = stride_X.  Host pointers; the
prefetch argument (a CPU software-prefetch schedule) is accepted and ignored.
Returns `nothing`, as the reference.
"""
function kernel!(pD::Ptr{Float64}, pA::Ptr{Float64}, pX::Ptr{Float64}, ::Kernel{Mₖ,Pₖ,stride_AD,stride_X,N},
                 pf = nothing) where {Mₖ,Pₖ,stride_AD,stride_X,N}
    check(:jb_kernel, ccall((:jb_kernel, libjblas), Cint,
        (Int64, Int64, Int64, Ptr{Float64}, Int64, Ptr{Float64}, Int64, Ptr{Float64}, Int64),
        Mₖ, N, Pₖ, pD, stride_AD, pA, stride_AD, pX, stride_X))
    nothing
end

"""
    initkernel!(pD::Ptr{Float64}, pA, pX, ::Kernel{Mₖ,Pₖ,stride_AD,stride_X,N}[, prefetch])

The reference's initialising microkernel (src/kernels.jl:273-275): D = A·X on
the tile, the first k-step a plain product, then fma (the jmul! recurrence).
"""
function initkernel!(pD::Ptr{Float64}, pA::Ptr{Float64}, pX::Ptr{Float64}, ::Kernel{Mₖ,Pₖ,stride_AD,stride_X,N},
                     pf = nothing) where {Mₖ,Pₖ,stride_AD,stride_X,N}
    check(:jb_initkernel, ccall((:jb_initkernel, libjblas), Cint,
        (Int64, Int64, Int64, Ptr{Float64}, Int64, Ptr{Float64}, Int64, Ptr{Float64}, Int64),
        Mₖ, N, Pₖ, pD, stride_AD, pA, stride_AD, pX, stride_X))
    nothing
end

"""
    nonfinite_check(stream = C_NULL)

Throws `JBError` (code 3) if an emulated call on device pointers (split-BF16,
split-INT8, 3xBF16) met Inf or NaN in an operand since the previous check
(those calls do not synchronise; the C elements such operands feed are NaN).
"""
nonfinite_check(stream::Ptr{Cvoid} = C_NULL) =
    check(:jb_nonfinite_check, ccall((:jb_nonfinite_check, libjblas), Cint, (Ptr{Cvoid},), stream))

"Non-mutating α*op(A)*op(B); differentiable through the ChainRulesCore extension."
function gemm(tA::AbstractChar, tB::AbstractChar, α::Real, A::StridedMatrix{T}, B::StridedMatrix{T};
              kw...) where {T<:Union{Float32,Float64}}
    m = tA == 'N' ? size(A, 1) : size(A, 2)
    n = tB == 'N' ? size(B, 2) : size(B, 1)
    gemm!(tA, tB, α, A, B, zero(T), Matrix{T}(undef, m, n); kw...)
end

"""
    pullback!(tA, tB, α, A, B, dC, dA, dB; method)

dA = α dC op(B)ᵀ (or its transpose for tA = 'T') and dB = α op(A)ᵀ dC (or
transposed for tB = 'T'), both in one grouped launch (jb_gemm_pullback).
Either of dA, dB may be `nothing`.
"""
function pullback!(tA::AbstractChar, tB::AbstractChar, α::Real, A::StridedMatrix{T}, B::StridedMatrix{T},
                   dC::StridedMatrix{T}, dA, dB; method::Symbol = :native,
                   kw...) where {T<:Union{Float32,Float64}}
    m, k = tA == 'N' ? size(A) : reverse(size(A))
    n = tB == 'N' ? size(B, 2) : size(B, 1)
    size(dC) == (m, n) || throw(DimensionMismatch("dC must be $m×$n"))
    mid = T === Float64 ? JB_F64_NATIVE : (method === :bf16x3 ? JB_F32_BF16X3 : JB_F32_NATIVE)
    pA = dA === nothing ? Ptr{T}(C_NULL) : pointer(dA)
    pB = dB === nothing ? Ptr{T}(C_NULL) : pointer(dB)
    GC.@preserve dA dB begin
        check(:jb_gemm_pullback, ccall((:jb_gemm_pullback, libjblas), Cint,
            (Cint, Cchar, Cchar, Int64, Int64, Int64, Float64, Ptr{T}, Int64, Ptr{T}, Int64, Ptr{T}, Int64,
             Ptr{T}, Int64, Ptr{T}, Int64),
            mid, tA, tB, m, n, k, α, A, ld(A), B, ld(B), dC, ld(dC),
            pA, dA === nothing ? 1 : ld(dA), pB, dB === nothing ? 1 : ld(dB)))
    end
    dA, dB
end

# ---- grouped launches on device memory ---------------------------------------
"C layout of jb_gemm_desc (include/jblas_hip.h)."
struct GemmDesc
    transA::Cchar
    transB::Cchar
    pad::NTuple{6,UInt8}
    m::Int64
    n::Int64
    k::Int64
    alpha::Float64
    beta::Float64
    A::Ptr{Cvoid}
    lda::Int64
    B::Ptr{Cvoid}
    ldb::Int64
    C::Ptr{Cvoid}
    ldc::Int64
end
GemmDesc(tA, tB, m, n, k, α, A, lda, B, ldb, β, C, ldc) =
    GemmDesc(Cchar(tA), Cchar(tB), ntuple(_ -> 0x00, 6), m, n, k, α, β, A, lda, B, ldb, C, ldc)

"A hipGraph-captured persistent launch of many GEMMs on device pointers."
mutable struct GroupedPlan
    handle::Ptr{Cvoid}
    descs::Vector{GemmDesc}
end

function GroupedPlan(method::Cint, descs::Vector{GemmDesc}; stream::Ptr{Cvoid} = C_NULL)
    h = Ref{Ptr{Cvoid}}(C_NULL)
    check(:jb_grouped_plan_create, ccall((:jb_grouped_plan_create, libjblas), Cint,
        (Cint, Int64, Ptr{GemmDesc}, Ptr{Cvoid}, Ref{Ptr{Cvoid}}), method, length(descs), descs, stream, h))
    p = GroupedPlan(h[], descs)
    finalizer(destroy!, p)
    p
end

launch!(p::GroupedPlan; stream::Ptr{Cvoid} = C_NULL) =
    check(:jb_grouped_plan_launch, ccall((:jb_grouped_plan_launch, libjblas), Cint, (Ptr{Cvoid}, Ptr{Cvoid}),
                                         p.handle, stream))

function destroy!(p::GroupedPlan)
    if p.handle != C_NULL
        ccall((:jb_grouped_plan_destroy, libjblas), Cint, (Ptr{Cvoid},), p.handle)
        p.handle = C_NULL
    end
    nothing
end

"All problems in one set of persistent launches, without a graph (synchronous on `stream`)."
gemm_grouped!(method::Cint, descs::Vector{GemmDesc}; stream::Ptr{Cvoid} = C_NULL) =
    check(:jb_gemm_grouped_dev, ccall((:jb_gemm_grouped_dev, libjblas), Cint,
                                      (Cint, Int64, Ptr{GemmDesc}, Ptr{Cvoid}), method, length(descs), descs, stream))

# ---- the autotuner (replaces pick_kernel_size / blocking_structure's static choice) ----
const TILE_NAMES = Dict(0 => "64x64", 1 => "32x32", 2 => "small", 3 => "128x128", 4 => "128x64", 5 => "64x128")

"Tune tile shape and split-K per size bucket of `descs`, timing the whole group (the C matrices are not touched)."
autotune!(method::Cint, descs::Vector{GemmDesc}; reps::Integer = 3, stream::Ptr{Cvoid} = C_NULL) =
    check(:jb_autotune_grouped, ccall((:jb_autotune_grouped, libjblas), Cint,
                                      (Cint, Int64, Ptr{GemmDesc}, Cint, Ptr{Cvoid}),
                                      method, length(descs), descs, Cint(reps), stream))
tune_save(path::AbstractString) = check(:jb_autotune_save, ccall((:jb_autotune_save, libjblas), Cint, (Cstring,), path))
tune_load(path::AbstractString) = check(:jb_autotune_load, ccall((:jb_autotune_load, libjblas), Cint, (Cstring,), path))
tune_reset() = check(:jb_autotune_reset, ccall((:jb_autotune_reset, libjblas), Cint, ()))

"(tile name, split-K) the table or the static rule gives one problem shape."
function tune_lookup(method::Cint, m::Integer, n::Integer, k::Integer)
    t, s = Ref{Cint}(0), Ref{Cint}(0)
    check(:jb_autotune_lookup, ccall((:jb_autotune_lookup, libjblas), Cint,
                                     (Cint, Int64, Int64, Int64, Ref{Cint}, Ref{Cint}), method, m, n, k, t, s))
    (TILE_NAMES[Int(t[])], Int(s[]))
end

"Per problem, the (tile name, split-K, k per split) a plan of `descs` would use now (`nothing` for m or n == 0)."
function grouped_choices(method::Cint, descs::Vector{GemmDesc})
    n = length(descs)
    t, s, kc = zeros(Cint, n), zeros(Cint, n), zeros(Int64, n)
    check(:jb_grouped_choices, ccall((:jb_grouped_choices, libjblas), Cint,
                                     (Cint, Int64, Ptr{GemmDesc}, Ptr{Cint}, Ptr{Cint}, Ptr{Int64}),
                                     method, n, descs, t, s, kc))
    [t[i] < 0 ? nothing : (TILE_NAMES[Int(t[i])], Int(s[i]), Int(kc[i])) for i in 1:n]
end

end # module
