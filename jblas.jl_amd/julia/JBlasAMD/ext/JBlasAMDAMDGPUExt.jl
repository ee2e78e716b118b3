module JBlasAMDAMDGPUExt

# Device-resident operands: ROCMatrix arguments go to the *_dev entry points with
# device pointers, on the HIP stream of the calling task (asynchronous, ordered
# with the caller's other work on that stream).  No host staging, no PCIe copy.

using JBlasAMD: JBlasAMD, check, libjblas
using AMDGPU: AMDGPU, ROCMatrix

devptr(A::ROCMatrix{T}) where {T} = reinterpret(Ptr{T}, pointer(A))
ldd(A::ROCMatrix) = max(1, size(A, 1))
hipstream() = AMDGPU.stream().stream

function JBlasAMD.jmul!(D::ROCMatrix{Float64}, A::ROCMatrix{Float64}, X::ROCMatrix{Float64},
                        ::Val = Val(7), ::Val = Val(7), ::Val = Val(3), ::Val = Val(3), ::Val = Val(3))
    M, N = size(A); N2, P = size(X)
    (N == N2 && size(D) == (M, P)) || throw(DimensionMismatch("D $(size(D)), A $(size(A)), X $(size(X))"))
    check(:jb_jmul_dev, ccall((:jb_jmul_dev, libjblas), Cint,
        (Int64, Int64, Int64, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Cvoid}),
        M, N, P, devptr(D), devptr(A), devptr(X), hipstream()))
    D
end

function JBlasAMD.jmul!(D::ROCMatrix{Float32}, A::ROCMatrix{Float32}, X::ROCMatrix{Float32},
                        ::Val = Val(7), ::Val = Val(7), ::Val = Val(3), ::Val = Val(3), ::Val = Val(3))
    M, N = size(A); N2, P = size(X)
    (N == N2 && size(D) == (M, P)) || throw(DimensionMismatch("D $(size(D)), A $(size(A)), X $(size(X))"))
    check(:jb_jmul_f32_dev, ccall((:jb_jmul_f32_dev, libjblas), Cint,
        (Int64, Int64, Int64, Ptr{Float32}, Ptr{Float32}, Ptr{Float32}, Ptr{Cvoid}),
        M, N, P, devptr(D), devptr(A), devptr(X), hipstream()))
    D
end

function JBlasAMD.fastmul!(D::ROCMatrix{Float64}, A::ROCMatrix{Float64}, X::ROCMatrix{Float64})
    M, N = size(A); N2, P = size(X)
    (N == N2 && size(D) == (M, P)) || throw(DimensionMismatch("D $(size(D)), A $(size(A)), X $(size(X))"))
    check(:jb_fastmul_dev, ccall((:jb_fastmul_dev, libjblas), Cint,
        (Int64, Int64, Int64, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Cvoid}),
        M, N, P, devptr(D), devptr(A), devptr(X), hipstream()))
    D
end

function JBlasAMD.fastmul!(D::ROCMatrix{Float32}, A::ROCMatrix{Float32}, X::ROCMatrix{Float32})
    M, N = size(A); N2, P = size(X)
    (N == N2 && size(D) == (M, P)) || throw(DimensionMismatch("D $(size(D)), A $(size(A)), X $(size(X))"))
    check(:jb_fastmul_f32_dev, ccall((:jb_fastmul_f32_dev, libjblas), Cint,
        (Int64, Int64, Int64, Ptr{Float32}, Ptr{Float32}, Ptr{Float32}, Ptr{Cvoid}),
        M, N, P, devptr(D), devptr(A), devptr(X), hipstream()))
    D
end

"D += A*X (kernel!, src/kernels.jl:212-241) on device matrices."
function JBlasAMD.kernel!(D::ROCMatrix{Float64}, A::ROCMatrix{Float64}, X::ROCMatrix{Float64})
    M, N = size(A); N2, P = size(X)
    (N == N2 && size(D) == (M, P)) || throw(DimensionMismatch("D $(size(D)), A $(size(A)), X $(size(X))"))
    check(:jb_kernel_dev, ccall((:jb_kernel_dev, libjblas), Cint,
        (Int64, Int64, Int64, Ptr{Float64}, Int64, Ptr{Float64}, Int64, Ptr{Float64}, Int64, Ptr{Cvoid}),
        M, N, P, devptr(D), ldd(D), devptr(A), ldd(A), devptr(X), ldd(X), hipstream()))
    D
end

"D = A*X (initkernel!, src/kernels.jl:242-275) on device matrices."
function JBlasAMD.initkernel!(D::ROCMatrix{Float64}, A::ROCMatrix{Float64}, X::ROCMatrix{Float64})
    M, N = size(A); N2, P = size(X)
    (N == N2 && size(D) == (M, P)) || throw(DimensionMismatch("D $(size(D)), A $(size(A)), X $(size(X))"))
    check(:jb_initkernel_dev, ccall((:jb_initkernel_dev, libjblas), Cint,
        (Int64, Int64, Int64, Ptr{Float64}, Int64, Ptr{Float64}, Int64, Ptr{Float64}, Int64, Ptr{Cvoid}),
        M, N, P, devptr(D), ldd(D), devptr(A), ldd(A), devptr(X), ldd(X), hipstream()))
    D
end

function _dims(tA, tB, A, B, C)
    m, k = tA == 'N' ? size(A) : reverse(size(A))
    k2, n = tB == 'N' ? size(B) : reverse(size(B))
    (k == k2 && size(C) == (m, n)) || throw(DimensionMismatch("op(A) $m×$k, op(B) $k2×$n, C $(size(C))"))
    m, n, k
end

"""
    gemm!(tA, tB, α, A::ROCMatrix, B::ROCMatrix, β, C::ROCMatrix; method, nslices, nprod)

LinearAlgebra.BLAS.gemm! argument order on device matrices; methods as for
host matrices.  The emulated methods do not synchronise: Inf/NaN operands make
the C elements they feed NaN and are reported by `JBlasAMD.nonfinite_check`.
"""
function JBlasAMD.gemm!(tA::AbstractChar, tB::AbstractChar, α::Real, A::ROCMatrix{Float64}, B::ROCMatrix{Float64},
                        β::Real, C::ROCMatrix{Float64}; method::Symbol = :native,
                        nslices::Integer = method === :ozaki_int8 ? 7 : 6)
    m, n, k = _dims(tA, tB, A, B, C)
    a, b, c, s = devptr(A), devptr(B), devptr(C), hipstream()
    if method === :native
        check(:jb_dgemm_dev, ccall((:jb_dgemm_dev, libjblas), Cint,
            (Cchar, Cchar, Int64, Int64, Int64, Float64, Ptr{Float64}, Int64, Ptr{Float64}, Int64, Float64,
             Ptr{Float64}, Int64, Ptr{Cvoid}), tA, tB, m, n, k, α, a, ldd(A), b, ldd(B), β, c, ldd(C), s))
    elseif method === :recurrence
        check(:jb_dgemm_recurrence_dev, ccall((:jb_dgemm_recurrence_dev, libjblas), Cint,
            (Cchar, Cchar, Int64, Int64, Int64, Float64, Ptr{Float64}, Int64, Ptr{Float64}, Int64, Float64,
             Ptr{Float64}, Int64, Ptr{Cvoid}), tA, tB, m, n, k, α, a, ldd(A), b, ldd(B), β, c, ldd(C), s))
    elseif method === :ozaki
        check(:jb_dgemm_ozaki_dev, ccall((:jb_dgemm_ozaki_dev, libjblas), Cint,
            (Cchar, Cchar, Int64, Int64, Int64, Float64, Ptr{Float64}, Int64, Ptr{Float64}, Int64, Float64,
             Ptr{Float64}, Int64, Cint, Ptr{Cvoid}), tA, tB, m, n, k, α, a, ldd(A), b, ldd(B), β, c, ldd(C),
            nslices, s))
    elseif method === :ozaki_int8
        check(:jb_dgemm_ozaki_i8_dev, ccall((:jb_dgemm_ozaki_i8_dev, libjblas), Cint,
            (Cchar, Cchar, Int64, Int64, Int64, Float64, Ptr{Float64}, Int64, Ptr{Float64}, Int64, Float64,
             Ptr{Float64}, Int64, Cint, Ptr{Cvoid}), tA, tB, m, n, k, α, a, ldd(A), b, ldd(B), β, c, ldd(C),
            nslices, s))
    else
        throw(ArgumentError("Float64 method must be :native, :recurrence, :ozaki or :ozaki_int8"))
    end
    C
end

function JBlasAMD.gemm!(tA::AbstractChar, tB::AbstractChar, α::Real, A::ROCMatrix{Float32}, B::ROCMatrix{Float32},
                        β::Real, C::ROCMatrix{Float32}; method::Symbol = :native, nprod::Integer = 6)
    m, n, k = _dims(tA, tB, A, B, C)
    a, b, c, s = devptr(A), devptr(B), devptr(C), hipstream()
    if method === :native
        check(:jb_sgemm_dev, ccall((:jb_sgemm_dev, libjblas), Cint,
            (Cchar, Cchar, Int64, Int64, Int64, Float32, Ptr{Float32}, Int64, Ptr{Float32}, Int64, Float32,
             Ptr{Float32}, Int64, Ptr{Cvoid}), tA, tB, m, n, k, α, a, ldd(A), b, ldd(B), β, c, ldd(C), s))
    elseif method === :bf16x3
        check(:jb_sgemm_bf16x3_dev, ccall((:jb_sgemm_bf16x3_dev, libjblas), Cint,
            (Cchar, Cchar, Int64, Int64, Int64, Float32, Ptr{Float32}, Int64, Ptr{Float32}, Int64, Float32,
             Ptr{Float32}, Int64, Cint, Ptr{Cvoid}), tA, tB, m, n, k, α, a, ldd(A), b, ldd(B), β, c, ldd(C),
            nprod, s))
    else
        throw(ArgumentError("Float32 method must be :native or :bf16x3"))
    end
    C
end

end # module
