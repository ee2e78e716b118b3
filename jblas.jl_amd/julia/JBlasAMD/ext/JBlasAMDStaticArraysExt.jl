module JBlasAMDStaticArraysExt

# jBLAS's own operand type: jmul!(D::MMatrix{M,P,T}, A::MMatrix{M,N,T}, X::MMatrix{N,P,T}, ::Val...)
# (src/gemm.jl:244-246) and fastmul!(D::MMatrix, A::MMatrix, X::MMatrix) (src/kernels.jl:202-208).
# An MMatrix is dense column-major storage with leading dimension M (src/gemm.jl:3-11); its
# pointer is taken under GC.@preserve, as the reference takes pointer(D) (src/gemm.jl:309-311).

using JBlasAMD: JBlasAMD, jmul_ptr!, fastmul_ptr!
using StaticArrays: MMatrix

const F = Union{Float32,Float64}

"""
    jmul!(D::MMatrix{M,P,T}, A::MMatrix{M,N,T}, X::MMatrix{N,P,T}, ::Val...) -> D

The reference's signature; shapes are checked by dispatch as in the reference
(a mismatch is a `MethodError`).  The `Val` prefetch knobs are accepted and
ignored.  T = Float64 runs on the FP64 MFMA kernel, T = Float32 on the
sequential-fma kernel; both give the reference recurrence's bits.
"""
function JBlasAMD.jmul!(D::MMatrix{M,P,T}, A::MMatrix{M,N,T}, X::MMatrix{N,P,T},
                        ::Val = Val(7), ::Val = Val(7), ::Val = Val(3), ::Val = Val(3),
                        ::Val = Val(3)) where {M,N,P,T<:F}
    GC.@preserve D A X jmul_ptr!(pointer(D), pointer(A), pointer(X), M, N, P)
    D
end

"fastmul!(D::MMatrix, A::MMatrix, X::MMatrix): one small-GEMM launch; returns `nothing` like the reference."
function JBlasAMD.fastmul!(D::MMatrix{M,P,T}, A::MMatrix{M,N,T}, X::MMatrix{N,P,T}) where {M,N,P,T<:F}
    GC.@preserve D A X fastmul_ptr!(pointer(D), pointer(A), pointer(X), M, N, P)
    nothing
end

end # module
