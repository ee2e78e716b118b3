module JBlasAMDjBLASExt

# Loaded when jBLAS (and StaticArrays) are loaded beside JBlasAMD: adds methods to
# jBLAS's own functions for the element types this library computes, so an
# unchanged `jmul!(D, A, X)` on Float64 MMatrix operands (src/gemm.jl:244-246)
# runs on the GPU.  These methods are more specific than jBLAS's generic
# `where {T,...}` ones (concrete element type), so dispatch picks them.

using JBlasAMD: JBlasAMD, jmul_ptr!, fastmul_ptr!
using StaticArrays: MMatrix
import jBLAS

const F = Union{Float32,Float64}

function jBLAS.jmul!(D::MMatrix{M,P,T}, A::MMatrix{M,N,T}, X::MMatrix{N,P,T},
                     ::Val = Val(7), ::Val = Val(7), ::Val = Val(3), ::Val = Val(3),
                     ::Val = Val(3)) where {M,N,P,T<:F}
    GC.@preserve D A X jmul_ptr!(pointer(D), pointer(A), pointer(X), M, N, P)
    D
end

function jBLAS.fastmul!(D::MMatrix{M,P,T}, A::MMatrix{M,N,T}, X::MMatrix{N,P,T}) where {M,N,P,T<:F}
    GC.@preserve D A X fastmul_ptr!(pointer(D), pointer(A), pointer(X), M, N, P)
    nothing
end

# kernel!/initkernel!(pD, pA, pX, ::jBLAS.Kernel{Mₖ,Pₖ,stride_AD,stride_X,N}[, pf])
# (src/kernels.jl:239-241, 273-275; descriptor src/kernel_structure.jl:8-9)
jBLAS.kernel!(pD::Ptr{Float64}, pA::Ptr{Float64}, pX::Ptr{Float64},
              ::jBLAS.Kernel{Mₖ,Pₖ,stride_AD,stride_X,N}) where {Mₖ,Pₖ,stride_AD,stride_X,N} =
    JBlasAMD.kernel!(pD, pA, pX, JBlasAMD.Kernel{Mₖ,Pₖ,stride_AD,stride_X,N}())
jBLAS.kernel!(pD::Ptr{Float64}, pA::Ptr{Float64}, pX::Ptr{Float64},
              ::jBLAS.Kernel{Mₖ,Pₖ,stride_AD,stride_X,N}, pf) where {Mₖ,Pₖ,stride_AD,stride_X,N} =
    JBlasAMD.kernel!(pD, pA, pX, JBlasAMD.Kernel{Mₖ,Pₖ,stride_AD,stride_X,N}(), pf)
jBLAS.initkernel!(pD::Ptr{Float64}, pA::Ptr{Float64}, pX::Ptr{Float64},
                  ::jBLAS.Kernel{Mₖ,Pₖ,stride_AD,stride_X,N}) where {Mₖ,Pₖ,stride_AD,stride_X,N} =
    JBlasAMD.initkernel!(pD, pA, pX, JBlasAMD.Kernel{Mₖ,Pₖ,stride_AD,stride_X,N}())
jBLAS.initkernel!(pD::Ptr{Float64}, pA::Ptr{Float64}, pX::Ptr{Float64},
                  ::jBLAS.Kernel{Mₖ,Pₖ,stride_AD,stride_X,N}, pf) where {Mₖ,Pₖ,stride_AD,stride_X,N} =
    JBlasAMD.initkernel!(pD, pA, pX, JBlasAMD.Kernel{Mₖ,Pₖ,stride_AD,stride_X,N}(), pf)

end # module
