module JBlasAMDChainRulesCoreExt

using ChainRulesCore
using JBlasAMD: JBlasAMD, gemm, pullback!

"""
rrule of `gemm(tA, tB, α, A, B)` = α op(A) op(B):
  Ā = α C̄ op(B)ᵀ (tA = 'N') or α op(B) C̄ᵀ (tA = 'T')
  B̄ = α op(A)ᵀ C̄ (tB = 'N') or α C̄ᵀ op(A) (tB = 'T')
  ᾱ = ⟨C̄, op(A) op(B)⟩
Ā and B̄ come from one grouped launch of the library's adjoint kernels.
"""
function ChainRulesCore.rrule(::typeof(gemm), tA::AbstractChar, tB::AbstractChar, α::Real,
                              A::StridedMatrix{T}, B::StridedMatrix{T}; kw...) where {T<:Union{Float32,Float64}}
    P = gemm(tA, tB, one(T), A, B; kw...)
    Y = T(α) .* P
    function gemm_pullback(Ȳ)
        C̄ = Matrix{T}(unthunk(Ȳ))
        Ā = similar(A)
        B̄ = similar(B)
        pullback!(tA, tB, α, A, B, C̄, Ā, B̄; kw...)
        ᾱ = sum(C̄ .* P)
        return (NoTangent(), NoTangent(), NoTangent(), ᾱ, Ā, B̄)
    end
    return Y, gemm_pullback
end

end # module
