# BandedMatricesHIPChainRulesExt -- reverse-mode rules for the device band
# solve and product, loaded (Project.toml [weakdeps]/[extensions], as the
# reference wires its extensions, Project.toml:11-17) when BandedMatrices,
# AMDGPU and ChainRulesCore are present.  The reference has no rrule; the
# adjoint solve reuses its transposed LU solve (src/banded/linalg.jl:41-47).
#
#   x = A \ b :  b̄ = A⁻ᵀ x̄ = λ (gbtrs 'T' on the same factors and pivots),
#                Ā = P_band(-λ xᵀ)
#   y = A * x :  x̄ = Aᵀ ȳ (gbmv 'T'),  Ā = P_band(ȳ xᵀ)
# P_band keeps the entries of A's band only: the cotangent of a banded matrix
# is banded with the same bandwidths (one fma per band entry on the device).
module BandedMatricesHIPChainRulesExt

using BandedMatrices, LinearAlgebra, AMDGPU, ChainRulesCore
import BandedMatrices: bandeddata, bandwidths, _BandedMatrix

const libbh = joinpath(@__DIR__, "..", "libbandedhip.so")
const HIPBanded{T} = BandedMatrix{T,<:ROCMatrix{T}}
dptr(a::ROCArray{T}) where {T} = reinterpret(Ptr{T}, pointer(a))

for (T, p) in ((Float64, :d), (Float32, :s))
    outer = Symbol(:bh_, p, :gb_band_outer)
    @eval function _band_outer(A::HIPBanded{$T}, α::$T, u::ROCVector{$T}, v::ROCVector{$T})
        l, k = bandwidths(A); m, n = size(A)
        G = ROCMatrix{$T}(undef, l + k + 1, n)
        r = ccall(($(QuoteNode(outer)), libbh), Cint,
                  (Int64, Int64, Int64, Int64, $T, Ptr{$T}, Ptr{$T}, $T, Ptr{$T}, Ptr{$T}, $T, Ptr{$T}, Int64),
                  m, n, l, k, α, dptr(u), dptr(v), zero($T), C_NULL, C_NULL, zero($T), dptr(G), stride(G, 2))
        r == 0 || error("$($(QuoteNode(outer))): status $r")
        _BandedMatrix(G, m, l, k)
    end
end

function ChainRulesCore.rrule(::typeof(\), A::HIPBanded{T}, b::ROCVector{T}) where {T<:Union{Float32,Float64}}
    F = lu(A)
    x = F \ b
    function ldiv_pullback(x̄)
        λ = transpose(F) \ ROCVector{T}(unthunk(x̄))      # the adjoint solve, same factors
        return NoTangent(), _band_outer(A, -one(T), λ, x), λ
    end
    return x, ldiv_pullback
end

function ChainRulesCore.rrule(::typeof(*), A::HIPBanded{T}, x::ROCVector{T}) where {T<:Union{Float32,Float64}}
    y = A * x
    function mul_pullback(ȳ)
        ȳd = ROCVector{T}(unthunk(ȳ))
        x̄ = similar(x)
        mul!(x̄, transpose(A), ȳd, one(T), zero(T))
        return NoTangent(), _band_outer(A, one(T), ȳd, x), x̄
    end
    return y, mul_pullback
end

end # module
