# BandedMatricesHIPExt -- routes BandedMatrices.jl's banded hot path for
# device-resident band data (AMDGPU.ROCArray) to libbandedhip.so.
#
# Loaded by Julia's package-extension mechanism (Project.toml [weakdeps] /
# [extensions], as the reference already does for its other extensions,
# Project.toml:11-17) when both BandedMatrices and AMDGPU are present.
# Every method below replaces one reference method for the device layout:
#   mul!(y, A, x, α, β)   materialize!(::MulAdd{BandedColumns,...})  src/generic/matmul.jl:61-64
#   A' * x, transpose(A)*x materialize!(::MulAdd{BandedRows,...})      src/generic/matmul.jl:66-92
#   mul!(Y, A, X, α, β)    dense multi-RHS                            src/generic/matmul.jl:243-256
#   A * B  (banded)        gbmm!                                      src/banded/gbmm.jl:207-340
#   lu / lu!               LAPACK.gbtrf!                              src/banded/BandedLU.jl:90-111
#   ldiv!(F, B), F' \ B    LAPACK.gbtrs!                              src/banded/linalg.jl:24-47
#   A \ b                  materialize!(::Ldiv{BandedColumns})        src/banded/linalg.jl:5-9,75
#   qr / ldiv!(QR, B)      _qr/_qr!, ldiv!(::QR), tbsv!               src/banded/bandedqr.jl:1-5,175-206
#   Q*B, Q'*B, B*Q, B*Q'   banded_qr_lmul!, banded_qr_rmul!           src/banded/bandedqr.jl:38-148,150-214
#   Tri(A) \ x, Tri(A) * x tbsv! / tbmv! on the triangle's band rows  src/tribanded.jl:19-26,47-54,75-96
#   lmul!/rmul!, axpy!     _banded_lmul!/_rmul!, _banded_axpy!         src/generic/broadcast.jl:942-1020
#   copyto!(B, A) banded   copyto_bandeddata!                          src/banded/BandedMatrix.jl:124-130
#   Symmetric(A) * x       BLAS.sbmv!                                  src/symbanded/symbanded.jl:72-91
#   cholesky / ldiv!       pbtrf! / pbtrs!                             src/symbanded/BandedCholesky.jl:1-79
#   ldlt! / ldiv!          banded LDLᵀ and its solve                   src/symbanded/ldlt.jl:8-69
module BandedMatricesHIPExt

using BandedMatrices, LinearAlgebra, ArrayLayouts, AMDGPU
import BandedMatrices: BandedColumns, bandeddata, bandwidths, BandedLU, _BandedMatrix
import ArrayLayouts: MemoryLayout, MulAdd, Ldiv, Lmul, TriangularLayout, triangulardata, materialize!
import LinearAlgebra: BlasInt

const libbh = joinpath(@__DIR__, "..", "libbandedhip.so")

# ---------------------------------------------------------------- layout hook
# Device column-major layout; bandedcolumns(::HIPColumnMajor) makes a
# BandedMatrix over a ROCMatrix dispatch to the methods below instead of the
# scalar-indexing fallback (src/generic/matmul.jl:100-110).
struct HIPColumnMajor <: ArrayLayouts.AbstractColumnMajor end
MemoryLayout(::Type{<:ROCArray{T,N}}) where {T,N} = HIPColumnMajor()
BandedMatrices.bandedcolumns(::HIPColumnMajor) = BandedColumns{HIPColumnMajor}()
const HIPBanded{T} = BandedMatrix{T,<:ROCMatrix{T}}

# ---------------------------------------------------------------- errors
# Return codes (include/bandedhip.h): 0 ok; -k: argument k invalid
# (ArgumentError, as LAPACK.chkargsok); BH_ERR_* <= -1000 runtime failures,
# -1003 BandError.  A singular pivot column comes back in `info`, not here.
const BH_ERR_BANDERROR = -1003
function chk(r::Integer, name)
    r == 0 && return nothing
    r == BH_ERR_BANDERROR && throw(BandedMatrices.BandError())
    r < 0 && r > -1000 && throw(ArgumentError("$name: invalid argument $(-r)"))
    msg = unsafe_string(ccall((:bh_last_error, libbh), Cstring, ()))
    error("$name: $msg")
end

dptr(a::ROCArray{T}) where {T} = reinterpret(Ptr{T}, pointer(a))

# Device copies of the pivots, keyed by the host ipiv vector a BandedLU holds
# (BandedLU.jl:12 stores ipiv on the host); gbtrs reads the device copy.
const IPIV_DEV = WeakKeyDict{Vector{BlasInt},ROCVector{Int64}}()

function __init__()
    chk(ccall((:bh_init, libbh), Cint, (Cint,), AMDGPU.device_id(AMDGPU.device()) - 1), :bh_init)
end

# Library calls go on AMDGPU's current stream so they order with other device work.
with_stream(f) = (ccall((:bh_set_stream, libbh), Cint, (Ptr{Cvoid},), AMDGPU.stream().stream); f())

for (T, p) in ((Float64, :d), (Float32, :s))
    gbmv = Symbol(:bh_, p, :gbmv); gbmv_multi = Symbol(:bh_, p, :gbmv_multi); gbmm = Symbol(:bh_, p, :gbmm)
    gbtrf = Symbol(:bh_, p, :gbtrf); gbtrs = Symbol(:bh_, p, :gbtrs); gbpad = Symbol(:bh_, p, :gbpad)
    gbqrf = Symbol(:bh_, p, :gbqrf); gbqrs = Symbol(:bh_, p, :gbqrs)
    gbqrm = Symbol(:bh_, p, :gbqrm); gbqrmr = Symbol(:bh_, p, :gbqrmr)
    tbsv = Symbol(:bh_, p, :tbsv); tbmv = Symbol(:bh_, p, :tbmv)
    gbscal = Symbol(:bh_, p, :gbscal); gbaxpy = Symbol(:bh_, p, :gbaxpy); gbcopy = Symbol(:bh_, p, :gbcopy)
    pbtrf = Symbol(:bh_, p, :pbtrf); pbtrs = Symbol(:bh_, p, :pbtrs); sbmv = Symbol(:bh_, p, :sbmv)
    sbldlt = Symbol(:bh_, p, :sbldlt); sbldlts = Symbol(:bh_, p, :sbldlts)
    @eval begin
        # ldlt! in place on the stored triangle (src/symbanded/ldlt.jl:8-37)
        function _sbldlt!(uplo::Char, A::HIPBanded{$T})
            l, u = bandwidths(A); d = bandeddata(A)
            k, off = uplo == 'U' ? (u, 0) : (l, u)
            with_stream() do
                chk(ccall(($(QuoteNode(sbldlt)), libbh), Cint, (UInt8, Int64, Int64, Ptr{$T}, Int64),
                          uplo, size(A, 1), k, dptr(d) + off * sizeof($T), stride(d, 2)), $(QuoteNode(sbldlt)))
            end
            A
        end
        # ldiv!(::LDLt, B) (src/symbanded/ldlt.jl:47-69)
        function _sbldlts!(uplo::Char, A::HIPBanded{$T}, B::ROCVecOrMat{$T})
            l, u = bandwidths(A); d = bandeddata(A)
            k, off = uplo == 'U' ? (u, 0) : (l, u)
            ldb = B isa ROCVector ? length(B) : stride(B, 2)
            with_stream() do
                chk(ccall(($(QuoteNode(sbldlts)), libbh), Cint,
                          (UInt8, Int64, Int64, Int64, Ptr{$T}, Int64, Ptr{$T}, Int64),
                          uplo, size(A, 1), k, size(B, 2), dptr(d) + off * sizeof($T), stride(d, 2), dptr(B), ldb),
                    $(QuoteNode(sbldlts)))
            end
            B
        end
        # pbtrf!(uplo, n, kd, AB) on the triangle's band rows (src/lapack.jl:268-292)
        function _pbtrf!(uplo::Char, n::Int, kd::Int, AB::ROCMatrix{$T}, off::Int)
            info = Ref{Int64}(0)
            with_stream() do
                chk(ccall(($(QuoteNode(pbtrf)), libbh), Cint, (UInt8, Int64, Int64, Ptr{$T}, Int64, Ref{Int64}),
                          uplo, n, kd, dptr(AB) + off * sizeof($T), stride(AB, 2), info), $(QuoteNode(pbtrf)))
            end
            info[]
        end
        # pbtrs!(uplo, n, kd, AB, B) (src/lapack.jl:297-333)
        function _pbtrs!(uplo::Char, n::Int, kd::Int, AB::ROCMatrix{$T}, off::Int, B::ROCVecOrMat{$T})
            ldb = B isa ROCVector ? length(B) : stride(B, 2)
            with_stream() do
                chk(ccall(($(QuoteNode(pbtrs)), libbh), Cint,
                          (UInt8, Int64, Int64, Int64, Ptr{$T}, Int64, Ptr{$T}, Int64),
                          uplo, n, kd, size(B, 2), dptr(AB) + off * sizeof($T), stride(AB, 2), dptr(B), ldb),
                    $(QuoteNode(pbtrs)))
            end
            B
        end
        # y <- α Symmetric(A) x + β y (BLAS.sbmv! on symbandeddata, src/symbanded/symbanded.jl:72-91)
        function _sbmv!(uplo::Char, α::$T, A::HIPBanded{$T}, x::ROCVector{$T}, β::$T, y::ROCVector{$T})
            l, u = bandwidths(A); d = bandeddata(A); n = size(A, 1)
            k, off = uplo == 'U' ? (u, 0) : (l, u)
            with_stream() do
                chk(ccall(($(QuoteNode(sbmv)), libbh), Cint,
                          (UInt8, Int64, Int64, $T, Ptr{$T}, Int64, Ptr{$T}, Int64, $T, Ptr{$T}, Int64),
                          uplo, n, k, α, dptr(d) + off * sizeof($T), stride(d, 2), dptr(x), 1, β, dptr(y), 1),
                    $(QuoteNode(sbmv)))
            end
            y
        end
        # lmul!(α, A) on the band data (src/generic/broadcast.jl:942-958)
        function _gbscal!(α::$T, A::HIPBanded{$T})
            l, u = bandwidths(A); d = bandeddata(A); m, n = size(A)
            with_stream() do
                chk(ccall(($(QuoteNode(gbscal)), libbh), Cint,
                          (Int64, Int64, Int64, Int64, $T, Ptr{$T}, Int64, Int64, Int64, Int64, Int64, Cint),
                          m, n, l, u, α, dptr(d), stride(d, 2), 0, m, 0, n, 0), $(QuoteNode(gbscal)))
            end
            A
        end
        # Y <- a X + Y (banded_generic_axpy!, src/generic/broadcast.jl:978-1020)
        function _gbaxpy!(a::$T, X::HIPBanded{$T}, Y::HIPBanded{$T})
            size(X) == size(Y) || throw(DimensionMismatch("X has size $(size(X)) but Y has size $(size(Y))"))
            (xl, xu), (yl, yu) = bandwidths(X), bandwidths(Y); dx, dy = bandeddata(X), bandeddata(Y)
            with_stream() do
                chk(ccall(($(QuoteNode(gbaxpy)), libbh), Cint,
                          (Int64, Int64, $T, Ptr{$T}, Int64, Int64, Int64, Ptr{$T}, Int64, Int64, Int64),
                          size(X, 1), size(X, 2), a, dptr(dx), stride(dx, 2), xl, xu, dptr(dy), stride(dy, 2), yl, yu),
                    $(QuoteNode(gbaxpy)))
            end
            Y
        end
        # copyto!(B, A) between banded matrices (copyto_bandeddata!, src/banded/BandedMatrix.jl:124-130)
        function _gbcopy!(B::HIPBanded{$T}, A::HIPBanded{$T})
            size(A) == size(B) || throw(DimensionMismatch("A has size $(size(A)) but B has size $(size(B))"))
            (al, au), (bl, bu) = bandwidths(A), bandwidths(B); da, db = bandeddata(A), bandeddata(B)
            with_stream() do
                chk(ccall(($(QuoteNode(gbcopy)), libbh), Cint,
                          (Int64, Int64, Ptr{$T}, Int64, Int64, Int64, Ptr{$T}, Int64, Int64, Int64),
                          size(A, 1), size(A, 2), dptr(da), stride(da, 2), al, au, dptr(db), stride(db, 2), bl, bu),
                    $(QuoteNode(gbcopy)))
            end
            B
        end
        # tbsv!/tbmv!(uplo, trans, diag, m, k, bandeddata(Tri(A)), x) (src/blas.jl:71-147): the
        # triangle's band is rows 1:u+1 ('U') or u+1:l+u+1 ('L') of A's band array
        # (src/tribanded.jl:19-26), passed as a pointer into it with A's leading dimension.
        function _tb!(op::Symbol, uplo::Char, trans::Char, diag::Char, A::HIPBanded{$T}, x::ROCVector{$T})
            l, u = bandwidths(A); d = bandeddata(A); n = size(A, 1)
            size(A, 2) == n || throw(DimensionMismatch("matrix is not square: dimensions are $(size(A))"))
            length(x) == n || throw(DimensionMismatch("size of A is $n != length(x) = $(length(x))"))
            k, off = uplo == 'U' ? (u, 0) : (l, u)
            fn = op === :tbsv ? $(QuoteNode(tbsv)) : $(QuoteNode(tbmv))
            with_stream() do
                r = op === :tbsv ?
                    ccall(($(QuoteNode(tbsv)), libbh), Cint,
                          (UInt8, UInt8, UInt8, Int64, Int64, Ptr{$T}, Int64, Ptr{$T}, Int64),
                          uplo, trans, diag, n, k, dptr(d) + off * sizeof($T), stride(d, 2), dptr(x), 1) :
                    ccall(($(QuoteNode(tbmv)), libbh), Cint,
                          (UInt8, UInt8, UInt8, Int64, Int64, Ptr{$T}, Int64, Ptr{$T}, Int64),
                          uplo, trans, diag, n, k, dptr(d) + off * sizeof($T), stride(d, 2), dptr(x), 1)
                chk(r, fn)
            end
            x
        end
        # y <- α A x + β y   (BLAS.gbmv! in banded_gbmv!, src/generic/matmul.jl:21-23)
        function _gbmv!(trans::Char, α::$T, A::HIPBanded{$T}, x::ROCVector{$T}, β::$T, y::ROCVector{$T})
            l, u = bandwidths(A); d = bandeddata(A); m, n = size(A)
            with_stream() do
                chk(ccall(($(QuoteNode(gbmv)), libbh), Cint,
                          (UInt8, Int64, Int64, Int64, Int64, $T, Ptr{$T}, Int64, Ptr{$T}, Int64, $T, Ptr{$T}, Int64),
                          trans, m, n, l, u, α, dptr(d), stride(d, 2), dptr(x), 1, β, dptr(y), 1), $(QuoteNode(gbmv)))
            end
            y
        end
        function _gbmv_multi!(trans::Char, α::$T, A::HIPBanded{$T}, X::ROCMatrix{$T}, β::$T, Y::ROCMatrix{$T})
            l, u = bandwidths(A); d = bandeddata(A); m, n = size(A)
            with_stream() do
                chk(ccall(($(QuoteNode(gbmv_multi)), libbh), Cint,
                          (UInt8, Int64, Int64, Int64, Int64, Int64, $T, Ptr{$T}, Int64, Ptr{$T}, Int64, $T, Ptr{$T}, Int64),
                          trans, m, n, l, u, size(X, 2), α, dptr(d), stride(d, 2), dptr(X), stride(X, 2), β,
                          dptr(Y), stride(Y, 2)), $(QuoteNode(gbmv_multi)))
            end
            Y
        end
        # C <- α A B + β C   (gbmm!, src/banded/gbmm.jl:207-340)
        function _gbmm!(α::$T, A::HIPBanded{$T}, B::HIPBanded{$T}, β::$T, C::HIPBanded{$T})
            (Al, Au), (Bl, Bu), (Cl, Cu) = bandwidths(A), bandwidths(B), bandwidths(C)
            a, b, c = bandeddata(A), bandeddata(B), bandeddata(C)
            with_stream() do
                chk(ccall(($(QuoteNode(gbmm)), libbh), Cint,
                          (Int64, Int64, Int64, $T, Ptr{$T}, Int64, Int64, Int64, Ptr{$T}, Int64, Int64, Int64,
                           $T, Ptr{$T}, Int64, Int64, Int64),
                          size(A, 1), size(A, 2), size(B, 2), α, dptr(a), stride(a, 2), Al, Au, dptr(b), stride(b, 2),
                          Bl, Bu, β, dptr(c), stride(c, 2), Cl, Cu), $(QuoteNode(gbmm)))
            end
            C
        end
        # qr!: padded copy (BandedMatrix{T}(A,(l,l+u)), bandedqr.jl:2) then the band
        # Householder QR (bandedqr.jl:8-35); R and the reflectors stay in the padded
        # band, τ on the device.
        function _qr(A::HIPBanded{$T})
            l, u = bandwidths(A); m, n = size(A); d = bandeddata(A)
            R = ROCMatrix{$T}(undef, 2l + u + 1, n)
            τ = ROCVector{$T}(undef, max(min(m, n), 1))
            with_stream() do
                chk(ccall(($(QuoteNode(gbpad)), libbh), Cint, (Int64, Int64, Int64, Int64, Ptr{$T}, Int64, Ptr{$T}, Int64),
                          m, n, l, u, dptr(d), stride(d, 2), dptr(R), stride(R, 2)), $(QuoteNode(gbpad)))
                chk(ccall(($(QuoteNode(gbqrf)), libbh), Cint, (Int64, Int64, Int64, Int64, Ptr{$T}, Int64, Ptr{$T}),
                          m, n, l, l + u, dptr(R), stride(R, 2), dptr(τ)), $(QuoteNode(gbqrf)))
            end
            QR(_BandedMatrix(R, m, l, l + u), τ)
        end
        # ldiv!(QR, B): Q'B then the upper band solve of R (bandedqr.jl:175-206)
        function _gbqrs!(F::QR{$T,<:HIPBanded{$T}}, B::ROCVecOrMat{$T})
            R = F.factors; m, n = size(R); l, u = bandwidths(R); d = bandeddata(R)
            n > m && error("Not implemented")  # bandedqr.jl:175-177
            size(B, 1) == m || throw(DimensionMismatch("B has $(size(B, 1)) rows, but needs $m"))
            ldb = B isa ROCVector ? length(B) : stride(B, 2)
            with_stream() do
                chk(ccall(($(QuoteNode(gbqrs)), libbh), Cint,
                          (Int64, Int64, Int64, Int64, Ptr{$T}, Int64, Ptr{$T}, Int64, Ptr{$T}, Int64),
                          m, n, l, u, dptr(d), stride(d, 2), dptr(F.τ), size(B, 2), dptr(B), ldb), $(QuoteNode(gbqrs)))
            end
            B
        end
        # lmul!(Q, B) / lmul!(Q', B) (banded_qr_lmul!, bandedqr.jl:38-94)
        function _gbqrm!(trans::Char, Q::LinearAlgebra.QRPackedQ{$T,<:HIPBanded{$T}}, B::ROCVecOrMat{$T})
            R = Q.factors; m, n = size(R); l, u = bandwidths(R); d = bandeddata(R)
            size(B, 1) == m || throw(DimensionMismatch("B has $(size(B, 1)) rows, Q is $m x $m"))
            ldb = B isa ROCVector ? length(B) : stride(B, 2)
            with_stream() do
                chk(ccall(($(QuoteNode(gbqrm)), libbh), Cint,
                          (UInt8, Int64, Int64, Int64, Int64, Ptr{$T}, Int64, Ptr{$T}, Int64, Ptr{$T}, Int64),
                          trans, m, n, l, u, dptr(d), stride(d, 2), dptr(Q.τ), size(B, 2), dptr(B), ldb),
                    $(QuoteNode(gbqrm)))
            end
            B
        end
        # rmul!(A, Q) / rmul!(A, Q') (banded_qr_rmul!, bandedqr.jl:96-148)
        function _gbqrmr!(trans::Char, A::ROCMatrix{$T}, Q::LinearAlgebra.QRPackedQ{$T,<:HIPBanded{$T}})
            R = Q.factors; m, n = size(R); l, u = bandwidths(R); d = bandeddata(R)
            size(A, 2) == m ||
                throw(DimensionMismatch("matrix A has dimensions $(size(A)) but matrix Q has dimensions ($m, $n)"))
            with_stream() do
                chk(ccall(($(QuoteNode(gbqrmr)), libbh), Cint,
                          (UInt8, Int64, Int64, Int64, Int64, Int64, Ptr{$T}, Int64, Ptr{$T}, Ptr{$T}, Int64),
                          trans, size(A, 1), m, n, l, u, dptr(d), stride(d, 2), dptr(Q.τ), dptr(A), stride(A, 2)),
                    $(QuoteNode(gbqrmr)))
            end
            A
        end
        # lu!: padded copy (BandedMatrix{T}(A,(l,l+u)), BandedLU.jl:110) then gbtrf
        # (BandedLU.jl:98).  ipiv is returned on the host, as BandedLU stores it.
        function _lu(A::HIPBanded{$T})
            l, u = bandwidths(A); m, n = size(A); d = bandeddata(A)
            AB = ROCMatrix{$T}(undef, 2l + u + 1, n)
            ipiv_d = ROCVector{Int64}(undef, min(m, n)); info = Ref{Int64}(0)
            with_stream() do
                chk(ccall(($(QuoteNode(gbpad)), libbh), Cint, (Int64, Int64, Int64, Int64, Ptr{$T}, Int64, Ptr{$T}, Int64),
                          m, n, l, u, dptr(d), stride(d, 2), dptr(AB), stride(AB, 2)), $(QuoteNode(gbpad)))
                chk(ccall(($(QuoteNode(gbtrf)), libbh), Cint,
                          (Int64, Int64, Int64, Int64, Ptr{$T}, Int64, Ptr{Int64}, Ref{Int64}),
                          m, n, l, u, dptr(AB), stride(AB, 2), dptr(ipiv_d), info), $(QuoteNode(gbtrf)))
            end
            # LAPACK.gbtrf! raises on a zero pivot (chklapackerror), so lu does too
            info[] > 0 && throw(LinearAlgebra.LAPACKException(BlasInt(info[])))
            F = BandedLU{$T}(_BandedMatrix(AB, m, l, l + u), Vector{BlasInt}(Array(ipiv_d)), BlasInt(0))
            IPIV_DEV[F.ipiv] = ipiv_d      # device mirror for gbtrs
            F
        end
        # ldiv!(F, B) / transpose(F) \ B   (LAPACK.gbtrs!, linalg.jl:24-47)
        function _gbtrs!(trans::Char, F::BandedLU{$T,<:HIPBanded{$T}}, B::ROCVecOrMat{$T})
            AB = F.factors; n = size(AB, 2); l, uu = bandwidths(AB); u = uu - l; d = bandeddata(AB)
            ipiv_d = get!(() -> ROCVector{Int64}(F.ipiv), IPIV_DEV, F.ipiv)
            ldb = B isa ROCVector ? length(B) : stride(B, 2)
            with_stream() do
                chk(ccall(($(QuoteNode(gbtrs)), libbh), Cint,
                          (UInt8, Int64, Int64, Int64, Int64, Ptr{$T}, Int64, Ptr{Int64}, Ptr{$T}, Int64),
                          trans, n, l, u, size(B, 2), dptr(d), stride(d, 2), dptr(ipiv_d), dptr(B), ldb),
                    $(QuoteNode(gbtrs)))
            end
            B
        end
    end
end

# ---------------------------------------------------------------- dispatch
const HIPBC = BandedColumns{HIPColumnMajor}
materialize!(M::MulAdd{HIPBC,HIPColumnMajor,HIPColumnMajor,<:Any,<:HIPBanded,<:ROCVector,<:ROCVector}) =
    _gbmv!('N', M.α, M.A, M.B, M.β, M.C)
materialize!(M::MulAdd{HIPBC,HIPColumnMajor,HIPColumnMajor,<:Any,<:HIPBanded,<:ROCMatrix,<:ROCMatrix}) =
    _gbmv_multi!('N', M.α, M.A, M.B, M.β, M.C)
materialize!(M::MulAdd{HIPBC,HIPBC,HIPBC}) = _gbmm!(M.α, M.A, M.B, M.β, M.C)
# transpose(A) * x and A' * x (real): the row-major band layout of the wrapper
# (BlasMatMulVecAdd{<:BandedRowMajor}, src/generic/matmul.jl:66-92)
const HIPBR = BandedMatrices.BandedRows{HIPColumnMajor}
materialize!(M::MulAdd{HIPBR,HIPColumnMajor,HIPColumnMajor,<:Any,<:Any,<:ROCVector,<:ROCVector}) =
    _gbmv!('T', M.α, parent(M.A), M.B, M.β, M.C)
LinearAlgebra.mul!(y::ROCVector{T}, A::Adjoint{T,<:HIPBanded{T}}, x::ROCVector{T}, α::Number, β::Number) where {T<:Union{Float32,Float64}} =
    _gbmv!('T', T(α), parent(A), x, T(β), y)
LinearAlgebra.mul!(y::ROCVector{T}, A::Transpose{T,<:HIPBanded{T}}, x::ROCVector{T}, α::Number, β::Number) where {T<:Union{Float32,Float64}} =
    _gbmv!('T', T(α), parent(A), x, T(β), y)
ArrayLayouts._lu(::HIPBC, axes, A::HIPBanded; kwds...) = _lu(A)
LinearAlgebra.ldiv!(F::BandedLU{T,<:HIPBanded{T}}, B::ROCVecOrMat{T}) where {T} = _gbtrs!('N', F, B)
LinearAlgebra.ldiv!(F::Transpose{T,<:BandedLU{T,<:HIPBanded{T}}}, B::ROCVecOrMat{T}) where {T} = _gbtrs!('T', parent(F), B)
LinearAlgebra.ldiv!(F::Adjoint{T,<:BandedLU{T,<:HIPBanded{T}}}, B::ROCVecOrMat{T}) where {T<:Real} = _gbtrs!('T', parent(F), B)
function materialize!(M::Ldiv{HIPBC})
    LinearAlgebra.checksquare(M.A)  # src/banded/linalg.jl:5-9
    ldiv!(lu(M.A), M.B)
end
# qr (bandedqr.jl:1-5) and the QR solve / Q products (bandedqr.jl:150-214)
ArrayLayouts._qr(::HIPBC, axes, A::HIPBanded; kwds...) = _qr(A)
ArrayLayouts._qr!(::HIPBC, axes, A::HIPBanded; kwds...) = _qr(A)
LinearAlgebra.ldiv!(F::QR{T,<:HIPBanded{T}}, B::ROCVecOrMat{T}) where {T} = _gbqrs!(F, B)
LinearAlgebra.lmul!(Q::LinearAlgebra.QRPackedQ{T,<:HIPBanded{T}}, B::ROCVecOrMat{T}) where {T} = _gbqrm!('N', Q, B)
LinearAlgebra.lmul!(Q::LinearAlgebra.AdjointQ{T,<:LinearAlgebra.QRPackedQ{T,<:HIPBanded{T}}}, B::ROCVecOrMat{T}) where {T<:Real} =
    _gbqrm!('T', parent(Q), B)
LinearAlgebra.rmul!(A::ROCMatrix{T}, Q::LinearAlgebra.QRPackedQ{T,<:HIPBanded{T}}) where {T} = _gbqrmr!('N', A, Q)
LinearAlgebra.rmul!(A::ROCMatrix{T}, Q::LinearAlgebra.AdjointQ{T,<:LinearAlgebra.QRPackedQ{T,<:HIPBanded{T}}}) where {T<:Real} =
    _gbqrmr!('T', A, parent(Q))
BandedMatrices._banded_lmul!(α::Number, A::HIPBanded{T}, ::HIPBC) where {T} = _gbscal!(T(α), A)
BandedMatrices._banded_rmul!(A::HIPBanded{T}, α::Number, ::HIPBC) where {T} = _gbscal!(T(α), A)
BandedMatrices._banded_axpy!(a::Number, X::HIPBanded{T}, Y::HIPBanded{T}, ::HIPBC, ::HIPBC) where {T} =
    _gbaxpy!(T(a), X, Y)
BandedMatrices.copyto_bandeddata!(B::HIPBanded{T}, ::HIPBC, A::HIPBanded{T}, ::HIPBC) where {T} = _gbcopy!(B, A)
# Symmetric views: sbmv for the product; banded_chol! / _ldiv! for the factorisation
# (src/symbanded/BandedCholesky.jl:1-11,66-79), on the parent's band rows in place
materialize!(M::MulAdd{<:ArrayLayouts.SymmetricLayout{HIPBC},HIPColumnMajor,HIPColumnMajor,<:Any,<:Any,<:ROCVector,<:ROCVector}) =
    _sbmv!(ArrayLayouts.symmetricuplo(M.A), M.α, parent(M.A), M.B, M.β, M.C)
function BandedMatrices.banded_chol!(::HIPBC, A::HIPBanded, ::Type{UpperTriangular})
    info = _pbtrf!('U', size(A, 1), bandwidth(A, 2), bandeddata(A), 0)
    UpperTriangular(A), info
end
function BandedMatrices.banded_chol!(::HIPBC, A::HIPBanded, ::Type{LowerTriangular})
    info = _pbtrf!('L', size(A, 1), bandwidth(A, 1), bandeddata(A), bandwidth(A, 2))
    LowerTriangular(A), info
end
function LinearAlgebra.ldlt!(F::Symmetric{T,<:HIPBanded{T}}) where {T<:Union{Float32,Float64}}
    _sbldlt!(F.uplo, F.data)
    LDLt{T,typeof(F)}(F)
end
LinearAlgebra.ldiv!(S::LDLt{T,<:Symmetric{T,<:HIPBanded{T}}}, B::ROCVecOrMat{T}) where {T<:Union{Float32,Float64}} =
    _sbldlts!(S.data.uplo, S.data.data, B)
function LinearAlgebra.ldiv!(C::Cholesky{T,<:HIPBanded{T}}, B::ROCVecOrMat{T}) where {T<:Union{Float32,Float64}}
    F = C.factors; l, u = bandwidths(F)
    C.uplo == 'U' ? _pbtrs!('U', size(F, 1), u, bandeddata(F), 0, B) : _pbtrs!('L', size(F, 1), l, bandeddata(F), u, B)
end
# Triangular views: the layouts triangularlayout(Tri, ::BandedColumns) gives
# (src/tribanded.jl:9-11), with UPLO in ('U','L') and UNIT in ('N','U').
materialize!(M::Ldiv{<:TriangularLayout{UPLO,UNIT,HIPBC},HIPColumnMajor,<:Any,<:ROCVector}) where {UPLO,UNIT} =
    _tb!(:tbsv, UPLO, 'N', UNIT, triangulardata(M.A), M.B)
materialize!(M::Lmul{<:TriangularLayout{UPLO,UNIT,HIPBC},HIPColumnMajor,<:Any,<:ROCVector}) where {UPLO,UNIT} =
    _tb!(:tbmv, UPLO, 'N', UNIT, triangulardata(M.A), M.B)

end # module
