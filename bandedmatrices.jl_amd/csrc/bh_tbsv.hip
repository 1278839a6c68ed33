// bh_tbsv.hip -- triangular band solve and product, op(A) x = b and
// x <- op(A) x (BLAS xTBSV / xTBMV), on gfx950.
//
// Replaces tbsv! and tbmv! (src/blas.jl:71-147), which triangular \ and * on
// banded data reach (src/tribanded.jl:47-54,75-96).  Storage as BLAS: upper
// A(i,j) at a[k + i - j + j*lda] (i = j-k..j, diagonal in row k), lower A(i,j)
// at a[i - j + j*lda] (i = j..j+k, diagonal in row 0).
//
// Solve:
//   'N' upper: j descending, x_j = b_j / a_jj, then b_i = fma(a_ij, -x_j, b_i), i = j-k..j-1;
//   'N' lower: j ascending,  the same with i = j+1..j+k;
//   'T' upper: j ascending,  x_j = (b_j - sum_i a_ij x_i) / a_jj over i = j-k..j-1;
//   'T' lower: j descending, the same over i = j+1..j+k.
// With diag = 'U' the division is skipped.  The 'N' sweeps keep the rounding
// of the column-ordered CPU routine (division, then one fma per entry), so
// they are bit-identical to it; the 'T' dots are summed across the wave with
// DPP and agree to rounding.
//
// Product: y_i = d_i + sum_{t=1..} a_(t) x_(t) with d_i = a_ii x_i (x_i for a
// unit diagonal) and the off-diagonal terms fused in order of their distance
// t from the diagonal.
#include "bh_tbsv.hpp"

namespace bh {
namespace {

constexpr i64 kTbArgMaxBytes = (1ll << 31) - (1ll << 24);

struct TbFlags {
    bool upper, trans, unit;
};

// BLAS argument order (uplo, trans, diag, n, k, A, lda, x, incx): a bad i-th
// argument returns -i.
template <typename T>
int tb_check(char uplo, char trans, char diag, i64 n, i64 k, i64 lda, i64 incx, TbFlags &f) {
    const bool up = (uplo == 'U' || uplo == 'u'), lo = (uplo == 'L' || uplo == 'l');
    const bool tn = (trans == 'N' || trans == 'n'), tt = (trans == 'T' || trans == 't' || trans == 'C' || trans == 'c');
    const bool unit = (diag == 'U' || diag == 'u'), nonunit = (diag == 'N' || diag == 'n');
    BH_ARG(up || lo, 1);
    BH_ARG(tn || tt, 2);
    BH_ARG(unit || nonunit, 3);
    BH_ARG(n >= 0, 4);
    BH_ARG(k >= 0, 5);
    BH_ARG(lda >= k + 1, 7);
    BH_ARG(incx >= 1, 9);  // positive strides only
    BH_ARG(n < (1ll << 30) && k < (1ll << 30) && lda * n * (i64)sizeof(T) < kTbArgMaxBytes, 4);
    BH_ARG(n == 0 || ((n - 1) * incx + 1) * (i64)sizeof(T) < kTbArgMaxBytes, 9);
    f = TbFlags{up, tt, unit};
    return BH_OK;
}

int tb_mode(const TbFlags &f) { return f.trans ? (f.upper ? 2 : 3) : (f.upper ? 0 : 1); }

template <typename T>
int tbsv_impl(char uplo, char trans, char diag, i64 n, i64 k, const T *a, i64 lda, T *x, i64 incx) {
    TbFlags f;
    const int c = tb_check<T>(uplo, trans, diag, n, k, lda, incx, f);
    if (c != BH_OK) return c;
    if (n == 0) return BH_OK;
    return tbsv_run(tb_mode(f), (int)n, (int)k, a, lda, x, (int)incx, 0, 1, f.unit);
}

// Columns of X (n x nrhs, leading dimension ldx) solved independently, one
// wave each, in one launch.
template <typename T>
int tbsv_multi_impl(char uplo, char trans, char diag, i64 n, i64 k, const T *a, i64 lda, i64 nrhs, T *X, i64 ldx) {
    TbFlags f;
    const int c = tb_check<T>(uplo, trans, diag, n, k, lda, 1, f);
    if (c != BH_OK) return c;
    BH_ARG(nrhs >= 0 && nrhs < (1ll << 31), 8);
    BH_ARG(ldx >= (n > 1 ? n : 1), 10);
    if (n == 0 || nrhs == 0) return BH_OK;
    return tbsv_run(tb_mode(f), (int)n, (int)k, a, lda, X, 1, ldx, (int)nrhs, f.unit);
}

template <typename T>
int tbmv_impl(char uplo, char trans, char diag, i64 n, i64 k, const T *a, i64 lda, T *x, i64 incx) {
    TbFlags f;
    const int c = tb_check<T>(uplo, trans, diag, n, k, lda, incx, f);
    if (c != BH_OK) return c;
    if (n == 0) return BH_OK;
    return tbmv_run(tb_mode(f), (int)n, (int)k, a, lda, x, (int)incx, f.unit);
}

}  // namespace
}  // namespace bh

using namespace bh;

extern "C" {

int bh_dtbsv(char uplo, char trans, char diag, bh_int n, bh_int k, const double *A, bh_int lda, double *x,
             bh_int incx) {
    return tbsv_impl<double>(uplo, trans, diag, n, k, A, lda, x, incx);
}
int bh_stbsv(char uplo, char trans, char diag, bh_int n, bh_int k, const float *A, bh_int lda, float *x,
             bh_int incx) {
    return tbsv_impl<float>(uplo, trans, diag, n, k, A, lda, x, incx);
}
int bh_dtbsv_multi(char uplo, char trans, char diag, bh_int n, bh_int k, const double *A, bh_int lda, bh_int nrhs,
                   double *X, bh_int ldx) {
    return tbsv_multi_impl<double>(uplo, trans, diag, n, k, A, lda, nrhs, X, ldx);
}
int bh_stbsv_multi(char uplo, char trans, char diag, bh_int n, bh_int k, const float *A, bh_int lda, bh_int nrhs,
                   float *X, bh_int ldx) {
    return tbsv_multi_impl<float>(uplo, trans, diag, n, k, A, lda, nrhs, X, ldx);
}
int bh_dtbmv(char uplo, char trans, char diag, bh_int n, bh_int k, const double *A, bh_int lda, double *x,
             bh_int incx) {
    return tbmv_impl<double>(uplo, trans, diag, n, k, A, lda, x, incx);
}
int bh_stbmv(char uplo, char trans, char diag, bh_int n, bh_int k, const float *A, bh_int lda, float *x,
             bh_int incx) {
    return tbmv_impl<float>(uplo, trans, diag, n, k, A, lda, x, incx);
}

}  // extern "C"
