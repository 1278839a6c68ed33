// bh_gbelem.hip -- element-wise operations on band storage (gfx950):
//   scale  A(i,j) <- alpha*A(i,j) over a row/column range      lmul!/rmul! on bandeddata
//          (src/generic/broadcast.jl:942-958) and the partial _fill_lmul! of the
//          band-pruned gbmm! (src/banded/gbmm.jl:231-279);
//   axpy   Y <- a*X + Y over the overlapping bands, BandError if X has a nonzero
//          outside Y's bands (banded_generic_axpy!, src/generic/broadcast.jl:978-1020);
//   copy   B <- A with other bandwidths, zero where A has no band
//          (copyto! / BandedMatrix(A, (l,u)), src/banded/BandedMatrix.jl:124-165).
// Band storage: A(i,j) at a[(u + i - j) + j*lda].  Only slots inside the m x n
// matrix are read or written; corner slots outside it are left alone.
// One thread per destination slot in storage order (slot row fastest), so a
// wave reads and writes contiguous memory: HBM-bound.
#include "bh_common.hpp"

namespace bh {
namespace {

constexpr int kElemBlock = 256;

// slot e of an h x n band array -> (t, j); exact for e < 2^53 via a double
// reciprocal and one correction step.
__device__ __forceinline__ void slot_of(i64 e, int h, double invh, i64 &t, i64 &j) {
    j = (i64)((double)e * invh);
    if (j * h > e) --j;
    else if ((j + 1) * h <= e) ++j;
    t = e - j * h;
}

// OP 0: scale (strong: alpha == 0 writes 0), 1: axpy, 2: copy.
template <typename T, int OP>
__global__ void __launch_bounds__(kElemBlock)
gb_elem_kernel(i64 m, i64 n, T alpha, const T *__restrict__ X, i64 ldx, i64 xl, i64 xu, T *__restrict__ Y, i64 ldy,
               i64 yl, i64 yu, i64 r0, i64 r1, i64 c0, i64 c1, int strong) {
    const int h = (int)(yl + yu + 1);
    const double invh = 1.0 / (double)h;
    const i64 total = (i64)h * (c1 - c0);
    for (i64 e = (i64)blockIdx.x * kElemBlock + threadIdx.x; e < total; e += (i64)gridDim.x * kElemBlock) {
        i64 t, jj;
        slot_of(e, h, invh, t, jj);
        const i64 j = c0 + jj;
        const i64 i = j - yu + t;  // row of this slot
        if (i < 0 || i >= m || i < r0 || i >= r1) continue;
        T *py = Y + t + j * ldy;
        if constexpr (OP == 0) {
            *py = (strong && alpha == T(0)) ? T(0) : alpha * *py;
        } else {
            const i64 d = i - j;  // inside X's band when -xu <= d <= xl
            const bool inx = (d >= -xu && d <= xl);
            if constexpr (OP == 1) {
                if (inx) {
                    const T p = alpha * X[(xu + d) + j * ldx];
                    *py = p + *py;  // a*x + y, two roundings
                }
            } else {
                *py = inx ? X[(xu + d) + j * ldx] : T(0);
            }
        }
    }
}

// Any nonzero of X (bands xl, xu) outside the bands (yl, yu) sets *flag.
template <typename T>
__global__ void __launch_bounds__(kElemBlock)
gb_outside_kernel(i64 m, i64 n, const T *__restrict__ X, i64 ldx, i64 xl, i64 xu, i64 yl, i64 yu, int *flag) {
    const int h = (int)(xl + xu + 1);
    const double invh = 1.0 / (double)h;
    const i64 total = (i64)h * n;
    const __amdgpu_buffer_rsrc_t rf = __builtin_amdgcn_make_buffer_rsrc(flag, 0, 4, 0x00020000);
    for (i64 e = (i64)blockIdx.x * kElemBlock + threadIdx.x; e < total; e += (i64)gridDim.x * kElemBlock) {
        i64 t, j;
        slot_of(e, h, invh, t, j);
        const i64 i = j - xu + t;
        const i64 d = i - j;
        if (i < 0 || i >= m || (d >= -yu && d <= yl)) continue;
        if (X[t + j * ldx] != T(0)) __builtin_amdgcn_raw_buffer_store_b32(1, rf, 0, 0, 0);  // vector store
    }
}

int elem_grid(i64 total) {
    const i64 b = cdiv(total, kElemBlock);
    return (int)(b < 8192 ? (b > 0 ? b : 1) : 8192);  // grid-stride beyond 2M threads
}

template <typename T, int OP>
int elem_launch(i64 m, i64 n, T alpha, const T *X, i64 ldx, i64 xl, i64 xu, T *Y, i64 ldy, i64 yl, i64 yu, i64 r0,
                i64 r1, i64 c0, i64 c1, int strong) {
    const i64 h = yl + yu + 1;
    if (h <= 0 || c1 <= c0 || m <= 0) return BH_OK;
    hipLaunchKernelGGL((gb_elem_kernel<T, OP>), dim3(elem_grid(h * (c1 - c0))), dim3(kElemBlock), 0, stream(), m, n,
                       alpha, X, ldx, xl, xu, Y, ldy, yl, yu, r0, r1, c0, c1, strong);
    BH_LAUNCH_CHECK();
    return BH_OK;
}

// BandError check: returns BH_ERR_BANDERROR when X has a nonzero outside (yl, yu).
template <typename T>
int outside_check(i64 m, i64 n, const T *X, i64 ldx, i64 xl, i64 xu, i64 yl, i64 yu) {
    if (!(xl > yl || xu > yu) || xl + xu + 1 <= 0 || n == 0 || m == 0) return BH_OK;
    hipStream_t st = stream();
    int *flag = nullptr;
    {
        i64 *cw = nullptr;
        const int cs = ctrl_words(&cw);
        if (cs) return cs;
        flag = reinterpret_cast<int *>(cw + 4);
    }
    int hflag = 0;
    hipError_t e = hipMemsetAsync(flag, 0, sizeof(int), st);
    if (e == hipSuccess) {
        hipLaunchKernelGGL((gb_outside_kernel<T>), dim3(elem_grid((xl + xu + 1) * n)), dim3(kElemBlock), 0, st, m, n,
                           X, ldx, xl, xu, yl, yu, flag);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(&hflag, flag, sizeof(int), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return hip_fail(e, "band check", __FILE__, __LINE__);
    if (hflag) {
        set_error("BandError: X has a nonzero outside the bands (%lld, %lld) of the destination", (long long)yl,
                  (long long)yu);
        return BH_ERR_BANDERROR;
    }
    return BH_OK;
}

template <typename T>
int gbscal_impl(i64 m, i64 n, i64 kl, i64 ku, T alpha, T *A, i64 lda, i64 r0, i64 r1, i64 c0, i64 c1, int strong) {
    BH_ARG(m >= 0, 1);
    BH_ARG(n >= 0, 2);
    BH_ARG(kl + ku + 1 >= 0, 3);
    BH_ARG(lda >= (kl + ku + 1 > 1 ? kl + ku + 1 : 1), 7);
    BH_ARG(r0 >= 0 && r0 <= r1 && r1 <= m, 8);
    BH_ARG(c0 >= 0 && c0 <= c1 && c1 <= n, 10);
    return elem_launch<T, 0>(m, n, alpha, (const T *)nullptr, 1, 0, 0, A, lda, kl, ku, r0, r1, c0, c1, strong);
}

template <typename T>
int gbaxpy_impl(i64 m, i64 n, T alpha, const T *X, i64 ldx, i64 xl, i64 xu, T *Y, i64 ldy, i64 yl, i64 yu) {
    BH_ARG(m >= 0, 1);
    BH_ARG(n >= 0, 2);
    BH_ARG(ldx >= (xl + xu + 1 > 1 ? xl + xu + 1 : 1), 5);
    BH_ARG(ldy >= (yl + yu + 1 > 1 ? yl + yu + 1 : 1), 9);
    if (xl + xu + 1 <= 0) return BH_OK;  // no bands in X
    const int c = outside_check<T>(m, n, X, ldx, xl, xu, yl, yu);
    if (c != BH_OK) return c;
    return elem_launch<T, 1>(m, n, alpha, X, ldx, xl, xu, Y, ldy, yl, yu, 0, m, 0, n, 0);
}

template <typename T>
int gbcopy_impl(i64 m, i64 n, const T *A, i64 lda, i64 al, i64 au, T *B, i64 ldb, i64 bl, i64 bu) {
    BH_ARG(m >= 0, 1);
    BH_ARG(n >= 0, 2);
    BH_ARG(lda >= (al + au + 1 > 1 ? al + au + 1 : 1), 4);
    BH_ARG(ldb >= (bl + bu + 1 > 1 ? bl + bu + 1 : 1), 8);
    const int c = outside_check<T>(m, n, A, lda, al, au, bl, bu);
    if (c != BH_OK) return c;
    // with no bands in A every slot of B is zero: read X nowhere
    const i64 xl = al + au + 1 > 0 ? al : -1, xu = al + au + 1 > 0 ? au : -1;
    return elem_launch<T, 2>(m, n, T(1), A, lda, xl, xu, B, ldb, bl, bu, 0, m, 0, n, 0);
}

}  // namespace
}  // namespace bh

using namespace bh;

extern "C" {

int bh_dgbscal(bh_int m, bh_int n, bh_int kl, bh_int ku, double alpha, double *A, bh_int lda, bh_int r0, bh_int r1,
               bh_int c0, bh_int c1, int strong_zero) {
    return gbscal_impl<double>(m, n, kl, ku, alpha, A, lda, r0, r1, c0, c1, strong_zero);
}
int bh_sgbscal(bh_int m, bh_int n, bh_int kl, bh_int ku, float alpha, float *A, bh_int lda, bh_int r0, bh_int r1,
               bh_int c0, bh_int c1, int strong_zero) {
    return gbscal_impl<float>(m, n, kl, ku, alpha, A, lda, r0, r1, c0, c1, strong_zero);
}
int bh_dgbaxpy(bh_int m, bh_int n, double alpha, const double *X, bh_int ldx, bh_int xl, bh_int xu, double *Y,
               bh_int ldy, bh_int yl, bh_int yu) {
    return gbaxpy_impl<double>(m, n, alpha, X, ldx, xl, xu, Y, ldy, yl, yu);
}
int bh_sgbaxpy(bh_int m, bh_int n, float alpha, const float *X, bh_int ldx, bh_int xl, bh_int xu, float *Y,
               bh_int ldy, bh_int yl, bh_int yu) {
    return gbaxpy_impl<float>(m, n, alpha, X, ldx, xl, xu, Y, ldy, yl, yu);
}
int bh_dgbcopy(bh_int m, bh_int n, const double *A, bh_int lda, bh_int al, bh_int au, double *B, bh_int ldb,
               bh_int bl, bh_int bu) {
    return gbcopy_impl<double>(m, n, A, lda, al, au, B, ldb, bl, bu);
}
int bh_sgbcopy(bh_int m, bh_int n, const float *A, bh_int lda, bh_int al, bh_int au, float *B, bh_int ldb,
               bh_int bl, bh_int bu) {
    return gbcopy_impl<float>(m, n, A, lda, al, au, B, ldb, bl, bu);
}

}  // extern "C"
