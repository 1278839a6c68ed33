// bh_pad.hip -- the padded copy that precedes LU and QR.
//
// BandedMatrix{T}(A, (l, l+u)) (src/banded/BandedMatrix.jl:222-232), used by
// _lu (src/banded/BandedLU.jl:108-111) and _banded_qr
// (src/banded/bandedqr.jl:2), copies the in-band entries of A into an array
// with l extra super-diagonals on top: exactly LAPACK's AB with
// LDAB >= 2l+u+1.  Here it is one coalesced pass (thread per stored entry of
// the destination), optionally converting FP64 -> FP32 for the mixed-precision
// factorisation.  Entries of the destination that lie outside A's band (the
// fill-in rows and the corners) are written as zero.
#include "bh_common.hpp"

namespace bh {

template <typename Ts, typename Td>
__global__ void gbpad_kernel(i64 m, i64 n, i64 kl, i64 ku, const Ts *__restrict__ A, i64 lda, Td *__restrict__ AB,
                             i64 ldab, i64 hrows) {
    const i64 e = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= hrows * n) return;
    const i64 j = e / hrows, r = e - j * hrows;
    // destination row r holds A(i,j) with i = j - (kl+ku) + r
    const i64 i = j - (kl + ku) + r;
    Td v = Td(0);
    if (i >= 0 && i < m && j - i <= ku && i - j <= kl) v = (Td)A[j * lda + (ku + i - j)];
    AB[j * ldab + r] = v;
}

template <typename Ts, typename Td>
static int gbpad_impl(i64 m, i64 n, i64 kl, i64 ku, const Ts *A, i64 lda, Td *AB, i64 ldab) {
    BH_ARG(m >= 0, 1);
    BH_ARG(n >= 0, 2);
    BH_ARG(kl >= 0, 3);
    BH_ARG(ku >= 0, 4);
    BH_ARG(lda >= kl + ku + 1, 6);
    BH_ARG(ldab >= 2 * kl + ku + 1, 8);
    const i64 hrows = ldab;
    if (n == 0) return BH_OK;
    hipLaunchKernelGGL((gbpad_kernel<Ts, Td>), dim3((unsigned)cdiv(hrows * n, 256)), dim3(256), 0, stream(), m, n, kl,
                       ku, A, lda, AB, ldab, hrows);
    BH_LAUNCH_CHECK();
    return BH_OK;
}

}  // namespace bh

using namespace bh;

extern "C" {

int bh_dgbpad(bh_int m, bh_int n, bh_int kl, bh_int ku, const double *A, bh_int lda, double *AB, bh_int ldab) {
    return gbpad_impl<double, double>(m, n, kl, ku, A, lda, AB, ldab);
}
int bh_sgbpad(bh_int m, bh_int n, bh_int kl, bh_int ku, const float *A, bh_int lda, float *AB, bh_int ldab) {
    return gbpad_impl<float, float>(m, n, kl, ku, A, lda, AB, ldab);
}
int bh_dsgbpad(bh_int m, bh_int n, bh_int kl, bh_int ku, const double *A, bh_int lda, float *AB, bh_int ldab) {
    return gbpad_impl<double, float>(m, n, kl, ku, A, lda, AB, ldab);
}

}  // extern "C"
