// bh_gbtrs.hip -- solves with the band LU factors on gfx950.
//
// Restates LAPACK xGBTRS as called by ldiv!(::BandedLU, B)
// (src/banded/linalg.jl:24-30, 'N') and its transposed form (:41-47, 'T'):
//   'N': for j: swap b_j <-> b_ipiv(j); b_{j+t} = fma(l_{j+t,j}, -b_j, b_{j+t});
//        then the upper band solve with kl+ku super-diagonals (xTBSV):
//        x_j = b_j / u_jj;  b_i = fma(u_ij, -x_j, b_i).
//   'T': transposed upper solve (dots accumulated in ascending order), then
//        the transposed unit-lower sweep in reverse with the interchanges.
// 'N' is bit-identical to the CPU routine.  One workgroup per right-hand
// side runs the sequential chain; the phases of a step are barrier-separated.
#include "bh_common.hpp"

namespace bh {

constexpr int kTrsThreads = 256;

template <typename T>
__device__ void gbtrs_n_col(i64 n, i64 kl, i64 ku, const T *__restrict__ ab, i64 ld, const i64 *__restrict__ ipiv,
                            T *b) {
    const i64 kd = kl + ku;
    __shared__ T sbj;
    if (kl > 0) {
        for (i64 j = 0; j < n - 1; ++j) {
            const i64 lm = imin(kl, n - 1 - j);
            if (threadIdx.x == 0) {
                const i64 p = ipiv[j] - 1;
                T bj = b[p];
                if (p != j) {
                    b[p] = b[j];
                    b[j] = bj;
                }
                sbj = -bj;
            }
            __syncthreads();
            const T nb = sbj;
            const T *l = ab + j * ld + kd;
            for (i64 t = 1 + threadIdx.x; t <= lm; t += blockDim.x) b[j + t] = dfma(l[t], nb, b[j + t]);
            __syncthreads();
        }
    }
    for (i64 j = n - 1; j >= 0; --j) {
        const T *u = ab + j * ld;  // u[kd + i - j] = U(i,j)
        if (threadIdx.x == 0) {
            const T xj = b[j] / u[kd];
            b[j] = xj;
            sbj = -xj;
        }
        __syncthreads();
        const T nx = sbj;
        for (i64 i = imax(0, j - kd) + threadIdx.x; i < j; i += blockDim.x) b[i] = dfma(u[kd + i - j], nx, b[i]);
        __syncthreads();
    }
}

template <typename T>
__device__ void gbtrs_t_col(i64 n, i64 kl, i64 ku, const T *__restrict__ ab, i64 ld, const i64 *__restrict__ ipiv,
                            T *b) {
    const i64 kd = kl + ku;
    // U^T x = b: x_j = (b_j - sum_{i<j} u_ij x_i) / u_jj, ascending dot order
    for (i64 j = 0; j < n; ++j) {
        if (threadIdx.x == 0) {
            const T *u = ab + j * ld;
            T acc = b[j];
            for (i64 i = imax(0, j - kd); i < j; ++i) acc = dfma(-u[kd + i - j], b[i], acc);
            b[j] = acc / u[kd];
        }
        __syncthreads();
    }
    if (kl > 0) {
        for (i64 j = n - 2; j >= 0; --j) {
            if (threadIdx.x == 0) {
                const i64 lm = imin(kl, n - 1 - j);
                const T *l = ab + j * ld + kd;
                T acc = T(0);
                for (i64 t = 1; t <= lm; ++t) acc = dfma(b[j + t], l[t], acc);
                b[j] = b[j] - acc;
                const i64 p = ipiv[j] - 1;
                if (p != j) {
                    T tmp = b[p];
                    b[p] = b[j];
                    b[j] = tmp;
                }
            }
            __syncthreads();
        }
    }
}

template <typename T>
__global__ void __launch_bounds__(kTrsThreads)
gbtrs_kernel(int trans_n, i64 n, i64 kl, i64 ku, const T *ab, i64 ld, const i64 *ipiv, T *B, i64 ldb) {
    T *b = B + (i64)blockIdx.x * ldb;
    if (trans_n)
        gbtrs_n_col<T>(n, kl, ku, ab, ld, ipiv, b);
    else
        gbtrs_t_col<T>(n, kl, ku, ab, ld, ipiv, b);
}

template <typename T>
int wave_gbtrs_launch(bool trans_n, i64 n, i64 kl, i64 ku, i64 nrhs, const T *ab, i64 ld, const i64 *ipiv, T *B,
                      i64 ldb, hipStream_t st);

template <typename T>
static int gbtrs_impl(char trans, i64 n, i64 kl, i64 ku, i64 nrhs, const T *ab, i64 ldab, const i64 *ipiv, T *B,
                      i64 ldb) {
    BH_ARG(trans == 'N' || trans == 'n' || trans == 'T' || trans == 't' || trans == 'C' || trans == 'c', 1);
    BH_ARG(n >= 0, 2);
    BH_ARG(kl >= 0, 3);
    BH_ARG(ku >= 0, 4);
    BH_ARG(nrhs >= 0, 5);
    BH_ARG(ldab >= 2 * kl + ku + 1, 7);
    BH_ARG(ldb >= imax(1, n), 10);
    if (n == 0 || nrhs == 0) return BH_OK;
    const int tn = (trans == 'N' || trans == 'n');
    // one wave per right-hand side with a register window when the band fits,
    // otherwise the barrier-stepped kernel working in global memory
    const int st = wave_gbtrs_launch<T>(tn != 0, n, kl, ku, nrhs, ab, ldab, ipiv, B, ldb, stream());
    if (st != -1) return st;
    hipLaunchKernelGGL(gbtrs_kernel<T>, dim3((unsigned)nrhs), dim3(kTrsThreads), 0, stream(), tn, n, kl, ku, ab, ldab,
                       ipiv, B, ldb);
    BH_LAUNCH_CHECK();
    return BH_OK;
}

}  // namespace bh

using namespace bh;

extern "C" {

int bh_dgbtrs(char trans, bh_int n, bh_int kl, bh_int ku, bh_int nrhs, const double *AB, bh_int ldab,
              const bh_int *ipiv, double *B, bh_int ldb) {
    return gbtrs_impl<double>(trans, n, kl, ku, nrhs, AB, ldab, ipiv, B, ldb);
}

int bh_sgbtrs(char trans, bh_int n, bh_int kl, bh_int ku, bh_int nrhs, const float *AB, bh_int ldab,
              const bh_int *ipiv, float *B, bh_int ldb) {
    return gbtrs_impl<float>(trans, n, kl, ku, nrhs, AB, ldab, ipiv, B, ldb);
}

}  // extern "C"
