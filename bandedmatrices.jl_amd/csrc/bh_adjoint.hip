// bh_adjoint.hip -- band-projected outer products for reverse-mode gradients.
//
// No reference counterpart (the reference defines no rrule; its transposed
// solve, src/banded/linalg.jl:41-47, is what the adjoint solve reuses).  For
// the loss L(A, b) = w' A^{-1} b + 1/2 ||A x0||^2 with x = A \ b,
// lambda = A^{-T} w and y0 = A x0:
//     dL/db = lambda,   dL/dA_ij = y0_i x0_j - lambda_i x_j   for (i,j) in the band,
// evaluated as fma(y0_i, x0_j, -(lambda_i x_j)) (one rounding for the product,
// one for the fused step).  bh_?gb_band_outer is the general rank-2 form
// G = beta G + alpha1 P(u1 v1') + alpha2 P(u2 v2') restricted to the band,
// the pullback of A\b (alpha = -1 with u = A^{-T} xbar, v = x) and of A*x0
// (alpha = 1 with u = ybar, v = x0).
//
// One thread per band entry, band storage column by column, so consecutive
// threads write consecutive addresses; HBM-bound.
#include "bh_common.hpp"

namespace bh {

namespace {

template <typename T>
__global__ void __launch_bounds__(256) band_grad_kernel(int m, int n, int kl, int ku, const T *__restrict__ lam,
                                                        const T *__restrict__ x, const T *__restrict__ y0,
                                                        const T *__restrict__ x0, T *__restrict__ G, i64 ldg) {
    const int h = kl + ku + 1;
    const i64 total = (i64)h * n;
    for (i64 f = (i64)blockIdx.x * blockDim.x + threadIdx.x; f < total; f += (i64)gridDim.x * blockDim.x) {
        const int j = (int)(f / h), t = (int)(f - (i64)j * h);
        const int i = j - ku + t;
        if (i < 0 || i >= m) continue;  // corner slots outside the matrix are left alone
        G[(i64)j * ldg + t] = dfma(y0[i], x0[j], -(lam[i] * x[j]));
    }
}

template <typename T>
__global__ void __launch_bounds__(256) band_outer_kernel(int m, int n, int kl, int ku, T alpha1, const T *__restrict__ u1,
                                                         const T *__restrict__ v1, T alpha2, const T *__restrict__ u2,
                                                         const T *__restrict__ v2, T beta, T *__restrict__ G, i64 ldg) {
    const int h = kl + ku + 1;
    const i64 total = (i64)h * n;
    for (i64 f = (i64)blockIdx.x * blockDim.x + threadIdx.x; f < total; f += (i64)gridDim.x * blockDim.x) {
        const int j = (int)(f / h), t = (int)(f - (i64)j * h);
        const int i = j - ku + t;
        if (i < 0 || i >= m) continue;
        T g = (beta == T(0)) ? T(0) : beta * G[(i64)j * ldg + t];
        if (u1) g = dfma(alpha1 * u1[i], v1[j], g);
        if (u2) g = dfma(alpha2 * u2[i], v2[j], g);
        G[(i64)j * ldg + t] = g;
    }
}

unsigned grid_for(i64 total) {
    const i64 blocks = cdiv(total, 256);
    return (unsigned)imin(imax(blocks, 1), 256 * 32);
}

template <typename T>
int band_grad_impl(i64 m, i64 n, i64 kl, i64 ku, const T *lam, const T *x, const T *y0, const T *x0, T *G, i64 ldg) {
    BH_ARG(m >= 0, 1);
    BH_ARG(n >= 0, 2);
    BH_ARG(kl >= 0, 3);
    BH_ARG(ku >= 0, 4);
    BH_ARG(ldg >= kl + ku + 1, 10);
    BH_ARG(m < (1ll << 31) && n < (1ll << 31), 1);
    if (m == 0 || n == 0) return BH_OK;
    hipLaunchKernelGGL(band_grad_kernel<T>, dim3(grid_for((kl + ku + 1) * n)), dim3(256), 0, stream(), (int)m,
                       (int)n, (int)kl, (int)ku, lam, x, y0, x0, G, ldg);
    BH_LAUNCH_CHECK();
    return BH_OK;
}

template <typename T>
int band_outer_impl(i64 m, i64 n, i64 kl, i64 ku, T alpha1, const T *u1, const T *v1, T alpha2, const T *u2,
                    const T *v2, T beta, T *G, i64 ldg) {
    BH_ARG(m >= 0, 1);
    BH_ARG(n >= 0, 2);
    BH_ARG(kl >= 0, 3);
    BH_ARG(ku >= 0, 4);
    BH_ARG((u1 == nullptr) == (v1 == nullptr), 6);
    BH_ARG((u2 == nullptr) == (v2 == nullptr), 9);
    BH_ARG(ldg >= kl + ku + 1, 13);
    BH_ARG(m < (1ll << 31) && n < (1ll << 31), 1);
    if (m == 0 || n == 0) return BH_OK;
    hipLaunchKernelGGL(band_outer_kernel<T>, dim3(grid_for((kl + ku + 1) * n)), dim3(256), 0, stream(), (int)m,
                       (int)n, (int)kl, (int)ku, alpha1, u1, v1, alpha2, u2, v2, beta, G, ldg);
    BH_LAUNCH_CHECK();
    return BH_OK;
}

}  // namespace
}  // namespace bh

using namespace bh;

extern "C" {

int bh_dgb_band_grad(bh_int m, bh_int n, bh_int kl, bh_int ku, const double *lam, const double *x, const double *y0,
                     const double *x0, double *G, bh_int ldg) {
    return band_grad_impl<double>(m, n, kl, ku, lam, x, y0, x0, G, ldg);
}
int bh_sgb_band_grad(bh_int m, bh_int n, bh_int kl, bh_int ku, const float *lam, const float *x, const float *y0,
                     const float *x0, float *G, bh_int ldg) {
    return band_grad_impl<float>(m, n, kl, ku, lam, x, y0, x0, G, ldg);
}
int bh_dgb_band_outer(bh_int m, bh_int n, bh_int kl, bh_int ku, double alpha1, const double *u1, const double *v1,
                      double alpha2, const double *u2, const double *v2, double beta, double *G, bh_int ldg) {
    return band_outer_impl<double>(m, n, kl, ku, alpha1, u1, v1, alpha2, u2, v2, beta, G, ldg);
}
int bh_sgb_band_outer(bh_int m, bh_int n, bh_int kl, bh_int ku, float alpha1, const float *u1, const float *v1,
                      float alpha2, const float *u2, const float *v2, float beta, float *G, bh_int ldg) {
    return band_outer_impl<float>(m, n, kl, ku, alpha1, u1, v1, alpha2, u2, v2, beta, G, ldg);
}

}  // extern "C"
