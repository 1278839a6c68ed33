// bh_mixed.hip -- FP32 band LU with FP64 iterative refinement (no reference
// counterpart: the reference only promotes eltypes, src/banded/BandedLU.jl:110).
//
// x = A \ B for an FP64 band matrix:
//   1. AB32 = padded FP32 copy of A; FP32 partial-pivot band LU (bh_sgbtrf);
//   2. x0 = FP32 solve of B (bh_sgbtrs), widened to FP64;
//   3. repeat: r = B - A x in FP64 (bh_dgbmv_multi, A read once for all
//      right-hand sides); stop when every column satisfies LAPACK dsgesv's
//      rule ||r||_inf <= sqrt(n) * eps64 * ||A||_inf * ||x||_inf; else
//      d = FP32 solve of r, x += d (FP64 accumulation);
//   4. after 30 iterations without convergence, or as soon as refinement
//      stagnates (the largest ratio ||r_c|| / ||x_c|| over the columns fails to
//      halve from one iteration to the next, twice running), or on an FP32
//      overflow or a zero FP32 pivot, fall back to FP64 LU + solve, reported
//      as iters < 0.
// Residuals, norms and the solution are FP64 throughout.
#include <math.h>

#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "bh_common.hpp"

namespace bh {

__global__ void d2s_kernel(i64 rows, i64 cols, const double *__restrict__ src, i64 lds, float *__restrict__ dst,
                           i64 ldd) {
    const i64 e = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= rows * cols) return;
    const i64 c = e / rows, r = e - c * rows;
    dst[c * ldd + r] = (float)src[c * lds + r];
}

__global__ void s2d_add_kernel(i64 rows, i64 cols, const float *__restrict__ d, i64 ldd, double *__restrict__ x,
                               i64 ldx, int assign) {
    const i64 e = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= rows * cols) return;
    const i64 c = e / rows, r = e - c * rows;
    const double v = (double)d[c * ldd + r];
    x[c * ldx + r] = assign ? v : x[c * ldx + r] + v;
}

// Non-negative doubles order like their bit patterns, so a max over
// magnitudes can use the 64-bit unsigned atomic max.
__device__ __forceinline__ void atomic_max_nonneg(double *addr, double v) {
    atomicMax(reinterpret_cast<unsigned long long *>(addr), (unsigned long long)__double_as_longlong(v));
}

// out[c] = max_r |v[r,c]| (out must start at 0); one atomic per workgroup
// (per-wave atomics on one address serialise)
__global__ void __launch_bounds__(256) colmax_kernel(i64 rows, i64 cols, const double *__restrict__ v, i64 ld,
                                                     double *out) {
    __shared__ double wm[4];
    const i64 c = blockIdx.y;
    double m = 0.0;
    for (i64 r = (i64)blockIdx.x * blockDim.x + threadIdx.x; r < rows; r += (i64)gridDim.x * blockDim.x) {
        const double a = fabs(v[c * ld + r]);
        m = (a > m || a != a) ? a : m;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const double o = __shfl_xor(m, off, 64);
        m = (o > m || o != o) ? o : m;
    }
    if ((threadIdx.x & 63) == 0) wm[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < 4; ++k) m = (wm[k] > m || wm[k] != wm[k]) ? wm[k] : m;
        atomic_max_nonneg(&out[c], m);
    }
}

// out = max_i sum_j |A(i,j)| over the band (||A||_inf); one atomic per workgroup
__global__ void __launch_bounds__(256) band_inf_norm_kernel(i64 n, i64 kl, i64 ku, const double *__restrict__ A,
                                                            i64 lda, double *out) {
    __shared__ double wm[4];
    const i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    double s = 0.0;
    if (i < n) {
        for (i64 j = imax(0, i - kl); j <= imin(n - 1, i + ku); ++j) s += fabs(A[j * lda + ku + i - j]);
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const double o = __shfl_xor(s, off, 64);
        s = (o > s || o != o) ? o : s;
    }
    if ((threadIdx.x & 63) == 0) wm[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < 4; ++k) s = (wm[k] > s || wm[k] != wm[k]) ? wm[k] : s;
        atomic_max_nonneg(out, s);
    }
}

__global__ void fill_zero_kernel(double *p, i64 cnt) {
    const i64 e = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < cnt) p[e] = 0.0;
}

static inline size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// The FP32 solves of the refinement use the blocked MFMA solve (its
// operands built once per factorisation) where the bands allow it, else
// bh_sgbtrs.  BANDEDHIP_MIXED_TRS=ring|blocked overrides the choice.
static bool mixed_blocked_trs(i64 n, i64 kl, i64 ku) {
    if (trs_blocked_bytes<float>(n, kl, ku) == 0) return false;
    const char *e = getenv("BANDEDHIP_MIXED_TRS");
    if (e && !strcmp(e, "ring")) return false;
    if (e && !strcmp(e, "blocked")) return true;
    return kl >= 16;
}

struct MixedWs {
    void *trs;  // blocked-solve operands (nullptr: bh_sgbtrs)
    float *ab32;
    i64 *ipiv;
    i64 *info;
    double *r;
    float *r32;
    double *norms;  // [0] = ||A||, [1..nrhs] = ||x_c||, [1+nrhs..] = ||r_c||
    size_t bytes;
};

static MixedWs mixed_layout(i64 n, i64 kl, i64 ku, i64 nrhs, void *base) {
    MixedWs w{};
    char *p = (char *)base;
    size_t off = 0;
    const i64 ldab = 2 * kl + ku + 1;
    w.ab32 = (float *)(p + off);
    off += align256((size_t)ldab * n * sizeof(float));
    w.ipiv = (i64 *)(p + off);
    off += align256((size_t)imax(n, 1) * sizeof(i64));
    w.info = (i64 *)(p + off);
    off += align256(sizeof(i64));
    w.r = (double *)(p + off);
    off += align256((size_t)n * nrhs * sizeof(double));
    w.r32 = (float *)(p + off);
    off += align256((size_t)n * nrhs * sizeof(float));
    w.norms = (double *)(p + off);
    off += align256((size_t)(1 + 2 * nrhs) * sizeof(double));
    w.trs = nullptr;
    if (mixed_blocked_trs(n, kl, ku)) {
        w.trs = p + off;
        off += align256(trs_blocked_bytes<float>(n, kl, ku));
    }
    w.bytes = off;
    return w;
}

static dim3 grid1(i64 cnt) { return dim3((unsigned)cdiv(cnt, 256)); }

// Bytes of the FP64 fallback (padded factors + pivots).
static size_t fallback_bytes(i64 n, i64 kl, i64 ku) {
    return align256((size_t)(2 * kl + ku + 1) * n * sizeof(double)) + (size_t)imax(n, 1) * sizeof(i64);
}

// ---- statically allocated scratch for bh_dsgbsv_noalloc.  The array lives in
// this code object on every device that loads it; its device address is
// looked up once per device by noalloc_setup (called from bh_init with the
// device current) and only that address is ever handed to kernels or copies.
constexpr size_t kNoAllocBytes = size_t(64) << 20;
constexpr int kMaxDevices = 64;
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];
static void *g_noalloc_ptr[kMaxDevices] = {};
static int g_noalloc_status[kMaxDevices] = {};

int noalloc_setup(int device) {
    if (device < 0 || device >= kMaxDevices) return BH_OK;  // the entry point reports it
    if (g_noalloc_ptr[device]) return BH_OK;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr) {
        // recorded, not fatal for bh_init: only bh_dsgbsv_noalloc needs it
        g_noalloc_status[device] = e != hipSuccess ? (int)e : (int)hipErrorInvalidSymbol;
        return BH_ERR_HIP;
    }
    g_noalloc_ptr[device] = p;
    return BH_OK;
}

// The mixed solve with its workspace at `work`; the FP64 fallback takes its
// factors from `fb` (nullptr: the library's growable scratch buffer).
static int dsgbsv_impl(i64 n, i64 kl, i64 ku, i64 nrhs, const double *A, i64 lda, const double *B, i64 ldb,
                       double *X, i64 ldx, void *work, size_t lwork, void *fb, i64 *ipiv32, i64 *iters_host,
                       i64 *info_host);

}  // namespace bh

using namespace bh;

extern "C" {

int bh_dsgbsv_workspace(bh_int n, bh_int kl, bh_int ku, bh_int nrhs, size_t *bytes) {
    BH_ARG(n >= 0, 1);
    BH_ARG(kl >= 0, 2);
    BH_ARG(ku >= 0, 3);
    BH_ARG(nrhs >= 0, 4);
    BH_ARG(bytes != nullptr, 5);
    *bytes = mixed_layout(n, kl, ku, nrhs, nullptr).bytes;
    return BH_OK;
}

int bh_dsgbsv(bh_int n, bh_int kl, bh_int ku, bh_int nrhs, const double *A, bh_int lda, const double *B, bh_int ldb,
              double *X, bh_int ldx, void *work, size_t lwork, bh_int *ipiv32, bh_int *iters_host, bh_int *info_host) {
    return dsgbsv_impl(n, kl, ku, nrhs, A, lda, B, ldb, X, ldx, work, lwork, nullptr, ipiv32, iters_host, info_host);
}

int bh_dsgbsv_noalloc(bh_int n, bh_int kl, bh_int ku, bh_int nrhs, const double *A, bh_int lda, const double *B,
                      bh_int ldb, double *X, bh_int ldx, bh_int *ipiv32, bh_int *iters_host, bh_int *info_host) {
    BH_ARG(n >= 0, 1);
    BH_ARG(kl >= 0, 2);
    BH_ARG(ku >= 0, 3);
    BH_ARG(nrhs >= 0, 4);
    BH_ARG(lda >= kl + ku + 1, 6);
    BH_ARG(ldb >= imax(1, n), 8);
    BH_ARG(ldx >= imax(1, n), 10);
    BH_ARG(iters_host != nullptr, 12);
    BH_ARG(info_host != nullptr, 13);
    const size_t need = std::max(mixed_layout(n, kl, ku, nrhs, nullptr).bytes, fallback_bytes(n, kl, ku));
    if (need > kNoAllocBytes) {
        set_error("bh_dsgbsv_noalloc: n=%lld kl=%lld ku=%lld nrhs=%lld needs %zu bytes of scratch; the static scratch "
                  "holds %zu (use bh_dsgbsv with a caller workspace)",
                  (long long)n, (long long)kl, (long long)ku, (long long)nrhs, need, kNoAllocBytes);
        return BH_ERR_WORKSPACE;
    }
    int dev = -1;
    BH_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices || g_noalloc_ptr[dev] == nullptr) {
        const int le = (dev >= 0 && dev < kMaxDevices) ? g_noalloc_status[dev] : 0;
        if (le)
            set_error("bh_dsgbsv_noalloc: hipGetSymbolAddress of the static scratch failed on device %d: %s", dev,
                      hipGetErrorString((hipError_t)le));
        else
            set_error("bh_dsgbsv_noalloc: no static scratch address for device %d (call bh_init(%d) first)", dev, dev);
        return BH_ERR_NODEVICE;
    }
    void *p = g_noalloc_ptr[dev];
    // the fallback overlays the refinement's workspace: by then it is dead
    return dsgbsv_impl(n, kl, ku, nrhs, A, lda, B, ldb, X, ldx, p, kNoAllocBytes, p, ipiv32, iters_host, info_host);
}

}  // extern "C"

namespace bh {

static int dsgbsv_impl(i64 n, i64 kl, i64 ku, i64 nrhs, const double *A, i64 lda, const double *B, i64 ldb,
                       double *X, i64 ldx, void *work, size_t lwork, void *fb, i64 *ipiv32, i64 *iters_host,
                       i64 *info_host) {
    BH_ARG(n >= 0, 1);
    BH_ARG(kl >= 0, 2);
    BH_ARG(ku >= 0, 3);
    BH_ARG(nrhs >= 0, 4);
    BH_ARG(lda >= kl + ku + 1, 6);
    BH_ARG(ldb >= imax(1, n), 8);
    BH_ARG(ldx >= imax(1, n), 10);
    BH_ARG(iters_host != nullptr, 14);
    BH_ARG(info_host != nullptr, 15);
    *iters_host = 0;
    *info_host = 0;
    if (n == 0 || nrhs == 0) return BH_OK;
    MixedWs w = mixed_layout(n, kl, ku, nrhs, work);
    if (work == nullptr || lwork < w.bytes) {
        set_error("bh_dsgbsv: workspace of %zu bytes needed, %zu given", w.bytes, lwork);
        return BH_ERR_WORKSPACE;
    }
    hipStream_t s = stream();
    const i64 ldab = 2 * kl + ku + 1;
    const double eps = 1.1102230246251565e-16;  // dlamch('Epsilon') = 2^-53
    const int itmax = 30;
    int st;

    // ||A||_inf
    hipLaunchKernelGGL(fill_zero_kernel, grid1(1 + 2 * nrhs), dim3(256), 0, s, w.norms, 1 + 2 * nrhs);
    hipLaunchKernelGGL(band_inf_norm_kernel, grid1(n), dim3(256), 0, s, n, kl, ku, A, lda, w.norms);
    BH_LAUNCH_CHECK();
    // FP32 factorisation
    st = bh_dsgbpad(n, n, kl, ku, A, lda, w.ab32, ldab);
    if (st) return st;
    i64 info32 = 0;
    st = bh_sgbtrf(n, n, kl, ku, w.ab32, ldab, w.ipiv, &info32);
    if (st) return st;
    if (ipiv32) BH_HIP(hipMemcpyAsync(ipiv32, w.ipiv, n * sizeof(i64), hipMemcpyDeviceToDevice, s));
    double anrm = 0.0;
    BH_HIP(hipMemcpyAsync(&anrm, w.norms, sizeof(double), hipMemcpyDeviceToHost, s));
    BH_HIP(hipStreamSynchronize(s));
    const double cte = anrm * eps * sqrt((double)n);
    bool fallback = (info32 != 0) || !(anrm <= 3.4028234663852886e38);
    int iter = 0;
    bool converged = false;
    if (!fallback) {
        // x0 = solve in FP32
        hipLaunchKernelGGL(d2s_kernel, grid1(n * nrhs), dim3(256), 0, s, n, nrhs, B, ldb, w.r32, n);
        BH_LAUNCH_CHECK();
        if (w.trs) {
            st = trs_blocked_prepare<float>(n, kl, ku, w.ab32, ldab, w.ipiv, w.trs, s);
            if (!st) st = trs_blocked_solve<float>(n, kl, ku, nrhs, w.ab32, ldab, w.trs, w.r32, n, s);
        } else {
            st = bh_sgbtrs('N', n, kl, ku, nrhs, w.ab32, ldab, w.ipiv, w.r32, n);
        }
        if (st) return st;
        hipLaunchKernelGGL(s2d_add_kernel, grid1(n * nrhs), dim3(256), 0, s, n, nrhs, w.r32, n, X, ldx, 1);
        BH_LAUNCH_CHECK();
        std::vector<double> nh(1 + 2 * nrhs);
        double prev_ratio = HUGE_VAL;
        int slow = 0;
        for (iter = 0; iter <= itmax; ++iter) {
            // r = B - A x (FP64)
            BH_HIP(hipMemcpy2DAsync(w.r, n * sizeof(double), B, ldb * sizeof(double), n * sizeof(double), nrhs,
                                    hipMemcpyDeviceToDevice, s));
            st = bh_dgbmv_multi('N', n, n, kl, ku, nrhs, -1.0, A, lda, X, ldx, 1.0, w.r, n);
            if (st) return st;
            hipLaunchKernelGGL(fill_zero_kernel, grid1(2 * nrhs), dim3(256), 0, s, w.norms + 1, 2 * nrhs);
            hipLaunchKernelGGL(colmax_kernel, dim3((unsigned)imin(cdiv(n, 2048), 64), (unsigned)nrhs), dim3(256), 0, s,
                               n, nrhs, X, ldx, w.norms + 1);
            hipLaunchKernelGGL(colmax_kernel, dim3((unsigned)imin(cdiv(n, 2048), 64), (unsigned)nrhs), dim3(256), 0, s,
                               n, nrhs, w.r, n, w.norms + 1 + nrhs);
            BH_LAUNCH_CHECK();
            BH_HIP(hipMemcpyAsync(nh.data(), w.norms, (1 + 2 * nrhs) * sizeof(double), hipMemcpyDeviceToHost, s));
            BH_HIP(hipStreamSynchronize(s));
            bool ok = true;
            double ratio = 0.0;
            for (i64 c = 0; c < nrhs; ++c) {
                const double xn = nh[1 + c], rn = nh[1 + nrhs + c];
                if (!(rn <= xn * cte)) ok = false;
                const double q = rn / (xn > 0.0 ? xn : 1.0);
                ratio = (q > ratio || q != q) ? q : ratio;
            }
            if (ok) {
                converged = true;
                break;
            }
            if (iter == itmax) break;
            // stagnation: the residual ratio must at least halve per correction
            slow = (iter >= 1 && !(ratio <= 0.5 * prev_ratio)) ? slow + 1 : 0;
            if (slow >= 2) break;
            prev_ratio = ratio;
            // d = A32 \ r ; x += d
            hipLaunchKernelGGL(d2s_kernel, grid1(n * nrhs), dim3(256), 0, s, n, nrhs, w.r, n, w.r32, n);
            BH_LAUNCH_CHECK();
            st = w.trs ? trs_blocked_solve<float>(n, kl, ku, nrhs, w.ab32, ldab, w.trs, w.r32, n, s)
                       : bh_sgbtrs('N', n, kl, ku, nrhs, w.ab32, ldab, w.ipiv, w.r32, n);
            if (st) return st;
            hipLaunchKernelGGL(s2d_add_kernel, grid1(n * nrhs), dim3(256), 0, s, n, nrhs, w.r32, n, X, ldx, 0);
            BH_LAUNCH_CHECK();
        }
    }
    if (converged) {
        *iters_host = iter;
        return BH_OK;
    }
    // FP64 fallback: LU + solve, the FP64 factors in the library's scratch
    // buffer (grown on the first fallback of a size, reused afterwards)
    double *ab64 = nullptr;
    i64 *ipiv64 = nullptr;
    {
        const size_t fa = align256((size_t)ldab * n * sizeof(double));
        void *sp = fb;
        if (sp == nullptr) {
            st = scratch_buffer(fallback_bytes(n, kl, ku), &sp);
            if (st) return st;
        }
        ab64 = (double *)sp;
        ipiv64 = (i64 *)((char *)sp + fa);
    }
    st = bh_dgbpad(n, n, kl, ku, A, lda, ab64, ldab);
    i64 info64 = 0;
    if (!st) st = bh_dgbtrf(n, n, kl, ku, ab64, ldab, ipiv64, &info64);
    if (!st && info64 == 0) {
        BH_HIP(hipMemcpy2DAsync(X, ldx * sizeof(double), B, ldb * sizeof(double), n * sizeof(double), nrhs,
                                hipMemcpyDeviceToDevice, s));
        st = bh_dgbtrs('N', n, kl, ku, nrhs, ab64, ldab, ipiv64, X, ldx);
    }
    BH_HIP(hipStreamSynchronize(s));
    *info_host = info64;
    *iters_host = -(iter + 1);
    return st;
}

}  // namespace bh
