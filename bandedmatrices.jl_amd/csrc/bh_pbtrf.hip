// bh_pbtrf.hip -- symmetric positive definite band Cholesky, its solve, and
// the symmetric band product (LAPACK xPBTRF / xPBTRS, BLAS xSBMV) on gfx950.
//
// Storage (LAPACK symmetric band): uplo 'U' keeps A(i,j), i <= j, at
// ab[kd + i - j + j*ldab]; 'L' keeps A(i,j), i >= j, at ab[i - j + j*ldab].
//
// Factorisation, restating xPBTF2 step for step (A = U'U or L L'): column j
// gets a_jj <- sqrt(a_jj) (stop with info = j+1 when a_jj <= 0; a NaN goes on), the kn
// = min(kd, n-1-j) entries of its row of U (column of L) are multiplied by the
// rounded reciprocal 1/a_jj, and the trailing kn x kn triangle takes the xSYR
// rank-1 update with alpha = -1: column q is skipped when v_q == 0, otherwise
// a_pq = fma(-v_q, v_p, a_pq).  Every entry sees the same operations in the
// same order as on the host, so factor and info are bit-identical.
//
// The same kernel runs the banded LDL' of src/symbanded/ldlt.jl:8-37 (no
// pivoting, no square root): d_j = a_jj, l = u / d_j by division, and the
// trailing update a_pq = fma(-l_p, u_q, a_pq) -- the right-looking form of the
// reference's left-looking loops, equal up to rounding.
//
// One workgroup walks the n steps with the kd+1 live columns in LDS (column c
// in slot c mod (kd+1)); per step: scale the row (one thread per entry), a
// barrier, the rank-1 update (one thread per column, which for column j+1
// also takes the next square root), write-out of the finished column j and
// refill of its slot with column j+kd+1, a barrier.  The barriers order LDS
// only.  Incoming columns arrive in chunks of B: a chunk is loaded into
// registers while the previous one is consumed and stored to LDS when it is
// done, so no step waits on HBM.
#include "bh_common.hpp"
#include "bh_tbsv.hpp"

namespace bh {
namespace {

template <typename T>
__device__ __forceinline__ T dsqrt(T x);
template <>
__device__ __forceinline__ double dsqrt<double>(double x) { return __builtin_sqrt(x); }
template <>
__device__ __forceinline__ float dsqrt<float>(float x) { return __builtin_sqrtf(x); }

constexpr int kPbRegs = 24;  // staged values per thread: chunk columns x (kd+1) <= kPbRegs x threads

template <typename T, bool UP, bool LDL>
__global__ void __launch_bounds__(256) pbtrf_lds_kernel(int n, int kd, T *__restrict__ ab, i64 ldab, i64 *__restrict__ info_out, int B) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int H = kd + 1;
    // slot stride: even, so that entries one row apart in neighbouring slots
    // (what neighbouring threads touch) sit an odd number of doubles apart and
    // fall in different LDS banks
    const int SL = H + (H & 1);
    T *W = reinterpret_cast<T *>(smem_raw);  // H slots of SL entries: the live columns
    T *V = W + (i64)H * SL;                  // v_1 .. v_kd of the current step (LDL: the unscaled u)
    T *VL = LDL ? V + H : V;                 // LDL: l = u / d; Cholesky: the same v
    T *S = V + (LDL ? 2 : 1) * H;            // the next B incoming columns (column j+H at S[(j mod B)*H])
    __shared__ T sR;                         // Cholesky: 1 / a_jj; LDL: the pivot d_j
    __shared__ int sInfo;
    const int tid = threadIdx.x, nt = blockDim.x;
    const int dr = UP ? kd : 0;  // diagonal row
    const int BH = B * H;
    auto slot = [&](int c) { return W + (i64)(c % H) * SL; };
    // first window (columns 0 .. kd) and first chunk (columns H .. H+B-1)
    for (int e = tid; e < H * H; e += nt) {
        const int c = e / H, r = e - c * H;
        if (c < n) W[(i64)c * SL + r] = ab[(i64)c * ldab + r];
    }
    for (int e = tid; e < BH; e += nt) {
        const int c = H + e / H, r = e % H;
        S[e] = c < n ? ab[(i64)c * ldab + r] : T(0);
    }
    if (tid == 0) sInfo = 0;
    __syncthreads();
    auto take_sqrt = [&](T *d, int info_if_bad) {
        const T v = *d;
        if (LDL) {  // no root and no check: the pivot itself (a zero pivot gives Inf/NaN, as on the host)
            sR = v;
        } else if (v <= T(0)) {
            sInfo = info_if_bad;
        } else {
            const T s = dsqrt(v);
            *d = s;
            sR = T(1) / s;
        }
    };
    if (tid == 0) take_sqrt(W + dr, 1);  // column 0 sits in slot 0
    lds_barrier();
    int done = 0;  // columns written back
    for (int j0 = 0; j0 < n && !sInfo; j0 += B) {
        // the chunk after this one, loaded now and stored into S when this chunk
        // is done: the registers are not touched in between, so no step waits
        T reg[kPbRegs];
#pragma unroll
        for (int i = 0; i < kPbRegs; ++i) {
            const int e = tid + i * nt;
            const int c = H + j0 + B + e / H, r = e % H;
            reg[i] = (e < BH && c < n) ? ab[(i64)c * ldab + r] : T(0);
        }
        const int jend = min(n, j0 + B);
        for (int j = j0; j < jend; ++j) {
            if (sInfo) break;  // uniform: read after a barrier
            const int kn = min(kd, n - 1 - j);
            // scale: v_c = U(j, j+c) (row kd-c of column j+c) or L(j+c, j) (row c of column j)
            for (int c = 1 + tid; c <= kn; c += nt) {
                T *pv = UP ? slot(j + c) + (kd - c) : slot(j) + c;
                if (LDL) {
                    const T u = *pv, l = u / sR;
                    *pv = l;
                    V[c] = u;
                    VL[c] = l;
                } else {
                    const T v = *pv * sR;
                    *pv = v;
                    V[c] = v;
                }
            }
            lds_barrier();
            // rank-1 update, thread per column q of the trailing triangle, in
            // chunks of 8 entries whose reads are issued before their writes
            for (int q = 1 + tid; q <= kn; q += nt) {
                T *col = slot(j + q);
                const T vq = V[q];
                if (vq != T(0)) {
                    const T tq = -vq;
                    const int p0 = UP ? 1 : q, p1 = UP ? q : kn;  // p0 .. p1 inclusive
                    T *a0 = UP ? col + (kd - q) : col - q;      // A(j+p, j+q) at a0[p]
                    int p = p0;
                    for (; p + 8 <= p1 + 1; p += 8) {
                        T vv[8], aa[8];
#pragma unroll
                        for (int u = 0; u < 8; ++u) {
                            vv[u] = VL[p + u];
                            aa[u] = a0[p + u];
                        }
#pragma unroll
                        for (int u = 0; u < 8; ++u) a0[p + u] = dfma(tq, vv[u], aa[u]);
                    }
                    for (; p <= p1; ++p) a0[p] = dfma(tq, VL[p], a0[p]);
                }
                if (q == 1) take_sqrt(col + dr, j + 2);  // column j+1's diagonal is final
            }
            // column j is final: write it out; its slot takes column j+kd+1 from S
            T *sj = slot(j);
            if (tid < H) {
                ab[(i64)j * ldab + tid] = sj[tid];
                if (j + H < n) sj[tid] = S[(j - j0) * H + tid];
            }
            // kd == 0: no update ran, and column j+1 has just entered slot(j)
            if (kd == 0 && tid == 0 && j + 1 < n) take_sqrt(sj, j + 2);
            done = j + 1;
            lds_barrier();
        }
#pragma unroll
        for (int i = 0; i < kPbRegs; ++i) {
            const int e = tid + i * nt;
            if (e < BH) S[e] = reg[i];
        }
        lds_barrier();
    }
    // stopped early: the columns still in the window go back as they stand
    if (done < n) {
        for (int c = done; c < min(n, done + H); ++c)
            for (int r = tid; r < H; r += nt) ab[(i64)c * ldab + r] = slot(c)[r];
    }
    if (tid == 0) *info_out = sInfo;
}

template <typename T, bool LDL>
int pbtrf_impl(char uplo, i64 n, i64 kd, T *ab, i64 ldab, i64 *info_host) {
    const bool up = (uplo == 'U' || uplo == 'u'), lo = (uplo == 'L' || uplo == 'l');
    BH_ARG(up || lo, 1);
    BH_ARG(n >= 0, 2);
    BH_ARG(kd >= 0, 3);
    BH_ARG(ldab >= kd + 1, 5);
    BH_ARG(info_host != nullptr, 6);
    *info_host = 0;
    if (n == 0) return BH_OK;
    const i64 H = kd + 1;
    const int nt = (int)imin(256, 64 * cdiv(H, 64));
    // LDS: the live window, v, and a chunk of B incoming columns (B as large as
    // the registers and 150 KiB allow, at most 32)
    const i64 fixed = (H * (H + (H & 1)) + (LDL ? 2 : 1) * H) * (i64)sizeof(T);
    const i64 budget = 150 * 1024;
    const i64 B = imin(imin(32, (i64)kPbRegs * nt / H), (budget - fixed) / (H * (i64)sizeof(T)));
    if (B < 4 || n >= (1ll << 31)) {
        set_error("pbtrf: bandwidth kd = %lld does not fit the LDS window (about 130 for FP64, 185 for FP32)",
                  (long long)kd);
        return -3;
    }
    const size_t lds = (size_t)(fixed + B * H * (i64)sizeof(T));
    hipStream_t st = stream();
    i64 *dinfo = nullptr;
    {
        const int cs = ctrl_words(&dinfo);
        if (cs) return cs;
        dinfo += 2;
    }
    auto kern = up ? pbtrf_lds_kernel<T, true, LDL> : pbtrf_lds_kernel<T, false, LDL>;
    hipError_t e = hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(kern, dim3(1), dim3(nt), lds, st, (int)n, (int)kd, ab, ldab, dinfo, (int)B);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(info_host, dinfo, sizeof(i64), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return hip_fail(e, "pbtrf", __FILE__, __LINE__);
    return BH_OK;
}

// xPBTRS: U'U x = b as U' y = b ('T' upper), U x = y ('N' upper); L L' x = b as
// L y = b ('N' lower), L' x = y ('T' lower), every column of B in one launch.
template <typename T>
int pbtrs_impl(char uplo, i64 n, i64 kd, i64 nrhs, const T *ab, i64 ldab, T *B, i64 ldb) {
    const bool up = (uplo == 'U' || uplo == 'u'), lo = (uplo == 'L' || uplo == 'l');
    BH_ARG(up || lo, 1);
    BH_ARG(n >= 0, 2);
    BH_ARG(kd >= 0, 3);
    BH_ARG(nrhs >= 0 && nrhs < (1ll << 31), 4);
    BH_ARG(ldab >= kd + 1, 6);
    BH_ARG(ldb >= (n > 1 ? n : 1), 8);
    BH_ARG(n < (1ll << 30) && kd < (1ll << 30) && ldab * n * (i64)sizeof(T) < (1ll << 31) - (1ll << 24), 2);
    if (n == 0 || nrhs == 0) return BH_OK;
    const int ni = (int)n, ki = (int)kd, nr = (int)nrhs;
    int st = tbsv_run(up ? 2 : 1, ni, ki, ab, ldab, B, 1, ldb, nr, false);
    if (st != BH_OK) return st;
    return tbsv_run(up ? 0 : 3, ni, ki, ab, ldab, B, 1, ldb, nr, false);
}

template <typename T>
__global__ void __launch_bounds__(256) diag_div_kernel(int n, int nrhs, const T *__restrict__ ab, i64 ldab, int dr,
                                                       T *__restrict__ B, i64 ldb) {
    const i64 e = (i64)blockIdx.x * 256 + threadIdx.x;
    if (e >= (i64)n * nrhs) return;
    const i64 c = e / n, j = e - c * n;
    B[c * ldb + j] = B[c * ldb + j] / ab[j * ldab + dr];
}

// ldiv!(::LDLt, B): L y = B (unit), y ./= d, L' x = y (unit), columns in one
// launch each; with 'U' storage L = U', so the sweeps are U' then U.
template <typename T>
int ldlts_impl(char uplo, i64 n, i64 kd, i64 nrhs, const T *ab, i64 ldab, T *B, i64 ldb) {
    const bool up = (uplo == 'U' || uplo == 'u'), lo = (uplo == 'L' || uplo == 'l');
    BH_ARG(up || lo, 1);
    BH_ARG(n >= 0, 2);
    BH_ARG(kd >= 0, 3);
    BH_ARG(nrhs >= 0 && nrhs < (1ll << 31), 4);
    BH_ARG(ldab >= kd + 1, 6);
    BH_ARG(ldb >= (n > 1 ? n : 1), 8);
    BH_ARG(n < (1ll << 30) && kd < (1ll << 30) && ldab * n * (i64)sizeof(T) < (1ll << 31) - (1ll << 24), 2);
    if (n == 0 || nrhs == 0) return BH_OK;
    const int ni = (int)n, ki = (int)kd, nr = (int)nrhs;
    int st = tbsv_run(up ? 2 : 1, ni, ki, ab, ldab, B, 1, ldb, nr, true);
    if (st != BH_OK) return st;
    hipLaunchKernelGGL(diag_div_kernel<T>, dim3((unsigned)cdiv(n * nrhs, 256)), dim3(256), 0, stream(), ni, nr, ab,
                       ldab, up ? ki : 0, B, ldb);
    BH_LAUNCH_CHECK();
    return tbsv_run(up ? 0 : 3, ni, ki, ab, ldab, B, 1, ldb, nr, true);
}

// ---- xSBMV: y <- alpha*A*x + beta*y, one thread per row; row i sums
// fma(A(i,j), x_j, t) for j = i-k .. i+k ascending from t = 0, then
// y_i = alpha*t (beta == 0) or alpha*t + beta*y_i.
template <typename T, bool UP>
__global__ void __launch_bounds__(256) sbmv_rows(int n, int k, T alpha, const T *__restrict__ ab, i64 ldab,
                                                 const T *__restrict__ x, int incx, T beta, T *__restrict__ y,
                                                 int incy) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    T t = T(0);
    const int j0 = max(0, i - k), j1 = min(n - 1, i + k);
    for (int j = j0; j <= j1; ++j) {
        const int r = min(i, j), c = max(i, j);  // the stored entry (r, c), r <= c
        const T a = UP ? ab[(i64)c * ldab + (k + r - c)] : ab[(i64)r * ldab + (c - r)];
        t = dfma(a, x[(i64)j * incx], t);
    }
    T *py = y + (i64)i * incy;
    *py = (beta == T(0)) ? alpha * t : alpha * t + beta * *py;
}

template <typename T>
int sbmv_impl(char uplo, i64 n, i64 k, T alpha, const T *ab, i64 ldab, const T *x, i64 incx, T beta, T *y,
              i64 incy) {
    const bool up = (uplo == 'U' || uplo == 'u'), lo = (uplo == 'L' || uplo == 'l');
    BH_ARG(up || lo, 1);
    BH_ARG(n >= 0, 2);
    BH_ARG(k >= 0, 3);
    BH_ARG(ldab >= k + 1, 6);
    BH_ARG(incx >= 1, 8);
    BH_ARG(incy >= 1, 11);
    BH_ARG(n < (1ll << 30) && k < (1ll << 30) && n * incx < (1ll << 31) && n * incy < (1ll << 31), 2);
    if (n == 0) return BH_OK;
    const int nb = (int)cdiv(n, 256);
    if (up)
        hipLaunchKernelGGL((sbmv_rows<T, true>), dim3(nb), dim3(256), 0, stream(), (int)n, (int)k, alpha, ab, ldab, x,
                           (int)incx, beta, y, (int)incy);
    else
        hipLaunchKernelGGL((sbmv_rows<T, false>), dim3(nb), dim3(256), 0, stream(), (int)n, (int)k, alpha, ab, ldab,
                           x, (int)incx, beta, y, (int)incy);
    BH_LAUNCH_CHECK();
    return BH_OK;
}

}  // namespace
}  // namespace bh

using namespace bh;

extern "C" {

int bh_dpbtrf(char uplo, bh_int n, bh_int kd, double *AB, bh_int ldab, bh_int *info_host) {
    return pbtrf_impl<double, false>(uplo, n, kd, AB, ldab, info_host);
}
int bh_spbtrf(char uplo, bh_int n, bh_int kd, float *AB, bh_int ldab, bh_int *info_host) {
    return pbtrf_impl<float, false>(uplo, n, kd, AB, ldab, info_host);
}
int bh_dsbldlt(char uplo, bh_int n, bh_int kd, double *AB, bh_int ldab) {
    bh_int info = 0;
    return pbtrf_impl<double, true>(uplo, n, kd, AB, ldab, &info);
}
int bh_ssbldlt(char uplo, bh_int n, bh_int kd, float *AB, bh_int ldab) {
    bh_int info = 0;
    return pbtrf_impl<float, true>(uplo, n, kd, AB, ldab, &info);
}
int bh_dsbldlts(char uplo, bh_int n, bh_int kd, bh_int nrhs, const double *AB, bh_int ldab, double *B, bh_int ldb) {
    return ldlts_impl<double>(uplo, n, kd, nrhs, AB, ldab, B, ldb);
}
int bh_ssbldlts(char uplo, bh_int n, bh_int kd, bh_int nrhs, const float *AB, bh_int ldab, float *B, bh_int ldb) {
    return ldlts_impl<float>(uplo, n, kd, nrhs, AB, ldab, B, ldb);
}
int bh_dpbtrs(char uplo, bh_int n, bh_int kd, bh_int nrhs, const double *AB, bh_int ldab, double *B, bh_int ldb) {
    return pbtrs_impl<double>(uplo, n, kd, nrhs, AB, ldab, B, ldb);
}
int bh_spbtrs(char uplo, bh_int n, bh_int kd, bh_int nrhs, const float *AB, bh_int ldab, float *B, bh_int ldb) {
    return pbtrs_impl<float>(uplo, n, kd, nrhs, AB, ldab, B, ldb);
}
int bh_dsbmv(char uplo, bh_int n, bh_int k, double alpha, const double *A, bh_int lda, const double *x, bh_int incx,
             double beta, double *y, bh_int incy) {
    return sbmv_impl<double>(uplo, n, k, alpha, A, lda, x, incx, beta, y, incy);
}
int bh_ssbmv(char uplo, bh_int n, bh_int k, float alpha, const float *A, bh_int lda, const float *x, bh_int incx,
             float beta, float *y, bh_int incy) {
    return sbmv_impl<float>(uplo, n, k, alpha, A, lda, x, incx, beta, y, incy);
}

}  // extern "C"
