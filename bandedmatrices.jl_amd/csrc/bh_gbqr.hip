// bh_gbqr.hip -- banded Householder QR, Q / Q' application and QR solve.
//
// Reference: src/banded/bandedqr.jl.  qr(A) copies A into a padded band
// matrix with bands (l, l+u) (bandedqr.jl:2, the same copy as the LU's) and
// runs _banded_qr! (bandedqr.jl:8-28): for k = 1 .. min(m-1, n) (real T) a
// reflector over rows k .. min(k+l, m) of column k (ArrayLayouts.reflector!:
// beta = -copysign(||x||, x1), x(2:end) /= x1 + copysign(||x||, x1),
// tau = (x1 + copysign(||x||,x1)) / copysign(||x||,x1); tau = 0 if ||x|| = 0),
// applied to the next u' = l+u columns (reflectorApply!: w = y1 + x(2:end)'y(2:end),
// w *= tau, y1 -= w, y(2:end) -= x(2:end) w).  R occupies the upper u'
// bands of the padded storage, the reflectors' v (v1 = 1 implicit) the
// lower l band, tau has min(m, n) entries (zero past the last reflector).
//
// In this file u is the padded upper bandwidth u' (the storage's own).
//
// Kernels:
//   * gbqrf_wave: one wave per system.  Lane c holds band column c of the
//     live window (columns k .. k+u, lane = column mod 64) with its rows
//     k .. k+l in registers; every lane computes the step-k reflector from
//     lane k mod 64's column (readlane), the other lanes apply it to their
//     columns with sequential dot products (the reference's order).  Rows
//     shift through registers; the row entering at the bottom and the column
//     entering at the right are prefetched PF steps ahead.
//   * gbqr_apply_wave: one wave per right-hand side; lanes hold b over the
//     reflector's rows; w = b_k + sum x_i b_i is a DPP tree sum.
//   * gbqrf_wide: bands beyond the wave kernel's window (l+1 > 32 or
//     u'+1 > 64, l+1 <= 512): one 1024-thread workgroup; per column wave 0
//     forms the reflector (norm as a wave-parallel sum), then every wave
//     applies it to every 16th later column of the band window, 64 lanes on
//     a column's rows with a DPP-reduced dot.  The sums run in a different
//     order than the reference's sequential loops: factors agree to rounding.
//   * gbqr_apply_wide: Q'B / QB for l+1 > 64: one wave per right-hand side,
//     the right-hand side resident in LDS, the same wave-parallel dot per
//     reflector (m up to kWideApplyRows).
//   * gbqrf_global / gbqr_apply_serial: any band, sequential sums in the
//     reference's order (fallback beyond those limits).
// The QR solve (bandedqr.jl:178-206) is Q'B followed by the upper-triangular
// band solve of R (tribanded.jl:75-83 -> tbsv), which is the back
// substitution of the band LU solve with no L part.
#include "bh_common.hpp"

namespace bh {

template <typename T>
int wave_gbtrs_launch(bool trans_n, i64 n, i64 kl, i64 ku, i64 nrhs, const T *ab, i64 ld, const i64 *ipiv, T *B,
                      i64 ldb, hipStream_t st);

namespace {

constexpr int QPF = 4;  // prefetch distance (steps)

__device__ __forceinline__ double qsqrt(double x) { return __builtin_sqrt(x); }
__device__ __forceinline__ float qsqrt(float x) { return __builtin_sqrtf(x); }
__device__ __forceinline__ double qcopysign(double a, double b) { return __builtin_copysign(a, b); }
__device__ __forceinline__ float qcopysign(float a, float b) { return __builtin_copysignf(a, b); }

// Element (i, j) of the band storage (diagonal at row u), or 0 outside.
template <typename T>
__device__ __forceinline__ T band_at(const T *D, i64 ld, int m, int n, int l, int u, int i, int j) {
    if (i < 0 || j < 0 || i >= m || j >= n || j - i > u || i - j > l) return T(0);
    return D[(i64)j * ld + (u + i - j)];
}

// ---------------------------------------------------------------- factor, one wave
template <typename T, int RL>
__global__ void __launch_bounds__(64) gbqrf_wave(int m, int n, int l, int u, T *D, i64 ld, T *tau) {
    const int lane = threadIdx.x;
    const int mn = m < n ? m : n;
    const int ncols = (m - 1 < n) ? m - 1 : n;
    for (int k = ncols + lane; k < mn; k += 64) tau[k] = T(0);
    if (ncols <= 0) return;
    // lane holds column col(lane) = the column c >= k with c = lane (mod 64)
    auto col_of = [&](int k) { return k + ((lane - k) % 64 + 64) % 64; };
    T y[RL];
    {
        const int c = col_of(0);
#pragma unroll
        for (int i = 0; i < RL; ++i) y[i] = (c <= u && i <= l) ? band_at(D, ld, m, n, l, u, i, c) : T(0);
    }
    // prefetch ring: entry of the row entering at step k+1 (row k+1+l) for this
    // lane's column at that time, and the entering column's rows (lane i <- row i)
    T rowin[QPF], colin[QPF];
    auto fetch = [&](int slot, int k) {
        // after step k the window is rows k+1 .. k+1+l, columns k+1 .. k+1+u
        const int c = col_of(k + 1);
        const int r = k + 1 + l;
        rowin[slot] = (c - (k + 1) <= u) ? band_at(D, ld, m, n, l, u, r, c) : T(0);
        const int cn = k + 1 + u;  // column entering at step k+1
        colin[slot] = (lane <= l) ? band_at(D, ld, m, n, l, u, k + 1 + lane, cn) : T(0);
    };
#pragma unroll
    for (int s = 0; s < QPF; ++s) fetch(s, s);
    for (int k0 = 0; k0 < ncols; k0 += QPF) {
#pragma unroll
        for (int s = 0; s < QPF; ++s) {
            const int k = k0 + s;
            if (k < ncols) {
                const int own = k & 63;  // lane of column k
                const int len = (l + 1 < m - k) ? l + 1 : m - k;
                // ---- reflector of column k (every lane, uniform values)
                T x[RL];
#pragma unroll
                for (int i = 0; i < RL; ++i) x[i] = lane_read(y[i], own);
                T nrm = x[0] * x[0];
#pragma unroll
                for (int i = 1; i < RL; ++i)
                    if (i < len) nrm = nrm + x[i] * x[i];
                T tk = T(0);
                if (nrm != T(0)) {
                    nrm = qsqrt(nrm);
                    const T nu = qcopysign(nrm, x[0]);
                    const T xi1 = x[0] + nu;
                    x[0] = -nu;
#pragma unroll
                    for (int i = 1; i < RL; ++i)
                        if (i < len) x[i] = x[i] / xi1;
                    tk = xi1 / nu;
                }
                // ---- apply to columns k+1 .. min(k+u, n-1)
                const int c = col_of(k);
                const int dc = c - k;
                if (lane == own) {
#pragma unroll
                    for (int i = 0; i < RL; ++i) y[i] = x[i];
                } else if (dc <= u && c < n) {
                    T w = y[0];
#pragma unroll
                    for (int i = 1; i < RL; ++i)
                        if (i < len) w = w + x[i] * y[i];
                    w = tk * w;
                    y[0] = y[0] - w;
#pragma unroll
                    for (int i = 1; i < RL; ++i)
                        if (i < len) y[i] = y[i] - x[i] * w;
                }
                // ---- outputs: column k (R diagonal + reflector), row k of the others
                if (lane == own) {
                    tau[k] = tk;
#pragma unroll
                    for (int i = 0; i < RL; ++i)
                        if (i < len) D[(i64)k * ld + u + i] = y[i];
                } else if (dc <= u && c < n) {
                    D[(i64)c * ld + (u - dc)] = y[0];
                }
                // ---- advance: rows shift up, the entering row / column arrive
#pragma unroll
                for (int i = 0; i + 1 < RL; ++i) y[i] = y[i + 1];
                y[RL - 1] = T(0);
                if (l < RL) {
                    // bottom row of the window: register l
#pragma unroll
                    for (int i = 0; i < RL; ++i)
                        if (i == l) y[i] = rowin[s];
                }
                const int lin = (k + 1 + u) & 63;  // lane taking the entering column
                const T cin[1] = {colin[s]};
#pragma unroll
                for (int i = 0; i < RL; ++i) {
                    const T v = lane_read(cin[0], i < 63 ? i : 63);
                    if (lane == lin && i <= l) y[i] = v;
                }
                fetch(s, k + QPF);
            }
        }
    }
    // Row ncols (if < m) of the columns still in the window is final too
    // (no later reflector touches it): write it, as the reference's in-place
    // storage holds it.
    {
        const int k = ncols;
        const int c = col_of(k);
        const int dc = c - k;
        if (k < m && dc <= u && c < n) D[(i64)c * ld + (u - dc)] = y[0];
    }
}

// ---------------------------------------------------------------- factor, wide bands
// RR rows per lane (l+1 <= 64 RR); 16 waves.
template <typename T, int RR>
__global__ void __launch_bounds__(1024) gbqrf_wide(int m, int n, int l, int u, T *D, i64 ld, T *tau) {
    __shared__ T sv[64 * RR];  // the current reflector v (v_0 = 1)
    __shared__ T stk;
    const int tid = threadIdx.x, lane = tid & 63, w = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int mn = m < n ? m : n;
    const int ncols = (m - 1 < n) ? m - 1 : n;
    for (int k = tid; k < mn; k += blockDim.x) tau[k] = T(0);
    __syncthreads();
    for (int k = 0; k < ncols; ++k) {
        const int len = (l + 1 < m - k) ? l + 1 : m - k;
        const int jmax = (u < n - 1 - k) ? u : n - 1 - k;
        if (w == 0) {
            T *x = D + (i64)k * ld + u;
            T xv[RR];
            T ss = T(0);
#pragma unroll
            for (int r = 0; r < RR; ++r) {
                const int i = lane + 64 * r;
                xv[r] = i < len ? x[i] : T(0);
                ss = ss + xv[r] * xv[r];
            }
            const T nrm2 = wave_sum_dpp(ss);
            T tk = T(0);
            if (nrm2 != T(0)) {
                const T nrm = qsqrt(nrm2);
                const T x0 = lane_read(xv[0], 0);
                const T nu = qcopysign(nrm, x0);
                const T xi1 = x0 + nu;
#pragma unroll
                for (int r = 0; r < RR; ++r) {
                    const int i = lane + 64 * r;
                    xv[r] = (i == 0) ? -nu : xv[r] / xi1;
                }
                tk = xi1 / nu;
            }
#pragma unroll
            for (int r = 0; r < RR; ++r) {
                const int i = lane + 64 * r;
                if (i < len) {
                    x[i] = xv[r];
                    sv[i] = (i == 0) ? T(1) : xv[r];
                }
            }
            if (lane == 0) {
                tau[k] = tk;
                stk = tk;
            }
        }
        __syncthreads();
        const T tk = stk;
        if (tk != T(0)) {
            T v[RR];
#pragma unroll
            for (int r = 0; r < RR; ++r) {
                const int i = lane + 64 * r;
                v[r] = i < len ? sv[i] : T(0);
            }
            // y <- y - v (tk v'y) for columns k+1 .. k+jmax, two columns in flight
            for (int jj = 1 + w; jj <= jmax; jj += 32) {
                const int jj2 = jj + 16;
                const bool two = jj2 <= jmax;  // wave-uniform
                T *y = D + (i64)(k + jj) * ld + (u - jj);
                T *y2 = D + (i64)(k + jj2) * ld + (u - jj2);
                T a[RR], b[RR];
                T s1 = T(0), s2 = T(0);
#pragma unroll
                for (int r = 0; r < RR; ++r) {
                    const int i = lane + 64 * r;
                    a[r] = i < len ? y[i] : T(0);
                    b[r] = (two && i < len) ? y2[i] : T(0);
                }
#pragma unroll
                for (int r = 0; r < RR; ++r) {
                    s1 = s1 + v[r] * a[r];
                    s2 = s2 + v[r] * b[r];
                }
                s1 = tk * wave_sum_dpp(s1);
                s2 = tk * wave_sum_dpp(s2);
#pragma unroll
                for (int r = 0; r < RR; ++r) {
                    const int i = lane + 64 * r;
                    if (i < len) {
                        y[i] = a[r] - v[r] * s1;
                        if (two) y2[i] = b[r] - v[r] * s2;
                    }
                }
            }
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------- factor, any band
template <typename T>
__global__ void __launch_bounds__(256) gbqrf_global(int m, int n, int l, int u, T *D, i64 ld, T *tau) {
    const int tid = threadIdx.x;
    const int mn = m < n ? m : n;
    const int ncols = (m - 1 < n) ? m - 1 : n;
    __shared__ T sh_tau;
    for (int k = tid; k < mn; k += blockDim.x) tau[k] = T(0);
    __syncthreads();
    for (int k = 0; k < ncols; ++k) {
        const int len = (l + 1 < m - k) ? l + 1 : m - k;
        T *x = D + (i64)k * ld + u;
        if (tid == 0) {
            T nrm = x[0] * x[0];
            for (int i = 1; i < len; ++i) nrm = nrm + x[i] * x[i];
            T tk = T(0);
            if (nrm != T(0)) {
                nrm = qsqrt(nrm);
                const T nu = qcopysign(nrm, x[0]);
                const T xi1 = x[0] + nu;
                x[0] = -nu;
                for (int i = 1; i < len; ++i) x[i] = x[i] / xi1;
                tk = xi1 / nu;
            }
            sh_tau = tk;
            tau[k] = tk;
        }
        __syncthreads();
        const T tk = sh_tau;
        const int jmax = (u < n - 1 - k) ? u : n - 1 - k;
        for (int jj = 1 + tid; jj <= jmax; jj += blockDim.x) {
            T *y = D + (i64)(k + jj) * ld + (u - jj);
            T w = y[0];
            for (int i = 1; i < len; ++i) w = w + x[i] * y[i];
            w = tk * w;
            y[0] = y[0] - w;
            for (int i = 1; i < len; ++i) y[i] = y[i] - x[i] * w;
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------- Q' B / Q B
// One wave per right-hand side: lanes 0..l hold b over rows k..k+l.
template <typename T>
__global__ void __launch_bounds__(64) gbqr_apply_wave(bool transpose, int m, int n, int l, int u, const T *D, i64 ld,
                                                       const T *tau, int kmax, T *B, i64 ldb) {
    const int lane = threadIdx.x;
    T *b = B + (i64)blockIdx.x * ldb;
    if (kmax <= 0) return;
    T xr[QPF], tr[QPF], br[QPF];
    // reflector k and the b entry entering the window after step k (row
    // k+1+l going forward, row k-1 going backward: untouched until then)
    auto fetch = [&](int slot, int k) {
        const bool ok = (k >= 0 && k < kmax);
        const int len = ok ? ((l + 1 < m - k) ? l + 1 : m - k) : 0;
        xr[slot] = (ok && lane >= 1 && lane < len) ? D[(i64)k * ld + u + lane] : T(0);
        tr[slot] = ok ? tau[k] : T(0);
        const int rin = transpose ? k + 1 + l : k - 1;
        br[slot] = (ok && rin >= 0 && rin < m) ? b[rin] : T(0);
    };
    if (transpose) {
        // k = 0 .. kmax-1; window rows k .. k+l, retire row k after step k
        T w = (lane <= l && lane < m) ? b[lane] : T(0);
#pragma unroll
        for (int s = 0; s < QPF; ++s) fetch(s, s);
        for (int k0 = 0; k0 < kmax; k0 += QPF) {
#pragma unroll
            for (int s = 0; s < QPF; ++s) {
                const int k = k0 + s;
                if (k < kmax) {
                    const T xv = (lane == 0) ? T(1) : xr[s];
                    T v = wave_sum_dpp(xv * w);
                    v = tr[s] * v;
                    w = w - xv * v;
                    if (lane == 0) b[k] = w;
                    w = wave_down1(w, T(0));
                    if (lane == l) w = br[s];
                    fetch(s, k + QPF);
                }
            }
        }
        // rows kmax .. kmax+l-1 still in the window
        const int r = kmax + lane;
        if (lane < l && r < m) b[r] = w;
    } else {
        // k = kmax-1 .. 0; window rows k .. k+l, retire row k+l after step k
        const int k1 = kmax - 1;
        T w = (lane <= l && k1 + lane < m) ? b[k1 + lane] : T(0);
#pragma unroll
        for (int s = 0; s < QPF; ++s) fetch(s, k1 - s);
        for (int k0 = k1; k0 >= 0; k0 -= QPF) {
#pragma unroll
            for (int s = 0; s < QPF; ++s) {
                const int k = k0 - s;
                if (k >= 0) {
                    const T xv = (lane == 0) ? T(1) : xr[s];
                    T v = wave_sum_dpp(xv * w);
                    v = tr[s] * v;
                    w = w - xv * v;
                    const int rout = k + l;
                    if (lane == l && rout < m) b[rout] = w;
                    w = wave_up1(w, br[s]);
                    fetch(s, k - QPF);
                }
            }
        }
        // rows 0 .. l-1 still in the window (window rows -1 .. l-1 after the last shift)
        const int r = lane - 1;
        if (lane >= 1 && lane <= l && r < m) b[r] = w;
    }
}

// Q' B / Q B for l + 1 > 64: one wave per right-hand side, b resident in LDS
// (m <= kWideApplyRows), reflector v of the next step loaded during this one.
constexpr int kWideApplyRows = 12288;
template <typename T, int RR>
__global__ void __launch_bounds__(64) gbqr_apply_wide(bool transpose, int m, int l, int u, const T *D, i64 ld,
                                                      const T *tau, int kmax, T *B, i64 ldb) {
    extern __shared__ __align__(16) unsigned char qa_raw[];
    T *bs = reinterpret_cast<T *>(qa_raw);
    const int lane = threadIdx.x;
    T *b = B + (i64)blockIdx.x * ldb;
    for (int i = lane; i < m; i += 64) bs[i] = b[i];
    auto load_v = [&](int t, T (&v)[RR], T &tk) {
        const int k = transpose ? t : kmax - 1 - t;
        const int len = (l + 1 < m - k) ? l + 1 : m - k;
        const T *x = D + (i64)k * ld + u;
#pragma unroll
        for (int r = 0; r < RR; ++r) {
            const int i = lane + 64 * r;
            v[r] = (i == 0) ? T(1) : (i < len ? x[i] : T(0));
        }
        tk = tau[k];
    };
    T v[RR], vn[RR], tk, tkn;
    load_v(0, v, tk);
    for (int t = 0; t < kmax; ++t) {
        if (t + 1 < kmax) load_v(t + 1, vn, tkn);
        const int k = transpose ? t : kmax - 1 - t;
        const int len = (l + 1 < m - k) ? l + 1 : m - k;
        T y[RR];
        T s = T(0);
#pragma unroll
        for (int r = 0; r < RR; ++r) {
            const int i = lane + 64 * r;
            y[r] = i < len ? bs[k + i] : T(0);
            s = s + v[r] * y[r];
        }
        s = tk * wave_sum_dpp(s);
#pragma unroll
        for (int r = 0; r < RR; ++r) {
            const int i = lane + 64 * r;
            if (i < len) bs[k + i] = y[r] - v[r] * s;
        }
        if (t + 1 < kmax) {
#pragma unroll
            for (int r = 0; r < RR; ++r) v[r] = vn[r];
            tk = tkn;
        }
    }
    for (int i = lane; i < m; i += 64) b[i] = bs[i];
}

// Q' B / Q B with reflectors k < kmax applied in sequence, for bands l + 1 > 64
// (thread 0 per right-hand side, slow path).
template <typename T>
__global__ void gbqr_apply_serial(bool transpose, int m, int n, int l, int u, const T *D, i64 ld, const T *tau,
                                  int kmax, int nrhs, T *B, i64 ldb) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nrhs) return;
    T *b = B + (i64)j * ldb;
    for (int t = 0; t < kmax; ++t) {
        const int k = transpose ? t : kmax - 1 - t;
        const int ihi = (k + l < m - 1) ? k + l : m - 1;
        const T *x = D + (i64)k * ld + u;
        T v = b[k];
        for (int i = k + 1; i <= ihi; ++i) v = v + x[i - k] * b[i];
        v = tau[k] * v;
        b[k] = b[k] - v;
        for (int i = k + 1; i <= ihi; ++i) b[i] = b[i] - x[i - k] * v;
    }
}

// A Q / A Q' (banded_qr_rmul!): one thread per row of A; the reflectors are
// applied in sequence (ascending for Q, descending for Q'), each to columns
// k .. min(k+l, mQ-1) of the thread's row, in the restated order of sums.
template <typename T>
__global__ void __launch_bounds__(256) gbqr_rmul_rows(bool transpose, int mA, int mQ, int kmax, int l, int u,
                                                      const T *__restrict__ D, i64 ld, const T *__restrict__ tau,
                                                      T *A, i64 lda) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= mA) return;
    T *a = A + i;
    for (int t = 0; t < kmax; ++t) {
        const int k = transpose ? kmax - 1 - t : t;
        const int jhi = (k + l < mQ - 1) ? k + l : mQ - 1;
        const T *x = D + (i64)k * ld + u - k;  // x[j] = v_j for j in k+1 .. jhi
        T v = a[(i64)k * lda];
        for (int j = k + 1; j <= jhi; ++j) v = v + a[(i64)j * lda] * x[j];
        v = v * tau[k];
        a[(i64)k * lda] = a[(i64)k * lda] - v;
        for (int j = k + 1; j <= jhi; ++j) a[(i64)j * lda] = a[(i64)j * lda] - v * x[j];
    }
}

template <typename T>
int gbqrmr_impl(char trans, i64 mA, i64 mQ, i64 nQ, i64 l, i64 u, const T *D, i64 ldd, const T *tau, T *A,
                i64 lda) {
    BH_ARG(trans == 'N' || trans == 'n' || trans == 'T' || trans == 't' || trans == 'C' || trans == 'c', 1);
    BH_ARG(mA >= 0, 2);
    BH_ARG(mQ >= 0, 3);
    BH_ARG(nQ >= 0, 4);
    BH_ARG(l >= 0, 5);
    BH_ARG(u >= 0, 6);
    BH_ARG(ldd >= l + u + 1, 8);
    BH_ARG(lda >= imax(1, mA), 11);
    BH_ARG(mA < (1ll << 31) && mQ < (1ll << 31), 2);
    const i64 kmax = imin(mQ, nQ);
    if (mA == 0 || kmax == 0) return BH_OK;
    const bool tr = !(trans == 'N' || trans == 'n');
    hipLaunchKernelGGL(gbqr_rmul_rows<T>, dim3((unsigned)cdiv(mA, 256)), dim3(256), 0, stream(), tr, (int)mA,
                       (int)mQ, (int)kmax, (int)l, (int)u, D, ldd, tau, A, lda);
    BH_LAUNCH_CHECK();
    return BH_OK;
}

template <typename T>
int gbqrf_impl(i64 m, i64 n, i64 l, i64 u, T *D, i64 ldd, T *tau) {
    BH_ARG(m >= 0, 1);
    BH_ARG(n >= 0, 2);
    BH_ARG(l >= 0, 3);
    BH_ARG(u >= 0, 4);
    BH_ARG(ldd >= l + u + 1, 6);
    BH_ARG(m < (1ll << 31) && n < (1ll << 31), 1);
    if (m == 0 || n == 0) return BH_OK;
    hipStream_t st = stream();
    const int mi = (int)m, ni = (int)n, li = (int)l, ui = (int)u;
#define BH_QR_CASE(RLv)                                                                      \
    if (l + 1 <= RLv && u + 1 <= 64) {                                                       \
        hipLaunchKernelGGL((gbqrf_wave<T, RLv>), dim3(1), dim3(64), 0, st, mi, ni, li, ui, D, ldd, tau); \
        BH_LAUNCH_CHECK();                                                                   \
        return BH_OK;                                                                        \
    }
    BH_QR_CASE(8)
    BH_QR_CASE(16)
    BH_QR_CASE(32)
#undef BH_QR_CASE
#define BH_QR_WIDE(RRv)                                                                       \
    if (l + 1 <= 64 * RRv) {                                                                  \
        hipLaunchKernelGGL((gbqrf_wide<T, RRv>), dim3(1), dim3(1024), 0, st, mi, ni, li, ui, D, ldd, tau); \
        BH_LAUNCH_CHECK();                                                                    \
        return BH_OK;                                                                         \
    }
    BH_QR_WIDE(1)
    BH_QR_WIDE(2)
    BH_QR_WIDE(4)
    BH_QR_WIDE(8)
#undef BH_QR_WIDE
    hipLaunchKernelGGL(gbqrf_global<T>, dim3(1), dim3(256), 0, st, mi, ni, li, ui, D, ldd, tau);
    BH_LAUNCH_CHECK();
    return BH_OK;
}

template <typename T>
int gbqrm_impl(char trans, i64 m, i64 n, i64 l, i64 u, const T *D, i64 ldd, const T *tau, i64 nrhs, T *B, i64 ldb) {
    BH_ARG(trans == 'N' || trans == 'n' || trans == 'T' || trans == 't' || trans == 'C' || trans == 'c', 1);
    BH_ARG(m >= 0, 2);
    BH_ARG(n >= 0, 3);
    BH_ARG(l >= 0, 4);
    BH_ARG(u >= 0, 5);
    BH_ARG(ldd >= l + u + 1, 7);
    BH_ARG(nrhs >= 0, 9);
    BH_ARG(ldb >= imax(1, m), 11);
    BH_ARG(m < (1ll << 31) && n < (1ll << 31), 2);
    if (m == 0 || n == 0 || nrhs == 0) return BH_OK;
    const bool tr = !(trans == 'N' || trans == 'n');
    const int kmax = (int)imin(m, n);
    hipStream_t st = stream();
    if (l + 1 <= 64) {
        hipLaunchKernelGGL(gbqr_apply_wave<T>, dim3((unsigned)nrhs), dim3(64), 0, st, tr, (int)m, (int)n, (int)l,
                           (int)u, D, ldd, tau, kmax, B, ldb);
    } else if (m <= kWideApplyRows && l + 1 <= 64 * 16) {
        const size_t lds = (size_t)m * sizeof(T);
        const int rr = (int)cdiv(l + 1, 64);
        auto launch = [&](auto kern) -> int {
            if (lds > 64 * 1024)
                BH_HIP(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
            hipLaunchKernelGGL(kern, dim3((unsigned)nrhs), dim3(64), lds, st, tr, (int)m, (int)l, (int)u, D, ldd, tau,
                               kmax, B, ldb);
            return BH_OK;
        };
        int ls;
        if (rr <= 2)
            ls = launch(gbqr_apply_wide<T, 2>);
        else if (rr <= 4)
            ls = launch(gbqr_apply_wide<T, 4>);
        else if (rr <= 8)
            ls = launch(gbqr_apply_wide<T, 8>);
        else
            ls = launch(gbqr_apply_wide<T, 16>);
        if (ls) return ls;
    } else {
        hipLaunchKernelGGL(gbqr_apply_serial<T>, dim3((unsigned)cdiv(nrhs, 64)), dim3(64), 0, st, tr, (int)m, (int)n,
                           (int)l, (int)u, D, ldd, tau, kmax, (int)nrhs, B, ldb);
    }
    BH_LAUNCH_CHECK();
    return BH_OK;
}

static int qr_tbsv_upper(i64 n, i64 k, const double *R, i64 ldr, i64 nrhs, double *B, i64 ldb) {
    return bh_dtbsv_multi('U', 'N', 'N', n, k, R, ldr, nrhs, B, ldb);
}
static int qr_tbsv_upper(i64 n, i64 k, const float *R, i64 ldr, i64 nrhs, float *B, i64 ldb) {
    return bh_stbsv_multi('U', 'N', 'N', n, k, R, ldr, nrhs, B, ldb);
}

template <typename T>
int gbqrs_impl(i64 m, i64 n, i64 l, i64 u, const T *D, i64 ldd, const T *tau, i64 nrhs, T *B, i64 ldb) {
    BH_ARG(m >= n, 1);
    BH_ARG(n >= 0, 2);
    BH_ARG(l >= 0, 3);
    BH_ARG(u >= 0, 4);
    BH_ARG(ldd >= l + u + 1, 6);
    BH_ARG(nrhs >= 0, 8);
    BH_ARG(ldb >= imax(1, m), 10);
    if (n == 0 || nrhs == 0) return BH_OK;
    int st = gbqrm_impl<T>('T', m, n, l, u, D, ldd, tau, nrhs, B, ldb);
    if (st != BH_OK) return st;
    // R \ (Q'B)[1:n, :]: back substitution with u super-diagonals, no L part;
    // bands beyond the lane-ring solve's reach take the triangular band solve
    // (the tbsv of the reference's ldiv!, src/banded/tribanded.jl:75-83)
    st = wave_gbtrs_launch<T>(true, n, 0, u, nrhs, D, ldd, nullptr, B, ldb, stream());
    if (st == -1) st = qr_tbsv_upper(n, u, D, ldd, nrhs, B, ldb);
    return st;
}

}  // namespace
}  // namespace bh

using namespace bh;

extern "C" {

int bh_dgbqrf(bh_int m, bh_int n, bh_int l, bh_int u, double *D, bh_int ldd, double *tau) {
    return gbqrf_impl<double>(m, n, l, u, D, ldd, tau);
}
int bh_sgbqrf(bh_int m, bh_int n, bh_int l, bh_int u, float *D, bh_int ldd, float *tau) {
    return gbqrf_impl<float>(m, n, l, u, D, ldd, tau);
}
int bh_dgbqrm(char trans, bh_int m, bh_int n, bh_int l, bh_int u, const double *D, bh_int ldd, const double *tau,
              bh_int nrhs, double *B, bh_int ldb) {
    return gbqrm_impl<double>(trans, m, n, l, u, D, ldd, tau, nrhs, B, ldb);
}
int bh_sgbqrm(char trans, bh_int m, bh_int n, bh_int l, bh_int u, const float *D, bh_int ldd, const float *tau,
              bh_int nrhs, float *B, bh_int ldb) {
    return gbqrm_impl<float>(trans, m, n, l, u, D, ldd, tau, nrhs, B, ldb);
}
int bh_dgbqrmr(char trans, bh_int mA, bh_int mQ, bh_int nQ, bh_int l, bh_int u, const double *D, bh_int ldd,
               const double *tau, double *A, bh_int lda) {
    return gbqrmr_impl<double>(trans, mA, mQ, nQ, l, u, D, ldd, tau, A, lda);
}
int bh_sgbqrmr(char trans, bh_int mA, bh_int mQ, bh_int nQ, bh_int l, bh_int u, const float *D, bh_int ldd,
               const float *tau, float *A, bh_int lda) {
    return gbqrmr_impl<float>(trans, mA, mQ, nQ, l, u, D, ldd, tau, A, lda);
}
int bh_dgbqrs(bh_int m, bh_int n, bh_int l, bh_int u, const double *D, bh_int ldd, const double *tau, bh_int nrhs,
              double *B, bh_int ldb) {
    return gbqrs_impl<double>(m, n, l, u, D, ldd, tau, nrhs, B, ldb);
}
int bh_sgbqrs(bh_int m, bh_int n, bh_int l, bh_int u, const float *D, bh_int ldd, const float *tau, bh_int nrhs,
              float *B, bh_int ldb) {
    return gbqrs_impl<float>(m, n, l, u, D, ldd, tau, nrhs, B, ldb);
}

}  // extern "C"
