// bh_gbtrs_wave.hip -- band LU solves with one wave per right-hand side.
//
// Restates LAPACK xGBTRS (ldiv!(::BandedLU, B), src/banded/linalg.jl:24-47).
// Each wave owns one column of B and walks the sequential chain with the
// active part of the vector held in registers across the 64 lanes ("window"):
// window position q lives in lane q%64 of register q/64.  Per step the wave
// reads one factor column (contiguous in band storage, so one coalesced load
// per register), updates the window with one fma per position, retires one
// entry and shifts the window by one lane.  Factor columns are prefetched
// PF steps ahead through a register ring (the step loop is unrolled by PF).
//
// 'N' forward:  swap b_j <-> b_p (p = ipiv_j - 1), b_{j+t} = fma(l_{j+t,j}, -b_j, b_{j+t})
// 'N' backward: x_j = b_j / u_jj, b_{j-q} = fma(u_{j-q,j}, -x_j, b_{j-q})
//   -- the rounding of xGBTRS/xTBSV 'N', so results are bit-identical.
// 'T' forward:  x_j = (b_j - sum_q u_{j-q,j} x_{j-q}) / u_jj
// 'T' backward: b_j -= sum_t l_{j+t,j} b_{j+t}; swap b_j <-> b_p
//   -- dot products are wave reductions (summation order differs from BLAS).
#include "bh_common.hpp"

#include <cstdlib>
#include <cstring>

namespace bh {

constexpr int PF = 4;  // prefetch depth (steps)

// Dynamic but wave-uniform register index (scalar branches).
template <int N, typename T>
__device__ __forceinline__ T reg_at(const T (&w)[N], int k) {
    T r = w[0];
#pragma unroll
    for (int i = 1; i < N; ++i)
        if (k == i) r = w[i];
    return r;
}

// ---------------------------------------------------------------- 'N' solve
template <typename T, int NF, int NB>
__global__ void __launch_bounds__(64)
gbtrs_n_wave(i64 n, i64 kl, i64 ku, const T *__restrict__ ab, i64 ld, const i64 *__restrict__ ipiv, T *B, i64 ldb) {
    const int lane = threadIdx.x;
    T *b = B + (i64)blockIdx.x * ldb;
    const i64 kd = kl + ku;
    // ------------------------------------------------ forward: L and pivots
    if (kl > 0 && n > 1) {
        constexpr int WF = 64 * NF;
        T w[NF];
#pragma unroll
        for (int k = 0; k < NF; ++k) {
            const i64 q = lane + 64 * k;
            w[k] = q < n ? b[q] : T(0);
        }
        // prefetch ring: multipliers of column j for window positions (lane+64k)
        T lr[PF][NF];
        i64 pr[PF];
        T vr[PF];
        auto fetch = [&](int slot, i64 j) {
#pragma unroll
            for (int k = 0; k < NF; ++k) {
                const i64 q = lane + 64 * k;
                lr[slot][k] = (j < n - 1 && q >= 1 && q <= imin(kl, n - 1 - j)) ? ab[j * ld + kd + q] : T(0);
            }
            pr[slot] = (j < n - 1) ? ipiv[j] - 1 : j;
            // raw b entering the top of the window after step j (never written before)
            vr[slot] = (j + WF < n) ? b[j + WF] : T(0);
        };
#pragma unroll
        for (int s = 0; s < PF; ++s) fetch(s, s);
        T outv = T(0);
        for (i64 j0 = 0; j0 < n - 1; j0 += PF) {
#pragma unroll
            for (int s = 0; s < PF; ++s) {
                const i64 j = j0 + s;
                if (j < n - 1) {
                    const int d = (int)(pr[s] - j);
                    T bj = lane_first(w[0]);
                    if (d != 0) {
                        const int kd_ = __builtin_amdgcn_readfirstlane(d / 64);
                        const int ld_ = __builtin_amdgcn_readfirstlane(d % 64);
                        const T bp = lane_read(reg_at(w, kd_), ld_);
                        w[0] = lane_write(w[0], bp, 0);
#pragma unroll
                        for (int k = 0; k < NF; ++k)
                            if (k == kd_) w[k] = lane_write(w[k], bj, ld_);
                        bj = bp;
                    }
                    const T nb = -bj;
                    const i64 lm = imin(kl, n - 1 - j);
#pragma unroll
                    for (int k = 0; k < NF; ++k) {
                        const i64 q = lane + 64 * k;
                        if (q >= 1 && q <= lm) w[k] = dfma(lr[s][k], nb, w[k]);
                    }
                    // retire position 0 (final b_j)
                    outv = lane_write(outv, bj, (int)(j & 63));
                    if ((j & 63) == 63) b[j - 63 + lane] = outv;
                    // shift the window down by one position; a raw b enters at the top
                    const T vin = vr[s];
#pragma unroll
                    for (int k = 0; k < NF; ++k) {
                        const T carry = (k + 1 < NF) ? lane_first(w[k + 1]) : vin;
                        w[k] = wave_down1(w[k], carry);
                    }
                    fetch(s, j + PF);
                }
            }
        }
        // flush: retired entries of the last partial group, then the window
        const i64 last = n - 1;  // first index not yet retired
        const i64 g0 = last & ~(i64)63;
        if (lane < (int)(last - g0)) b[g0 + lane] = outv;
        __builtin_amdgcn_s_waitcnt(0);
#pragma unroll
        for (int k = 0; k < NF; ++k) {
            const i64 q = last + lane + 64 * k;
            if (q < n) b[q] = w[k];
        }
        __threadfence_block();
    }
    // ------------------------------------------------ backward: U, kd super-diagonals
    {
        constexpr int WB = 64 * NB;
        T w[NB];
        const i64 top = n - 1;
#pragma unroll
        for (int k = 0; k < NB; ++k) {
            const i64 q = lane + 64 * k;  // position q <-> b[j - q]
            w[k] = (top - q >= 0) ? b[top - q] : T(0);
        }
        T ur[PF][NB];
        T dr[PF], vr[PF];
        auto fetch = [&](int slot, i64 j) {
#pragma unroll
            for (int k = 0; k < NB; ++k) {
                const i64 q = lane + 64 * k;
                ur[slot][k] = (j >= 0 && q >= 1 && q <= imin(kd, j)) ? ab[j * ld + kd - q] : T(0);
            }
            dr[slot] = (j >= 0) ? ab[j * ld + kd] : T(1);
            // b entering the top of the window after step j: final forward output
            vr[slot] = (j - WB >= 0) ? b[j - WB] : T(0);
        };
#pragma unroll
        for (int s = 0; s < PF; ++s) fetch(s, top - s);
        T outv = T(0);
        for (i64 j0 = top; j0 >= 0; j0 -= PF) {
#pragma unroll
            for (int s = 0; s < PF; ++s) {
                const i64 j = j0 - s;
                if (j >= 0) {
                    const T xj = lane_first(w[0]) / dr[s];
                    const T nx = -xj;
                    const i64 qm = imin(kd, j);
#pragma unroll
                    for (int k = 0; k < NB; ++k) {
                        const i64 q = lane + 64 * k;
                        if (q >= 1 && q <= qm) w[k] = dfma(ur[s][k], nx, w[k]);
                    }
                    const i64 r = top - j;  // retire count
                    outv = lane_write(outv, xj, (int)(r & 63));
                    if ((r & 63) == 63) b[j + 63 - lane] = outv;
                    const T vin = vr[s];
#pragma unroll
                    for (int k = 0; k < NB; ++k) {
                        const T carry = (k + 1 < NB) ? lane_first(w[k + 1]) : vin;
                        w[k] = wave_down1(w[k], carry);
                    }
                    fetch(s, j - PF);
                }
            }
        }
        const i64 cnt = n;  // all retired; flush the last partial group
        const i64 g0 = cnt & ~(i64)63;
        // entries retired with r in [g0, cnt) are outv lanes 0..cnt-g0-1, x index top - r
        if (lane < (int)(cnt - g0)) b[top - (g0 + lane)] = outv;
    }
}

// ---------------------------------------------------------------- 'T' solve
template <typename T, int NF, int NB>
__global__ void __launch_bounds__(64)
gbtrs_t_wave(i64 n, i64 kl, i64 ku, const T *__restrict__ ab, i64 ld, const i64 *__restrict__ ipiv, T *B, i64 ldb) {
    const int lane = threadIdx.x;
    T *b = B + (i64)blockIdx.x * ldb;
    const i64 kd = kl + ku;
    // ------------------------------------------------ U^T x = b (ascending)
    {
        T w[NB];  // position q' <-> x_{j-1-q'}
#pragma unroll
        for (int k = 0; k < NB; ++k) w[k] = T(0);
        T ur[PF][NB];
        T dr[PF], br[PF];
        auto fetch = [&](int slot, i64 j) {
#pragma unroll
            for (int k = 0; k < NB; ++k) {
                const i64 q = lane + 64 * k + 1;
                ur[slot][k] = (j < n && q <= imin(kd, j)) ? ab[j * ld + kd - q] : T(0);
            }
            dr[slot] = (j < n) ? ab[j * ld + kd] : T(1);
            br[slot] = (j < n) ? b[j] : T(0);
        };
#pragma unroll
        for (int s = 0; s < PF; ++s) fetch(s, s);
        for (i64 j0 = 0; j0 < n; j0 += PF) {
#pragma unroll
            for (int s = 0; s < PF; ++s) {
                const i64 j = j0 + s;
                if (j < n) {
                    T part = T(0);
#pragma unroll
                    for (int k = 0; k < NB; ++k) part = dfma(ur[s][k], w[k], part);
                    const T xj = (br[s] - wave_sum_dpp(part)) / dr[s];
                    if (lane == 0) b[j] = xj;
#pragma unroll
                    for (int k = NB - 1; k >= 0; --k) {
                        const T carry = (k > 0) ? lane_read(w[k - 1], 63) : xj;
                        w[k] = wave_up1(w[k], carry);
                    }
                    fetch(s, j + PF);
                }
            }
        }
        __threadfence_block();
    }
    // ------------------------------------------------ L^T, pivots (descending)
    if (kl > 0 && n > 1) {
        T w[NF];  // position t <-> b_{j+t}
        const i64 jstart = n - 1;
#pragma unroll
        for (int k = 0; k < NF; ++k) {
            const i64 t = lane + 64 * k;
            w[k] = (jstart + t < n) ? b[jstart + t] : T(0);
        }
        T lr[PF][NF];
        i64 pr[PF];
        T vr[PF];
        auto fetch = [&](int slot, i64 j) {
#pragma unroll
            for (int k = 0; k < NF; ++k) {
                const i64 t = lane + 64 * k;
                lr[slot][k] = (j >= 0 && t >= 1 && t <= imin(kl, n - 1 - j)) ? ab[j * ld + kd + t] : T(0);
            }
            pr[slot] = (j >= 0) ? ipiv[j] - 1 : 0;
            vr[slot] = (j >= 0) ? b[j] : T(0);  // U^T solution entering at position 0
        };
#pragma unroll
        for (int s = 0; s < PF; ++s) fetch(s, n - 2 - s);
        for (i64 j0 = n - 2; j0 >= 0; j0 -= PF) {
#pragma unroll
            for (int s = 0; s < PF; ++s) {
                const i64 j = j0 - s;
                if (j >= 0) {
                    // shift up: b_j enters at position 0, the top position retires
                    const T top = lane_read(w[NF - 1], 63);
                    const i64 tq = j + 1 + 64 * NF - 1;  // index leaving the window
                    if (lane == 0 && tq < n) b[tq] = top;
                    const T vin = vr[s];
#pragma unroll
                    for (int k = NF - 1; k >= 0; --k) {
                        const T carry = (k > 0) ? lane_read(w[k - 1], 63) : vin;
                        w[k] = wave_up1(w[k], carry);
                    }
                    T part = T(0);
#pragma unroll
                    for (int k = 0; k < NF; ++k) part = dfma(lr[s][k], w[k], part);
                    const T sum = wave_sum_dpp(part);
                    T bj = lane_first(w[0]) - sum;
                    const int d = (int)(pr[s] - j);
                    if (d != 0) {
                        const int kd_ = __builtin_amdgcn_readfirstlane(d / 64);
                        const int ld_ = __builtin_amdgcn_readfirstlane(d % 64);
                        const T bp = lane_read(reg_at(w, kd_), ld_);
#pragma unroll
                        for (int k = 0; k < NF; ++k)
                            if (k == kd_) w[k] = lane_write(w[k], bj, ld_);
                        bj = bp;
                    }
                    w[0] = lane_write(w[0], bj, 0);
                    fetch(s, j - PF);
                }
            }
        }
        __threadfence_block();
#pragma unroll
        for (int k = 0; k < NF; ++k) {
            const i64 t = lane + 64 * k;
            if (t < n) b[t] = w[k];
        }
    }
}

template <typename T>
int ring_gbtrs_launch(i64 n, i64 kl, i64 ku, i64 nrhs, const T *ab, i64 ld, const i64 *ipiv, T *B, i64 ldb,
                      hipStream_t st);
template <typename T>
int ring_gbtrs_t_launch(i64 n, i64 kl, i64 ku, i64 nrhs, const T *ab, i64 ld, const i64 *ipiv, T *B, i64 ldb,
                        hipStream_t st);

template <typename T>
int wave_gbtrs_launch(bool trans_n, i64 n, i64 kl, i64 ku, i64 nrhs, const T *ab, i64 ld, const i64 *ipiv, T *B,
                      i64 ldb, hipStream_t st) {
    // the lane-ring kernels unless BANDEDHIP_TRS_KERNEL=shift asks for this file's
    {
        const char *e = getenv("BANDEDHIP_TRS_KERNEL");
        if (!(e && !strcmp(e, "shift"))) {
            const int r = trans_n ? ring_gbtrs_launch<T>(n, kl, ku, nrhs, ab, ld, ipiv, B, ldb, st)
                                  : ring_gbtrs_t_launch<T>(n, kl, ku, nrhs, ab, ld, ipiv, B, ldb, st);
            if (r != -1) return r;
        }
    }
    const i64 kd = kl + ku;
    const int nf = (int)cdiv(kl + 1, 64), nb = (int)cdiv(kd + 1, 64);
    if (nf > 3 || nb > 5) return -1;
    dim3 grid((unsigned)nrhs), block(64);
#define BH_TRS_CASE(NFv, NBv)                                                                                   \
    if (nf <= NFv && nb <= NBv) {                                                                               \
        if (trans_n)                                                                                            \
            hipLaunchKernelGGL((gbtrs_n_wave<T, NFv, NBv>), grid, block, 0, st, n, kl, ku, ab, ld, ipiv, B, ldb); \
        else                                                                                                    \
            hipLaunchKernelGGL((gbtrs_t_wave<T, NFv, NBv>), grid, block, 0, st, n, kl, ku, ab, ld, ipiv, B, ldb); \
        BH_LAUNCH_CHECK();                                                                                      \
        return BH_OK;                                                                                           \
    }
    BH_TRS_CASE(1, 1)
    BH_TRS_CASE(1, 2)
    BH_TRS_CASE(2, 3)
    BH_TRS_CASE(3, 5)
#undef BH_TRS_CASE
    return -1;
}

template int wave_gbtrs_launch<double>(bool, i64, i64, i64, i64, const double *, i64, const i64 *, double *, i64,
                                       hipStream_t);
template int wave_gbtrs_launch<float>(bool, i64, i64, i64, i64, const float *, i64, const i64 *, float *, i64,
                                      hipStream_t);

}  // namespace bh
