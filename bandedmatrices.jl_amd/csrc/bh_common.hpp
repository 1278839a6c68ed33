// bh_common.hpp -- shared helpers for the banded HIP library (gfx950 only).
//
// Band storage everywhere is the BandedMatrices.jl layout
// (src/banded/BandedMatrix.jl:16-28): entry (i,j), 0-based, of a matrix with
// bandwidths (kl,ku) is at data[(ku + i - j) + j*ld].  LU factors use the
// LAPACK AB layout of the padded copy (BandedMatrix.jl:222-232): ld >= 2kl+ku+1,
// diagonal at row kl+ku.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>

#include "../../include/bandedhip.h"

namespace bh {

using i64 = int64_t;

// Library-wide state: one current stream per host thread is enough for the
// LAPACK-like calling convention (the caller serialises calls on a handle).
hipStream_t stream();
// Library-owned device memory: 32 control words (info, flags) and a scratch
// buffer of at least `bytes` (grown, never shrunk; shared by every call on
// the library stream).
int ctrl_words(i64 **p);
int scratch_buffer(size_t bytes, void **p);
// Looks up, with `device` current, the device address of the statically
// allocated scratch used by bh_dsgbsv_noalloc (once per device).
int noalloc_setup(int device);
// The blocked band solve (bh_gbtrs_mfma.hip): workspace bytes (0: not
// applicable to these bands), the once-per-factorisation pre-pass, and the
// forward + backward sweeps for nrhs columns of B.
template <typename T>
size_t trs_blocked_bytes(i64 n, i64 kl, i64 ku);
template <typename T>
int trs_blocked_prepare(i64 n, i64 kl, i64 ku, const T *ab, i64 ldab, const i64 *ipiv, void *work, hipStream_t st);
template <typename T>
int trs_blocked_solve(i64 n, i64 kl, i64 ku, i64 nrhs, const T *ab, i64 ldab, const void *work, T *B, i64 ldb,
                      hipStream_t st);
void set_error(const char *fmt, ...);
int hip_fail(hipError_t e, const char *what, const char *file, int line);

#define BH_HIP(call)                                                        \
    do {                                                                    \
        hipError_t e_ = (call);                                             \
        if (e_ != hipSuccess) return bh::hip_fail(e_, #call, __FILE__, __LINE__); \
    } while (0)

#define BH_LAUNCH_CHECK()                                                   \
    do {                                                                    \
        hipError_t e_ = hipGetLastError();                                  \
        if (e_ != hipSuccess) return bh::hip_fail(e_, "kernel launch", __FILE__, __LINE__); \
    } while (0)

// Argument checks in LAPACK style: a bad i-th argument returns -i.
#define BH_ARG(cond, idx)                                                   \
    do {                                                                    \
        if (!(cond)) {                                                      \
            bh::set_error("%s: argument %d is invalid (%s)", __func__, (int)(idx), #cond); \
            return -(idx);                                                  \
        }                                                                   \
    } while (0)

__host__ __device__ __forceinline__ i64 imin(i64 a, i64 b) { return a < b ? a : b; }
__host__ __device__ __forceinline__ i64 imax(i64 a, i64 b) { return a > b ? a : b; }
__host__ __device__ __forceinline__ i64 cdiv(i64 a, i64 b) { return (a + b - 1) / b; }

// Fused multiply-add with the precision of T (no contraction games: all
// device code is compiled with -ffp-contract=off so only these fuse).
__device__ __forceinline__ double dfma(double a, double b, double c) { return __builtin_fma(a, b, c); }
__device__ __forceinline__ float dfma(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

// Workgroup barrier that orders LDS only: it waits for this wave's LDS
// traffic, not for outstanding global loads/stores (which __syncthreads()
// would, stalling every step on HBM latency).  Use it where no thread reads
// global data another thread of the workgroup wrote.
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// ---- cross-lane moves without LDS (DPP / readlane), for float and double.
// DPP controls used (lane layout checked on gfx950):
//   0x130 wave_shl1: lane i <- lane i+1      0x138 wave_shr1: lane i <- lane i-1
//   0x111.. row_shr:k: lane i <- lane i-k within a 16-lane row
//   0x142 row_bcast15 / 0x143 row_bcast31: last lane of a row to the next row(s)
template <int CTRL, int ROWMASK = 0xF>
__device__ __forceinline__ float dpp_mov(float old, float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(__builtin_bit_cast(int, old),
                                                                 __builtin_bit_cast(int, v), CTRL, ROWMASK, 0xF, false));
}
template <int CTRL, int ROWMASK = 0xF>
__device__ __forceinline__ double dpp_mov(double old, double v) {
    const long long o = __builtin_bit_cast(long long, old), x = __builtin_bit_cast(long long, v);
    const int lo = __builtin_amdgcn_update_dpp((int)o, (int)x, CTRL, ROWMASK, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp((int)(o >> 32), (int)(x >> 32), CTRL, ROWMASK, 0xF, false);
    return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned int)lo);
}
__device__ __forceinline__ float lane_read(float v, int l) {
    return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), l));
}
__device__ __forceinline__ double lane_read(double v, int l) {
    const long long x = __builtin_bit_cast(long long, v);
    const int lo = __builtin_amdgcn_readlane((int)x, l), hi = __builtin_amdgcn_readlane((int)(x >> 32), l);
    return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned int)lo);
}
__device__ __forceinline__ float lane_first(float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, v)));
}
__device__ __forceinline__ double lane_first(double v) {
    const long long x = __builtin_bit_cast(long long, v);
    const int lo = __builtin_amdgcn_readfirstlane((int)x), hi = __builtin_amdgcn_readfirstlane((int)(x >> 32));
    return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned int)lo);
}
// v with lane l replaced by the (uniform) value x
template <typename T>
__device__ __forceinline__ T lane_write(T v, T x, int l) {
    return ((int)__lane_id() == l) ? x : v;
}
// lane i <- lane i+1, lane 63 <- carry (uniform)
template <typename T>
__device__ __forceinline__ T wave_down1(T v, T carry) {
    return lane_write(dpp_mov<0x130>(v, v), carry, 63);
}
// lane i <- lane i-1, lane 0 <- carry (uniform)
template <typename T>
__device__ __forceinline__ T wave_up1(T v, T carry) {
    return lane_write(dpp_mov<0x138>(v, v), carry, 0);
}
// sum over the 64 lanes (result uniform); DPP partial sums, lane 63 holds the total
template <typename T>
__device__ __forceinline__ T wave_sum_dpp(T x) {
    x += dpp_mov<0x111>(T(0), x);
    x += dpp_mov<0x112>(T(0), x);
    x += dpp_mov<0x114>(T(0), x);
    x += dpp_mov<0x118>(T(0), x);
    x += dpp_mov<0x142, 0xA>(T(0), x);
    x += dpp_mov<0x143, 0xC>(T(0), x);
    return lane_read(x, 63);
}

__device__ __forceinline__ double dabs(double x) { return __builtin_fabs(x); }
__device__ __forceinline__ float dabs(float x) { return __builtin_fabsf(x); }

}  // namespace bh
