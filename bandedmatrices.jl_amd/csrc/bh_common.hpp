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

__device__ __forceinline__ double dabs(double x) { return __builtin_fabs(x); }
__device__ __forceinline__ float dabs(float x) { return __builtin_fabsf(x); }

}  // namespace bh
