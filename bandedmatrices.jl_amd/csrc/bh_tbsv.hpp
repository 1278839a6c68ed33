// bh_tbsv.hpp -- kernels of the triangular band solve and product; compiled
// once per element type (bh_tbsv_d.hip, bh_tbsv_s.hip) with BH_TB_T set, and
// included without it by bh_tbsv.hip for the declarations.
#pragma once
#include "bh_common.hpp"

namespace bh {

namespace tbk {

template <int K>
struct TC {
    static constexpr int value = K;
};
template <int N>
struct TUnroll {
    template <class F>
    __device__ __forceinline__ static void run(F &&f) {
        TUnroll<N - 1>::run(f);
        f(TC<N - 1>{});
    }
};
template <>
struct TUnroll<0> {
    template <class F>
    __device__ __forceinline__ static void run(F &&) {}
};

constexpr int kTbWord3 = 0x00020000;    // word 3 of a raw buffer resource (gfx9 family)
constexpr int kTbPastEnd = 0x7f800000;  // a byte offset past the end of any buffer taken here

__device__ __forceinline__ __amdgpu_buffer_rsrc_t tb_rsrc(const void *p, i64 bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), 0, (int)bytes, kTbWord3);
}
__device__ __forceinline__ double tb_ld(__amdgpu_buffer_rsrc_t r, int off, double) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
}
__device__ __forceinline__ float tb_ld(__amdgpu_buffer_rsrc_t r, int off, float) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0));
}
__device__ __forceinline__ void tb_st(__amdgpu_buffer_rsrc_t r, int off, double v) {
    typedef unsigned v2u __attribute__((ext_vector_type(2)));
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(v2u, v), r, off, 0, 0);
}
__device__ __forceinline__ void tb_st(__amdgpu_buffer_rsrc_t r, int off, float v) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, off, 0, 0);
}
__device__ __forceinline__ int tmin(int a, int b) { return a < b ? a : b; }

// MODE 0: 'N' upper, 1: 'N' lower, 2: 'T' upper, 3: 'T' lower.
template <int MODE>
struct TbMode {
    static constexpr bool ASC = (MODE == 1 || MODE == 2);     // steps j ascending
    static constexpr bool AXPY = (MODE <= 1);                 // column updates ('N') or dots ('T')
    static constexpr bool FUTURE = (MODE == 1 || MODE == 3);  // the entries a step touches lie after j
    static constexpr bool UPPER = (MODE == 0 || MODE == 2);   // diagonal in row k, entry t above it at row k-t
};

// One wave walks the chain; b lives in a ring of 64-lane registers (entry e
// in lane e mod 64 of the register of its 64-entry block), and the factor
// column of each step is fetched PFD steps ahead with buffer loads whose
// out-of-range lanes read 0 (the update itself is masked).
// NW = 1 + ceil(k/64) window registers; PFD divides 64.
template <typename T, int MODE, int NW, int PFD>
__global__ void __launch_bounds__(64) tbsv_ring(int n, int k, const T *__restrict__ a, i64 lda, T *x, int incx,
                                                i64 ldx, int unit) {
    x += (i64)blockIdx.x * ldx;  // one wave per right-hand side
    static_assert(64 % PFD == 0, "prefetch distance divides the block");
    using M = TbMode<MODE>;
    constexpr int SZ = (int)sizeof(T);
    const int lane = threadIdx.x;
    const int dr = M::UPPER ? k : 0;
    // the band holds rows 0..k of each column: the last column ends k+1 past its start
    const __amdgpu_buffer_rsrc_t ra = tb_rsrc(a, ((i64)(n - 1) * lda + k + 1) * SZ);
    const __amdgpu_buffer_rsrc_t rx = tb_rsrc(x, ((i64)(n - 1) * incx + 1) * SZ);
    auto xoff = [&](int e) { return (e >= 0 && e < n) ? e * incx * SZ : -1; };
    // W[r] holds block g + r (FUTURE) or g - r.  For register r, lane L and
    // step j = 64g + s, the entry is t = 64r + L - s (FUTURE) or s - L + 64r
    // rows away from the diagonal, stored at j*lda + dr - t (upper) or
    // j*lda + dr + t (lower).
    auto tof = [&](int r, int ln, int s) { return M::FUTURE ? 64 * r + ln - s : s - ln + 64 * r; };
    const int G = (n + 63) >> 6;
    T W[NW];
    T fp[PFD][NW];
    auto fetch = [&](int slot, int j, int s, int ln) {
        const int base = (j >= 0 && j < n) ? (int)(((i64)j * lda + dr) * SZ) : kTbPastEnd;
#pragma unroll
        for (int r = 0; r < NW; ++r) {
            const int t = tof(r, ln, s);
            fp[slot][r] = tb_ld(ra, base + (M::UPPER ? -t : t) * SZ, T());
        }
    };
    const int g0 = M::ASC ? 0 : G - 1;
#pragma unroll
    for (int r = 0; r < NW; ++r) W[r] = tb_ld(rx, xoff(64 * (M::FUTURE ? g0 + r : g0 - r) + lane), T());
    TUnroll<PFD>::run([&](auto qc) {
        constexpr int s = M::ASC ? decltype(qc)::value : 63 - decltype(qc)::value;
        fetch(s % PFD, 64 * g0 + s, s, lane);
    });
    // The ring moves with the sweep: when the window's far end leads
    // (FUTURE ascending, past descending) the next block enters at W[NW-1],
    // otherwise (the dot sweeps) it enters at W[0] and the oldest drops off.
    constexpr bool LEAD = (M::FUTURE == M::ASC);
    for (int g = g0; M::ASC ? g < G : g >= 0; g += M::ASC ? 1 : -1) {
        const int j0 = 64 * g;
        const int gin = LEAD ? (M::ASC ? g + NW : g - NW) : (M::ASC ? g + 1 : g - 1);
        const T wnext = tb_ld(rx, xoff(64 * gin + lane), T());
        int ln = lane;
        asm volatile("" : "+v"(ln));  // per-step lane offsets are recomputed, not hoisted into registers
        int unitv = unit;
        asm volatile("" : "+v"(unitv));  // held per lane, so selecting on it needs no branch
        // 64 steps unrolled with compile-time slots: the compiler then counts
        // the loads in flight exactly and waits only for the slot it reads.
        TUnroll<64>::run([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            constexpr int s = M::ASC ? q : 63 - q;
            constexpr int slot = s % PFD;
            const int j = j0 + s;
            // Steps past the end (j >= n, only in the first or last block) run
            // with an empty range: they write lane s of an entry that is never
            // stored or read.  No branch on the chain.
            const int tm = j < n ? (M::FUTURE ? tmin(k, n - 1 - j) : tmin(k, j)) : -1;
            const T bj = lane_read(W[0], s);
            // a unit diagonal divides by 1 (exact): a per-lane select, no branch
            const T djj = lane_read(unitv ? T(1) : fp[slot][0], s);
            if constexpr (M::AXPY) {
                const T xj = bj / djj;
                W[0] = lane_write(W[0], xj, s);
                const T nx = -xj;
#pragma unroll
                for (int r = 0; r < NW; ++r) {
                    const int t = tof(r, ln, s);
                    const T u = dfma(fp[slot][r], nx, W[r]);
                    W[r] = (t >= 1 && t <= tm) ? u : W[r];
                }
            } else {
                T part = T(0);
#pragma unroll
                for (int r = 0; r < NW; ++r) {
                    const int t = tof(r, ln, s);
                    const T qv = dfma(fp[slot][r], W[r], part);
                    part = (t >= 1 && t <= tm) ? qv : part;
                }
                W[0] = lane_write(W[0], (bj - wave_sum_dpp(part)) / djj, s);
            }
            fetch(slot, M::ASC ? j + PFD : j - PFD, M::ASC ? (s + PFD) & 63 : (s - PFD) & 63, ln);
        });
        // block g is final once the sweep leaves it
        tb_st(rx, xoff(j0 + lane), W[0]);
        if constexpr (LEAD) {
#pragma unroll
            for (int r = 0; r + 1 < NW; ++r) W[r] = W[r + 1];
            W[NW - 1] = wnext;
        } else {
#pragma unroll
            for (int r = NW - 1; r > 0; --r) W[r] = W[r - 1];
            W[0] = wnext;
        }
    }
}

// Any bandwidth: one workgroup of 256 threads, b in memory, two barriers per
// step.  Used above the ring's register window (k > 256).
constexpr int kTbBlock = 256;

template <typename T, int MODE, bool UNIT>
__global__ void __launch_bounds__(kTbBlock) tbsv_block(int n, int k, const T *__restrict__ a, i64 lda, T *x,
                                                       int incx, i64 ldx) {
    x += (i64)blockIdx.x * ldx;  // one workgroup per right-hand side
    using M = TbMode<MODE>;
    __shared__ T red[kTbBlock / 64];
    const int tid = threadIdx.x;
    const int dr = M::UPPER ? k : 0;
    for (int q = 0; q < n; ++q) {
        const int j = M::ASC ? q : n - 1 - q;
        const int tm = M::FUTURE ? tmin(k, n - 1 - j) : tmin(k, j);
        const T *col = a + (i64)j * lda + dr;
        if constexpr (M::AXPY) {
            T xj = x[(i64)j * incx];
            if (!UNIT) xj = xj / col[0];
            __syncthreads();  // every thread holds b_j before it is overwritten
            if (tid == 0) x[(i64)j * incx] = xj;
            const T nx = -xj;
            for (int t = 1 + tid; t <= tm; t += kTbBlock) {
                const int i = M::FUTURE ? j + t : j - t;
                T *xi = x + (i64)i * incx;
                *xi = dfma(col[M::UPPER ? -t : t], nx, *xi);
            }
            __syncthreads();
        } else {
            T part = T(0);
            for (int t = 1 + tid; t <= tm; t += kTbBlock) {
                const int i = M::FUTURE ? j + t : j - t;
                part = dfma(col[M::UPPER ? -t : t], x[(i64)i * incx], part);
            }
            part = wave_sum_dpp(part);
            if ((tid & 63) == 0) red[tid >> 6] = part;
            __syncthreads();
            if (tid == 0) {
                T s = red[0];
#pragma unroll
                for (int w = 1; w < kTbBlock / 64; ++w) s += red[w];
                const T v = x[(i64)j * incx] - s;
                x[(i64)j * incx] = UNIT ? v : v / col[0];
            }
            __syncthreads();
        }
    }
}

template <typename T, int MODE>
int tbsv_launch(int n, int k, const T *a, i64 lda, T *x, int incx, i64 ldx, int nrhs, bool unit, hipStream_t st) {
    const int nw = 1 + (int)cdiv(imin(k, n - 1), 64);  // no step reaches further than n-1 entries
    constexpr int PFD = sizeof(T) == 4 ? 32 : 16;
#define BH_TB_CASE(NWv)                                                                                             \
    if (nw <= NWv) {                                                                                                \
        hipLaunchKernelGGL((tbsv_ring<T, MODE, NWv, PFD>), dim3(nrhs), dim3(64), 0, st, n, k, a, lda, x, incx, \
                           ldx, (int)unit);                                                                 \
        BH_LAUNCH_CHECK();                                                                                          \
        return BH_OK;                                                                                               \
    }
    BH_TB_CASE(2)
    BH_TB_CASE(3)
    BH_TB_CASE(5)
#undef BH_TB_CASE
    if (unit)
        hipLaunchKernelGGL((tbsv_block<T, MODE, true>), dim3(nrhs), dim3(kTbBlock), 0, st, n, k, a, lda, x, incx, ldx);
    else
        hipLaunchKernelGGL((tbsv_block<T, MODE, false>), dim3(nrhs), dim3(kTbBlock), 0, st, n, k, a, lda, x, incx, ldx);
    BH_LAUNCH_CHECK();
    return BH_OK;
}

// ---- product: one thread per output entry into y, then y -> x.
template <typename T, int MODE, bool UNIT>
__global__ void __launch_bounds__(256) tbmv_rows(int n, int k, const T *__restrict__ a, i64 lda,
                                                 const T *__restrict__ x, int incx, T *__restrict__ y) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const int dr = TbMode<MODE>::UPPER ? k : 0;
    const T xi = x[(i64)i * incx];
    T acc = UNIT ? xi : a[(i64)i * lda + dr] * xi;
    if constexpr (MODE == 0) {  // row i of U: A(i, i+t) at (i+t)*lda + k - t
        const int tm = tmin(k, n - 1 - i);
        for (int t = 1; t <= tm; ++t) acc = dfma(a[(i64)(i + t) * lda + k - t], x[(i64)(i + t) * incx], acc);
    } else if constexpr (MODE == 1) {  // row i of L: A(i, i-t) at (i-t)*lda + t
        const int tm = tmin(k, i);
        for (int t = 1; t <= tm; ++t) acc = dfma(a[(i64)(i - t) * lda + t], x[(i64)(i - t) * incx], acc);
    } else if constexpr (MODE == 2) {  // column i of U: A(i-t, i) at i*lda + k - t
        const int tm = tmin(k, i);
        for (int t = 1; t <= tm; ++t) acc = dfma(a[(i64)i * lda + k - t], x[(i64)(i - t) * incx], acc);
    } else {  // column i of L: A(i+t, i) at i*lda + t
        const int tm = tmin(k, n - 1 - i);
        for (int t = 1; t <= tm; ++t) acc = dfma(a[(i64)i * lda + t], x[(i64)(i + t) * incx], acc);
    }
    y[i] = acc;
}

template <typename T>
__global__ void __launch_bounds__(256) tb_scatter(int n, const T *__restrict__ y, T *__restrict__ x, int incx) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) x[(i64)i * incx] = y[i];
}

template <typename T, int MODE, bool UNIT>
int tbmv_launch(int n, int k, const T *a, i64 lda, T *x, int incx, hipStream_t st) {
    T *y = nullptr;
    {
        const int cs = scratch_buffer((size_t)n * sizeof(T), (void **)&y);
        if (cs) return cs;
    }
    const int nb = (int)cdiv(n, 256);
    hipLaunchKernelGGL((tbmv_rows<T, MODE, UNIT>), dim3(nb), dim3(256), 0, st, n, k, a, lda, (const T *)x, incx, y);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) {
        hipLaunchKernelGGL((tb_scatter<T>), dim3(nb), dim3(256), 0, st, n, (const T *)y, x, incx);
        e = hipGetLastError();
    }
    if (e != hipSuccess) return hip_fail(e, "kernel launch", __FILE__, __LINE__);
    return BH_OK;
}

template <typename T>
int tbsv_dispatch(int mode, int n, int k, const T *a, i64 lda, T *x, int incx, i64 ldx, int nrhs, bool unit,
                  hipStream_t st) {
    switch (mode) {
        case 0: return tbsv_launch<T, 0>(n, k, a, lda, x, incx, ldx, nrhs, unit, st);
        case 1: return tbsv_launch<T, 1>(n, k, a, lda, x, incx, ldx, nrhs, unit, st);
        case 2: return tbsv_launch<T, 2>(n, k, a, lda, x, incx, ldx, nrhs, unit, st);
        default: return tbsv_launch<T, 3>(n, k, a, lda, x, incx, ldx, nrhs, unit, st);
    }
}

template <typename T>
int tbmv_dispatch(int mode, int n, int k, const T *a, i64 lda, T *x, int incx, bool unit, hipStream_t st) {
#define BH_TBMV_MODE(Mv)                                                                                   \
    if (mode == Mv)                                                                                        \
        return unit ? tbmv_launch<T, Mv, true>(n, k, a, lda, x, incx, st) : tbmv_launch<T, Mv, false>(n, k, a, lda, x, incx, st);
    BH_TBMV_MODE(0)
    BH_TBMV_MODE(1)
    BH_TBMV_MODE(2)
    BH_TBMV_MODE(3)
#undef BH_TBMV_MODE
    return BH_ERR_HIP;
}

}  // namespace tbk

// mode 0: 'N' upper, 1: 'N' lower, 2: 'T' upper, 3: 'T' lower
int tbsv_run(int mode, int n, int k, const double *a, i64 lda, double *x, int incx, i64 ldx, int nrhs, bool unit);
int tbsv_run(int mode, int n, int k, const float *a, i64 lda, float *x, int incx, i64 ldx, int nrhs, bool unit);
int tbmv_run(int mode, int n, int k, const double *a, i64 lda, double *x, int incx, bool unit);
int tbmv_run(int mode, int n, int k, const float *a, i64 lda, float *x, int incx, bool unit);

#ifdef BH_TB_T
int tbsv_run(int mode, int n, int k, const BH_TB_T *a, i64 lda, BH_TB_T *x, int incx, i64 ldx, int nrhs, bool unit) {
    return tbk::tbsv_dispatch<BH_TB_T>(mode, n, k, a, lda, x, incx, ldx, nrhs, unit, stream());
}
int tbmv_run(int mode, int n, int k, const BH_TB_T *a, i64 lda, BH_TB_T *x, int incx, bool unit) {
    return tbk::tbmv_dispatch<BH_TB_T>(mode, n, k, a, lda, x, incx, unit, stream());
}
#endif

}  // namespace bh
