// bh_tbsv.hip -- triangular band solve and product, op(A) x = b and
// x <- op(A) x (BLAS xTBSV / xTBMV), on gfx950.
//
// Replaces tbsv! and tbmv! (src/blas.jl:71-147), which triangular \ and * on
// banded data reach (src/tribanded.jl:47-54,75-96).  Storage as BLAS: upper
// A(i,j) at a[k + i - j + j*lda] (i = j-k..j, diagonal in row k), lower A(i,j)
// at a[i - j + j*lda] (i = j..j+k, diagonal in row 0).
//
// Solve:
//   'N' upper: j descending, x_j = b_j / a_jj, then b_i = fma(a_ij, -x_j, b_i), i = j-k..j-1;
//   'N' lower: j ascending,  the same with i = j+1..j+k;
//   'T' upper: j ascending,  x_j = (b_j - sum_i a_ij x_i) / a_jj over i = j-k..j-1;
//   'T' lower: j descending, the same over i = j+1..j+k.
// With diag = 'U' the division is skipped.  The 'N' sweeps keep the rounding
// of the column-ordered CPU routine (division, then one fma per entry), so
// they are bit-identical to it; the 'T' dots are summed across the wave with
// DPP and agree to rounding.
//
// Product: y_i = d_i + sum_{t=1..} a_(t) x_(t) with d_i = a_ii x_i (x_i for a
// unit diagonal) and the off-diagonal terms fused in order of their distance
// t from the diagonal.
#include "bh_common.hpp"

namespace bh {

namespace {

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
constexpr i64 kTbMaxBytes = (1ll << 31) - (1ll << 24);

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
                                                int unit) {
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
    // step s of a block uses slot s mod PFD: q (ascending) or PFD-1-q (descending)
    auto slot_of = [](auto qc) { return M::ASC ? decltype(qc)::value : PFD - 1 - decltype(qc)::value; };
    const int g0 = M::ASC ? 0 : G - 1;
#pragma unroll
    for (int r = 0; r < NW; ++r) W[r] = tb_ld(rx, xoff(64 * (M::FUTURE ? g0 + r : g0 - r) + lane), T());
    TUnroll<PFD>::run([&](auto qc) {
        const int s = M::ASC ? decltype(qc)::value : 63 - decltype(qc)::value;
        fetch(slot_of(qc), 64 * g0 + s, s, lane);
    });
    // The ring moves with the sweep: when the window's far end leads
    // (FUTURE ascending, past descending) the next block enters at W[NW-1],
    // otherwise (the dot sweeps) it enters at W[0] and the oldest drops off.
    constexpr bool LEAD = (M::FUTURE == M::ASC);
    for (int g = g0; M::ASC ? g < G : g >= 0; g += M::ASC ? 1 : -1) {
        const int j0 = 64 * g;
        const int gin = LEAD ? (M::ASC ? g + NW : g - NW) : (M::ASC ? g + 1 : g - 1);
        const T wnext = tb_ld(rx, xoff(64 * gin + lane), T());
        for (int h = 0; h < 64; h += PFD) {
            int ln = lane;
            asm volatile("" : "+v"(ln));  // per-step lane offsets are recomputed, not hoisted into registers
            TUnroll<PFD>::run([&](auto qc) {
                constexpr int q = decltype(qc)::value;
                const int s = M::ASC ? h + q : 63 - h - q;  // uniform
                const int slot = slot_of(qc);
                const int j = j0 + s;
                if (j < n) {
                    const int tm = M::FUTURE ? tmin(k, n - 1 - j) : tmin(k, j);
                    const T bj = lane_read(W[0], s);
                    if constexpr (M::AXPY) {
                        const T xj = unit ? bj : bj / lane_read(fp[slot][0], s);
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
                        const T v = bj - wave_sum_dpp(part);
                        W[0] = lane_write(W[0], unit ? v : v / lane_read(fp[slot][0], s), s);
                    }
                }
                fetch(slot, M::ASC ? j + PFD : j - PFD, M::ASC ? (s + PFD) & 63 : (s - PFD) & 63, ln);
            });
        }
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
                                                       int incx) {
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
int tbsv_launch(int n, int k, const T *a, i64 lda, T *x, int incx, bool unit, hipStream_t st) {
    const int nw = 1 + (int)cdiv(imin(k, n - 1), 64);  // no step reaches further than n-1 entries
    constexpr int PFD = sizeof(T) == 4 ? 32 : 16;
#define BH_TB_CASE(NWv)                                                                                             \
    if (nw <= NWv) {                                                                                                \
        hipLaunchKernelGGL((tbsv_ring<T, MODE, NWv, PFD>), dim3(1), dim3(64), 0, st, n, k, a, lda, x, incx,     \
                           (int)unit);                                                                      \
        BH_LAUNCH_CHECK();                                                                                          \
        return BH_OK;                                                                                               \
    }
    BH_TB_CASE(2)
    BH_TB_CASE(3)
    BH_TB_CASE(5)
#undef BH_TB_CASE
    if (unit)
        hipLaunchKernelGGL((tbsv_block<T, MODE, true>), dim3(1), dim3(kTbBlock), 0, st, n, k, a, lda, x, incx);
    else
        hipLaunchKernelGGL((tbsv_block<T, MODE, false>), dim3(1), dim3(kTbBlock), 0, st, n, k, a, lda, x, incx);
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
    BH_HIP(hipMallocAsync((void **)&y, (size_t)n * sizeof(T), st));
    const int nb = (int)cdiv(n, 256);
    hipLaunchKernelGGL((tbmv_rows<T, MODE, UNIT>), dim3(nb), dim3(256), 0, st, n, k, a, lda, (const T *)x, incx, y);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) {
        hipLaunchKernelGGL((tb_scatter<T>), dim3(nb), dim3(256), 0, st, n, (const T *)y, x, incx);
        e = hipGetLastError();
    }
    const hipError_t ef = hipFreeAsync(y, st);
    if (e != hipSuccess) return hip_fail(e, "kernel launch", __FILE__, __LINE__);
    BH_HIP(ef);
    return BH_OK;
}

struct TbFlags {
    bool upper, trans, unit;
};

// BLAS argument order (uplo, trans, diag, n, k, A, lda, x, incx): a bad i-th
// argument returns -i.
template <typename T>
int tb_check(const char *fn, char uplo, char trans, char diag, i64 n, i64 k, i64 lda, i64 incx, TbFlags &f) {
    (void)fn;
    const bool up = (uplo == 'U' || uplo == 'u'), lo = (uplo == 'L' || uplo == 'l');
    const bool tn = (trans == 'N' || trans == 'n'), tt = (trans == 'T' || trans == 't' || trans == 'C' || trans == 'c');
    const bool unit = (diag == 'U' || diag == 'u'), nonunit = (diag == 'N' || diag == 'n');
    BH_ARG(up || lo, 1);
    BH_ARG(tn || tt, 2);
    BH_ARG(unit || nonunit, 3);
    BH_ARG(n >= 0, 4);
    BH_ARG(k >= 0, 5);
    BH_ARG(lda >= k + 1, 7);
    BH_ARG(incx >= 1, 9);  // positive strides only
    BH_ARG(n < (1ll << 30) && k < (1ll << 30) && lda * n * (i64)sizeof(T) < kTbMaxBytes, 4);
    BH_ARG(n == 0 || ((n - 1) * incx + 1) * (i64)sizeof(T) < kTbMaxBytes, 9);
    f = TbFlags{up, tt, unit};
    return BH_OK;
}

template <typename T>
int tbsv_impl(char uplo, char trans, char diag, i64 n, i64 k, const T *a, i64 lda, T *x, i64 incx) {
    TbFlags f;
    const int c = tb_check<T>("tbsv", uplo, trans, diag, n, k, lda, incx, f);
    if (c != BH_OK) return c;
    if (n == 0) return BH_OK;
    const int mode = f.trans ? (f.upper ? 2 : 3) : (f.upper ? 0 : 1);
    const int ni = (int)n, ki = (int)k, ix = (int)incx;
    hipStream_t st = stream();
#define BH_TB_MODE(Mv)                                                         \
    if (mode == Mv) return tbsv_launch<T, Mv>(ni, ki, a, lda, x, ix, f.unit, st);
    BH_TB_MODE(0)
    BH_TB_MODE(1)
    BH_TB_MODE(2)
    BH_TB_MODE(3)
#undef BH_TB_MODE
    return BH_ERR_HIP;
}

template <typename T>
int tbmv_impl(char uplo, char trans, char diag, i64 n, i64 k, const T *a, i64 lda, T *x, i64 incx) {
    TbFlags f;
    const int c = tb_check<T>("tbmv", uplo, trans, diag, n, k, lda, incx, f);
    if (c != BH_OK) return c;
    if (n == 0) return BH_OK;
    const int mode = f.trans ? (f.upper ? 2 : 3) : (f.upper ? 0 : 1);
    const int ni = (int)n, ki = (int)k, ix = (int)incx;
    hipStream_t st = stream();
#define BH_TB_MODE(Mv)                                                         \
    if (mode == Mv)                                                            \
        return f.unit ? tbmv_launch<T, Mv, true>(ni, ki, a, lda, x, ix, st)    \
                      : tbmv_launch<T, Mv, false>(ni, ki, a, lda, x, ix, st);
    BH_TB_MODE(0)
    BH_TB_MODE(1)
    BH_TB_MODE(2)
    BH_TB_MODE(3)
#undef BH_TB_MODE
    return BH_ERR_HIP;
}

}  // namespace
}  // namespace bh

using namespace bh;

extern "C" {

int bh_dtbsv(char uplo, char trans, char diag, bh_int n, bh_int k, const double *A, bh_int lda, double *x,
             bh_int incx) {
    return tbsv_impl<double>(uplo, trans, diag, n, k, A, lda, x, incx);
}
int bh_stbsv(char uplo, char trans, char diag, bh_int n, bh_int k, const float *A, bh_int lda, float *x,
             bh_int incx) {
    return tbsv_impl<float>(uplo, trans, diag, n, k, A, lda, x, incx);
}
int bh_dtbmv(char uplo, char trans, char diag, bh_int n, bh_int k, const double *A, bh_int lda, double *x,
             bh_int incx) {
    return tbmv_impl<double>(uplo, trans, diag, n, k, A, lda, x, incx);
}
int bh_stbmv(char uplo, char trans, char diag, bh_int n, bh_int k, const float *A, bh_int lda, float *x,
             bh_int incx) {
    return tbmv_impl<float>(uplo, trans, diag, n, k, A, lda, x, incx);
}

}  // extern "C"
