// bh_lu_lane.hpp -- narrow-band LU (and fused solve) with one wave per
// system and the subdiagonal count fixed at compile time.
//
// Arithmetic of LAPACK xGBTF2 (LAPACK.gbtrf!, src/banded/BandedLU.jl:98) and
// xGBTRS 'N' (ldiv!, src/banded/linalg.jl:24-30), bit for bit: IDAMAX pivot,
// interchange over the live columns, multipliers a * (1/pivot), updates
// fma(l, -u, a); forward solve fma(l, -b_j, b_i) after the interchange, back
// substitution x_j = b_j / u_jj, fma(u_ij, -x_j, b_i).
//
// Layout: lane c mod 64 holds band column c of the window (columns
// j..j+kl+ku) with its rows j..j+kl in registers; the lane of the retired
// column j takes, when the window reaches it, column j+64.
//
// The kernels are instantiated per element type in bh_lu_lane_d.hip and
// bh_lu_lane_s.hip (define BH_LANE_T before including this header); other
// files see only the descriptor and the launchers.
#pragma once
#include "bh_common.hpp"

#include <type_traits>

namespace bh {

// One system of a batch, in schedule order.
struct BatchDesc {
    i64 ab_off;   // element offset of the system's padded AB
    i64 vec_off;  // element offset of its b / ipiv
    int n, kl, ku, cls, sys;
};

// Single-system gbtrf through the rotating-register kernel (kl <= 8 or
// kl = 16, kl + ku <= 63, band storage under 2 GiB), or -1.
int lane_gbtrf_launch(i64 m, i64 n, i64 kl, i64 ku, double *ab, i64 ldab, i64 *ipiv, i64 *dinfo, hipStream_t st);
int lane_gbtrf_launch(i64 m, i64 n, i64 kl, i64 ku, float *ab, i64 ldab, i64 *ipiv, i64 *dinfo, hipStream_t st);
// The persistent batched gbsv kernel over a schedule of nsys descriptors.
int lane_batch_launch(unsigned blocks, unsigned threads, hipStream_t st, const BatchDesc *desc, int nsys, double *AB,
                      double *B, i64 *ipiv, i64 *info);
int lane_batch_launch(unsigned blocks, unsigned threads, hipStream_t st, const BatchDesc *desc, int nsys, float *AB,
                      float *B, i64 *ipiv, i64 *info);

#ifdef BH_LANE_T
namespace {

template <int K>
struct SK {
    static constexpr int value = K;
};
template <int N>
struct SUnroll {
    template <class F>
    __device__ __forceinline__ static void run(F &&f) {
        SUnroll<N - 1>::run(f);
        f(SK<N - 1>{});
    }
};
template <>
struct SUnroll<0> {
    template <class F>
    __device__ __forceinline__ static void run(F &&) {}
};

__device__ __forceinline__ int smin(int a, int b) { return a < b ? a : b; }
__device__ __forceinline__ int smax(int a, int b) { return a > b ? a : b; }

// ---------------------------------------------------------------- rotating-register form
// The same arithmetic as small_gbsv_wave, restructured so that a step is a
// short straight-line sequence:
//  * KL (the subdiagonal count) is a compile-time constant and the window
//    has exactly R = KL + 1 rows; the steps are unrolled in blocks of U (a
//    multiple of R), and row i of step s lives in register (i + s) mod R, so
//    rows never move between registers;
//  * pivot: every lane takes max |a| over its column's rows (v_max_f64,
//    which passes over NaN), the pivot column and its max are broadcast to
//    scalars, and the first row whose |a| equals the max is found with scalar
//    integer compares -- the reference IDAMAX choice (first largest; a NaN
//    is never larger, and a NaN first entry is never beaten, so row 0);
//  * multipliers are formed from the broadcast column, so every lane holds
//    them without a second broadcast;
//  * memory goes through buffer resources: a lane with nothing to load or
//    store passes an offset past the end (the hardware returns 0 / drops the
//    store), so no access sits in a branch and the memory counter waits stay
//    exact; the incoming row of each step is loaded U steps ahead.
constexpr int kBufWord3 = 0x00020000;  // word 3 of a raw buffer resource (gfx9 family)

__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const void *p, i64 bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), 0, (int)bytes, kBufWord3);
}
__device__ __forceinline__ double buf_ld(__amdgpu_buffer_rsrc_t r, int off, double) {
    typedef unsigned v2u __attribute__((ext_vector_type(2)));
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
}
__device__ __forceinline__ float buf_ld(__amdgpu_buffer_rsrc_t r, int off, float) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0));
}
__device__ __forceinline__ void buf_st(__amdgpu_buffer_rsrc_t r, int off, double v) {
    typedef unsigned v2u __attribute__((ext_vector_type(2)));
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(v2u, v), r, off, 0, 0);
}
__device__ __forceinline__ void buf_st(__amdgpu_buffer_rsrc_t r, int off, float v) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, off, 0, 0);
}
__device__ __forceinline__ void buf_st64(__amdgpu_buffer_rsrc_t r, int off, long long v) {
    typedef unsigned v2u __attribute__((ext_vector_type(2)));
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(v2u, v), r, off, 0, 0);
}
// |x| as an unsigned integer (order-preserving for non-NaN values)
__device__ __forceinline__ unsigned long long abs_bits(double x) {
    return __builtin_bit_cast(unsigned long long, x) & 0x7fffffffffffffffull;
}
__device__ __forceinline__ unsigned abs_bits(float x) { return __builtin_bit_cast(unsigned, x) & 0x7fffffffu; }
__device__ __forceinline__ unsigned long long inf_bits(double) { return 0x7ff0000000000000ull; }
// value of lane l in every lane, into a vector register (ds_bpermute)
__device__ __forceinline__ float lane_bcast(float v, int l) {
    return __builtin_bit_cast(float, __builtin_amdgcn_ds_bpermute(l << 2, __builtin_bit_cast(int, v)));
}
__device__ __forceinline__ double lane_bcast(double v, int l) {
    const long long x = __builtin_bit_cast(long long, v);
    const int lo = __builtin_amdgcn_ds_bpermute(l << 2, (int)x), hi = __builtin_amdgcn_ds_bpermute(l << 2, (int)(x >> 32));
    return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned int)lo);
}
__device__ __forceinline__ unsigned inf_bits(float) { return 0x7f800000u; }

template <typename T, int KL, bool SOLVE, int UMIN>
__device__ __forceinline__ int lane_gbsv_wave(int m, int n, int ku, T *__restrict__ ab, i64 ld, i64 *__restrict__ ipiv,
                              T *__restrict__ b) {
    constexpr int R = KL + 1;
    constexpr int U = R * ((UMIN - 1 + R) / R);  // steps per unrolled block (>= UMIN, a multiple of R)
    constexpr int SZ = (int)sizeof(T);
    const int lane = (int)__lane_id();
    const int kv = KL + ku;
    const int mn = smin(m, n);
    if (mn == 0) return 0;
    const __amdgpu_buffer_rsrc_t rab = buf_rsrc(ab, ld * (i64)n * SZ);
    const __amdgpu_buffer_rsrc_t rip = buf_rsrc(ipiv, (i64)mn * 8);
    const __amdgpu_buffer_rsrc_t rb = buf_rsrc(b, SOLVE ? (i64)n * SZ : 0);
    // byte offset of A(r, c) in AB, or -1 (past the end) when it is not stored there
    auto off_a = [&](int r, int c) -> int {
        const bool ok = r >= 0 && r < m && c >= 0 && c < n && c - r <= ku && r - c <= KL;
        return ok ? (int)(((i64)c * ld + kv + r - c) * SZ) : -1;
    };
    // the row entering at the end of step j: row j+1+KL of the lane's column then
    auto off_in = [&](int j) -> int {
        const int c = (j + 1) + ((lane - (j + 1)) & 63);
        return j < mn ? off_a(j + 1 + KL, c) : -1;
    };
    auto off_bin = [&](int j) -> int {
        const int e = j + 1 + KL;
        return (SOLVE && j < mn && e < n) ? e * SZ : -1;
    };

    T y[R], pre[U];
    T bw[R], bpre[U];
#pragma unroll
    for (int i = 0; i < R; ++i) y[i] = buf_ld(rab, off_a(i, lane), T());
    SUnroll<U>::run([&](auto sc) {
        constexpr int s = decltype(sc)::value;
        pre[s] = buf_ld(rab, off_in(s), T());
    });
    if (SOLVE) {
#pragma unroll
        for (int i = 0; i < R; ++i) bw[i] = buf_ld(rb, i < n ? i * SZ : -1, T());
        SUnroll<U>::run([&](auto sc) {
            constexpr int s = decltype(sc)::value;
            bpre[s] = buf_ld(rb, off_bin(s), T());
        });
    }
    // A lane takes column j+64 when column j retires; rows of column j it still
    // holds when j+64 enters the window must read as zero.  Outside the
    // window a lane's rows never move, so one is replaced by an incoming row
    // per step, all of them before j+64 enters unless 2*KL + ku > 63.
    const bool clear_retired = 2 * KL + ku > 63;
    int ju = 0, info = 0, j0 = 0;

    auto step = [&](auto sc, auto tc) {
        constexpr int s = decltype(sc)::value;
        constexpr bool TAIL = decltype(tc)::value;
        const int j = j0 + s;
        if (TAIL && j >= mn) return;
        const int own = j & 63;
        const int km = TAIL ? smin(KL, m - 1 - j) : KL;
        auto reg = [](int i) { return (i + s) % R; };
        // pivot: max |a| of the lane's column rows 0..km, then the first row of
        // column j whose |a| equals column j's max
        T mx = dabs(y[reg(0)]);
#pragma unroll
        for (int i = 1; i < R; ++i)
            if (!TAIL || i <= km) mx = __builtin_fmax(mx, dabs(y[reg(i)]));
        int jp = 0;
        T pv;
        T l[R];
        bool nz;
        if constexpr (R <= 3) {
            // narrow windows: column j broadcast to scalars, scalar compares
            T xs[R];
#pragma unroll
            for (int i = 0; i < R; ++i) xs[i] = lane_read(y[reg(i)], own);
            const auto mb = abs_bits(lane_read(mx, own));
#pragma unroll
            for (int i = R - 1; i >= 1; --i)
                if ((!TAIL || i <= km) && abs_bits(xs[i]) == mb) jp = i;
            // row 0 also when its |a| is NaN: nothing compares greater than it
            if (abs_bits(xs[0]) == mb || abs_bits(xs[0]) > inf_bits(T())) jp = 0;
            pv = xs[0];
#pragma unroll
            for (int i = 1; i < R; ++i)
                if (jp == i) pv = xs[i];
            nz = abs_bits(pv) != 0;  // NaN counts as nonzero, as in LAPACK's test
            const T rcp = T(1) / pv;
            // multipliers of the interchanged column: a * (1/pivot)
#pragma unroll
            for (int i = 1; i < R; ++i) {
                const T xp = (jp == i) ? xs[0] : xs[i];
                l[i] = nz ? xp * rcp : xp;
            }
        } else {
            // wide windows (too many values for scalar registers): every lane
            // finds the pivot row of its own column and forms its own
            // multipliers; lane own's are broadcast into vector registers
            const auto mbl = abs_bits(mx);
            int jpl = 0;
#pragma unroll
            for (int i = R - 1; i >= 1; --i)
                if ((!TAIL || i <= km) && abs_bits(y[reg(i)]) == mbl) jpl = i;
            if (abs_bits(y[reg(0)]) == mbl || abs_bits(y[reg(0)]) > inf_bits(T())) jpl = 0;
            jp = __builtin_amdgcn_readlane(jpl, own);
            T ysel = y[reg(0)];
#pragma unroll
            for (int i = 1; i < R; ++i)
                if (jp == i) ysel = y[reg(i)];
            pv = lane_read(ysel, own);
            nz = abs_bits(pv) != 0;
            const T rcp = T(1) / pv;
#pragma unroll
            for (int i = 1; i < R; ++i) {
                const T xp = (jp == i) ? y[reg(0)] : y[reg(i)];
                l[i] = lane_bcast(nz ? xp * rcp : xp, own);
            }
        }
        if (nz)
            ju = smax(ju, smin(j + jp + ku, n - 1));
        else if (info == 0)
            info = j + 1;
        // interchange rows 0 and jp over columns j..ju only: a lane outside
        // them keeps its rows in place, so the rows a retired column leaves
        // behind are overwritten by incoming rows within R steps
        const int c = j + ((lane - j) & 63);
        if (jp != 0 && c <= ju) {
#pragma unroll
            for (int i = 1; i < R; ++i)
                if (jp == i) {
                    const T t0 = y[reg(0)];
                    y[reg(0)] = y[reg(i)];
                    y[reg(i)] = t0;
                }
        }
        // rank-1 update of columns j+1..ju, rows 1..km
        if (nz && c > j && c <= ju) {
            const T nu = -y[reg(0)];
#pragma unroll
            for (int i = 1; i < R; ++i)
                if (!TAIL || i <= km) y[reg(i)] = dfma(l[i], nu, y[reg(i)]);
        }
        // U row j (lanes of columns j..j+kv; the pivot from lane own), then
        // column j's multipliers and the pivot index
        buf_st(rab, (c - j <= kv && c < n) ? (int)(((i64)c * ld + kv - (c - j)) * SZ) : -1, y[reg(0)]);
#pragma unroll
        for (int i = 1; i < R; ++i)
            if (!TAIL || i <= km) buf_st(rab, lane == 0 ? (int)(((i64)j * ld + kv + i) * SZ) : -1, l[i]);
        buf_st64(rip, lane == 0 ? j * 8 : -1, (long long)(j + jp + 1));
        if (SOLVE) {
            // forward substitution step j (xGBTRS: interchange, then the L column)
            if (jp != 0) {
#pragma unroll
                for (int i = 1; i < R; ++i)
                    if (jp == i) {
                        const T t0 = bw[reg(0)];
                        bw[reg(0)] = bw[reg(i)];
                        bw[reg(i)] = t0;
                    }
            }
            const T nb = -bw[reg(0)];
            const int kb = TAIL ? smin(KL, n - 1 - j) : KL;
#pragma unroll
            for (int i = 1; i < R; ++i)
                if (!TAIL || i <= kb) bw[reg(i)] = dfma(l[i], nb, bw[reg(i)]);
            buf_st(rb, lane == 0 ? j * SZ : -1, bw[reg(0)]);
            bw[reg(0)] = bpre[s];
            bpre[s] = buf_ld(rb, off_bin(j + U), T());
        }
        if (clear_retired && lane == own) {
#pragma unroll
            for (int i = 0; i < R; ++i) y[i] = T(0);
        }
        y[reg(0)] = pre[s];
        pre[s] = buf_ld(rab, off_in(j + U), T());
    };
    const int jmain = smin(mn, m - KL);  // steps before it have km == KL
    for (; j0 + U <= jmain; j0 += U)
        SUnroll<U>::run([&](auto sc) { step(sc, std::false_type{}); });
    for (; j0 < mn; j0 += U)
        SUnroll<U>::run([&](auto sc) { step(sc, std::true_type{}); });

    // Fill-in rows LAPACK zeroes that no U row covered (rows >= m).
    {
        const int c_lo = smax(ku + 1, m + ku), c_hi = smin(n, mn + kv);
        for (int c = c_lo + lane; c < c_hi; c += 64) {
            for (int q = smax(0, kv - c); q < KL; ++q)
                if (c - kv + q >= m) ab[(i64)c * ld + q] = T(0);
        }
    }
    if (SOLVE) {
        // back substitution: lane t holds b_{j-t}, t = 0 .. kl+ku
        T w = buf_ld(rb, (lane <= kv && n - 1 - lane >= 0) ? (n - 1 - lane) * SZ : -1, T());
        T ur[U], br[U];
        auto off_u = [&](int j) { return (j >= 0 && lane <= smin(kv, j)) ? (int)(((i64)j * ld + kv - lane) * SZ) : -1; };
        auto off_br = [&](int j) { return (j >= 0 && j - 1 - kv >= 0) ? (j - 1 - kv) * SZ : -1; };
        SUnroll<U>::run([&](auto sc) {
            constexpr int s = decltype(sc)::value;
            ur[s] = buf_ld(rab, off_u(n - 1 - s), T());
            br[s] = buf_ld(rb, off_br(n - 1 - s), T());
        });
        for (int jb = n - 1; jb >= 0; jb -= U) {
            SUnroll<U>::run([&](auto sc) {
                constexpr int s = decltype(sc)::value;
                const int j = jb - s;
                if (j >= 0) {
                    const T xj = lane_read(w, 0) / lane_read(ur[s], 0);
                    buf_st(rb, lane == 0 ? j * SZ : -1, xj);
                    if (lane >= 1 && lane <= smin(kv, j)) w = dfma(ur[s], -xj, w);
                    w = wave_down1(w, T(0));
                    if (lane == kv) w = br[s];
                    ur[s] = buf_ld(rab, off_u(j - U), T());
                    br[s] = buf_ld(rb, off_br(j - U), T());
                }
            });
        }
    }
    return info;
}

template <typename T, int KL>
__global__ void __launch_bounds__(64) gbtrf_lane_kernel(int m, int n, int ku, T *ab, i64 ld, i64 *ipiv, i64 *info) {
    const int r = lane_gbsv_wave<T, KL, false, 16>(m, n, ku, ab, ld, ipiv, nullptr);
    if (threadIdx.x == 0) *info = r;
}


// One body per subdiagonal count (inlined; the kernel's register budget is
// that of the widest).
// One outlined body per subdiagonal count, so the batched kernel is a switch
// over calls rather than nine inlined copies.
template <typename T, int KL>
__device__ __noinline__ int batch_sys(int n, int ku, T *ab, i64 ld, i64 *ip, T *b) {
    return lane_gbsv_wave<T, KL, true, 8>(n, n, ku, ab, ld, ip, b);
}

template <typename T>
__device__ __forceinline__ int batch_one(const BatchDesc &d, T *AB, T *Bv, i64 *IP) {
    const i64 ld = 2 * d.kl + d.ku + 1;
    T *ab = AB + d.ab_off;
    T *b = Bv + d.vec_off;
    i64 *ip = IP + d.vec_off;
    switch (d.kl) {
#define BH_BATCH_CASE(KLv) \
    case KLv:              \
        return batch_sys<T, KLv>(d.n, d.ku, ab, ld, ip, b);
        BH_BATCH_CASE(0)
        BH_BATCH_CASE(1)
        BH_BATCH_CASE(2)
        BH_BATCH_CASE(3)
        BH_BATCH_CASE(4)
        BH_BATCH_CASE(5)
        BH_BATCH_CASE(6)
        BH_BATCH_CASE(7)
        default:
            return batch_sys<T, 8>(d.n, d.ku, ab, ld, ip, b);
#undef BH_BATCH_CASE
    }
}

// Persistent: wave w of the W resident waves solves systems w, w + W,
// w + 2W, ... of the schedule (the plan orders the systems so that this
// visit order balances the work).  All loop bounds are wave-uniform
// (readfirstlane), so the step loops stay scalar loops.
template <typename T>
__global__ void __launch_bounds__(256) gbsv_batched_kernel(const BatchDesc *__restrict__ desc, int nsys, T *AB,
                                                           T *Bv, i64 *IP, i64 *info) {
    auto uni64 = [](i64 v) {
        const int lo = __builtin_amdgcn_readfirstlane((int)(v & 0xffffffff));
        const int hi = __builtin_amdgcn_readfirstlane((int)(v >> 32));
        return (i64)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
    };
    const int nw = (int)(gridDim.x * (blockDim.x >> 6));
    const int w0 = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)));
    for (int q = w0; q < nsys; q += nw) {
        const BatchDesc dv = desc[q];
        BatchDesc d;
        d.ab_off = uni64(dv.ab_off);
        d.vec_off = uni64(dv.vec_off);
        d.n = __builtin_amdgcn_readfirstlane(dv.n);
        d.kl = __builtin_amdgcn_readfirstlane(dv.kl);
        d.ku = __builtin_amdgcn_readfirstlane(dv.ku);
        d.cls = __builtin_amdgcn_readfirstlane(dv.cls);
        d.sys = __builtin_amdgcn_readfirstlane(dv.sys);
        const int r = batch_one<T>(d, AB, Bv, IP);
        if (d.sys >= 0 && __lane_id() == 0) info[d.sys] = r;
    }
}

}  // namespace

int lane_gbtrf_launch(i64 m, i64 n, i64 kl, i64 ku, BH_LANE_T *ab, i64 ldab, i64 *ipiv, i64 *dinfo, hipStream_t st) {
    using T = BH_LANE_T;
    if (kl + ku > 63 || m >= (1ll << 30) || n >= (1ll << 30) || ldab * n * (i64)sizeof(T) >= (1ll << 31)) return -1;
    const int mi = (int)m, ni = (int)n, ui = (int)ku;
    switch ((int)kl) {
#define BH_LANE_CASE(KLv)                                                                                       \
    case KLv:                                                                                                   \
        hipLaunchKernelGGL((gbtrf_lane_kernel<T, KLv>), dim3(1), dim3(64), 0, st, mi, ni, ui, ab, ldab, ipiv, \
                           dinfo);                                                                              \
        break;
        BH_LANE_CASE(0) BH_LANE_CASE(1) BH_LANE_CASE(2) BH_LANE_CASE(3) BH_LANE_CASE(4)
        BH_LANE_CASE(5) BH_LANE_CASE(6) BH_LANE_CASE(7) BH_LANE_CASE(8) BH_LANE_CASE(16)
#undef BH_LANE_CASE
        default:
            return -1;
    }
    BH_LAUNCH_CHECK();
    return BH_OK;
}

int lane_batch_launch(unsigned blocks, unsigned threads, hipStream_t st, const BatchDesc *desc, int nsys,
                      BH_LANE_T *AB, BH_LANE_T *B, i64 *ipiv, i64 *info) {
    hipLaunchKernelGGL(gbsv_batched_kernel<BH_LANE_T>, dim3(blocks), dim3(threads), 0, st, desc, nsys, AB, B, ipiv,
                       info);
    BH_LAUNCH_CHECK();
    return BH_OK;
}
#endif  // BH_LANE_T

}  // namespace bh
