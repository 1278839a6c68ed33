// bh_gbtrf_mfma.hip -- blocked partial-pivot band LU with the trailing
// update on the matrix cores (gfx950 MFMA), FP64 and FP32.
//
// This is the counterpart of LAPACK's blocked xGBTRF, the path the reference
// takes in lu! (LAPACK.gbtrf!, src/banded/BandedLU.jl:98) when ku > 64 and
// kl >= 32: ILAENV gives a block size of 32 there and xGBTRF leaves xGBTF2.
// The pivots are chosen exactly as xGBTF2/xGBTRF choose them (first row of
// the largest |a| among rows j..j+km of the current column, the column
// having received every earlier update); the factors are the same matrices
// up to rounding: the updates of a panel reach the trailing band as one
// rank-16 product on MFMA instead of 16 rank-1 steps, as xGBTRF's DGEMM
// does.  Output layout, ipiv and info follow LAPACK exactly.
//
// One workgroup factors one system.  Window of panel P (j0 = 16P): rows
// j0 .. j0+R-1 (R = kl + 16) -- every row the panel's pivots and
// multipliers can touch.  Three kinds of waves:
//   * the PANEL WAVE factors the 16 panel columns in registers, rows spread
//     three (PR) to a lane, in xGBTF2's arithmetic (reciprocal-scaled
//     multipliers, fma rank-1 updates in step order).  Row interchanges only
//     relabel rows ("slots" keep their data; each slot carries its current
//     position), so a step moves no data.  The multipliers go to AB at once
//     in xGBTF2's order; at the end of the panel the slot contents and the
//     final position -> slot map are published in LDS.
//   * the HELPER WAVE follows the panel wave one step behind and builds
//     L11^{-1} (16x16 unit lower) row by row from the pivot rows'
//     multipliers; it also writes the panel's U11 block to AB.
//   * ten UPDATE WAVES apply the panel to the trailing block columns (16
//     columns each) as xGBTRF does, on MFMA 16x16x4:
//         U12 = L11^{-1} (P A)[0:16]            (4 MFMA)
//         A22 = (P A)[16:R] - L21 U12           (4 MFMA per 16-row tile)
//     reading the rows of P A straight from AB through the slot map (a
//     gather: the permutation costs nothing) and writing U12 and A22 back
//     in final order.  A workgroup's waves go to the four SIMDs in turn, so
//     waves 4 and 8 share the panel wave's SIMD: they take look-ahead tiles
//     only, and the other eight update waves the block columns, block
//     column B always the one of rank B mod 8, so its loads see its own
//     earlier stores.
// Look-ahead: block column P+1 (the next panel) goes through LDS only.  In
// the phase where the panel wave factors panel P, the update waves apply
// panel P-1 to block columns P+1.. (block column P+1 into the LDS
// look-ahead buffer); after barrier B1 all update waves apply panel P to
// the look-ahead buffer, one 16-row tile each, into the next panel's input
// buffer; barrier B2 releases the panel wave onto panel P+1.  All
// communication between waves is through LDS (barriers order LDS only).
//
// Caveats (documented beside the entry point): on finite inputs the pivots
// equal LAPACK's unless two candidates differ only in rounding; the update
// sums in a different order than xGBTRF's DGEMM, so factors agree to
// rounding, not bit for bit.
#include "bh_common.hpp"
#include "bh_mfma.hpp"

namespace bh {

namespace {

constexpr int QB = 16;       // panel width = block-column width
constexpr int QUW = 10;      // update waves
constexpr int QP1 = 8;       // of which take block columns (the rest only look-ahead tiles)
constexpr int QWAVES = QUW + 2;
constexpr int QBIG = 1 << 29;  // position of a slot that holds no row

// LDS row stride (elements) of the [rows][16] buffers: 80 and 144 bytes
template <typename T>
struct QStride {
    static constexpr int value = sizeof(T) == 4 ? 20 : 18;
};

template <typename T>
struct QArgs {
    T *ab;
    i64 *ipiv;
    i64 *info;  // info[0]: LAPACK info; info[1]: nonzero if a wait gave up
    int m, n, kl, ku;
    i64 ldab;
    int R, RT, RP;  // window rows, 16-row tiles, padded rows (16*RT)
};

template <typename T, int PR>
struct QCol {
    typedef T type __attribute__((ext_vector_type(PR)));
};

template <int K>
struct QIntC {
    static constexpr int value = K;
};
template <int N>
struct QFor {
    template <class F>
    __device__ __forceinline__ static void run(F &&f) {
        QFor<N - 1>::run(f);
        f(QIntC<N - 1>{});
    }
};
template <>
struct QFor<0> {
    template <class F>
    __device__ __forceinline__ static void run(F &&) {}
};

__device__ __forceinline__ int qmin(int a, int b) { return a < b ? a : b; }
__device__ __forceinline__ int qmax(int a, int b) { return a > b ? a : b; }

// Pivot-search keys: larger means a better pivot; NaN ranks below every
// number; 0 marks "not a candidate".
__device__ __forceinline__ unsigned qkey(float v) {
    const unsigned b = __builtin_bit_cast(unsigned, v) & 0x7fffffffu;
    return b > 0x7f800000u ? 1u : b + 2u;
}
__device__ __forceinline__ unsigned long long qkey(double v) {
    const unsigned long long b = __builtin_bit_cast(unsigned long long, v) & 0x7fffffffffffffffull;
    return b > 0x7ff0000000000000ull ? 1ull : b + 2ull;
}
template <int CTRL, int ROWMASK = 0xF>
__device__ __forceinline__ unsigned qdpp0(unsigned v) {
    return (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, ROWMASK, 0xF, true);
}
__device__ __forceinline__ unsigned qumax(unsigned a, unsigned b) { return a > b ? a : b; }
// maximum over the wave (uniform), DPP with zero fill (0 is below every key)
__device__ __forceinline__ unsigned wave_max_key(unsigned x) {
    x = qumax(x, qdpp0<0x111>(x));
    x = qumax(x, qdpp0<0x112>(x));
    x = qumax(x, qdpp0<0x114>(x));
    x = qumax(x, qdpp0<0x118>(x));
    x = qumax(x, qdpp0<0x142, 0xA>(x));
    x = qumax(x, qdpp0<0x143, 0xC>(x));
    return (unsigned)__builtin_amdgcn_readlane((int)x, 63);
}
__device__ __forceinline__ unsigned long long wave_max_key(unsigned long long x) {
    auto step = [](unsigned long long v, unsigned long long o) { return v > o ? v : o; };
#define QDPP64(CTRL, RM)                                                                              \
    {                                                                                                 \
        const unsigned lo = qdpp0<CTRL, RM>((unsigned)x), hi = qdpp0<CTRL, RM>((unsigned)(x >> 32)); \
        x = step(x, ((unsigned long long)hi << 32) | lo);                                             \
    }
    QDPP64(0x111, 0xF)
    QDPP64(0x112, 0xF)
    QDPP64(0x114, 0xF)
    QDPP64(0x118, 0xF)
    QDPP64(0x142, 0xA)
    QDPP64(0x143, 0xC)
#undef QDPP64
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)x, 63);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(x >> 32), 63);
    return ((unsigned long long)hi << 32) | lo;
}
// High word of |v| as an unsigned integer, monotone in |v| over numbers and
// infinities; ~0u for a NaN.
__device__ __forceinline__ unsigned qhiword(float v) {
    const unsigned b = __builtin_bit_cast(unsigned, v) & 0x7fffffffu;
    return b > 0x7f800000u ? 0xffffffffu : b;
}
__device__ __forceinline__ unsigned qhiword(double v) {
    const unsigned long long b = __builtin_bit_cast(unsigned long long, v) & 0x7fffffffffffffffull;
    return b > 0x7ff0000000000000ull ? 0xffffffffu : (unsigned)(b >> 32);
}
// Reciprocal of a nonzero pivot: the hardware estimate refined by Newton
// steps (error within an ulp or so; the blocked factorisation promises
// agreement with LAPACK to rounding only).
__device__ __forceinline__ float qrcp(float x) {
    float r = __builtin_amdgcn_rcpf(x);
    const float e = __builtin_fmaf(-x, r, 1.0f);
    return __builtin_fmaf(r, e, r);
}
__device__ __forceinline__ double qrcp(double x) {
    double r = __builtin_amdgcn_rcp(x);
    double e = __builtin_fma(-x, r, 1.0);
    r = __builtin_fma(r, e, r);
    e = __builtin_fma(-x, r, 1.0);
    r = __builtin_fma(r, e, r);
    e = __builtin_fma(-x, r, 1.0);
    return __builtin_fma(r, e, r);
}
// maximum over the wave (uniform) of non-NaN values; lanes without a source
// keep their own value
__device__ __forceinline__ float wave_fmax(float x) {
    x = __builtin_fmaxf(x, dpp_mov<0x111>(x, x));
    x = __builtin_fmaxf(x, dpp_mov<0x112>(x, x));
    x = __builtin_fmaxf(x, dpp_mov<0x114>(x, x));
    x = __builtin_fmaxf(x, dpp_mov<0x118>(x, x));
    x = __builtin_fmaxf(x, dpp_mov<0x142, 0xA>(x, x));
    x = __builtin_fmaxf(x, dpp_mov<0x143, 0xC>(x, x));
    return lane_read(x, 63);
}
__device__ __forceinline__ double wave_fmax(double x) {
    x = __builtin_fmax(x, dpp_mov<0x111>(x, x));
    x = __builtin_fmax(x, dpp_mov<0x112>(x, x));
    x = __builtin_fmax(x, dpp_mov<0x114>(x, x));
    x = __builtin_fmax(x, dpp_mov<0x118>(x, x));
    x = __builtin_fmax(x, dpp_mov<0x142, 0xA>(x, x));
    x = __builtin_fmax(x, dpp_mov<0x143, 0xC>(x, x));
    return lane_read(x, 63);
}
// minimum of non-negative ints over the wave (uniform)
__device__ __forceinline__ int wave_min_int(int v) {
    // max of the complement
    unsigned x = 0x7fffffffu - (unsigned)v;
    x = wave_max_key(x);
    return (int)(0x7fffffffu - x);
}

// Pivot of one elimination step, searched over the panel wave's slots: slot r
// of a lane holds value col[r] at position pos[r]; the candidates are the
// positions k .. k+km.  The pivot is the largest |a| (NaN below every number,
// zero above NaN), the smallest position among equals -- xGBTF2's IxAMAX
// choice on finite data.  Branch-free: one wave maximum of the 32-bit keys
// (two for the 64-bit keys of FP64), then one of the tied candidates'
// complemented (position, slot, lane) labels.  Uniform results: the pivot's
// lane lp, slot rp and position p.
template <int PR>
__device__ __forceinline__ void qpivot(const float (&col)[PR], const int (&pos)[PR], int k, int km, int lane,
                                       int &lp, int &rp, int &p) {
    unsigned key[PR], kmax = 0;
    bool cand[PR];
#pragma unroll
    for (int r = 0; r < PR; ++r) {
        cand[r] = (unsigned)(pos[r] - k) <= (unsigned)km;
        key[r] = cand[r] ? qkey(col[r]) : 0u;
        kmax = qumax(kmax, key[r]);
    }
    const unsigned M = wave_max_key(kmax);
    unsigned sel = 0;
#pragma unroll
    for (int r = 0; r < PR; ++r) {
        const unsigned lab = ((unsigned)pos[r] << 8) | (unsigned)((r << 6) | lane);
        sel = qumax(sel, (cand[r] && key[r] == M) ? 0x10000u - lab : 0u);
    }
    const unsigned lab = 0x10000u - wave_max_key(sel);
    lp = (int)(lab & 63u);
    rp = (int)((lab >> 6) & 3u);
    p = (int)(lab >> 8);
}
template <int PR>
__device__ __forceinline__ void qpivot(const double (&col)[PR], const int (&pos)[PR], int k, int km, int lane,
                                       int &lp, int &rp, int &p) {
    unsigned hi[PR], lo[PR], hmax = 0;
    bool cand[PR];
#pragma unroll
    for (int r = 0; r < PR; ++r) {
        cand[r] = (unsigned)(pos[r] - k) <= (unsigned)km;
        const unsigned long long kk = cand[r] ? qkey(col[r]) : 0ull;
        hi[r] = (unsigned)(kk >> 32);
        lo[r] = (unsigned)kk;
        hmax = qumax(hmax, hi[r]);
    }
    const unsigned MH = wave_max_key(hmax);
    unsigned lmax = 0;
#pragma unroll
    for (int r = 0; r < PR; ++r) lmax = qumax(lmax, (cand[r] && hi[r] == MH) ? lo[r] : 0u);
    const unsigned ML = wave_max_key(lmax);
    unsigned sel = 0;
#pragma unroll
    for (int r = 0; r < PR; ++r) {
        const unsigned lab = ((unsigned)pos[r] << 8) | (unsigned)((r << 6) | lane);
        sel = qumax(sel, (cand[r] && hi[r] == MH && lo[r] == ML) ? 0x10000u - lab : 0u);
    }
    const unsigned lab = 0x10000u - wave_max_key(sel);
    lp = (int)(lab & 63u);
    rp = (int)((lab >> 6) & 3u);
    p = (int)(lab >> 8);
}

__device__ __forceinline__ int lds_load_relaxed(const int *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ void lds_store_relaxed(int *p, int v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

template <typename T, int PR>
__global__ void __launch_bounds__(64 * QWAVES) gbtrf_mfma_kernel(QArgs<T> s) {
    using v4 = typename Mf<T>::v4;
    constexpr int QS = QStride<T>::value;
    const int m = s.m, n = s.n, kl = s.kl, ku = s.ku, kv = kl + ku;
    const int R = s.R, RT = s.RT, RP = s.RP;
    const i64 ld = s.ldab;
    const int ldi = (int)ld;  // 16 columns of AB fit 2^31 bytes (checked at launch)
    T *__restrict__ ab = s.ab;
    const int tid = threadIdx.x, wv = tid >> 6, lane = tid & 63;
    const int mn = qmin(m, n);
    const int NP = (mn + QB - 1) / QB;

    // ---- LDS
    extern __shared__ __align__(16) unsigned char q_raw[];
    // rows of 16 entries at a padded stride QS (no LDS bank conflicts when the
    // 16 rows of a tile are read column by column)
    T *pbuf = reinterpret_cast<T *>(q_raw);  // [2][RP][QS] panel slots after factoring
    T *pan = pbuf + 2 * RP * QS;              // [2][RP][QS] next panel input (position order)
    T *lab = pan + 2 * RP * QS;               // [RP][QS]    look-ahead block column (slot order)
    T *colb = lab + RP * QS;                  // [QB][64*PR] multipliers per step
    T *linv = colb + QB * 64 * PR;            // [2][QB][QB] L11^{-1}, row major
    __shared__ int s_src[2][256];             // final position -> slot
    __shared__ int s_sig[QB];                 // pivot slot of each step
    __shared__ int s_land[2][64];             // stores of the lanes that do not signal
    __shared__ int s_ju[2];
    __shared__ int s_step;
    __shared__ int s_info;
    __shared__ int s_err;

    if (mn == 0) {
        if (tid == 0) {
            s.info[0] = 0;
            s.info[1] = 0;
        }
        return;
    }

    // entry (r, c) of AB as it stands (0 outside the matrix or the stored band)
    auto gload = [&](int r, int c) -> T {
        if (r < 0 || r >= m || c >= n || c - r > kv || r - c > kl) return T(0);
        return ab[(i64)c * ld + (kv + r - c)];
    };
    auto gstore = [&](int r, int c, T v) {
        if (r >= 0 && r < m && c < n && c - r <= kv && r - c <= kl) ab[(i64)c * ld + (kv + r - c)] = v;
    };

    // ---- prologue: panel 0 input, look-ahead buffer of panel 0, next panel's new rows
    if (tid == 0) {
        s_step = 0;
        s_info = 0;
        s_err = 0;
    }
    if (wv >= 2) {
        const int t0 = tid - 128, nt = 64 * QUW;
        for (int e = t0; e < RP * QB; e += nt) {
            const int q = e / QB, c = e - q * QB;
            pan[q * QS + c] = q < R ? gload(q, c) : T(0);
            lab[q * QS + c] = q < R ? gload(q, QB + c) : T(0);
        }
        for (int e = t0; e < QB * QB; e += nt) {
            const int i = e / QB, c = e - i * QB;
            pan[RP * QS + (R - QB + i) * QS + c] = gload(R + i, QB + c);
        }
    }
    lds_barrier();  // B0

    if (wv == 0) {
        // ================================================================ panel wave
        __builtin_amdgcn_s_setprio(3);
        int ju = 0, info = 0;
        for (int P = 0; P < NP; ++P) {
            const int j0 = P * QB, np = qmin(QB, mn - j0), b = P & 1;
            const T *pin = pan + b * RP * QS;
            // slot values by column: pp[c][r] is column c of slot r (a short vector
            // per column, so the pivot slot's value is picked by a uniform index)
            typename QCol<T, PR>::type pp[QB];
            int pos[PR];
            // 16-byte LDS reads (a row's 16 entries are contiguous, rows 16-byte aligned);
            // slots past the window read a clamped row and take zeros
            typedef T QV16 __attribute__((ext_vector_type(16 / sizeof(T))));
            constexpr int VW = 16 / (int)sizeof(T);
#pragma unroll
            for (int r = 0; r < PR; ++r) {
                const int q = lane + 64 * r;
                pos[r] = (q < R && j0 + q < m) ? q : QBIG;
                const T *row = pin + qmin(q, RP - 1) * QS;
#pragma unroll
                for (int c = 0; c < QB; c += VW) {
                    const QV16 v = *reinterpret_cast<const QV16 *>(row + c);
#pragma unroll
                    for (int e = 0; e < VW; ++e) pp[c + e][r] = (q < R) ? v[e] : T(0);
                }
            }
            int myp = 0;
            int slab[PR];  // (slot << 6) | lane, the label the pivot search returns
#pragma unroll
            for (int r = 0; r < PR; ++r) slab[r] = (r << 6) | lane;
            // multipliers of a step go to AB column j through a buffer resource over
            // the panel's columns (an inactive slot's store lands out of range)
            const int ncol = qmin(QB, n - j0);
            const __amdgpu_buffer_rsrc_t rpan = __builtin_amdgcn_make_buffer_rsrc(
                ab + (i64)j0 * ld, 0, (int)((i64)ncol * ld * (i64)sizeof(T)), kQRsrcWord3);
            // Each step finds the next step's pivot as soon as the next column has
            // its update, and updates the remaining columns after that: the
            // pivot search's wave reductions and the rank-1 update are independent
            // and interleave in one branch-free block.
            int lp, rp, p;  // pivot of the current step (uniform)
            {
                T c0v[PR];
#pragma unroll
                for (int r = 0; r < PR; ++r) c0v[r] = pp[0][r];
                qpivot<PR>(c0v, pos, 0, qmin(kl, m - 1 - j0), lane, lp, rp, p);
            }
            QFor<QB>::run([&](auto kc) {
                constexpr int k = decltype(kc)::value;
                if (k < np) {
                    const int j = j0 + k;
                    const int km = qmin(kl, m - 1 - j);
                    // ---- the pivot row, columns k .. QB-1 (uniform): the pivot slot's
                    // value picked per column, read from the pivot lane
                    T u[QB];
#pragma unroll
                    for (int c = 0; c < QB; ++c) u[c] = T(0);
#pragma unroll
                    for (int c = k; c < QB; ++c) u[c] = lane_read(pp[c][rp], lp);
                    const T pv = u[k];
                    const bool nz = (pv != T(0));
                    // interchange = relabel (nonzero pivot): the slot at position k
                    // takes position p, the pivot slot takes position k
                    ju = nz ? qmax(ju, qmin(j + ku + (p - k), n - 1)) : ju;
                    info = (!nz && info == 0) ? j + 1 : info;
                    // a zero pivot relabels nothing: no slot matches the label then,
                    // and position k maps to itself
                    const int pn = nz ? p : k;
                    const int labn = nz ? ((rp << 6) | lp) : -1;
#pragma unroll
                    for (int r = 0; r < PR; ++r) {
                        int L = (pos[r] == k) ? pn : pos[r];
                        pos[r] = (slab[r] == labn) ? k : L;
                    }
                    // a zero pivot leaves the column as it is: reciprocal 1, and no
                    // slot takes a nonzero multiplier
                    const T rcp = nz ? qrcp(pv) : T(1);
                    const int kmz = nz ? km : 0;
                    T *cb = colb + k * (64 * PR);
                    T le[PR];
#pragma unroll
                    for (int r = 0; r < PR; ++r) {
                        const int L = pos[r];
                        const bool act = (unsigned)(L - k - 1) < (unsigned)km;  // k < L <= k + km
                        const T raw = pp[k][r];
                        const T l = raw * rcp;
                        const int off = act ? (k * (ldi - 1) + kv + L) * (int)sizeof(T) : kQOob;
                        qbstore<T>(rpan, off, l);
                        pp[k][r] = act ? l : raw;
                        // inactive slots take multiplier 0 (a signed zero can lose its
                        // sign in their rank-1 update)
                        le[r] = ((unsigned)(L - k - 1) < (unsigned)kmz) ? l : T(0);
                        cb[lane + 64 * r] = le[r];
                    }
                    const int lp0 = lp, rp0 = rp, p0 = p;
                    if constexpr (k + 1 < QB) {
                        // column k+1 first, then the next pivot (its result is unused
                        // past the panel's last column)
                        T nx[PR];
#pragma unroll
                        for (int r = 0; r < PR; ++r) {
                            pp[k + 1][r] = dfma(le[r], -u[k + 1], pp[k + 1][r]);
                            nx[r] = pp[k + 1][r];
                        }
                        qpivot<PR>(nx, pos, k + 1, qmin(kl, m - 2 - j), lane, lp, rp, p);
                    }
                    // rank-1 update of the panel's later columns
#pragma unroll
                    for (int r = 0; r < PR; ++r) {
#pragma unroll
                        for (int c = k + 2; c < QB; ++c) pp[c][r] = dfma(le[r], -u[c], pp[c][r]);
                    }
                    myp = (lane == k) ? p0 : myp;
                    // the step's signal, from lane 0.  FP32 and FP64 with up to two slots: the
                    // other lanes store to their
                    // own landing words, so there is no branch and the step stays one block
                    // (its reductions interleave with the update); FP64 with three or four
                    // slots branches, which keeps the step's live registers within budget.
                    // LDS operations of a wave complete in order: a reader that sees the
                    // step count also sees this step's multipliers.
                    if constexpr (sizeof(T) == 4 || PR <= 2) {
                        *(lane == 0 ? &s_sig[k] : &s_land[0][lane]) = lp0 + 64 * rp0;
                        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                        lds_store_relaxed(lane == 0 ? &s_step : &s_land[1][lane], P * QB + k + 1);
                    } else if (lane == 0) {
                        s_sig[k] = lp0 + 64 * rp0;
                        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                        lds_store_relaxed(&s_step, P * QB + k + 1);
                    }
                }
            });
            // ---- publish: pivots, slot contents, final position -> slot
            if (lane < np) s.ipiv[j0 + lane] = (i64)(j0 + myp + 1);
            T *pb = pbuf + b * RP * QS;
#pragma unroll
            for (int r = 0; r < PR; ++r) {
                const int q = lane + 64 * r;
                if (q < RP) {
#pragma unroll
                    for (int c = 0; c < QB; ++c) pb[q * QS + c] = pp[c][r];
                    const int f = pos[r] < QBIG ? pos[r] : q;
                    s_src[b][f] = q;
                }
            }
            if (lane == 0) {
                s_ju[b] = ju;
                s_info = info;
            }
            lds_barrier();  // B1
            lds_barrier();  // B2
        }
    } else if (wv == 1) {
        // ================================================================ helper wave
        const int c16 = lane & 15;
        auto write_u11 = [&](int Pp) {
            // U11 of panel Pp: U[k][c] = row of the pivot slot of step k, column c >= k
            const int j0 = Pp * QB, np = qmin(QB, mn - j0), b = Pp & 1;
            const T *pb = pbuf + b * RP * QS;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int k = (lane >> 4) + 4 * i;
                if (k < np && k <= c16) {
                    const int q = s_src[b][k];
                    gstore(j0 + k, j0 + c16, pb[q * QS + c16]);
                }
            }
        };
        for (int P = 0; P < NP; ++P) {
            const int j0 = P * QB, np = qmin(QB, mn - j0), b = P & 1;
            // X = L11^{-1}; lane j (< 16) holds column j, X[i] = X(i, j)
            T X[QB];
#pragma unroll
            for (int i = 0; i < QB; ++i) X[i] = (i == lane) ? T(1) : T(0);
            QFor<QB>::run([&](auto kc) {
                constexpr int k = decltype(kc)::value;
                if (k < np) {
                    const int target = P * QB + k + 1;
                    int it = 0;
                    while (lds_load_relaxed(&s_step) < target) {
                        if (++it > (1 << 22)) {
                            s_err = 1;
                            break;
                        }
                        __builtin_amdgcn_s_sleep(1);
                    }
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
                    const int q = s_sig[k];
                    // row k: X(k, j) = [k == j] - sum_{k' < k} L11(k, k') X(k', j); colb holds l
                    T x = (lane == k) ? T(1) : T(0);
#pragma unroll
                    for (int kk = 0; kk < k; ++kk) x = dfma(-colb[kk * (64 * PR) + q], X[kk], x);
                    X[k] = x;
                }
            });
            if (lane < QB) {
                T *li = linv + b * QB * QB;
#pragma unroll
                for (int i = 0; i < QB; ++i) li[i * QB + lane] = X[i];
            }
            lds_barrier();  // B1
            // U11 of the previous panel while the update waves work on the
            // look-ahead block (its slots and slot map are replaced only after B2)
            if (P >= 1) write_u11(P - 1);
            lds_barrier();  // B2
        }
        write_u11(NP - 1);
        if (lane == 0) {
            // the last wave to finish the panel loop reports
            s.info[0] = (i64)s_info;
            s.info[1] = (i64)s_err;
        }
    } else {
        // ================================================================ update waves
        const int w = wv - 2;
        // rank among the waves that take block columns (-1: look-ahead tiles only)
        const int w1 = (wv & 3) == 0 ? -1 : w - (w > 2 ? 1 : 0) - (w > 6 ? 1 : 0);
        const int g = lane >> 4, li = lane & 15;
        // Byte offset of element (r, c0 + x) inside block column c0, x*ld + kv + r - c0 - x
        // elements, as min(t, kQOob - x*ld*ES) + x*ld*ES with t = (kv + r - c0)*ES - x*ES
        // taken unsigned: a band row kv + r - c below 0 (c - r > kv) wraps t past 2^31 and
        // the offset becomes kQOob, as does a row r >= m (row term -2^30); columns past the
        // matrix lie beyond the resource's columns.  The band's lower edge (r - c <= kl)
        // holds for every row a block column right of the look-ahead one is updated with.
        constexpr int ES = (int)sizeof(T);
        const unsigned ldes = (unsigned)ldi * (unsigned)ES;
        unsigned xe[4], xl[4], xc[4];
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const unsigned x = (unsigned)drow<T>(g, v);
            xe[v] = x * ES;
            xl[v] = x * ldes;
            xc[v] = (unsigned)kQOob - xl[v];
        }
        const unsigned ye = (unsigned)li * ES, yl = (unsigned)li * ldes, yc = (unsigned)kQOob - yl;
        auto rterm = [&](int r, int c0) -> int { return r < m ? (kv + r - c0) * ES : -(1 << 30); };
        auto boff = [](int rt_, unsigned e_, unsigned l_, unsigned c_) -> int {
            const unsigned t = (unsigned)(rt_ - (int)e_);
            return (int)((t < c_ ? t : c_) + l_);
        };
        // Apply panel Pp to block column B, reading and writing AB (mode 0) or
        // writing the rows the next panel needs to the look-ahead buffer (tola).
        // Every load of the block column is issued before the first is used
        // (one memory round trip per block column); loads and stores go
        // through a buffer resource over the block column's 16 AB columns, an
        // out-of-band element reading as 0 and its store being dropped.
        auto update_glb = [&](int Pp, int B, bool tola) {
            constexpr int MT = 4 * PR - 1;  // most 16-row tiles below the U rows
            const int j0p = Pp * QB, npp = qmin(QB, mn - j0p), b = Pp & 1;
            const int c0 = B * QB, ncols = qmin(QB, n - c0);
            const int *src = s_src[b];
            const T *pb = pbuf + b * RP * QS;
            const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
                ab + (i64)c0 * ld, 0, (int)((i64)ncols * ld * (i64)sizeof(T)), kQRsrcWord3);
            // element (r, c0+x) of AB sits (x*(ld-1) + kv + r - c0) elements into the
            // block column; it is stored iff r < m, r - kl - c0 <= x <= r + kv - c0
            // and c0 + x < n
            const int cmax = n - 1 - c0;
            auto offs = [&](int r, int c) -> int {
                const int x = c - c0;
                const bool ok = r < m && x >= r - kl - c0 && x <= qmin(r + kv - c0, cmax);
                return ok ? (int)__umul24((unsigned)x, (unsigned)((ldi - 1) * (int)sizeof(T))) +
                                (kv + r - c0) * (int)sizeof(T)
                          : kQOob;
            };
            T braw[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int f = 4 * g + i;
                braw[i] = qbload<T>(rs, f < npp ? boff(rterm(j0p + src[f], c0), ye, yl, yc) : kQOob);
            }
            const T *lv = linv + b * QB * QB + li * QB + 4 * g;
            v4 u = {T(0), T(0), T(0), T(0)}, nu;
            // tiles in chunks of CH (FP64 keeps fewer loads in flight to stay
            // within the register budget); a chunk's stores never overwrite a
            // row a later chunk reads (those come from the row itself or from
            // the U rows' positions, written last)
            constexpr int CH = sizeof(T) == 8 ? 6 : MT;
            constexpr int NCH = (MT + CH - 1) / CH;
#pragma unroll
            for (int ch = 0; ch < NCH; ++ch) {
                T acc[CH][4];
                int qv[CH];
#pragma unroll
                for (int t = 0; t < CH; ++t) {
                    const int rt = ch * CH + t + 1;
                    if (rt < RT) {
                        const int q = src[16 * rt + li];
                        qv[t] = q;
                        const int rq = rterm(j0p + q, c0);
#pragma unroll
                        for (int v = 0; v < 4; ++v) acc[t][v] = qbload<T>(rs, boff(rq, xe[v], xl[v], xc[v]));
                    }
                }
                if (ch == 0) {
#pragma unroll
                    for (int i = 0; i < 4; ++i) u = Mf<T>::mma(lv[i], braw[i], u);
#pragma unroll
                    for (int i = 0; i < 4; ++i) nu[i] = -u[i];
                }
#pragma unroll
                for (int t = 0; t < CH; ++t) {
                    const int rt = ch * CH + t + 1;
                    if (rt < RT) {
                        const int f = 16 * rt + li;
                        v4 a = {acc[t][0], acc[t][1], acc[t][2], acc[t][3]};
                        const T *lr = pb + qv[t] * QS;
#pragma unroll
                        for (int i = 0; i < 4; ++i) a = Mf<T>::mma(nu[i], lr[drow<T>(g, i)], a);
                        if (f < R) {
                            if (tola) {
#pragma unroll
                                for (int v = 0; v < 4; ++v) lab[(f - QB) * QS + drow<T>(g, v)] = a[v];
                            } else {
                                const int rf = rterm(j0p + f, c0);
#pragma unroll
                                for (int v = 0; v < 4; ++v) qbstore<T>(rs, boff(rf, xe[v], xl[v], xc[v]), a[v]);
                            }
                        }
                    }
                }
            }
            // U rows, after every tile has read its rows (the rows displaced by
            // this panel's pivots are read from the positions the U rows go to)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int f = drow<T>(g, v);
                qbstore<T>(rs, f < npp ? boff(rterm(j0p + f, c0), ye, yl, yc) : kQOob, u[v]);
            }
            if (tola) {
                // rows entering: look-ahead buffer positions R-16.. and the next
                // panel's input positions R-16.. (original rows, untouched so far)
                const int j0 = Pp * QB + QB;
                T *pn = pan + ((Pp + 2) & 1) * RP * QS;
                T e0[4], e1[4];
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int e = lane + 64 * i, r = e >> 4, c = e & 15;
                    e0[i] = qbload<T>(rs, offs(j0 + R - QB + r, c0 + c));
                    e1[i] = qbload<T>(rs, offs(j0 + R + r, c0 + c));
                }
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int e = lane + 64 * i, r = e >> 4, c = e & 15;
                    lab[(R - QB + r) * QS + c] = e0[i];
                    pn[(R - QB + r) * QS + c] = e1[i];
                }
            }
        };
        // Apply panel Pp to the look-ahead block column Pp+1 (LDS to LDS):
        // tiles rt0, rt0 + QUW, ...; the U rows are stored when store_u.
        auto update_la = [&](int Pp, int rt0, bool store_u) {
            const int j0p = Pp * QB, npp = qmin(QB, mn - j0p), b = Pp & 1;
            const int c0 = (Pp + 1) * QB;
            const int *src = s_src[b];
            const T *pb = pbuf + b * RP * QS;
            T braw[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int f = 4 * g + i;
                braw[i] = (f < npp) ? lab[src[f] * QS + li] : T(0);
            }
            const T *lv = linv + b * QB * QB + li * QB + 4 * g;
            v4 u = {T(0), T(0), T(0), T(0)};
#pragma unroll
            for (int i = 0; i < 4; ++i) u = Mf<T>::mma(lv[i], braw[i], u);
            v4 nu;
#pragma unroll
            for (int i = 0; i < 4; ++i) nu[i] = -u[i];
            T *dstp = pan + ((Pp + 1) & 1) * RP * QS;
            for (int rt = rt0; rt < RT; rt += QUW) {
                const int f = 16 * rt + li;
                const int q = src[f];
                v4 a;
#pragma unroll
                for (int v = 0; v < 4; ++v) a[v] = lab[q * QS + drow<T>(g, v)];
                const T *lr = pb + q * QS;
#pragma unroll
                for (int i = 0; i < 4; ++i) a = Mf<T>::mma(nu[i], lr[drow<T>(g, i)], a);
                if (f < R) {
#pragma unroll
                    for (int v = 0; v < 4; ++v) dstp[(f - QB) * QS + drow<T>(g, v)] = a[v];
                }
            }
            if (store_u) {
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int f = drow<T>(g, v);
                    if (f < npp) gstore(j0p + f, c0 + li, u[v]);
                }
            }
        };
        for (int P = 0; P <= NP; ++P) {
            // ---- phase 1: panel P-1 onto block columns P+1 ..
            if (P >= 1) {
                const int Pp = P - 1;
                const int jup = s_ju[Pp & 1];
                const int Bmax = qmin((n - 1) / QB, jup / QB);
                const bool la = (P < NP) && ((P + 1) * QB < n);
                const int Bend = la ? qmax(Bmax, P + 1) : Bmax;
                for (int B = P + 1; B <= Bend; ++B) {
                    if (B % QP1 != w1) continue;
                    update_glb(Pp, B, la && B == P + 1);
                }
            }
            if (P == NP) break;
            lds_barrier();  // B1
            // ---- phase 2: panel P onto the look-ahead block column P+1 (waves
            // without a tile stay out of the LDS traffic)
            if ((P + 1) * QB < n && (1 + w < RT || w == 0)) update_la(P, 1 + w, w == 0);
            lds_barrier();  // B2
        }
    }
}

template <typename T, int PR>
int mfma_launch_one(const QArgs<T> &a, hipStream_t st) {
    const size_t lds = sizeof(T) * ((size_t)5 * a.RP * QStride<T>::value + (size_t)QB * 64 * PR + 2 * QB * QB);
    auto kern = gbtrf_mfma_kernel<T, PR>;
    if (lds > 64 * 1024)
        BH_HIP(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL(kern, dim3(1), dim3(64 * QWAVES), lds, st, a);
    BH_LAUNCH_CHECK();
    return BH_OK;
}

// LAPACK zeroes the fill-in rows of columns ku+1 .. kv-1 before the first
// step and those of column j+kv at step j (j < min(m,n)); zero them all up
// front (rows whose row index is >= 0 only).
template <typename T>
__global__ void gbtrf_zero_fill(T *ab, i64 ld, int n, int mn, int kl, int ku) {
    const int kv = kl + ku;
    const int c = blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
    const int cend = qmin(n, mn + kv);
    if (c <= ku || c >= cend) return;
    for (int r = qmax(0, kv - c) + (threadIdx.x & 63); r < kl; r += 64) ab[(i64)c * ld + r] = T(0);
}

}  // namespace

// LDS bytes the kernel needs for a band, or 0 when it does not apply.
size_t mfma_lds_bytes(i64 kl, i64 elem) {
    if (kl < 1 || kl > 240) return 0;
    const i64 R = kl + QB, RP = 16 * ((R + 15) / 16), PR = (R + 63) / 64;
    const i64 QSe = elem == 4 ? 20 : 18;
    return (size_t)elem * ((size_t)5 * RP * QSe + (size_t)QB * 64 * PR + 2 * QB * QB);
}

// The blocked MFMA band LU, or -1 when the band does not fit it (kl + 16 rows
// in at most four slots per lane and the LDS buffers, with the kernel's
// static words (under 4 KiB), within 160 KiB).
template <typename T>
int mfma_gbtrf_launch(i64 m, i64 n, i64 kl, i64 ku, T *ab, i64 ldab, i64 *ipiv, i64 *dinfo, hipStream_t st) {
    const size_t lds = mfma_lds_bytes(kl, sizeof(T));
    if (lds == 0 || lds + 4096 > 160 * 1024 || m >= (1ll << 29) || n >= (1ll << 29) || kl + ku >= (1 << 20) ||
        (i64)QB * ldab * (i64)sizeof(T) >= (1ll << 31) || ldab * (i64)sizeof(T) >= (1ll << 24))
        return -1;
    const int R = (int)kl + QB, RT = (R + 15) / 16, RP = 16 * RT;
    QArgs<T> a{ab, ipiv, dinfo, (int)m, (int)n, (int)kl, (int)ku, ldab, R, RT, RP};
    const int mn = (int)imin(m, n);
    const int cend = (int)imin(n, (i64)mn + kl + ku);
    if (cend > ku + 1) {
        const int per = 4;  // columns per 256-thread block
        hipLaunchKernelGGL(gbtrf_zero_fill<T>, dim3((cend + per - 1) / per), dim3(64 * per), 0, st, ab, ldab, (int)n,
                           mn, (int)kl, (int)ku);
        BH_LAUNCH_CHECK();
    }
    const int PR = (R + 63) / 64;
    switch (PR) {
        case 1:
            return mfma_launch_one<T, 1>(a, st);
        case 2:
            return mfma_launch_one<T, 2>(a, st);
        case 3:
            return mfma_launch_one<T, 3>(a, st);
        case 4:
            return mfma_launch_one<T, 4>(a, st);
        default:
            return -1;
    }
}

template int mfma_gbtrf_launch<double>(i64, i64, i64, i64, double *, i64, i64 *, i64 *, hipStream_t);
template int mfma_gbtrf_launch<float>(i64, i64, i64, i64, float *, i64, i64 *, i64 *, hipStream_t);

}  // namespace bh
