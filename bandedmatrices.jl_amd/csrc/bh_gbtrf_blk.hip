// bh_gbtrf_blk.hip -- blocked partial-pivot band LU for wide bands (gfx950).
//
// One workgroup factors one system, NB = 16 columns ("a panel") at a time,
// with the arithmetic of LAPACK xGBTF2 (the routine behind LAPACK.gbtrf!,
// src/banded/BandedLU.jl:98): IDAMAX pivots (first index of the largest
// |a|), multipliers a * (1/pivot), and every entry receives its rank-1
// updates fma(l, -u, a) one step at a time in step order.  Inside the live
// window the updates are applied without LAPACK's row/column range masks,
// using a multiplier of 0 for rows a step does not touch and u = +0 for
// columns beyond its reach.  On finite inputs that contain no -0.0 this adds
// exact zeros only, so factors and pivots are bit-identical to xGBTF2; a
// -0.0 input entry can come out +0.0, and a NaN or Inf spreads to more
// entries of the window than LAPACK's would.
//
// Waves of the workgroup:
//   * NW WINDOW WAVES hold the live window in registers: row slots (wave w
//     owns slots w*RRW .. w*RRW+RRW-1, the same rows in every lane) by the
//     kv = kl+ku trailing columns [j0+NB, j0+NB+kv), column c in ring slot
//     c mod kv, lane l owning ring slots l + 64*cc.  A row interchange
//     relabels two slots; no data moves.
//   * the PANEL WAVE (raised issue priority) factors the 16 panel columns
//     [j0, j0+NB), taken from LDS, alone and without barriers: DPP max
//     reductions pick the pivot, readlane broadcasts the pivot row, each lane
//     updates its row slots.  It writes multipliers and pivots to AB / ipiv.
// Per panel P (j0 = 16P), once the panel wave is done (barrier B1):
//   2. window waves copy the pivot rows' trailing entries to LDS (U12) and
//      the next panel's 16 columns (as they stand) to the panel buffer;
//   3. a unit-lower solve with the panel's multipliers turns U12 into the U
//      rows' trailing part, one thread per column;
//   4. the next panel's columns receive this panel's rank-16 update in LDS,
//      so the panel wave can start on panel P+1 at once;
//   5. while it does, the window waves apply the rank-16 update to the rest
//      of the window, retire the pivot rows (their slots take the rows
//      entering the band) and shift the column ring; panel P's U rows are
//      written to AB and the rows entering next prefetched early in it.
// All barriers order LDS only.
#include "bh_common.hpp"

namespace bh {

namespace {

constexpr int NB = 16;

template <typename T>
struct BlkArgs {
    T *ab;
    i64 *ipiv;
    i64 *info;
    int m, n, kl, ku;
    i64 ldab;
};

__device__ __forceinline__ int imin32(int a, int b) { return a < b ? a : b; }
__device__ __forceinline__ int imax32(int a, int b) { return a > b ? a : b; }

// Compile-time loop: f(IntC<0>{}), ..., f(IntC<N-1>{}) -- register arrays
// indexed by the step stay in registers however large the body is.
template <int K>
struct IntC {
    static constexpr int value = K;
};
template <int N>
struct StaticFor {
    template <class F>
    __device__ __forceinline__ static void run(F &&f) {
        StaticFor<N - 1>::run(f);
        f(IntC<N - 1>{});
    }
};
template <>
struct StaticFor<0> {
    template <class F>
    __device__ __forceinline__ static void run(F &&) {}
};

// Unsigned maximum over the wave (uniform): DPP partial maxima with zero
// fill for lanes without a source (0 is the identity: all keys are >= 0).
template <int CTRL, int ROWMASK = 0xF>
__device__ __forceinline__ unsigned dpp_zero(unsigned v) {
    return (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, ROWMASK, 0xF, true);
}
__device__ __forceinline__ unsigned umax32(unsigned a, unsigned b) { return a > b ? a : b; }
__device__ __forceinline__ unsigned wave_umax(unsigned x) {
    x = umax32(x, dpp_zero<0x111>(x));
    x = umax32(x, dpp_zero<0x112>(x));
    x = umax32(x, dpp_zero<0x114>(x));
    x = umax32(x, dpp_zero<0x118>(x));
    x = umax32(x, dpp_zero<0x142, 0xA>(x));
    x = umax32(x, dpp_zero<0x143, 0xC>(x));
    return (unsigned)__builtin_amdgcn_readlane((int)x, 63);
}
// Maximum of doubles over the wave (uniform); inputs are never NaN.
__device__ __forceinline__ double wave_dmax(double x) {
    x = __builtin_fmax(x, dpp_mov<0x111>(x, x));
    x = __builtin_fmax(x, dpp_mov<0x112>(x, x));
    x = __builtin_fmax(x, dpp_mov<0x114>(x, x));
    x = __builtin_fmax(x, dpp_mov<0x118>(x, x));
    x = __builtin_fmax(x, dpp_mov<0x142, 0xA>(x, x));
    x = __builtin_fmax(x, dpp_mov<0x143, 0xC>(x, x));
    return lane_read(x, 63);
}

// Pivot search key of a candidate: ordered like |v| with NaN below every
// number; 0 means "not a candidate".
__device__ __forceinline__ unsigned pivot_key(float v) {
    const unsigned b = __builtin_bit_cast(unsigned, v) & 0x7fffffffu;
    return b > 0x7f800000u ? 1u : b + 2u;
}

template <typename T, int NW, int RRW, int CCL>
struct BlkGeom {
    static_assert(RRW % 2 == 0, "row pairs");
    static_assert(CCL <= 4, "ring groups");
    static constexpr int RS = NW * RRW;          // row slots of the window
    static constexpr int PR = (RS + 63) / 64;    // row slots per lane of the panel wave
    static constexpr int RSP = 64 * PR;          // padded slot count (panel layout)
    static constexpr int KVP = 64 * CCL;         // ring slot capacity (>= kv)
    static constexpr int NWT = 64 * NW;          // window threads
    static constexpr int RRWP = (RRW + 3) & ~3;  // padded rows per wave (multiplier layout)
    static constexpr int NPRE = ((NB + KVP) * NB + NWT - 1) / NWT;  // prefetch registers
    // dynamic LDS in elements of T
    static size_t lds_elems(int kv) {
        return (size_t)NB * RSP + 2 * (size_t)NW * NB * RRWP + (size_t)NB * KVP + NB * NB + 2 * NB * NB + NB * NB +
               (size_t)NB * (kv + 1);
    }
};

template <typename T, int NW, int RRW, int CCL>
__global__ void __launch_bounds__(64 * (NW + 1)) gbtrf_blk_kernel(BlkArgs<T> s) {
    using G = BlkGeom<T, NW, RRW, CCL>;
    constexpr int RS = G::RS, PR = G::PR, RSP = G::RSP, KVP = G::KVP, NWT = G::NWT, RRWP = G::RRWP;
    constexpr int NPRE = G::NPRE;
    const int m = s.m, n = s.n, kl = s.kl, ku = s.ku, kv = kl + ku, H = kv + 1;
    const i64 ld = s.ldab;
    T *__restrict__ ab = s.ab;
    const int tid = threadIdx.x, wv = tid >> 6, lane = tid & 63;
    const bool panel = (wv == NW);
    const int mn = imin32(m, n);

    extern __shared__ __align__(16) unsigned char dyn_raw[];
    T *Pn = reinterpret_cast<T *>(dyn_raw);  // [NB][RSP]     panel columns by slot
    T *Lm = Pn + NB * RSP;                     // [2][NW][NB][RRWP] multipliers by (wave, step, row)
    T *U12 = Lm + 2 * NW * NB * RRWP;          // [NB][KVP]     U rows' trailing part by ring slot
    T *U12n = U12 + NB * KVP;                  // [NB][NB]      the same for the next panel's columns
    T *Up = U12n + NB * NB;                    // [2][NB][NB]   U rows' panel part
    T *L11 = Up + 2 * NB * NB;                 // [NB][NB]      multipliers of the pivot rows
    T *rows = L11 + NB * NB;                   // [NB][H]       rows entering at the next retirement
    __shared__ int labA[RSP], labB[RSP];
    __shared__ int sh_kmin[4];
    __shared__ int sh_nzm, sh_info;

    if (mn == 0) {
        if (tid == 0) *s.info = 0;
        return;
    }
    // original entry (r, c) of the matrix (0 outside the band / matrix)
    auto orig = [&](int r, int c) -> T {
        if (r < 0 || r >= m || c < 0 || c >= n || c - r > ku || r - c > kl) return T(0);
        return ab[(i64)c * ld + (kv + r - c)];
    };
    const int NP = (mn + NB - 1) / NB;

    if (panel) {
        // ============================================================ panel wave
        __builtin_amdgcn_s_setprio(3);
        lds_barrier();  // B0
        int ju = 0, info = 0;
        for (int P = 0; P < NP; ++P) {
            const int j0 = P * NB, np = imin32(NB, mn - j0), buf = P & 1;
            T pp[PR][NB];
            int lb[PR];
#pragma unroll
            for (int r = 0; r < PR; ++r) {
                const int sl = lane + 64 * r;
                lb[r] = labB[sl];
#pragma unroll
                for (int x = 0; x < NB; ++x) pp[r][x] = Pn[x * RSP + sl];
            }
            int nz_mask = 0, myp = 0, myju = 0;
            T *colp = ab + (i64)j0 * ld + kv - j0;  // colp[L]: multiplier slot of row L in column j
            StaticFor<NB>::run([&](auto kc) {
                constexpr int k = decltype(kc)::value;
                if (k < np) {
                    const int j = j0 + k;
                    const int km = imin32(kl, m - 1 - j);
                    // ---- pivot: largest |a| over rows j..j+km, first row on ties
                    int bl = 0x7fffffff, br = 0;
                    bool top;
                    if constexpr (sizeof(T) == 4) {
                        unsigned bk = 0;
#pragma unroll
                        for (int r = 0; r < PR; ++r) {
                            const int L = lb[r];
                            const unsigned key =
                                ((unsigned)(L - j) <= (unsigned)km) ? pivot_key((float)pp[r][k]) : 0u;
                            const bool better = key > bk || (key == bk && L < bl);
                            bk = better ? key : bk;
                            bl = better ? L : bl;
                            br = better ? r : br;
                        }
                        top = (bk == wave_umax(bk));
                    } else {
                        T best = T(-1);
#pragma unroll
                        for (int r = 0; r < PR; ++r) {
                            const int L = lb[r];
                            T v = dabs(pp[r][k]);
                            if (!(v == v)) v = T(-0.5);  // NaN: a candidate that loses to any number
                            v = ((unsigned)(L - j) <= (unsigned)km) ? v : T(-1);
                            const bool better = v > best || (v == best && L < bl);
                            best = better ? v : best;
                            bl = better ? L : bl;
                            br = better ? r : br;
                        }
                        top = (best == wave_dmax(best));
                    }
                    // lanes holding the maximum: one (the usual case) or a tie
                    // broken by the smallest row
                    unsigned long long tops = (unsigned long long)__ballot(top);
                    int p, ol;
                    if (__popcll(tops) == 1) {
                        ol = __builtin_amdgcn_readfirstlane((int)__ffsll(tops) - 1);
                        p = __builtin_amdgcn_readlane(bl, ol);
                    } else {
                        p = 0x7fffffff - (int)wave_umax(top ? (unsigned)(0x7fffffff - bl) : 0u);
                        tops = (unsigned long long)__ballot(top && bl == p);
                        ol = __builtin_amdgcn_readfirstlane((int)__ffsll(tops) - 1);
                    }
                    const int pr_ = __builtin_amdgcn_readlane(br, ol);
                    // ---- the pivot row (uniform), columns k .. NB-1
                    T u[NB];
#pragma unroll
                    for (int c = 0; c < NB; ++c) u[c] = T(0);
#pragma unroll
                    for (int r = 0; r < PR; ++r) {
                        if (pr_ == r) {
#pragma unroll
                            for (int c = k; c < NB; ++c) u[c] = lane_read(pp[r][c], ol);
                        }
                    }
                    const T pv = u[k];
                    const bool nz = (pv != T(0));
                    if (nz) {
                        nz_mask |= 1 << k;
                        ju = imax32(ju, imin32(p + ku, n - 1));
                        if (p != j) {
#pragma unroll
                            for (int r = 0; r < PR; ++r)
                                if (lb[r] == j) lb[r] = p;
                        }
                    } else if (info == 0) {
                        info = j + 1;
                    }
                    if (lane == k) {
                        myp = p;
                        myju = ju;
                    }
                    const T rcp = T(1) / pv;
#pragma unroll
                    for (int r = 0; r < PR; ++r) {
                        const bool own = (r == pr_) && (lane == ol);
                        if (own) lb[r] = j;
                        const int L = lb[r];
                        const bool in = (L > j) && (L <= j + km);
                        const T raw = pp[r][k];
                        const T lv = nz ? raw * rcp : raw;
                        if (in) colp[L] = lv;
                        const T lm = (in && nz) ? lv : T(0);
#pragma unroll
                        for (int c = k + 1; c < NB; ++c) pp[r][c] = dfma(lm, -u[c], pp[r][c]);
                        // column k of a slot: the pivot (own), the U entry of an
                        // earlier pivot row of this panel (kept), or the multiplier
                        const bool keep = (L >= j0) && (L < j);
                        pp[r][k] = own ? pv : (keep ? raw : lm);
                    }
                    colp += ld - 1;
                }
            });
            // ---- publish: multipliers, U rows' panel part, pivots, labels
#pragma unroll
            for (int r = 0; r < PR; ++r) {
                const int sl = lane + 64 * r;
                if (sl < RS) {
                    const int w = sl / RRW, i = sl - w * RRW;
                    T *dst = Lm + ((buf * NW + w) * NB) * RRWP + i;
#pragma unroll
                    for (int k = 0; k < NB; ++k) dst[k * RRWP] = pp[r][k];
                }
                labA[sl] = lb[r];
                const int L = lb[r];
                if (L >= j0 && L < j0 + np) {
                    const int kk = L - j0;
#pragma unroll
                    for (int k = 0; k < NB; ++k) {
                        L11[kk * NB + k] = (k < kk) ? pp[r][k] : T(0);
                        Up[(buf * NB + kk) * NB + k] = pp[r][k];
                    }
                }
            }
            if (lane < np) s.ipiv[j0 + lane] = (i64)myp + 1;
            {
                // first step whose update reaches the smallest column of each
                // group of 64 ring slots (window waves skip the earlier steps)
                const int base = (j0 + NB) % kv;
                const bool stepok = lane < np && ((nz_mask >> lane) & 1);
#pragma unroll
                for (int cc = 0; cc < CCL; ++cc) {
                    const int cs0 = 64 * cc, cs1 = imin32(cs0 + 64, kv);
                    int km_ = NB;
                    if (cs0 < kv) {
                        int off = (base >= cs0 && base < cs1) ? 0 : cs0 - base;
                        if (off < 0) off += kv;
                        const unsigned long long reach = (unsigned long long)__ballot(stepok && myju >= j0 + NB + off);
                        if (reach) km_ = (int)__ffsll(reach) - 1;
                    }
                    if (lane == 0) sh_kmin[cc] = km_;
                }
            }
            if (lane == 0) {
                sh_nzm = nz_mask;
                sh_info = info;
            }
            lds_barrier();  // B1
            lds_barrier();  // B2
            lds_barrier();  // B3
            if (P + 1 < NP) lds_barrier();  // B4
        }
        lds_barrier();  // end
    } else {
        // ============================================================ window waves
        // rows entering at the retirement of panel q: e = 16q + kl + NB + b,
        // gathered column by column (consecutive rows are consecutive in AB)
        T pre[NPRE];
        auto prefetch = [&](int q) {
            const int jq = q * NB;
#pragma unroll
            for (int r = 0; r < NPRE; ++r) {
                const int f = tid + r * NWT;
                const int b = f & (NB - 1), ci = f >> 4;
                const int t = ci - b;
                T v = T(0);
                if (ci < NB + kv && t >= 0 && t <= kv) v = orig(jq + kl + NB + b, jq + NB + ci);
                pre[r] = v;
            }
        };
        auto commit = [&]() {
#pragma unroll
            for (int r = 0; r < NPRE; ++r) {
                const int f = tid + r * NWT;
                const int b = f & (NB - 1), ci = f >> 4;
                const int t = ci - b;
                if (ci < NB + kv && t >= 0 && t <= kv) rows[b * H + t] = pre[r];
            }
        };

        T a[RRW][CCL];
        int labv[RRW];
        int kmin[CCL];
        int nzm = 0;
        {
            const int base = NB % kv;
#pragma unroll
            for (int i = 0; i < RRW; ++i) {
                const int rs = wv * RRW + i;
                labv[i] = (rs < kl + NB && rs < m) ? rs : -1;
#pragma unroll
                for (int cc = 0; cc < CCL; ++cc) {
                    const int cs = lane + 64 * cc;
                    int off = cs - base;
                    if (off < 0) off += kv;
                    a[i][cc] = (cs < kv) ? orig(labv[i], NB + off) : T(0);
                }
            }
            for (int f = tid; f < NB * RSP; f += NWT) {
                const int x = f / RSP, sl = f - x * RSP;
                const int L = (sl < RS && sl < kl + NB && sl < m) ? sl : -1;
                Pn[f] = orig(L, x);
            }
            for (int f = tid; f < RSP; f += NWT) labB[f] = (f < RS && f < kl + NB && f < m) ? f : -1;
            prefetch(0);
        }
#pragma unroll
        for (int cc = 0; cc < CCL; ++cc) kmin[cc] = NB;
        lds_barrier();  // B0

        // ---- phase 5 of panel q
        auto finish_panel = [&](int q, bool more) {
            const int jq = q * NB, npq = imin32(NB, mn - jq), bq = q & 1;
            const int base = (jq + NB) % kv;
            // U rows jq .. jq+npq-1 to AB (contiguous band rows per column);
            // stores and the prefetch go first so they complete during the update
            for (int f = tid; f < (npq + kv) * NB; f += NWT) {
                const int jj = f & (NB - 1), ci = f >> 4;
                const int c = jq + ci, t = ci - jj;
                if (jj < npq && t >= 0 && t <= kv && c < n) {
                    T v;
                    if (ci < NB) {
                        v = Up[(bq * NB + jj) * NB + ci];
                    } else {
                        int cs = base + ci - NB;
                        if (cs >= kv) cs -= kv;
                        v = U12[jj * KVP + cs];
                    }
                    ab[(i64)c * ld + (kv - t)] = v;
                }
            }
            // (FP64 prefetches after the update: its registers would not fit beside it)
            if (sizeof(T) == 4 && more) prefetch(q + 1);
            // rank-16 update of the window (multiplier 0 / u = +0 outside LAPACK's ranges)
            for (int k = 0; k < npq; ++k) {
                if (!((nzm >> k) & 1)) continue;
                T lr[RRW];
                const T *lp = Lm + (((bq * NW + wv) * NB + k) * RRWP);
#pragma unroll
                for (int i = 0; i < RRW; ++i) lr[i] = lp[i];
#pragma unroll
                for (int cc = 0; cc < CCL; ++cc) {
                    if (k < kmin[cc]) continue;
                    const T nu = -U12[k * KVP + lane + 64 * cc];
#pragma unroll
                    for (int i = 0; i < RRW; ++i) a[i][cc] = dfma(lr[i], nu, a[i][cc]);
                }
            }
            if (sizeof(T) == 8 && more) prefetch(q + 1);
            // retire the pivot rows: their slots take the entering rows
            int col[CCL];
#pragma unroll
            for (int cc = 0; cc < CCL; ++cc) {
                int off = lane + 64 * cc - base;
                if (off < 0) off += kv;
                col[cc] = jq + NB + off;
            }
#pragma unroll
            for (int i = 0; i < RRW; ++i) {
                const int L = labv[i];
                if (L >= jq && L < jq + npq) {
                    const int b = L - jq, e = L + kl + NB;
#pragma unroll
                    for (int cc = 0; cc < CCL; ++cc) {
                        const int c = col[cc], t = c - (e - kl);
                        a[i][cc] =
                            (lane + 64 * cc < kv && e < m && c < n && t >= 0 && t <= kv) ? rows[b * H + t] : T(0);
                    }
                    labv[i] = e < m ? e : -1;
                }
            }
            // the next panel's columns left the ring (to LDS); their slots
            // take the columns entering the window
#pragma unroll
            for (int cc = 0; cc < CCL; ++cc) {
                if (lane + 64 * cc < kv && col[cc] < jq + 2 * NB) {
                    const int c2 = col[cc] + kv;
#pragma unroll
                    for (int i = 0; i < RRW; ++i) {
                        const int e = labv[i], b = e - (jq + kl + NB), t = c2 - (e - kl);
                        a[i][cc] = (b >= 0 && b < NB && c2 < n && t >= 0 && t <= kv) ? rows[b * H + t] : T(0);
                    }
                }
            }
        };

        for (int P = 0; P < NP; ++P) {
            const int j0 = P * NB, np = imin32(NB, mn - j0), buf = P & 1;
            const bool more = P + 1 < NP;
            if (P > 0) finish_panel(P - 1, true);
            lds_barrier();  // B1: panel P factored, panel P-1 finished
            const int base = (j0 + NB) % kv;
            {
                // 2. labels, column ranges, staged rows, pivot rows' trailing
                //    entries, the next panel's columns as they stand
                nzm = __builtin_amdgcn_readfirstlane(sh_nzm);
#pragma unroll
                for (int i = 0; i < RRW; ++i) labv[i] = __builtin_amdgcn_readfirstlane(labA[wv * RRW + i]);
#pragma unroll
                for (int cc = 0; cc < CCL; ++cc) kmin[cc] = __builtin_amdgcn_readfirstlane(sh_kmin[cc]);
                commit();
#pragma unroll
                for (int i = 0; i < RRW; ++i) {
                    const int L = labv[i];
                    if (L >= j0 && L < j0 + np) {
                        const int kk = L - j0;
#pragma unroll
                        for (int cc = 0; cc < CCL; ++cc) U12[kk * KVP + lane + 64 * cc] = a[i][cc];
                    }
                }
                if (more) {
#pragma unroll
                    for (int cc = 0; cc < CCL; ++cc) {
                        const int cs = lane + 64 * cc;
                        int off = cs - base;
                        if (off < 0) off += kv;
                        if (cs < kv && off < NB) {
#pragma unroll
                            for (int i = 0; i < RRW; ++i) Pn[off * RSP + wv * RRW + i] = a[i][cc];
                        }
                    }
                    if (lane == 0) {
#pragma unroll
                        for (int i = 0; i < RRW; ++i) {
                            const int L = labv[i];
                            const int e = L + kl + NB;
                            labB[wv * RRW + i] = (L >= j0 && L < j0 + np) ? (e < m ? e : -1) : L;
                        }
                    }
                }
            }
            lds_barrier();  // B2
            if (tid < kv) {
                // 3. U12 <- L11^{-1} U12, column tid, steps in order
                T u[NB];
#pragma unroll
                for (int k = 0; k < NB; ++k) u[k] = (k < np) ? U12[k * KVP + tid] : T(0);
#pragma unroll
                for (int k = 1; k < NB; ++k) {
                    if (k < np) {
                        T lrow[NB];
#pragma unroll
                        for (int k2 = 0; k2 < NB; ++k2) lrow[k2] = L11[k * NB + k2];
#pragma unroll
                        for (int k2 = 0; k2 < k; ++k2) u[k] = dfma(lrow[k2], -u[k2], u[k]);
                        U12[k * KVP + tid] = u[k];
                    }
                }
                int off = tid - base;
                if (off < 0) off += kv;
                if (more && off < NB) {
#pragma unroll
                    for (int k = 0; k < NB; ++k) U12n[k * NB + off] = u[k];
                }
            }
            lds_barrier();  // B3
            if (more) {
                // 4. rank-16 update of the next panel's columns in LDS: one
                //    thread per (row pair, 4 columns)
                constexpr int HP = RRW / 2;
                if (tid < NW * HP * 4) {
                    const int xq = tid & 3, rp = tid >> 2;
                    const int w = rp / HP, i0 = 2 * (rp - w * HP);
                    const int sl0 = w * RRW + i0;
                    int Lr[2];
                    bool piv[2];
                    T v[2][4];
#pragma unroll
                    for (int h = 0; h < 2; ++h) {
                        Lr[h] = labA[sl0 + h];
                        piv[h] = Lr[h] >= j0 && Lr[h] < j0 + np;
#pragma unroll
                        for (int c = 0; c < 4; ++c) v[h][c] = Pn[(4 * xq + c) * RSP + sl0 + h];
                    }
                    if (!(piv[0] && piv[1])) {
                        const T *lp = Lm + ((buf * NW + w) * NB) * RRWP + i0;
#pragma unroll
                        for (int k = 0; k < NB; ++k) {
                            if (k < np && ((nzm >> k) & 1)) {
                                const T l0 = lp[k * RRWP], l1 = lp[k * RRWP + 1];
                                T uu[4];
#pragma unroll
                                for (int c = 0; c < 4; ++c) uu[c] = -U12n[k * NB + 4 * xq + c];
#pragma unroll
                                for (int c = 0; c < 4; ++c) {
                                    v[0][c] = dfma(l0, uu[c], v[0][c]);
                                    v[1][c] = dfma(l1, uu[c], v[1][c]);
                                }
                            }
                        }
                    }
#pragma unroll
                    for (int h = 0; h < 2; ++h) {
                        if (piv[h]) {
                            const int b = Lr[h] - j0, e = Lr[h] + kl + NB;
#pragma unroll
                            for (int c = 0; c < 4; ++c) {
                                const int x = 4 * xq + c, t = x - b;
                                v[h][c] = (e < m && j0 + NB + x < n && t >= 0 && t <= kv) ? rows[b * H + t] : T(0);
                            }
                        }
#pragma unroll
                        for (int c = 0; c < 4; ++c) Pn[(4 * xq + c) * RSP + sl0 + h] = v[h][c];
                    }
                }
                lds_barrier();  // B4
            }
        }
        finish_panel(NP - 1, false);
        lds_barrier();  // end
    }
    // Fill-in rows LAPACK zeroes that no U row covered (rows >= m).
    {
        const int c_lo = imax32(ku + 1, m + ku), c_hi = imin32(n, mn + kv);
        for (int c = c_lo; c < c_hi; ++c) {
            const int q0 = imax32(0, kv - c);
            for (int q = q0 + tid; q < kl; q += blockDim.x) {
                const int r = c - kv + q;
                if (r >= m) ab[(i64)c * ld + q] = T(0);
            }
        }
    }
    if (tid == 0) *s.info = sh_info;
}

template <typename T, int NW, int RRW, int CCL>
int blk_launch_one(const BlkArgs<T> &a, hipStream_t st) {
    using G = BlkGeom<T, NW, RRW, CCL>;
    const size_t lds = G::lds_elems(a.kl + a.ku) * sizeof(T);
    auto kern = gbtrf_blk_kernel<T, NW, RRW, CCL>;
    if (lds > 64 * 1024)
        BH_HIP(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL(kern, dim3(1), dim3(64 * (NW + 1)), lds, st, a);
    BH_LAUNCH_CHECK();
    return BH_OK;
}

}  // namespace

// Blocked instance for a band, or -1.  Needs kl + 16 row slots and kv ring
// slots, kv >= 16 (the ring holds the next panel's columns).
int blk_shape_for(i64 kl, i64 ku) {
    const i64 kv = kl + ku;
    if (kv < NB || kl < 1) return -1;
    if (kl + NB <= 150 && kv <= 256) return 0;
    return -1;
}

template <typename T>
int blk_gbtrf_launch(i64 m, i64 n, i64 kl, i64 ku, T *ab, i64 ldab, i64 *ipiv, i64 *dinfo, hipStream_t st) {
    if (blk_shape_for(kl, ku) < 0 || m >= (1ll << 30) || n >= (1ll << 30)) return -1;
    BlkArgs<T> a{ab, ipiv, dinfo, (int)m, (int)n, (int)kl, (int)ku, ldab};
    return blk_launch_one<T, 15, 10, 4>(a, st);
}

template int blk_gbtrf_launch<double>(i64, i64, i64, i64, double *, i64, i64 *, i64 *, hipStream_t);
template int blk_gbtrf_launch<float>(i64, i64, i64, i64, float *, i64, i64 *, i64 *, hipStream_t);

}  // namespace bh
