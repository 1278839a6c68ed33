// bh_gbtrs_part.hpp -- partitioned band LU solve: A X = B or A^T X = B with
// the factors of gbtrf, the rows split into blocks that are solved at once
// across the whole GPU.
//
// Same job as LAPACK.gbtrs! in ldiv!(::BandedLU, B) and
// ldiv!(transpose(::BandedLU), B) (src/banded/linalg.jl:24-47); the result
// equals bh_?gbtrs's to rounding, not bit for bit (bh_?gbtrs keeps the
// bit-exact sequential chain).
//
// Each triangular sweep ('N': L with the interchanges forward, then U
// backward; 'T': U^T forward, then L^T backward with the interchanges) is a
// recurrence over the rows whose state is a window of KW values (KW > the
// factor's bandwidth): the rows the next steps update (axpy form, L and U) or
// the last values produced (dot form, U^T and L^T).  The state a block hands
// on is linear in the state it takes over and in its own right-hand side:
// w_out = A_i w_in + d_i.  Three launches per sweep:
//   1. all blocks at once, one lane per run: KW runs from the unit states
//      (the columns of A_i) and one run per right-hand side from the zero
//      state (d_i);
//   2. one workgroup walks the blocks in order, w_next = A_i w_i + d_i;
//   3. all blocks at once, one lane per right-hand side, from the true state
//      w_i, writing the block's results.
// A block's factor columns, pivots and right-hand-side rows are staged in LDS
// a chunk at a time by the whole wave (coalesced); the window lives in
// registers and is renamed, not moved, from step to step (KW steps unrolled).
#pragma once

#include "bh_common.hpp"

namespace bh {

namespace part {

enum { kLN = 0, kUN = 1, kUT = 2, kLT = 3 };
constexpr int kPartNR = 16;  // right-hand sides per wave

template <int OP>
struct POp {
    static constexpr bool asc = OP == kLN || OP == kUT;     // rows in increasing order
    static constexpr bool axpy = OP == kLN || OP == kUN;    // window of rows ahead, renamed leftwards
    static constexpr bool lower = OP == kLN || OP == kLT;   // L: interchanges, unit diagonal
};

template <typename T>
struct PartArgs {
    const T *ab;
    const i64 *ipiv;
    const T *bin;  // right-hand sides as they enter the sweep
    T *yout;       // the sweep's results (launch 3)
    T *amat;       // [p][KW][KW]: amat[(i*KW + q)*KW + t] = entry t of the state block i hands on from e_q
    T *dvec;       // [p][nrhs][KW]: the state block i hands on from the zero state
    T *cvec;       // [p][nrhs][KW]: the state block i takes over
    unsigned *growth;  // largest |entry| of the A_i (FP32 bits, NaN above all), per sweep
    i64 ld, ldb, ldy;
    int n, kl, kv, nrhs, M, p;
};

template <int K>
struct PC {
    static constexpr int value = K;
};
template <int N>
struct PUnroll {
    template <class F>
    __device__ __forceinline__ static void run(F &&f) {
        PUnroll<N - 1>::run(f);
        f(PC<N - 1>{});
    }
};
template <>
struct PUnroll<0> {
    template <class F>
    __device__ __forceinline__ static void run(F &&) {}
};
template <int OP>
__device__ __forceinline__ int prow(int js, int je, int s) {
    return POp<OP>::asc ? js + s : je - 1 - s;
}

// Launches 1 (FINAL = false) and 3 (FINAL = true).  Block i = blockIdx.x;
// launch 1: blockIdx.y == 0 runs the KW unit states (lane q: e_q), blockIdx.y
// = 1 + g the right-hand sides 16g .. 16g+15 from the zero state; launch 3:
// blockIdx.y = g the right-hand sides 16g .. from the true state.
template <typename T, int KW, int OP, bool FINAL>
__global__ void __launch_bounds__(64) part_run(PartArgs<T> a) {
    constexpr int G = (64 + KW - 1) / KW;
    constexpr int CHS = G * KW;  // steps per staged chunk (whole groups of KW)
    const int lane = threadIdx.x;
    const int n = a.n, kl = a.kl, kv = a.kv;
    const int i = blockIdx.x;
    const int js = i * a.M, je = (i == a.p - 1) ? n : js + a.M, len = je - js;
    const bool unit = !FINAL && blockIdx.y == 0;
    const int r0 = kPartNR * (FINAL ? (int)blockIdx.y : (int)blockIdx.y - 1);
    const int nr = unit ? 0 : (int)imin(kPartNR, a.nrhs - r0);
    const bool io = lane < nr;  // this lane carries right-hand side r0 + lane

    extern __shared__ __align__(16) unsigned char praw[];
    T *coef = reinterpret_cast<T *>(praw);  // [CHS][KW]
    T *ent = coef + CHS * KW;               // [CHS][kPartNR] entering values, then results
    T *wsw = ent + CHS * kPartNR;           // [KW][64] the window, for interchanges (lower factors)
    int *piv = reinterpret_cast<int *>(wsw + (POp<OP>::lower ? KW * 64 : 0));  // [CHS] interchange offsets

    T P[KW];
#pragma unroll
    for (int t = 0; t < KW; ++t) {
        T v = T(0);
        if (unit) v = (lane == t) ? T(1) : T(0);
        if (FINAL && io) v = a.cvec[((i64)i * a.nrhs + r0 + lane) * KW + t];
        P[t] = v;
    }

    for (int c0 = 0; c0 < len; c0 += CHS) {
        const int cn = imin(CHS, len - c0);
        // ---- stage the chunk: factor columns, interchanges, entering rows
        for (int e = lane; e < cn * KW; e += 64) {
            const int s = e / KW, t = e - s * KW;
            const int j = prow<OP>(js, je, c0 + s);
            const T *col = a.ab + (i64)j * a.ld + kv;
            T v = T(0);
            if constexpr (POp<OP>::lower) {
                if (t >= 1 && t <= kl && j + t < n) v = col[t];
            } else {
                if (t == 0)
                    v = col[0];
                else if (t <= kv && j - t >= 0)
                    v = col[-t];
            }
            coef[s * KW + t] = v;
        }
        if constexpr (POp<OP>::lower) {
            for (int e = lane; e < cn; e += 64) {
                const int j = prow<OP>(js, je, c0 + e);
                int d = 0;
                if (j < n - 1) {
                    const i64 q = a.ipiv[j] - 1 - j;
                    d = (q >= 0 && q <= kl) ? (int)q : 0;
                }
                piv[e] = d;
            }
        }
        if (!unit) {
            for (int e = lane; e < cn * kPartNR; e += 64) {
                const int r = e / cn, s = e - r * cn;
                const int j = prow<OP>(js, je, c0 + s);
                const int row = OP == kLN ? j + KW : OP == kUN ? j - KW : j;
                T v = T(0);
                if (r < nr && row >= 0 && row < n) v = a.bin[(i64)row + (i64)(r0 + r) * a.ldb];
                ent[s * kPartNR + r] = v;
            }
        }
        __syncthreads();
        // ---- the steps, KW at a time so the window's register names cycle
        for (int g = 0; g < cn; g += KW) {
            PUnroll<KW>::run([&](auto sc) {
                constexpr int S = decltype(sc)::value;
                const int sg = g + S;
                if (sg < cn) {
                    T cf[KW];
#pragma unroll
                    for (int t = 0; t < KW; ++t) cf[t] = coef[sg * KW + t];
                    T *slot = ent + sg * kPartNR + lane;
                    const T e = io ? *slot : T(0);
                    if constexpr (POp<OP>::axpy) {
                        // logical t sits in P[(t + S) % KW]
                        auto L = [](int t) { return (t + S) % KW; };
                        T y;
                        if constexpr (OP == kLN) {
                            const int d = __builtin_amdgcn_readfirstlane(piv[sg]);
                            if (d != 0) {
                                // interchange logical 0 and d through LDS (the
                                // registers keep compile-time names)
#pragma unroll
                                for (int t = 0; t < KW; ++t) wsw[t * 64 + lane] = P[L(t)];
                                const T xd = wsw[d * 64 + lane];
                                wsw[d * 64 + lane] = P[L(0)];
#pragma unroll
                                for (int t = 1; t < KW; ++t) P[L(t)] = wsw[t * 64 + lane];
                                P[L(0)] = xd;
                            }
                            y = P[L(0)];
                        } else {
                            y = P[L(0)] / cf[0];
                        }
                        if (FINAL && io) *slot = y;
#pragma unroll
                        for (int t = 1; t < KW; ++t) P[L(t)] = dfma(-cf[t], y, P[L(t)]);
                        P[L(0)] = e;  // the entering row, logical KW-1 from the next step on
                    } else {
                        // logical t sits in P[(t - S) mod KW]
                        auto R = [](int t) { return (t - S + KW) % KW; };
                        T acc = e;
#pragma unroll
                        for (int t = 1; t < KW; ++t) acc = dfma(-cf[t], P[R(t - 1)], acc);
                        if constexpr (OP == kUT) {
                            const T y = acc / cf[0];
                            if (FINAL && io) *slot = y;
                            P[R(KW - 1)] = y;  // logical 0 from the next step on
                        } else {
                            const int d = __builtin_amdgcn_readfirstlane(piv[sg]);
                            if (d != 0) {
                                // interchange position j (acc) and j + d (logical d-1)
                                // through LDS (the registers keep compile-time names)
#pragma unroll
                                for (int t = 0; t < KW; ++t) wsw[t * 64 + lane] = P[R(t)];
                                const T x = wsw[(d - 1) * 64 + lane];
                                wsw[(d - 1) * 64 + lane] = acc;
#pragma unroll
                                for (int t = 0; t < KW; ++t) P[R(t)] = wsw[t * 64 + lane];
                                acc = x;
                            }
                            if (FINAL && io) *slot = P[R(KW - 1)];  // position j + KW leaves the window
                            P[R(KW - 1)] = acc;
                        }
                    }
                }
            });
        }
        __syncthreads();
        if (FINAL) {
            for (int e = lane; e < cn * kPartNR; e += 64) {
                const int r = e / cn, s = e - r * cn;
                const int j = prow<OP>(js, je, c0 + s);
                const int row = OP == kLT ? j + KW : j;
                if (r < nr && row < n) a.yout[(i64)row + (i64)(r0 + r) * a.ldy] = ent[s * kPartNR + r];
            }
            __syncthreads();
        }
    }

    // ---- the window in logical order (its names are rotated by len mod KW)
    const bool need = FINAL ? (OP == kLT && i == 0) : true;
    if (need) {
        T *wsc = coef;  // [KW][64]
#pragma unroll
        for (int t = 0; t < KW; ++t) wsc[t * 64 + lane] = P[t];
        __syncthreads();
        const int rot = len % KW;
        for (int t = 0; t < KW; ++t) {
            const int ph = POp<OP>::axpy ? (t + rot) % KW : (t - rot + KW) % KW;
            const T v = wsc[ph * 64 + lane];
            if (FINAL) {
                // block 0 of the L^T sweep: positions 0 .. KW-1 are final
                if (io && t < n) a.yout[(i64)t + (i64)(r0 + lane) * a.ldy] = v;
            } else if (unit) {
                if (lane < KW) a.amat[((i64)i * KW + lane) * KW + t] = v;
            } else if (io) {
                a.dvec[((i64)i * a.nrhs + r0 + lane) * KW + t] = v;
            }
        }
    }
}

// Launch 2: w for every block, in the sweep's order.  256 threads: thread
// (t, r) keeps entry t of right-hand side r0 + r (blockIdx.x picks the group
// of RB right-hand sides); all threads together copy A_i into LDS.  The copies
// run PD blocks ahead through registers (one register set per block in
// flight, so the loop is unrolled PD times and every set has a fixed name).
template <typename T, int KW>
struct PChain {
    static constexpr int MPER = (KW * KW + 255) / 256;  // A entries per thread and block
    static constexpr int PD = KW <= 24 ? 8 : KW <= 40 ? 6 : 4;  // blocks in flight
    static constexpr int AS = KW + 2;  // LDS row stride of A^T (entries)
};

template <typename T, int KW, int OP>
__global__ void __launch_bounds__(256) part_chain(PartArgs<T> a, int RB) {
    using C = PChain<T, KW>;
    constexpr int MPER = C::MPER, PD = C::PD, AS = C::AS;
    const int tid = threadIdx.x;
    const int t = tid % KW, r = tid / KW;
    const int rr = (int)blockIdx.x * RB + r;
    const bool on = r < RB && rr < a.nrhs;
    const int n = a.n, p = a.p;
    extern __shared__ __align__(16) unsigned char craw[];
    T *at = reinterpret_cast<T *>(craw);  // [2][KW][AS]: A_i transposed, at[t][q] = A_i[q][t]
    T *cw = at + 2 * KW * AS;             // [2][RB][KW] the states
    auto blk = [&](int k) { return POp<OP>::asc ? k : p - 1 - k; };
    T stg[PD][MPER], dq[PD];
    auto issue = [&](int k, T (&dst)[MPER], T &d) {
        d = (on && k < p - 1) ? a.dvec[((i64)blk(k) * a.nrhs + rr) * KW + t] : T(0);
        if (k < p - 1) {
            const T *src = a.amat + (i64)blk(k) * KW * KW;
#pragma unroll
            for (int m = 0; m < MPER; ++m) {
                const int e = tid + 256 * m;
                dst[m] = e < KW * KW ? src[e] : T(0);
            }
        }
    };
    auto park = [&](int slot, const T (&v)[MPER]) {
#pragma unroll
        for (int m = 0; m < MPER; ++m) {
            const int e = tid + 256 * m;
            if (e < KW * KW) {
                const int q = e / KW, tt = e - q * KW;  // amat is [q][t]
                at[slot * KW * AS + tt * AS + q] = v[m];
            }
        }
    };
    T w = T(0);
    if (on) {
        if (OP == kLN && t < n) w = a.bin[(i64)t + (i64)rr * a.ldb];
        if (OP == kUN && n - 1 - t >= 0) w = a.bin[(i64)(n - 1 - t) + (i64)rr * a.ldb];
    }
    if (r < RB) cw[r * KW + t] = w;
    // A_0 straight into slot 0; A_1 .. A_PD in flight
    T dn;
    {
        T v0[MPER];
        issue(0, v0, dn);
        park(0, v0);
    }
    PUnroll<PD>::run([&](auto sc) {
        constexpr int J = decltype(sc)::value;
        issue(1 + J, stg[J], dq[J]);
    });
    __syncthreads();
    float g = 0.0f;
    const int rc = r < RB ? r : 0;
    for (int k0 = 0; k0 < p; k0 += PD) {
        bool done = false;
        PUnroll<PD>::run([&](auto sc) {
            constexpr int J = decltype(sc)::value;
            const int k = k0 + J;
            if (done || k >= p) {
                done = true;
                return;
            }
            const int i = blk(k);
            const T *cur = cw + (k & 1) * RB * KW + rc * KW;
            if (on) a.cvec[((i64)i * a.nrhs + rr) * KW + t] = cur[t];
            if (k == p - 1) {
                done = true;
                return;
            }
            // A_{k+1} (loaded PD blocks ago) into the other slot; A_{k+1+PD} into its registers
            park((k + 1) & 1, stg[J]);
            const T dnext = dq[J];
            issue(k + 1 + PD, stg[J], dq[J]);
            const T *arow = at + (k & 1) * KW * AS + t * AS;
            T acc0 = dn, acc1 = T(0), acc2 = T(0), acc3 = T(0);
#pragma unroll
            for (int q = 0; q < KW; q += 4) {
                acc0 = dfma(arow[q], cur[q], acc0);
                if (q + 1 < KW) acc1 = dfma(arow[q + 1], cur[q + 1], acc1);
                if (q + 2 < KW) acc2 = dfma(arow[q + 2], cur[q + 2], acc2);
                if (q + 3 < KW) acc3 = dfma(arow[q + 3], cur[q + 3], acc3);
            }
#pragma unroll
            for (int q = 0; q < KW; ++q) {
                const float v = __builtin_fabsf((float)arow[q]);
                if (!(v <= g) && g == g) g = v;  // a NaN stays
            }
            if (r < RB) cw[((k + 1) & 1) * RB * KW + r * KW + t] = (acc0 + acc1) + (acc2 + acc3);
            dn = dnext;
            __syncthreads();
        });
        if (done) break;
    }
    if (tid < KW * RB) atomicMax(a.growth, __builtin_bit_cast(unsigned, g));
}

static __global__ void part_clear(unsigned *g) {
    if (threadIdx.x < 2) g[threadIdx.x] = 0u;
}

constexpr int kPartKW[] = {8, 16, 24, 32, 40, 48, 64};

template <int KW>
constexpr int part_chs() {
    return ((64 + KW - 1) / KW) * KW;
}

struct PartPlan {
    int KW, M, p;
};

inline PartPlan part_plan(i64 n, int need, int maxkw) {
    PartPlan pl{0, 0, 0};
    for (int kw : kPartKW)
        if (kw >= need && kw <= maxkw) {
            pl.KW = kw;
            break;
        }
    if (pl.KW == 0) return pl;
    const int chs = ((64 + pl.KW - 1) / pl.KW) * pl.KW;
    // blocks of about sqrt(n) rows (balances the block length against the
    // chain's length), whole chunks; the last block takes the remainder
    i64 m = 1;
    while (m * m < n) ++m;
    i64 M = chs * imax(1, (m + chs / 2) / chs);
    pl.M = (int)M;
    pl.p = (int)imax(1, n / M);
    return pl;
}

template <typename T>
size_t part_bytes(i64 n, i64 kl, i64 ku, i64 nrhs) {
    const PartPlan a = part_plan(n, (int)kl + 1, 32), b = part_plan(n, (int)(kl + ku) + 1, 64);
    if (a.KW == 0 || b.KW == 0) return 0;
    auto one = [&](const PartPlan &pl) {
        return (size_t)pl.p * pl.KW * pl.KW + 2 * (size_t)pl.p * nrhs * pl.KW;
    };
    const size_t e = (one(a) > one(b) ? one(a) : one(b)) + (size_t)n * nrhs;
    return ((e * sizeof(T) + 255) & ~(size_t)255) + 256;  // + the two growth words
}

template <typename T, int KW, int OP>
int part_sweep_kw(PartArgs<T> a, bool head, hipStream_t st) {
    constexpr int CHS = part_chs<KW>();
    const size_t lds = sizeof(T) * ((size_t)CHS * KW + (size_t)CHS * kPartNR + (POp<OP>::lower ? (size_t)KW * 64 : 0)) +
                       sizeof(int) * CHS;
    static_assert(sizeof(T) * CHS * KW >= sizeof(T) * KW * 64, "window scratch fits the coefficient buffer");
    const int gr = (int)cdiv(a.nrhs, kPartNR);
    if (head) {
        if (a.p > 1) {
            hipLaunchKernelGGL((part_run<T, KW, OP, false>), dim3(a.p, 1 + gr), dim3(64), lds, st, a);
            BH_LAUNCH_CHECK();
        }
        int RB = 256 / KW;
        if (RB > a.nrhs) RB = a.nrhs;
        const size_t lc = sizeof(T) * (2 * (size_t)KW * PChain<T, KW>::AS + 2 * (size_t)RB * KW);
        hipLaunchKernelGGL((part_chain<T, KW, OP>), dim3((unsigned)cdiv(a.nrhs, RB)), dim3(256), lc, st, a, RB);
        BH_LAUNCH_CHECK();
    } else {
        hipLaunchKernelGGL((part_run<T, KW, OP, true>), dim3(a.p, gr), dim3(64), lds, st, a);
        BH_LAUNCH_CHECK();
    }
    return BH_OK;
}

// The 64-wide U sweeps are instantiated in units of their own
// (bh_gbtrs_part_up64_{d,s}.hip) to spread the compile.
#ifndef BH_PART_KW64
extern template int part_sweep_kw<double, 64, kUN>(PartArgs<double>, bool, hipStream_t);
extern template int part_sweep_kw<double, 64, kUT>(PartArgs<double>, bool, hipStream_t);
extern template int part_sweep_kw<float, 64, kUN>(PartArgs<float>, bool, hipStream_t);
extern template int part_sweep_kw<float, 64, kUT>(PartArgs<float>, bool, hipStream_t);
#endif

// launches 1 and 2 (head) or 3 of one sweep
template <typename T, int OP>
int part_sweep(const PartPlan &pl, PartArgs<T> a, bool head, hipStream_t st) {
    a.M = pl.M;
    a.p = pl.p;
    switch (pl.KW) {
        case 8:
            return part_sweep_kw<T, 8, OP>(a, head, st);
        case 16:
            return part_sweep_kw<T, 16, OP>(a, head, st);
        case 24:
            return part_sweep_kw<T, 24, OP>(a, head, st);
        case 32:
            return part_sweep_kw<T, 32, OP>(a, head, st);
        default:
            break;
    }
    if constexpr (!POp<OP>::lower) {
        switch (pl.KW) {
            // the L sweeps take windows up to 32 wide (kl <= 31)
            case 40:
                if constexpr (!POp<OP>::lower) return part_sweep_kw<T, 40, OP>(a, head, st);
                break;
            case 48:
                if constexpr (!POp<OP>::lower) return part_sweep_kw<T, 48, OP>(a, head, st);
                break;
            case 64:
                if constexpr (!POp<OP>::lower) return part_sweep_kw<T, 64, OP>(a, head, st);
                break;
            default:
                break;
        }
    }
    return BH_ERR_INTERNAL;
}

// The sweeps are instantiated in their own units (bh_gbtrs_part_{lo,up}_{d,s}.hip,
// compiled in parallel); the units that only call them see these declarations.
#ifndef BH_PART_INSTANTIATE
extern template int part_sweep<double, kLN>(const PartPlan &, PartArgs<double>, bool, hipStream_t);
extern template int part_sweep<double, kLT>(const PartPlan &, PartArgs<double>, bool, hipStream_t);
extern template int part_sweep<double, kUN>(const PartPlan &, PartArgs<double>, bool, hipStream_t);
extern template int part_sweep<double, kUT>(const PartPlan &, PartArgs<double>, bool, hipStream_t);
extern template int part_sweep<float, kLN>(const PartPlan &, PartArgs<float>, bool, hipStream_t);
extern template int part_sweep<float, kLT>(const PartPlan &, PartArgs<float>, bool, hipStream_t);
extern template int part_sweep<float, kUN>(const PartPlan &, PartArgs<float>, bool, hipStream_t);
extern template int part_sweep<float, kUT>(const PartPlan &, PartArgs<float>, bool, hipStream_t);
#endif

// Largest |entry| of the blocks' transfer matrices A_i above which the
// partitioned result is not trusted (its rounding error grows with them) and
// the sequential solve runs instead.
template <typename T>
constexpr float kPartGrowthMax = sizeof(T) == 8 ? 1e8f : 1e3f;

template <typename T>
int part_solve(char trans, i64 n, i64 kl, i64 ku, i64 nrhs, const T *ab, i64 ldab, const i64 *ipiv, T *B, i64 ldb,
               void *work, hipStream_t st) {
    const PartPlan pl = part_plan(n, (int)kl + 1, 32), pu = part_plan(n, (int)(kl + ku) + 1, 64);
    auto one = [&](const PartPlan &p) { return (size_t)p.p * p.KW * p.KW + 2 * (size_t)p.p * nrhs * p.KW; };
    const size_t e1 = one(pl) > one(pu) ? one(pl) : one(pu);
    T *w = reinterpret_cast<T *>(work);
    PartArgs<T> a{};
    a.ab = ab;
    a.ipiv = ipiv;
    a.ld = ldab;
    a.n = (int)n;
    a.kl = (int)kl;
    a.kv = (int)(kl + ku);
    a.nrhs = (int)nrhs;
    T *Y = w + e1;  // n x nrhs, the first sweep's results
    auto carve = [&](const PartPlan &p) {
        a.amat = w;
        a.dvec = w + (size_t)p.p * p.KW * p.KW;
        a.cvec = a.dvec + (size_t)p.p * nrhs * p.KW;
    };
    const bool tr = trans != 'N';
    a.growth = reinterpret_cast<unsigned *>(Y + (size_t)n * nrhs);
    hipLaunchKernelGGL(part_clear, dim3(1), dim3(64), 0, st, a.growth);
    BH_LAUNCH_CHECK();
    // first sweep: B -> Y
    a.bin = B;
    a.ldb = ldb;
    a.yout = Y;
    a.ldy = n;
    int stt = BH_OK;
    // kl = 0 (a triangular U): the L sweeps have no multipliers and no
    // interchanges (an interchange stays within rows j .. j + kl, as LAPACK's
    // xGBTRS skips L when kl = 0), so they run as plain copies and leave their
    // growth word at zero
    const bool copy_l = kl == 0;
    if (!tr && copy_l) {
        BH_HIP(hipMemcpy2DAsync(Y, (size_t)n * sizeof(T), B, (size_t)ldb * sizeof(T), (size_t)n * sizeof(T),
                                (size_t)nrhs, hipMemcpyDeviceToDevice, st));
    } else if (!tr) {
        carve(pl);
        stt = part_sweep<T, kLN>(pl, a, true, st);
        if (!stt) stt = part_sweep<T, kLN>(pl, a, false, st);
    } else {
        carve(pu);
        stt = part_sweep<T, kUT>(pu, a, true, st);
        if (!stt) stt = part_sweep<T, kUT>(pu, a, false, st);
    }
    if (stt) return stt;
    // second sweep: Y -> B, its results written only once both sweeps' growth
    // is known to be small (B still holds the right-hand sides otherwise)
    a.bin = Y;
    a.ldb = n;
    a.yout = B;
    a.ldy = ldb;
    a.growth += 1;
    const PartPlan &p2 = tr ? pl : pu;
    carve(p2);
    if (!(tr && copy_l)) stt = tr ? part_sweep<T, kLT>(p2, a, true, st) : part_sweep<T, kUN>(p2, a, true, st);
    if (stt) return stt;
    unsigned gh[2];
    BH_HIP(hipMemcpyAsync(gh, a.growth - 1, sizeof gh, hipMemcpyDeviceToHost, st));
    BH_HIP(hipStreamSynchronize(st));
    const float g0 = __builtin_bit_cast(float, gh[0]), g1 = __builtin_bit_cast(float, gh[1]);
    if (!(g0 <= kPartGrowthMax<T>) || !(g1 <= kPartGrowthMax<T>)) {
        // the blocks amplify too much for the partitioned result: the sequential solve
        if constexpr (sizeof(T) == 8)
            return bh_dgbtrs(tr ? 'T' : 'N', n, kl, ku, nrhs, (const double *)ab, ldab, ipiv, (double *)B, ldb);
        else
            return bh_sgbtrs(tr ? 'T' : 'N', n, kl, ku, nrhs, (const float *)ab, ldab, ipiv, (float *)B, ldb);
    }
    if (tr && copy_l) {
        BH_HIP(hipMemcpy2DAsync(B, (size_t)ldb * sizeof(T), Y, (size_t)n * sizeof(T), (size_t)n * sizeof(T),
                                (size_t)nrhs, hipMemcpyDeviceToDevice, st));
        return BH_OK;
    }
    return tr ? part_sweep<T, kLT>(p2, a, false, st) : part_sweep<T, kUN>(p2, a, false, st);
}

}  // namespace part

}  // namespace bh

using namespace bh;
using namespace bh::part;

template <typename T>
int gbtrs_part_impl(char trans, i64 n, i64 kl, i64 ku, i64 nrhs, const T *ab, i64 ldab, const i64 *ipiv,
                           T *B, i64 ldb, void *work, size_t lwork) {
    BH_ARG(trans == 'N' || trans == 'T' || trans == 'C', 1);
    BH_ARG(n >= 0, 2);
    BH_ARG(kl >= 0, 3);
    BH_ARG(ku >= 0, 4);
    BH_ARG(nrhs >= 0, 5);
    BH_ARG(ldab >= 2 * kl + ku + 1, 7);
    BH_ARG(ldb >= imax(1, n), 10);
    if (n == 0 || nrhs == 0) return BH_OK;
    if (n >= (1ll << 30) || nrhs >= (1ll << 20)) {
        set_error("gbtrs_part: n = %lld, nrhs = %lld outside the partitioned solve's range", (long long)n,
                  (long long)nrhs);
        return BH_ERR_WORKSPACE;
    }
    const size_t need = part_bytes<T>(n, kl, ku, nrhs);
    if (need == 0) {
        set_error("gbtrs_part: bands (%lld,%lld) outside the partitioned solve's kl <= 31, kl + ku <= 63; use bh_?gbtrs",
                  (long long)kl, (long long)ku);
        return BH_ERR_WORKSPACE;
    }
    if (work == nullptr || lwork < need) {
        set_error("gbtrs_part: workspace of %zu bytes needed, %zu given", need, lwork);
        return BH_ERR_WORKSPACE;
    }
    return part_solve<T>(trans, n, kl, ku, nrhs, ab, ldab, ipiv, B, ldb, work, stream());
}

