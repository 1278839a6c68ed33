// bh_gbtrf_reg.hip -- register-resident partial-pivot band LU (one workgroup
// per system) for gfx950.
//
// Same arithmetic as LAPACK xGBTF2 (the path of LAPACK.gbtrf!,
// src/banded/BandedLU.jl:98): per column j, IDAMAX pivot (first index of the
// largest |a|), interchange, scale by the rounded reciprocal of the pivot,
// rank-1 update a = fma(l, -u, a) over columns j+1..ju.  Factors and pivots
// are therefore bit-identical to that routine.
//
// Layout of the live window (rows j..j+kl x columns j..j+kl+ku) in registers:
//   * threads form an RT x CT grid, tid = ct*RT + rt;
//   * ROW SLOTS: thread rt holds row slots rs = rt + i*RT (i < RR).  A slot
//     carries one matrix row (its "position" pos[i]); a row interchange is a
//     relabelling of two slots, so no data moves.  When row j retires its slot
//     takes the row entering at the bottom of the window (j+kl+1);
//   * COLUMN RING: thread ct holds column slots cs = ct + k*CT (k < CC); column
//     c lives in slot c mod W, W = kl+ku.  At step j column j leaves its slot
//     (it is the pivot column) and column j+kl+ku enters the same slot.
// Global memory is touched only at chunk boundaries (every D steps): the
// retired U rows, multipliers and pivots of the last chunk, buffered in LDS,
// are flushed as contiguous per-column segments, and the rows entering during
// the next chunk -- gathered from AB one chunk ahead into registers -- are
// written to an LDS ring.  Inside a chunk the steps use LDS and registers
// only, and barriers order LDS only, so no step waits on HBM latency.
#include "bh_common.hpp"

namespace bh {

template <typename T>
struct RegLuArgs {
    T *ab;
    i64 *ipiv;
    i64 *info;
    i64 m, n, kl, ku, ldab;
};

// Copy column register KK of every row slot (compile-time index after the
// switch below, so no select chains).
template <int KK, typename T, int RR, int CC>
__device__ __forceinline__ void col_get(const T (&a)[RR][CC], T (&v)[RR]) {
#pragma unroll
    for (int i = 0; i < RR; ++i) v[i] = a[i][KK];
}
template <int KK, typename T, int RR, int CC>
__device__ __forceinline__ void col_set(T (&a)[RR][CC], const T (&v)[RR]) {
#pragma unroll
    for (int i = 0; i < RR; ++i) a[i][KK] = v[i];
}
template <int I, typename T, int RR, int CC>
__device__ __forceinline__ void row_get(const T (&a)[RR][CC], T (&v)[CC]) {
#pragma unroll
    for (int k = 0; k < CC; ++k) v[k] = a[I][k];
}
template <int I, typename T, int RR, int CC>
__device__ __forceinline__ void row_set(T (&a)[RR][CC], const T (&v)[CC]) {
#pragma unroll
    for (int k = 0; k < CC; ++k) a[I][k] = v[k];
}
// Uniform-index dispatch: one switch case per register index.
#define BH_DISPATCH16(idx, CALL)                                         \
    switch (idx) {                                                        \
        case 0: CALL(0); break;   case 1: CALL(1); break;                \
        case 2: CALL(2); break;   case 3: CALL(3); break;                \
        case 4: CALL(4); break;   case 5: CALL(5); break;                \
        case 6: CALL(6); break;   case 7: CALL(7); break;                \
        case 8: CALL(8); break;   case 9: CALL(9); break;                \
        case 10: CALL(10); break; case 11: CALL(11); break;              \
        case 12: CALL(12); break; case 13: CALL(13); break;              \
        case 14: CALL(14); break; case 15: CALL(15); break;              \
        default: break;                                                   \
    }

// Argmax over the RT consecutive lanes of a group (RT <= 16, groups aligned in
// a DPP row): larger |value| wins, equal values keep the smaller position.
// After the scan the group's last lane holds the winner.
template <int RT, typename T>
__device__ __forceinline__ void group_argmax(T &best, int &bpos, int &bslot, T &bval) {
    const int g = (int)__lane_id() % RT;
#define BH_ARGMAX_STEP(K, CTRL)                                                             \
    if (RT > K) {                                                                           \
        const T ob = dpp_mov<CTRL>(best, best), ov = dpp_mov<CTRL>(bval, bval);             \
        const int op = __builtin_amdgcn_update_dpp(bpos, bpos, CTRL, 0xF, 0xF, false);     \
        const int os = __builtin_amdgcn_update_dpp(bslot, bslot, CTRL, 0xF, 0xF, false);   \
        if (g >= K && (ob > best || (ob == best && op < bpos))) {                           \
            best = ob; bpos = op; bslot = os; bval = ov;                                    \
        }                                                                                   \
    }
    BH_ARGMAX_STEP(1, 0x111)
    BH_ARGMAX_STEP(2, 0x112)
    BH_ARGMAX_STEP(4, 0x114)
    BH_ARGMAX_STEP(8, 0x118)
#undef BH_ARGMAX_STEP
}

// D = steps per I/O chunk: long enough to hide a global load behind a chunk
// of steps, short enough that the staged rows fit in registers.
template <typename T, int RT, int CT, int RR, int CC, int D>
__device__ void reg_gbtf2(const RegLuArgs<T> &s) {
    constexpr int NT = RT * CT;
    constexpr int RB = D + 1;                                 // entering-row ring (rows)
    constexpr int NPRE = (RB * (CT * CC + 1) + NT - 1) / NT;  // prefetch registers per thread
    const i64 m = s.m, n = s.n, kl = s.kl, ku = s.ku, kv = kl + ku, ld = s.ldab;
    T *__restrict__ ab = s.ab;
    const int W = (int)kv;  // ring size (kv >= 1 is required by the host)
    const int H = W + 1;    // entries per staged row / U row
    const int tid = threadIdx.x;
    const int rt = tid % RT, ct = tid / RT;

    __shared__ T lbuf[2][RT * RR];
    __shared__ T ubuf[2][CT * CC];
    __shared__ T sh_pv[2];
    __shared__ int sh_p[2], sh_ps[2];
    __shared__ i64 outP[D];
    extern __shared__ __align__(16) unsigned char dyn_raw[];
    T *rows = reinterpret_cast<T *>(dyn_raw);  // [RB][H]: row e at (e mod RB)
    T *outU = rows + RB * H;                    // [D][H]: U row j at (j mod D)
    T *outL = outU + D * H;                     // [D][kl+1]: multipliers of column j

    const i64 mn = imin(m, n);
    if (mn == 0) {
        if (tid == 0) *s.info = 0;
        return;
    }

    // entry t of row e (columns e-kl .. e+ku): original band data
    auto row_entry = [&](i64 e, i64 t) -> T {
        const i64 c = e - kl + t;
        if (e >= m || c >= n || c < 0) return T(0);
        return ab[c * ld + (kv + kl - t)];
    };
    // rows staged for chunk q: e in [q*D + kl, q*D + D + kl]
    T pre[NPRE];
    auto prefetch = [&](i64 q) {
#pragma unroll
        for (int r = 0; r < NPRE; ++r) {
            const int f = tid + r * NT;  // flat index over RB x H
            T v = T(0);
            if (f < RB * H) {
                const int rr = f / H, t = f - rr * H;
                v = row_entry(q * D + kl + rr, t);
            }
            pre[r] = v;
        }
    };
    auto commit = [&](i64 q) {
#pragma unroll
        for (int r = 0; r < NPRE; ++r) {
            const int f = tid + r * NT;
            if (f < RB * H) {
                const int rr = f / H, t = f - rr * H;
                const i64 e = q * D + kl + rr;
                rows[(int)(e % RB) * H + t] = pre[r];
            }
        }
    };
    // write the U rows, multipliers and pivots of steps [j0, j1) to AB/ipiv
    auto flush = [&](i64 j0, i64 j1) {
        const i64 cnt = j1 - j0;
        if (cnt <= 0) return;
        // U: for column c the rows j in [j0, j1) with 0 <= c - j <= kv are
        // consecutive band rows kv + j - c: contiguous per column
        const i64 ncols = cnt + kv;
        for (i64 f = tid; f < ncols * cnt; f += NT) {
            const i64 cc = f / cnt, jj = f - cc * cnt;
            const i64 c = j0 + cc, j = j0 + jj, t = c - j;
            if (t >= 0 && t <= kv && c < n) ab[c * ld + (kv - t)] = outU[(int)(j % D) * H + (int)t];
        }
        if (kl > 0) {
            for (i64 f = tid; f < cnt * kl; f += NT) {
                const i64 jj = f / kl, t = f - jj * kl + 1;
                const i64 j = j0 + jj;
                if (t <= imin(kl, m - 1 - j)) ab[j * ld + kv + t] = outL[(int)(j % D) * (kl + 1) + (int)t];
            }
        }
        for (i64 f = tid; f < cnt; f += NT) s.ipiv[j0 + f] = outP[(int)((j0 + f) % D)];
    };

    T a[RR][CC];
    int pos[RR];
    int cs_[CC];
#pragma unroll
    for (int k = 0; k < CC; ++k) cs_[k] = ct + k * CT;

    // ---- initial window: positions 0..min(kl,m-1), columns 0..min(kv,n)-1
#pragma unroll
    for (int i = 0; i < RR; ++i) {
        const int rs = rt + i * RT;
        pos[i] = (rs <= kl && rs < m) ? rs : -1;
#pragma unroll
        for (int k = 0; k < CC; ++k) {
            const i64 c = cs_[k];
            T v = T(0);
            if (pos[i] >= 0 && cs_[k] < W && c < n) {
                const i64 q = kv + pos[i] - c;
                if (q >= kl && q <= kv + kl) v = ab[c * ld + q];
            }
            a[i][k] = v;
        }
    }
    prefetch(0);
    commit(0);
    prefetch(1);
    lds_barrier();

    i64 ju = 0;
    i64 info = 0;
    int jmodW = 0;  // j mod W
    for (i64 j = 0; j < mn; ++j) {
        if (j > 0 && j % D == 0) {
            // chunk boundary: stage rows of chunk q (loaded a chunk ago), flush
            // the previous chunk, start loading chunk q+1
            const i64 q = j / D;
            lds_barrier();
            commit(q);
            flush(j - D, j);
            prefetch(q + 1);
            lds_barrier();
        }
        const int buf = (int)(j & 1);
        const int jd = (int)(j % D);
        const int j32 = (int)j;
        const int km = (int)imin(kl, m - 1 - j);
        const int ctj = jmodW % CT, kkj = jmodW / CT;
        // ------------------------------------------------------------ phase A
        // owners of column j: pivot search, multipliers, recycle the slot
        if (ct == ctj) {
            const int kk = __builtin_amdgcn_readfirstlane(kkj);
            T colv[RR];
#define BH_COLGET(K) if (K < CC) col_get<(K < CC ? K : 0)>(a, colv)
            BH_DISPATCH16(kk, BH_COLGET)
#undef BH_COLGET
            T best = T(-1);
            int bpos = 0x7fffffff, bslot = -1;
            T bval = T(0);
#pragma unroll
            for (int i = 0; i < RR; ++i) {
                const int p = pos[i];
                if (p >= j32 && p <= j32 + km) {
                    const T cv = colv[i];
                    T v = dabs(cv);
                    if (!(v == v)) v = T(-0.5);  // NaN: a candidate that loses to any number
                    if (v > best || (v == best && p < bpos)) {
                        best = v;
                        bpos = p;
                        bslot = rt + i * RT;
                        bval = cv;
                    }
                }
            }
            group_argmax<RT>(best, bpos, bslot, bval);
            {
                // the group's result sits in its last lane; broadcast it
                const int last = __builtin_amdgcn_readfirstlane(((ctj * RT) & 63) + RT - 1);
                bpos = __builtin_amdgcn_readlane(bpos, last);
                bslot = __builtin_amdgcn_readlane(bslot, last);
                bval = lane_read(bval, last);
            }
            const int p = bpos, ps = bslot;
            const T pv = bval;
            const bool nz = (pv != T(0));
            const T r = nz ? T(1) / pv : T(0);
            // multipliers (positions after the interchange)
#pragma unroll
            for (int i = 0; i < RR; ++i) {
                const int rs = rt + i * RT;
                int np = pos[i];
                if (nz && rs != ps && np == j32) np = p;
                if (rs != ps && np > j32 && np <= j32 + km) {
                    const T cv = colv[i];
                    const T l = nz ? cv * r : cv;
                    lbuf[buf][rs] = l;
                    outL[jd * (kl + 1) + (np - j32)] = l;
                }
            }
            if (rt == 0) {
                sh_p[buf] = p;
                sh_ps[buf] = ps;
                sh_pv[buf] = pv;
                outP[jd] = (i64)p + 1;
                outU[jd * H] = pv;
            }
            // column j leaves the slot; column j+kv enters it (fill-in zeros
            // except the entry of row j+kl, staged with that row)
            const bool newcol = (j + kv < n);
            const T ev = rows[(int)((j + kl) % RB) * H + W];
#pragma unroll
            for (int i = 0; i < RR; ++i) colv[i] = (newcol && pos[i] == j32 + (int)kl) ? ev : T(0);
#define BH_COLSET(K) if (K < CC) col_set<(K < CC ? K : 0)>(a, colv)
            BH_DISPATCH16(kk, BH_COLSET)
#undef BH_COLSET
        }
        lds_barrier();  // B1
        // ------------------------------------------------------------ phase B
        const int p = sh_p[buf];
        const int ps = sh_ps[buf];
        const T pv = sh_pv[buf];
        const bool nz = (pv != T(0));
        if (nz) {
            ju = imax(ju, imin((i64)p + ku, n - 1));
            if (p != j) {
#pragma unroll
                for (int i = 0; i < RR; ++i) {
                    const int rs = rt + i * RT;
                    if (rs != ps && pos[i] == j32) pos[i] = p;
                }
            }
        } else if (info == 0) {
            info = j + 1;
        }
        // the pivot's slot now holds position j (retiring row)
        const int jm1 = (jmodW + 1 == W) ? 0 : jmodW + 1;  // (j+1) mod W
        int colk[CC];
#pragma unroll
        for (int k = 0; k < CC; ++k) {
            int d = cs_[k] - jm1;
            if (d < 0) d += W;
            colk[k] = (cs_[k] < W) ? (int)(j + 1 + d) : 0x7fffffff;  // column in slot, in [j+1, j+kv]
        }
        if (rt == ps % RT) {
            const int i0 = __builtin_amdgcn_readfirstlane(ps / RT);
            T urow[CC];
#define BH_ROWGET(I) if (I < RR) row_get<(I < RR ? I : 0)>(a, urow)
            BH_DISPATCH16(i0, BH_ROWGET)
#undef BH_ROWGET
#pragma unroll
            for (int k = 0; k < CC; ++k) {
                const T u = urow[k];
                const i64 c = colk[k];
                if (c < n) {
                    outU[jd * H + (int)(c - j)] = u;
                    if (nz && c <= ju) ubuf[buf][cs_[k]] = u;
                }
            }
        }
        lds_barrier();  // B2
        // ------------------------------------------------------------ phase C
        if (nz) {
            T uk[CC];
            bool ck[CC];
#pragma unroll
            for (int k = 0; k < CC; ++k) {
                ck[k] = colk[k] <= ju;
                uk[k] = ck[k] ? -ubuf[buf][cs_[k]] : T(0);
            }
#pragma unroll
            for (int i = 0; i < RR; ++i) {
                const int pi = pos[i];
                const int rs = rt + i * RT;
                if (rs != ps && pi > j32 && pi <= j32 + km) {
                    const T l = lbuf[buf][rs];
#pragma unroll
                    for (int k = 0; k < CC; ++k)
                        if (ck[k]) a[i][k] = dfma(l, uk[k], a[i][k]);
                }
            }
        }
        // retire row j: its slot takes row e = j+kl+1
        if (rt == ps % RT) {
            const int i0 = __builtin_amdgcn_readfirstlane(ps / RT);
            const i64 e = j + kl + 1;
            const T *rbe = rows + (int)(e % RB) * H;
            T nv[CC];
#pragma unroll
            for (int k = 0; k < CC; ++k) {
                const i64 c = colk[k];
                nv[k] = (e < m && c < n) ? rbe[c - j - 1] : T(0);
            }
#pragma unroll
            for (int i = 0; i < RR; ++i)
                if (i0 == i) pos[i] = (e < m) ? (int)e : -1;
#define BH_ROWSET(I) if (I < RR) row_set<(I < RR ? I : 0)>(a, nv)
            BH_DISPATCH16(i0, BH_ROWSET)
#undef BH_ROWSET
        }
        jmodW = jm1;
    }
    lds_barrier();
    flush(((mn - 1) / D) * D, mn);
    // Fill-in rows LAPACK zeroes that no U row covered (rows >= m).
    if (m < n + kl) {
        for (i64 c = ku + 1; c < imin(n, mn + kv); ++c) {
            const i64 q0 = imax(0, kv - c);
            for (i64 q = q0 + tid; q < kl; q += NT) {
                const i64 r = c - kv + q;
                if (r >= m) ab[c * ld + q] = T(0);
            }
        }
    }
    if (tid == 0) *s.info = info;
}

template <typename T, int RT, int CT, int RR, int CC, int D>
__global__ void __launch_bounds__(RT * CT) reg_gbtf2_kernel(RegLuArgs<T> s) {
    reg_gbtf2<T, RT, CT, RR, CC, D>(s);
}

// Instances: (RT, CT, RR, CC) -> capacity kl+1 <= RT*RR, kl+ku <= CT*CC.
struct RegShape {
    int RT, CT, RR, CC;
};
static const RegShape kRegShapes[] = {
    {8, 8, 1, 1},    // 64 thr:  kl <= 7,   kv <= 8
    {4, 16, 3, 1},   // 64 thr:  kl <= 11,  kv <= 16
    {8, 32, 3, 1},   // 256 thr: kl <= 23,  kv <= 32
    {8, 32, 5, 2},   // 256 thr: kl <= 39,  kv <= 64
    {16, 32, 5, 4},  // 512:     kl <= 79,  kv <= 128
    {16, 32, 9, 8},  // 512:     kl <= 143, kv <= 256
};

// Dynamic LDS: entering-row ring + U rows + multipliers.
template <typename T>
static size_t reg_lds_bytes(i64 kl, i64 ku, int D) {
    const i64 H = kl + ku + 1;
    return ((size_t)(D + 1) * H + (size_t)D * H + (size_t)D * (kl + 1)) * sizeof(T);
}

template <typename T, int RT, int CT, int RR, int CC, int D>
static int reg_launch_one(const RegLuArgs<T> &a, hipStream_t st) {
    const size_t lds = reg_lds_bytes<T>(a.kl, a.ku, D);
    auto kern = reg_gbtf2_kernel<T, RT, CT, RR, CC, D>;
    if (lds > 64 * 1024)
        BH_HIP(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL(kern, dim3(1), dim3(RT * CT), lds, st, a);
    BH_LAUNCH_CHECK();
    return BH_OK;
}

template <typename T>
static int reg_launch(int which, const RegLuArgs<T> &a, hipStream_t st) {
    switch (which) {
        case 0: return reg_launch_one<T, 8, 8, 1, 1, 32>(a, st);
        case 1: return reg_launch_one<T, 4, 16, 3, 1, 32>(a, st);
        case 2: return reg_launch_one<T, 8, 32, 3, 1, 32>(a, st);
        case 3: return reg_launch_one<T, 8, 32, 5, 2, 16>(a, st);
        case 4: return reg_launch_one<T, 16, 32, 5, 4, 16>(a, st);
        case 5: return reg_launch_one<T, 16, 32, 9, 8, 8>(a, st);
        default: return -1;
    }
}

// Smallest instance that holds the window, or -1 (caller falls back).
int reg_shape_for(i64 kl, i64 ku) {
    const i64 kv = kl + ku;
    if (kv < 1) return -1;
    for (int w = 0; w < (int)(sizeof(kRegShapes) / sizeof(kRegShapes[0])); ++w) {
        const RegShape &r = kRegShapes[w];
        if (kl + 1 <= (i64)r.RT * r.RR && kv <= (i64)r.CT * r.CC) return w;
    }
    return -1;
}

template <typename T>
int reg_gbtrf_launch(i64 m, i64 n, i64 kl, i64 ku, T *ab, i64 ldab, i64 *ipiv, i64 *dinfo, hipStream_t st) {
    const int w = reg_shape_for(kl, ku);
    if (w < 0) return -1;
    RegLuArgs<T> a{ab, ipiv, dinfo, m, n, kl, ku, ldab};
    return reg_launch<T>(w, a, st);
}

template int reg_gbtrf_launch<double>(i64, i64, i64, i64, double *, i64, i64 *, i64 *, hipStream_t);
template int reg_gbtrf_launch<float>(i64, i64, i64, i64, float *, i64, i64 *, i64 *, hipStream_t);

}  // namespace bh
