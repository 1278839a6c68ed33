// bh_gbtrs_ring.hpp -- band LU solve 'N' and 'T' with one wave per right-hand
// side, the vector held in a lane ring (no per-step lane shifts).  The
// launchers are instantiated one per unit (bh_gbtrs_ring_{n,t}_{d,s}.hip) so
// the fully unrolled kernels compile in parallel.
#pragma once
//
// Restates LAPACK xGBTRS 'N' (ldiv!(::BandedLU, B), src/banded/linalg.jl:24-30)
// with its rounding, so results are bit-identical:
//   forward  (j = 0 .. n-2): swap b_j <-> b_p (p = ipiv_j - 1),
//            b_{j+t} = fma(l_{j+t,j}, -b_j, b_{j+t}), t = 1 .. min(kl, n-1-j);
//   backward (j = n-1 .. 0): x_j = b_j / u_jj (a true division),
//            b_{j-t} = fma(u_{j-t,j}, -x_j, b_{j-t}), t = 1 .. min(kl+ku, j).
//
// Layout: entry e of b lives in lane e mod 64 of a register that holds one
// 64-entry block.  The wave walks the steps block by block (64 steps,
// unrolled, so lane and register of every operand are compile-time); the
// registers rotate once per block (the finished block is stored, the block
// entering the window is loaded one block ahead).  The factor columns are
// read straight from AB into registers PFD steps ahead of their use -- a lone
// wave on its CU can afford PFD * (window registers) VGPRs, which is what
// hides the memory latency; the step itself is register-only.
#include "bh_common.hpp"

namespace bh {

namespace {

template <int K>
struct SC {
    static constexpr int value = K;
};
template <int N>
struct Unroll {
    template <class F>
    __device__ __forceinline__ static void run(F &&f) {
        Unroll<N - 1>::run(f);
        f(SC<N - 1>{});
    }
};
template <>
struct Unroll<0> {
    template <class F>
    __device__ __forceinline__ static void run(F &&) {}
};

__device__ __forceinline__ int mini(int a, int b) { return a < b ? a : b; }

constexpr int kRingWord3 = 0x00020000;  // word 3 of a raw buffer resource (gfx9 family)
// A byte offset past the end of any AB the ring kernel takes (those are under
// 2 GiB - 16 MiB), with room for +-64 * 5 entries either side.
constexpr int kRingPastEnd = 0x7f800000;

__device__ __forceinline__ __amdgpu_buffer_rsrc_t ring_rsrc(const void *p, i64 bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), 0, (int)bytes, kRingWord3);
}
__device__ __forceinline__ double ring_ld(__amdgpu_buffer_rsrc_t r, int off, double) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
}
__device__ __forceinline__ float ring_ld(__amdgpu_buffer_rsrc_t r, int off, float) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0));
}
__device__ __forceinline__ void ring_st(__amdgpu_buffer_rsrc_t r, int off, double v) {
    typedef unsigned v2u __attribute__((ext_vector_type(2)));
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(v2u, v), r, off, 0, 0);
}
__device__ __forceinline__ void ring_st(__amdgpu_buffer_rsrc_t r, int off, float v) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, off, 0, 0);
}
// NF = 1 + ceil(kl/64) forward registers, NB = 1 + ceil((kl+ku)/64) backward
// registers, PFD = prefetch distance in steps (divides 64).  The window
// registers are only ever indexed with compile-time indices (an interchange
// selects with uniform conditions and replaces single lanes), and all memory
// goes through buffer resources: a lane with nothing to load passes an
// offset past the end and reads 0, so no access sits in a branch.
template <typename T, int NF, int NB, int PFD>
__global__ void __launch_bounds__(64) gbtrs_n_ring(int n, int kl, int ku, const T *__restrict__ ab, i64 ld,
                                                   const i64 *__restrict__ ipiv, T *B, i64 ldb) {
    static_assert(64 % PFD == 0, "prefetch distance divides the block");
    constexpr int SZ = (int)sizeof(T);
    const int lane = threadIdx.x;
    T *b = B + (i64)blockIdx.x * ldb;
    const int kd = kl + ku;
    const __amdgpu_buffer_rsrc_t rab = ring_rsrc(ab, ld * (i64)n * SZ);
    const __amdgpu_buffer_rsrc_t rb = ring_rsrc(b, (i64)n * SZ);
    const __amdgpu_buffer_rsrc_t rip = ring_rsrc(ipiv, (i64)n * 8);
    auto bload = [&](int e) { return ring_ld(rb, (e >= 0 && e < n) ? e * SZ : -1, T()); };

    // ------------------------------------------------------------ forward: L and pivots
    if (kl > 0 && n > 1) {
        const int nsteps = n - 1;
        const int G = (nsteps + 63) >> 6;
        T w[NF];  // w[r]: block g + r
#pragma unroll
        for (int r = 0; r < NF; ++r) w[r] = bload(64 * r + lane);
        T lp[PFD][NF];
        // multipliers of step j for window register r, lane = entry j + t
        // Lanes outside rows 1..lm load whatever the band storage holds there
        // (the update below is masked); only offsets off either end of AB
        // matter, and those read 0.  A step past the end gets a base offset
        // past the end of AB.
        auto fetch = [&](int slot, int j, int s, int ln) {
            const int base = j < nsteps ? (int)((i64)j * ld + kd) * SZ : kRingPastEnd;
#pragma unroll
            for (int r = 0; r < NF; ++r) lp[slot][r] = ring_ld(rab, base + (64 * r + ln - s) * SZ, T());
        };
        Unroll<PFD>::run([&](auto sc) {
            constexpr int s = decltype(sc)::value;
            fetch(s, s, s, lane);
        });
        auto piv = [&](int e) { return (int)__builtin_amdgcn_raw_buffer_load_b32(rip, e < nsteps ? e * 8 : -1, 0, 0) - 1; };
        int pvc = piv(lane);
        for (int g = 0; g < G; ++g) {
            const int j0 = 64 * g;
            // the lane index goes through an empty asm once per block, so the
            // per-step lane offsets are recomputed (one add) instead of being
            // hoisted out of the loop as 64 * NF live registers
            int ln = lane;
            asm volatile("" : "+v"(ln));
            const int pvn = piv(j0 + 64 + lane);
            const T bnext = bload(64 * (g + NF) + lane);
            Unroll<64>::run([&](auto sc) {
                constexpr int s = decltype(sc)::value;
                constexpr int slot = s % PFD;
                const int j = j0 + s;
                if (j < nsteps) {
                    const int lm = mini(kl, n - 1 - j);
                    const int p = __builtin_amdgcn_readlane(pvc, s);
                    if (p != j) {
                        const int pos = s + (p - j);
                        const int rr = pos >> 6, pl = pos & 63;
                        T wsel = w[0];
#pragma unroll
                        for (int r = 1; r < NF; ++r)
                            if (rr == r) wsel = w[r];
                        const T bp = lane_read(wsel, pl);
                        const T bj0 = lane_read(w[0], s);
                        w[0] = lane_write(w[0], bp, s);
#pragma unroll
                        for (int r = 0; r < NF; ++r)
                            if (r == rr) w[r] = lane_write(w[r], bj0, pl);
                    }
                    const T nb = -lane_read(w[0], s);
#pragma unroll
                    for (int r = 0; r < NF; ++r) {
                        const int t = 64 * r + ln - s;
                        const T u = dfma(lp[slot][r], nb, w[r]);
                        w[r] = (t >= 1 && t <= lm) ? u : w[r];
                    }
                }
                fetch(slot, j + PFD, (s + PFD) & 63, ln);
            });
            ring_st(rb, (j0 + lane < n) ? (j0 + lane) * SZ : -1, w[0]);
#pragma unroll
            for (int r = 0; r + 1 < NF; ++r) w[r] = w[r + 1];
            w[NF - 1] = bnext;
            pvc = pvn;
        }
#pragma unroll
        for (int r = 0; r < NF; ++r) {
            const int e = 64 * (G + r) + lane;
            ring_st(rb, e < n ? e * SZ : -1, w[r]);
        }
    }

    // the back substitution reads what the forward pass stored
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    // ------------------------------------------------------------ backward: U, kl+ku super-diagonals
    {
        const int G0 = (n - 1) >> 6;
        T w[NB];  // w[r]: block g - r
#pragma unroll
        for (int r = 0; r < NB; ++r) w[r] = bload(64 * (G0 - r) + lane);
        T up[PFD][NB];
        // column j of U for window register r, lane = entry j - t (t = 0: diagonal)
        auto fetch = [&](int slot, int j, int s, int ln) {
            const int base = (j >= 0 && j < n) ? (int)((i64)j * ld + kd) * SZ : kRingPastEnd;
#pragma unroll
            for (int r = 0; r < NB; ++r) up[slot][r] = ring_ld(rab, base - (s - ln + 64 * r) * SZ, T());
        };
        Unroll<PFD>::run([&](auto sc) {
            constexpr int s = 63 - decltype(sc)::value;
            fetch(s % PFD, 64 * G0 + s, s, lane);
        });
        for (int g = G0; g >= 0; --g) {
            const int j0 = 64 * g;
            int ln = lane;
            asm volatile("" : "+v"(ln));
            const T bnext = bload(64 * (g - NB) + lane);
            Unroll<64>::run([&](auto sc) {
                constexpr int s = 63 - decltype(sc)::value;
                constexpr int slot = s % PFD;
                const int j = j0 + s;
                if (j <= n - 1) {
                    const T ujj = lane_read(up[slot][0], s);
                    const T xj = lane_read(w[0], s) / ujj;
                    w[0] = lane_write(w[0], xj, s);
                    const T nx = -xj;
                    const int tm = mini(kd, j);
#pragma unroll
                    for (int r = 0; r < NB; ++r) {
                        const int t = s - ln + 64 * r;
                        const T u = dfma(up[slot][r], nx, w[r]);
                        w[r] = (t >= 1 && t <= tm) ? u : w[r];
                    }
                }
                fetch(slot, j - PFD, (s - PFD) & 63, ln);
            });
            ring_st(rb, (j0 + lane < n) ? (j0 + lane) * SZ : -1, w[0]);
#pragma unroll
            for (int r = 0; r + 1 < NB; ++r) w[r] = w[r + 1];
            w[NB - 1] = bnext;
        }
    }
}

// 'T' solve (transpose(F) \ B, src/banded/linalg.jl:41-47): U' y = b with
// j ascending, x_j = (b_j - sum_t u_{j-t,j} x_{j-t}) / u_jj, then L' with the
// interchanges in reverse, b_j -= sum_t l_{j+t,j} b_{j+t} followed by the swap
// b_j <-> b_{ipiv_j}.  Each step is a dot product over the band entries of
// one factor column (contiguous in band storage, so one coalesced load per
// window register) with the finished entries held in the lane ring, summed
// across the wave with DPP.  The dots are summed in a different order than
// the CPU routine's, so results agree to rounding, not bit for bit.
template <typename T, int NF, int NB, int PFD>
__global__ void __launch_bounds__(64) gbtrs_t_ring(int n, int kl, int ku, const T *__restrict__ ab, i64 ld,
                                                   const i64 *__restrict__ ipiv, T *B, i64 ldb) {
    static_assert(64 % PFD == 0, "prefetch distance divides the block");
    constexpr int SZ = (int)sizeof(T);
    const int lane = threadIdx.x;
    T *b = B + (i64)blockIdx.x * ldb;
    const int kd = kl + ku;
    const __amdgpu_buffer_rsrc_t rab = ring_rsrc(ab, ld * (i64)n * SZ);
    const __amdgpu_buffer_rsrc_t rb = ring_rsrc(b, (i64)n * SZ);
    const __amdgpu_buffer_rsrc_t rip = ring_rsrc(ipiv, (i64)n * 8);
    auto bload = [&](int e) { return ring_ld(rb, (e >= 0 && e < n) ? e * SZ : -1, T()); };

    // ------------------------------------------------------------ U' (ascending)
    {
        const int G = (n + 63) >> 6;
        T P[NB];  // P[r]: block g - r; entries below j are finished
#pragma unroll
        for (int r = 0; r < NB; ++r) P[r] = bload(64 * (0 - r) + lane);
        T up[PFD][NB];
        // column j of U: lane L of register r holds u_{64(g-r)+L, j}, at t = s - L + 64r above the diagonal
        auto fetch = [&](int slot, int j, int s, int ln) {
            const int base = j < n ? (int)((i64)j * ld + kd) * SZ : kRingPastEnd;
#pragma unroll
            for (int r = 0; r < NB; ++r) up[slot][r] = ring_ld(rab, base - (s - ln + 64 * r) * SZ, T());
        };
        Unroll<PFD>::run([&](auto sc) {
            constexpr int s = decltype(sc)::value;
            fetch(s, s, s, lane);
        });
        for (int g = 0; g < G; ++g) {
            const int j0 = 64 * g;
            int ln = lane;
            asm volatile("" : "+v"(ln));
            const T bnext = bload(64 * (g + 1) + lane);
            Unroll<64>::run([&](auto sc) {
                constexpr int s = decltype(sc)::value;
                constexpr int slot = s % PFD;
                const int j = j0 + s;
                if (j < n) {
                    const int tm = mini(kd, j);
                    T part = T(0);
#pragma unroll
                    for (int r = 0; r < NB; ++r) {
                        const int t = s - ln + 64 * r;
                        const T q = dfma(up[slot][r], P[r], part);
                        part = (t >= 1 && t <= tm) ? q : part;
                    }
                    const T dot = wave_sum_dpp(part);
                    const T xj = (lane_read(P[0], s) - dot) / lane_read(up[slot][0], s);
                    P[0] = lane_write(P[0], xj, s);
                }
                fetch(slot, j + PFD, (s + PFD) & 63, ln);
            });
            ring_st(rb, (j0 + lane < n) ? (j0 + lane) * SZ : -1, P[0]);
#pragma unroll
            for (int r = NB - 1; r >= 1; --r) P[r] = P[r - 1];
            P[0] = bnext;
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    // ------------------------------------------------------------ L' and the interchanges (descending)
    if (kl > 0 && n > 1) {
        const int g0 = (n - 2) >> 6;
        T F[NF];  // F[r]: block g + r; entries above j are finished
#pragma unroll
        for (int r = 0; r < NF; ++r) F[r] = bload(64 * (g0 + r) + lane);
        T lp[PFD][NF];
        // column j of L: lane L of register r holds l_{64(g+r)+L, j}, at t = 64r + L - s below the diagonal
        auto fetch = [&](int slot, int j, int s, int ln) {
            const int base = (j >= 0 && j < n - 1) ? (int)((i64)j * ld + kd) * SZ : kRingPastEnd;
#pragma unroll
            for (int r = 0; r < NF; ++r) lp[slot][r] = ring_ld(rab, base + (64 * r + ln - s) * SZ, T());
        };
        Unroll<PFD>::run([&](auto sc) {
            constexpr int s = 63 - decltype(sc)::value;
            fetch(s % PFD, 64 * g0 + s, s, lane);
        });
        auto piv = [&](int e) {
            return (int)__builtin_amdgcn_raw_buffer_load_b32(rip, (e >= 0 && e < n - 1) ? e * 8 : -1, 0, 0) - 1;
        };
        for (int g = g0; g >= 0; --g) {
            const int j0 = 64 * g;
            int ln = lane;
            asm volatile("" : "+v"(ln));
            const int pvb = piv(j0 + lane);
            const T bprev = bload(64 * (g - 1) + lane);
            Unroll<64>::run([&](auto sc) {
                constexpr int s = 63 - decltype(sc)::value;
                constexpr int slot = s % PFD;
                const int j = j0 + s;
                if (j <= n - 2) {
                    const int lm = mini(kl, n - 1 - j);
                    T part = T(0);
#pragma unroll
                    for (int r = 0; r < NF; ++r) {
                        const int t = 64 * r + ln - s;
                        const T q = dfma(lp[slot][r], F[r], part);
                        part = (t >= 1 && t <= lm) ? q : part;
                    }
                    const T dot = wave_sum_dpp(part);
                    T bj = lane_read(F[0], s) - dot;
                    const int p = __builtin_amdgcn_readlane(pvb, s);
                    if (p != j) {
                        const int pos = s + (p - j);
                        const int rr = pos >> 6, pl = pos & 63;
                        T fsel = F[0];
#pragma unroll
                        for (int r = 1; r < NF; ++r)
                            if (rr == r) fsel = F[r];
                        const T bp = lane_read(fsel, pl);
#pragma unroll
                        for (int r = 0; r < NF; ++r)
                            if (r == rr) F[r] = lane_write(F[r], bj, pl);
                        bj = bp;
                    }
                    F[0] = lane_write(F[0], bj, s);
                }
                fetch(slot, j - PFD, (s - PFD) & 63, ln);
            });
            // the top block is finished: no later step reaches it
            const int et = 64 * (g + NF - 1) + lane;
            ring_st(rb, et < n ? et * SZ : -1, F[NF - 1]);
#pragma unroll
            for (int r = NF - 1; r >= 1; --r) F[r] = F[r - 1];
            F[0] = bprev;
        }
#pragma unroll
        for (int r = 1; r < NF; ++r) {
            const int e = 64 * (r - 1) + lane;
            ring_st(rb, e < n ? e * SZ : -1, F[r]);
        }
    }
}

}  // namespace

// 'N' solve for bands with kl <= 128 and kl + ku <= 256; -1 otherwise.
template <typename T>
int ring_gbtrs_launch(i64 n, i64 kl, i64 ku, i64 nrhs, const T *ab, i64 ld, const i64 *ipiv, T *B, i64 ldb,
                      hipStream_t st) {
    const i64 kd = kl + ku;
    if (n >= (1ll << 30) || kl > 128 || kd > 256 || ld * n * (i64)sizeof(T) >= (1ll << 31) - (1ll << 24)) return -1;
    const int nf = 1 + (int)cdiv(kl, 64), nb = 1 + (int)cdiv(kd, 64);
    constexpr int PFD = sizeof(T) == 4 ? 32 : 16;
    dim3 grid((unsigned)nrhs), block(64);
#define BH_RING_CASE(NFv, NBv)                                                                               \
    if (nf <= NFv && nb <= NBv) {                                                                            \
        hipLaunchKernelGGL((gbtrs_n_ring<T, NFv, NBv, PFD>), grid, block, 0, st, (int)n, (int)kl, (int)ku, ab, \
                           ld, ipiv, B, ldb);                                                                \
        BH_LAUNCH_CHECK();                                                                                   \
        return BH_OK;                                                                                        \
    }
    BH_RING_CASE(2, 2)
    BH_RING_CASE(2, 3)
    BH_RING_CASE(3, 3)
    BH_RING_CASE(3, 5)
#undef BH_RING_CASE
    return -1;
}

// 'T' solve for bands with kl <= 128 and kl + ku <= 256; -1 otherwise.
template <typename T>
int ring_gbtrs_t_launch(i64 n, i64 kl, i64 ku, i64 nrhs, const T *ab, i64 ld, const i64 *ipiv, T *B, i64 ldb,
                        hipStream_t st) {
    const i64 kd = kl + ku;
    if (n >= (1ll << 30) || kl > 128 || kd > 256 || ld * n * (i64)sizeof(T) >= (1ll << 31) - (1ll << 24)) return -1;
    const int nf = 1 + (int)cdiv(kl, 64), nb = 1 + (int)cdiv(kd, 64);
    constexpr int PFD = sizeof(T) == 4 ? 32 : 16;
    dim3 grid((unsigned)nrhs), block(64);
#define BH_RING_T_CASE(NFv, NBv)                                                                               \
    if (nf <= NFv && nb <= NBv) {                                                                              \
        hipLaunchKernelGGL((gbtrs_t_ring<T, NFv, NBv, PFD>), grid, block, 0, st, (int)n, (int)kl, (int)ku, ab, \
                           ld, ipiv, B, ldb);                                                                  \
        BH_LAUNCH_CHECK();                                                                                     \
        return BH_OK;                                                                                          \
    }
    BH_RING_T_CASE(2, 2)
    BH_RING_T_CASE(2, 3)
    BH_RING_T_CASE(3, 3)
    BH_RING_T_CASE(3, 5)
#undef BH_RING_T_CASE
    return -1;
}

}  // namespace bh
