// bh_gbtrs_ring.hip -- band LU solve 'N' with one wave per right-hand side,
// the vector held in a lane ring (no per-step lane shifts).
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

// Wave-uniform register index (scalar branches).
template <int N, typename T>
__device__ __forceinline__ T reg_pick(const T (&w)[N], int k) {
    T r = w[0];
#pragma unroll
    for (int i = 1; i < N; ++i)
        if (k == i) r = w[i];
    return r;
}

// NF = 1 + ceil(kl/64) forward registers, NB = 1 + ceil((kl+ku)/64) backward
// registers, PFD = prefetch distance in steps (divides 64).
template <typename T, int NF, int NB, int PFD>
__global__ void __launch_bounds__(64) gbtrs_n_ring(int n, int kl, int ku, const T *__restrict__ ab, i64 ld,
                                                   const i64 *__restrict__ ipiv, T *B, i64 ldb) {
    static_assert(64 % PFD == 0, "prefetch distance divides the block");
    const int lane = threadIdx.x;
    T *b = B + (i64)blockIdx.x * ldb;
    const int kd = kl + ku;

    // ------------------------------------------------------------ forward: L and pivots
    if (kl > 0 && n > 1) {
        const int nsteps = n - 1;
        const int G = (nsteps + 63) >> 6;
        T w[NF];  // w[r]: block g + r
#pragma unroll
        for (int r = 0; r < NF; ++r) {
            const int e = 64 * r + lane;
            w[r] = e < n ? b[e] : T(0);
        }
        T lp[PFD][NF];
        // multipliers of step j for window register r, lane = entry j + t
        auto fetch = [&](int slot, int j, int s) {
            const int lm = mini(kl, n - 1 - j);
            const T *col = ab + (i64)j * ld + kd;
#pragma unroll
            for (int r = 0; r < NF; ++r) {
                const int t = 64 * r + lane - s;
                lp[slot][r] = (j < nsteps && t >= 1 && t <= lm) ? col[t] : T(0);
            }
        };
        Unroll<PFD>::run([&](auto sc) {
            constexpr int s = decltype(sc)::value;
            fetch(s, s, s);
        });
        int pvc = lane < nsteps ? (int)(ipiv[lane] - 1) : 0;
        for (int g = 0; g < G; ++g) {
            const int j0 = 64 * g;
            const int pvn = (j0 + 64 + lane < nsteps) ? (int)(ipiv[j0 + 64 + lane] - 1) : 0;
            const int ein = 64 * (g + NF) + lane;
            const T bnext = ein < n ? b[ein] : T(0);
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
                        const T bp = lane_read(reg_pick(w, rr), pl);
                        const T bj0 = lane_read(w[0], s);
                        w[0] = lane_write(w[0], bp, s);
#pragma unroll
                        for (int r = 0; r < NF; ++r)
                            if (r == rr) w[r] = lane_write(w[r], bj0, pl);
                    }
                    const T nb = -lane_read(w[0], s);
#pragma unroll
                    for (int r = 0; r < NF; ++r) {
                        const int t = 64 * r + lane - s;
                        if (t >= 1 && t <= lm) w[r] = dfma(lp[slot][r], nb, w[r]);
                    }
                }
                fetch(slot, j + PFD, (s + PFD) & 63);
            });
            if (j0 + lane < n) b[j0 + lane] = w[0];
#pragma unroll
            for (int r = 0; r + 1 < NF; ++r) w[r] = w[r + 1];
            w[NF - 1] = bnext;
            pvc = pvn;
        }
#pragma unroll
        for (int r = 0; r < NF; ++r) {
            const int e = 64 * (G + r) + lane;
            if (e < n) b[e] = w[r];
        }
    }

    // ------------------------------------------------------------ backward: U, kl+ku super-diagonals
    {
        const int G0 = (n - 1) >> 6;
        T w[NB];  // w[r]: block g - r
#pragma unroll
        for (int r = 0; r < NB; ++r) {
            const int e = 64 * (G0 - r) + lane;
            w[r] = (e >= 0 && e < n) ? b[e] : T(0);
        }
        T up[PFD][NB];
        // column j of U for window register r, lane = entry j - t (t = 0: diagonal)
        auto fetch = [&](int slot, int j, int s) {
            const int tm = mini(kd, j);
            const T *col = ab + (i64)j * ld + kd;
#pragma unroll
            for (int r = 0; r < NB; ++r) {
                const int t = s - lane + 64 * r;
                up[slot][r] = (j >= 0 && j < n && t >= 0 && t <= tm) ? col[-t] : T(0);
            }
        };
        Unroll<PFD>::run([&](auto sc) {
            constexpr int s = 63 - decltype(sc)::value;
            fetch(s % PFD, 64 * G0 + s, s);
        });
        for (int g = G0; g >= 0; --g) {
            const int j0 = 64 * g;
            const int ein = 64 * (g - NB) + lane;
            const T bnext = ein >= 0 ? b[ein] : T(0);
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
                        const int t = s - lane + 64 * r;
                        if (t >= 1 && t <= tm) w[r] = dfma(up[slot][r], nx, w[r]);
                    }
                }
                fetch(slot, j - PFD, (s - PFD) & 63);
            });
            if (j0 + lane < n) b[j0 + lane] = w[0];
#pragma unroll
            for (int r = 0; r + 1 < NB; ++r) w[r] = w[r + 1];
            w[NB - 1] = bnext;
        }
    }
}

}  // namespace

// 'N' solve for bands with kl <= 128 and kl + ku <= 256; -1 otherwise.
template <typename T>
int ring_gbtrs_launch(i64 n, i64 kl, i64 ku, i64 nrhs, const T *ab, i64 ld, const i64 *ipiv, T *B, i64 ldb,
                      hipStream_t st) {
    const i64 kd = kl + ku;
    if (n >= (1ll << 30) || kl > 128 || kd > 256) return -1;
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

template int ring_gbtrs_launch<double>(i64, i64, i64, i64, const double *, i64, const i64 *, double *, i64,
                                       hipStream_t);
template int ring_gbtrs_launch<float>(i64, i64, i64, i64, const float *, i64, const i64 *, float *, i64,
                                      hipStream_t);

}  // namespace bh
