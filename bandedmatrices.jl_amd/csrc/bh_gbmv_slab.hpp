// bh_gbmv_slab.hpp -- y <- alpha*A*x + beta*y ('N') for unit strides and
// contiguous band storage (lda == kl+ku+1), the layout of BandedMatrix.data.
// The launcher is instantiated once per precision in bh_gbmv_slab_{d,s}.hip,
// so the two compile in parallel.
#pragma once
//
// A workgroup owns R consecutive rows.  The band columns those rows touch,
// [r0-kl, r0+R+ku), are one contiguous range of A; it is copied into LDS with
// 16-byte loads (the LDS image keeps the global address modulo 16, so every
// 16-byte chunk stays aligned on both sides), alpha*x over the same columns
// next to it.  Then each thread walks its rows' band from LDS in ascending j
// with y_i = fma(alpha*x_j, a_ij, y_i) -- the rounding of the column-ordered
// BLAS gbmv that BandedMatrices.jl calls (src/generic/matmul.jl:21-23), so
// the result is bit-identical to it.  A is read from HBM once (plus the
// kl+ku columns of halo per R rows); with several right-hand sides the slab
// is reused for all of them.
#include <type_traits>

#include "bh_common.hpp"

#include <cstdlib>
#include <cstdio>

namespace bh {

namespace {

template <int K>
struct CIdx {
    static constexpr int value = K;
};
// f(CIdx<0>{}) ... f(CIdx<N-1>{}): loop indices that are compile-time
// constants from the first pass on, so register arrays stay in registers.
template <int N>
struct CFor {
    template <class F>
    __device__ __forceinline__ static void run(F &&f) {
        CFor<N - 1>::run(f);
        f(CIdx<N - 1>{});
    }
};
template <>
struct CFor<0> {
    template <class F>
    __device__ __forceinline__ static void run(F &&) {}
};

// Block b of a grid of nb runs on XCD b mod 8; give each XCD one contiguous
// range of tiles so that the halo columns two neighbouring tiles share are
// read through the same L2.
__device__ __forceinline__ int xcd_tile(int b, int nb) {
    const int per = nb >> 3, rem = nb & 7, x = b & 7, k = b >> 3;
    return x < rem ? x * (per + 1) + k : rem * (per + 1) + (x - rem) * per + k;
}

// right-hand sides per step of gbmv_n_slab
constexpr int kSlabRP = 4;

template <typename T, int NT, int RPT, int NPF>
__global__ void __launch_bounds__(NT) gbmv_n_slab(int m, int n, int kl, int ku, int nrhs, T alpha,
                                                  const T *__restrict__ A, i64 a_elems, const T *__restrict__ X,
                                                  i64 ldx, T beta, T *__restrict__ Y, i64 ldy, int ax_off) {
    constexpr int R = NT * RPT;
    const int h = kl + ku + 1;
    const int tid = threadIdx.x;
    const int r0 = xcd_tile(blockIdx.x, gridDim.x) * R;
    const int c0 = r0 - kl > 0 ? r0 - kl : 0;
    const int c1 = (r0 + R + ku < n) ? r0 + R + ku : n;
    const int nc = c1 - c0;
    const bool interior = r0 >= kl && r0 + R + ku <= n && r0 + R <= m;
    extern __shared__ __align__(16) unsigned char lds[];

    // byte range of columns [c0, c1) relative to A, widened to 16-byte
    // aligned absolute addresses; chunks reaching outside A (its first or
    // last bytes) are read dword by dword within A only.  All of a thread's
    // chunk loads (and its x loads) are issued before any LDS store, so a
    // workgroup waits for one memory latency, not one per chunk.
    const i64 gs = (i64)c0 * h * (i64)sizeof(T), ge = (i64)c1 * h * (i64)sizeof(T);
    const int mis = (int)(reinterpret_cast<unsigned long long>(A) & 15);
    const i64 ga = ((gs + mis) & ~(i64)15) - mis;
    const int shift = (int)(gs - ga);
    const int nchunks = (int)((ge - ga + 15) >> 4);
    const i64 abytes = a_elems * (i64)sizeof(T);
    const unsigned char *gsrc = reinterpret_cast<const unsigned char *>(A) + ga;
    uint4 pre[NPF];
    // Every chunk inside A (all workgroups but those at A's two ends): one
    // unconditional 16-byte buffer load per chunk, chunks past the slab read
    // as zero, so the compiler has no branch to merge and waits only once.
    typedef unsigned sv4u __attribute__((ext_vector_type(4)));
    const bool inside = ga >= 0 && ga + 16 * (i64)nchunks <= abytes && abytes < (1ll << 31) - (1ll << 24);
    if (inside) {
        const __amdgpu_buffer_rsrc_t ra =
            __builtin_amdgcn_make_buffer_rsrc(const_cast<T *>(A), 0, (int)abytes, 0x00020000);
#pragma unroll
        for (int q = 0; q < NPF; ++q) {
            const int k = tid + q * NT;
            const int o = k < nchunks ? (int)(ga + 16 * (i64)k) : 0x7f800000;  // past the end: reads 0
            const sv4u v = __builtin_amdgcn_raw_buffer_load_b128(ra, o, 0, 0);
            pre[q] = make_uint4(v.x, v.y, v.z, v.w);
        }
    } else {
#pragma unroll
        for (int q = 0; q < NPF; ++q) {
            const int k = tid + q * NT;
            const i64 off = ga + 16 * (i64)k;
            uint4 v = make_uint4(0, 0, 0, 0);
            if (k < nchunks) {
                if (off >= 0 && off + 16 <= abytes) {
                    v = *reinterpret_cast<const uint4 *>(gsrc + 16 * (i64)k);
                } else {
                    unsigned w[4] = {0, 0, 0, 0};
                    for (int b = 0; b < 4; ++b)
                        if (off + 4 * b >= 0 && off + 4 * b + 4 <= abytes)
                            w[b] = *reinterpret_cast<const unsigned *>(gsrc + 16 * (i64)k + 4 * b);
                    v = make_uint4(w[0], w[1], w[2], w[3]);
                }
            }
            pre[q] = v;
        }
    }
    constexpr int XPT = RPT + 1;  // x values per thread (nc <= R + kl + ku <= XPT * NT when kl+ku <= NT)
    constexpr int XD = 8;         // right-hand sides whose x is in flight
    constexpr int RP = kSlabRP;   // right-hand sides per step (they share each slab read)
    constexpr int CH = 8 / (RP > 2 ? 2 : 1);  // band entries whose LDS reads go out before their fmas
    // x of right-hand sides r .. r+XD-1 are loaded into registers XD
    // right-hand sides ahead of their use, so with many right-hand sides the
    // x loads overlap the products instead of waiting once per right-hand side
    T xq[XD][XPT];
    // Each right-hand side's x comes through a buffer resource covering just
    // its nc slab entries (none past the last right-hand side): every load is
    // unconditional, lanes past the range read 0, and the compiler can count
    // the loads in flight instead of waiting for all of them.
    auto load_x = [&](T(&xv)[XPT], int r) __attribute__((always_inline)) {
        const bool live = r < nrhs;
        int xb = live ? nc * (int)sizeof(T) : 0;
        asm volatile("" : "+s"(xb));
        const __amdgpu_buffer_rsrc_t rx = __builtin_amdgcn_make_buffer_rsrc(
            const_cast<T *>(X) + (live ? (i64)r * ldx + c0 : 0), 0, xb, 0x00020000);
        CFor<XPT>::run([&](auto qi) {
            constexpr int q = decltype(qi)::value;
            const int c = tid + q * NT;
            if constexpr (sizeof(T) == 8)
                xv[q] = __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(rx, c * 8, 0, 0));
            else
                xv[q] = __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(rx, c * 4, 0, 0));
        });
    };
    CFor<XD>::run([&](auto di) {
        constexpr int d = decltype(di)::value;
        load_x(xq[d], d);
    });
#pragma unroll
    for (int q = 0; q < NPF; ++q) {
        const int k = tid + q * NT;
        if (k < nchunks) *reinterpret_cast<uint4 *>(lds + 16 * k) = pre[q];
    }
    const T *slab = reinterpret_cast<const T *>(lds + shift);  // slab[c*h + t]: band row t of column c0+c
    // alpha*x of the step's RP right-hand sides: row k at ax + k * (RPT + 1) * NT
    T *ax = reinterpret_cast<T *>(lds + ax_off);
    constexpr int AXS = (RPT + 1) * NT;

    // RP right-hand sides per step: alpha*x of each to LDS, their x loads XD
    // right-hand sides ahead, the rows' products, the stores.  Each band entry
    // read from the slab serves all RP products (LDS bandwidth bounds this
    // kernel: the slab reads per output drop from 2h to (1 + 1/RP)h), and
    // each output keeps its own ascending fma order.  The workgroup's
    // interior/edge choice and the last, partial group of XD are compile-time
    // variants, so in the steady state every step issues the same memory
    // operations (2 RP x loads and RP RPT stores, no early exit) and the
    // compiler's wait for an x value issued XD right-hand sides earlier counts
    // every operation issued since, instead of draining the whole prefetch.
    int rb = 0;
    auto step = [&](auto pi, auto interiorc, auto tailc) __attribute__((always_inline)) {
        constexpr int d0 = RP * decltype(pi)::value;
        constexpr bool inner = decltype(interiorc)::value, tail = decltype(tailc)::value;
        const int r = rb + d0;
        int live = RP;  // right-hand sides of this step that exist (uniform)
        if constexpr (tail) {
            if (r >= nrhs) return;  // uniform
            live = nrhs - r < RP ? nrhs - r : RP;
        }
        CFor<XPT>::run([&](auto qi) {
            constexpr int q = decltype(qi)::value;
            const int c = tid + q * NT;
            if (c < nc) {
#pragma unroll
                for (int k = 0; k < RP; ++k) ax[k * AXS + c] = alpha * xq[d0 + k][q];
            }
        });
        __syncthreads();
#pragma unroll
        for (int k = 0; k < RP; ++k) load_x(xq[d0 + k], r + k + XD);
        if constexpr (inner) {
            // every row has all h band entries inside the matrix: a uniform
            // trip count, so the LDS reads of CH steps are issued before their
            // fmas consume them (same ascending-j fma order)
#pragma unroll
            for (int q = 0; q < RPT; ++q) {
                const int li = tid + q * NT;
                T *yp = Y + (i64)r * ldy + r0 + li;
                T acc[RP];
#pragma unroll
                for (int k = 0; k < RP; ++k) acc[k] = (beta == T(0) || k >= live) ? T(0) : beta * yp[k * ldy];
                const T *ap = slab + li * h + (kl + ku);  // entry (i, i-kl+d) at ap[d*(h-1)]
                const T *xp = ax + li;                    // alpha*x_{i-kl+d} of step row k at xp[k*AXS + d]
                int e = 0;
                for (; e + CH <= h; e += CH) {
                    T av[CH], xw[RP][CH];
#pragma unroll
                    for (int u = 0; u < CH; ++u) {
                        av[u] = ap[(e + u) * (h - 1)];
#pragma unroll
                        for (int k = 0; k < RP; ++k) xw[k][u] = xp[k * AXS + e + u];
                    }
#pragma unroll
                    for (int u = 0; u < CH; ++u) {
#pragma unroll
                        for (int k = 0; k < RP; ++k) acc[k] = dfma(xw[k][u], av[u], acc[k]);
                    }
                }
                for (; e < h; ++e) {
                    const T a = ap[e * (h - 1)];
#pragma unroll
                    for (int k = 0; k < RP; ++k) acc[k] = dfma(xp[k * AXS + e], a, acc[k]);
                }
#pragma unroll
                for (int k = 0; k < RP; ++k)
                    if (k < live) yp[k * ldy] = acc[k];
            }
        } else {
#pragma unroll
            for (int q = 0; q < RPT; ++q) {
                const int i = r0 + tid + q * NT;
                if (i < m) {
                    T *yp = Y + (i64)r * ldy + i;
                    T acc[RP];
#pragma unroll
                    for (int k = 0; k < RP; ++k) acc[k] = (beta == T(0) || k >= live) ? T(0) : beta * yp[k * ldy];
                    const int jlo = i - kl > 0 ? i - kl : 0;
                    const int jhi = (i + ku < n - 1) ? i + ku : n - 1;
                    const T *ap = slab + (jlo - c0) * h + (ku + i - jlo);
                    const T *xp = ax + (jlo - c0);
                    for (int j = jlo; j <= jhi; ++j) {
#pragma unroll
                        for (int k = 0; k < RP; ++k) acc[k] = dfma(xp[k * AXS], *ap, acc[k]);
                        ++xp;
                        ap += h - 1;
                    }
#pragma unroll
                    for (int k = 0; k < RP; ++k)
                        if (k < live) yp[k * ldy] = acc[k];
                }
            }
        }
        __syncthreads();
    };
    static_assert(XD % RP == 0, "a step's right-hand sides divide the prefetch depth");
    auto all = [&](auto interiorc) __attribute__((always_inline)) {
        for (; rb + XD <= nrhs; rb += XD)
            CFor<XD / RP>::run([&](auto pi) { step(pi, interiorc, std::false_type{}); });
        if (rb < nrhs) CFor<XD / RP>::run([&](auto pi) { step(pi, interiorc, std::true_type{}); });
    };
    if (interior)
        all(std::true_type{});
    else
        all(std::false_type{});
}

template <typename T, int NT, int RPT, int NPF>
int slab_launch_npf(int m, int n, int kl, int ku, int nrhs, T alpha, const T *A, i64 a_elems, const T *X, i64 ldx,
                    T beta, T *Y, i64 ldy, int slab_cap, hipStream_t st) {
    constexpr int R = NT * RPT;
    const size_t lds = (size_t)slab_cap + kSlabRP * (size_t)(RPT + 1) * NT * sizeof(T);  // slab + alpha*x rows
    auto kern = gbmv_n_slab<T, NT, RPT, NPF>;
    if (lds > 64 * 1024)
        BH_HIP(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL(kern, dim3((unsigned)cdiv(m, R)), dim3(NT), lds, st, m, n, kl, ku, nrhs, alpha, A, a_elems, X,
                       ldx, beta, Y, ldy, slab_cap);
    BH_LAUNCH_CHECK();
    return BH_OK;
}

// Smallest register batch (16-byte chunks per thread) that holds the slab,
// or -1.  The LDS image is sized to the slab itself (chunks of the widest
// tile, with 16 bytes of alignment slack at either end), so narrow bands
// leave room for several workgroups per CU.
template <typename T, int NT, int RPT>
int slab_launch(int m, int n, int kl, int ku, int nrhs, T alpha, const T *A, i64 a_elems, const T *X, i64 ldx,
                T beta, T *Y, i64 ldy, hipStream_t st) {
    constexpr int R = NT * RPT;
    if (kl + ku > NT) return -1;
    const int h = kl + ku + 1;
    const i64 span = (i64)(R + kl + ku) * h * (i64)sizeof(T) + 32;
    const i64 slab_cap = cdiv(span, 16) * 16;
    if (slab_cap + kSlabRP * (i64)(RPT + 1) * NT * (i64)sizeof(T) > 160 * 1024) return -1;
    const i64 need = cdiv(span, 16 * (i64)NT);
#define BH_NPF(P)                                                                                               \
    if (need <= P)                                                                                              \
        return slab_launch_npf<T, NT, RPT, P>(m, n, kl, ku, nrhs, alpha, A, a_elems, X, ldx, beta, Y, ldy, \
                                              (int)slab_cap, st);
    BH_NPF(3)
    BH_NPF(5)
    BH_NPF(7)
    BH_NPF(9)
    BH_NPF(11)
    BH_NPF(13)
    BH_NPF(17)
#undef BH_NPF
    return -1;
}

typedef unsigned v4u __attribute__((ext_vector_type(4)));

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Word 3 of a raw (stride 0) buffer resource on gfx9-family parts.
constexpr int kRsrcWord3 = 0x00020000;

template <typename T>
__device__ __forceinline__ T buffer_load_elem(__amdgpu_buffer_rsrc_t r, int off);
template <>
__device__ __forceinline__ double buffer_load_elem<double>(__amdgpu_buffer_rsrc_t r, int off) {
    typedef unsigned v2u __attribute__((ext_vector_type(2)));
    const v2u v = __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0);
    return __builtin_bit_cast(double, v);
}
template <>
__device__ __forceinline__ float buffer_load_elem<float>(__amdgpu_buffer_rsrc_t r, int off) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0));
}

// Persistent single-vector form, beta = 0, for 16-byte aligned storage whose
// size is a multiple of 16 bytes (so every chunk of a slab lies inside A) and
// kl + ku <= 64 and m <= n (every tile then has columns).  One wave per workgroup, R = 64*RPL rows per tile, LDS
// private to the wave (no workgroup barriers).  Workgroup b serves XCD b mod
// 8, which owns one contiguous eighth of the tiles (neighbouring tiles share
// kl+ku halo columns through that XCD's L2); the XCD's waves take its tiles
// in turn.  NS tiles' chunks are held in registers (stages): a tile's chunks
// go to LDS, the loads of the tile NS ahead are issued into the registers
// just freed, then the tile is computed -- NS-1 to NS tiles' loads are in
// flight through every compute phase.  Loads are buffer loads through a
// resource sized to the tile (lanes past it read 0 without touching memory;
// past the wave's last tile the size is 0), issued on every path, and LDS
// stores are unconditional (chunks past a slab land in the unused tail of the
// LDS image): the memory counter waits stay exact.
template <typename T, int RPL, int NPF, int NS>
__global__ void __launch_bounds__(64) gbmv_n_wave(int m, int n, int kl, int ku, T alpha, const T *__restrict__ A,
                                                  const T *__restrict__ x, T *__restrict__ y, int ntiles, int ax_off) {
    constexpr int R = 64 * RPL;
    constexpr int XQ = RPL + 1;  // x values per lane: R + kl + ku <= 64 * XQ
    extern __shared__ __align__(16) unsigned char lds[];
    T *ax = reinterpret_cast<T *>(lds + ax_off);
    const int lane = threadIdx.x;
    const int h = kl + ku + 1;
    const int g8 = gridDim.x >> 3, xcd = blockIdx.x & 7, w = blockIdx.x >> 3;
    const int lo = (int)((i64)ntiles * xcd / 8), hi = (int)((i64)ntiles * (xcd + 1) / 8);
    v4u cs[NS][NPF];  // chunks of the NS tiles in flight
    T xs[NS][XQ];     // and their x values

    auto issue = [&](v4u(&c)[NPF], T(&xv)[XQ], int t) __attribute__((always_inline)) {
        // the bound goes through an empty asm so that the compiler cannot tie
        // it to the loop exit and drop (or branch around) loads it would
        // prove unused there: every path then issues the same loads
        int hb = hi;
        asm volatile("" : "+s"(hb));
        const bool live = t < hb;
        const int r0 = t * R;
        const int c0 = r0 - kl > 0 ? r0 - kl : 0;
        const int c1 = (r0 + R + ku < n) ? r0 + R + ku : n;
        const i64 gs = (i64)c0 * h * (i64)sizeof(T), ge = (i64)c1 * h * (i64)sizeof(T);
        const i64 ga = gs & ~(i64)15;
        const int nbytes = live ? (int)(((ge - ga + 15) >> 4) << 4) : 0;
        const int xbytes = live ? (c1 - c0) * (int)sizeof(T) : 0;
        const __amdgpu_buffer_rsrc_t ra = __builtin_amdgcn_make_buffer_rsrc(
            const_cast<unsigned char *>(reinterpret_cast<const unsigned char *>(A)) + (live ? ga : 0), 0, nbytes,
            kRsrcWord3);
        const __amdgpu_buffer_rsrc_t rx =
            __builtin_amdgcn_make_buffer_rsrc(const_cast<T *>(x) + (live ? c0 : 0), 0, xbytes, kRsrcWord3);
        CFor<NPF>::run([&](auto qi) {
            constexpr int q = decltype(qi)::value;
            c[q] = __builtin_amdgcn_raw_buffer_load_b128(ra, 16 * (lane + 64 * q), 0, 0);
        });
        CFor<XQ>::run([&](auto qi) {
            constexpr int q = decltype(qi)::value;
            xv[q] = buffer_load_elem<T>(rx, (int)sizeof(T) * (lane + 64 * q));
        });
    };
    auto stage = [&](const v4u(&c)[NPF], const T(&xv)[XQ]) __attribute__((always_inline)) {
        CFor<NPF>::run([&](auto qi) {
            constexpr int q = decltype(qi)::value;
            *reinterpret_cast<v4u *>(lds + 16 * (lane + 64 * q)) = c[q];
        });
        CFor<XQ>::run([&](auto qi) {
            constexpr int q = decltype(qi)::value;
            ax[lane + 64 * q] = alpha * xv[q];
        });
    };
    auto compute = [&](int t) __attribute__((always_inline)) {
        const int r0 = t * R;
        const int c0 = r0 - kl > 0 ? r0 - kl : 0;
        const int shift = (int)(((i64)c0 * h * (i64)sizeof(T)) & 15);
        const T *slab = reinterpret_cast<const T *>(lds + shift);  // slab[c*h + t]: band row t of column c0+c
        if (r0 >= kl && r0 + R + ku <= n && r0 + R <= m) {
            CFor<RPL>::run([&](auto qi) {
                constexpr int q = decltype(qi)::value;
                const int li = lane + 64 * q;
                const T *ap = slab + li * h + (kl + ku);  // entry (i, i-kl+d) at ap[d*(h-1)]
                const T *xp = ax + li;                    // alpha*x_{i-kl+d} at xp[d]
                T acc = T(0);
                int d = 0;
                for (; d + 8 <= h; d += 8) {
                    T av[8], xw[8];
#pragma unroll
                    for (int u = 0; u < 8; ++u) {
                        av[u] = ap[(d + u) * (h - 1)];
                        xw[u] = xp[d + u];
                    }
#pragma unroll
                    for (int u = 0; u < 8; ++u) acc = dfma(xw[u], av[u], acc);
                }
                for (; d < h; ++d) acc = dfma(xp[d], ap[d * (h - 1)], acc);
                y[r0 + li] = acc;
            });
        } else {
            for (int q = 0; q < RPL; ++q) {
                const int i = r0 + lane + 64 * q;
                if (i < m) {
                    T acc = T(0);
                    const int jlo = i - kl > 0 ? i - kl : 0;
                    const int jhi = (i + ku < n - 1) ? i + ku : n - 1;
                    const T *ap = slab + (jlo - c0) * h + (ku + i - jlo);
                    const T *xp = ax + (jlo - c0);
                    for (int j = jlo; j <= jhi; ++j) {
                        acc = dfma(*xp, *ap, acc);
                        ++xp;
                        ap += h - 1;
                    }
                    y[i] = acc;
                }
            }
        }
    };

    int t = lo + w;
    if (t >= hi) return;  // uniform over the wave
    CFor<NS>::run([&](auto si) {
        constexpr int sg = decltype(si)::value;
        issue(cs[sg], xs[sg], t + sg * g8);
    });
    // every stage is staged in every pass (only the compute of a tile past
    // the end is skipped), so no load is unused on any path through the loop
    for (;;) {
        CFor<NS>::run([&](auto si) {
            constexpr int sg = decltype(si)::value;
            const int ts = t + sg * g8;
            stage(cs[sg], xs[sg]);
            wave_sync();
            issue(cs[sg], xs[sg], ts + NS * g8);
            if (sg == 0 || ts < hi) compute(ts);
            wave_sync();
        });
        t += NS * g8;
        if (t >= hi) break;
    }
}

template <typename T, int RPL, int NPF, int NS>
int wave_launch_npf(int m, int n, int kl, int ku, T alpha, const T *A, const T *x, T *y, hipStream_t st) {
    constexpr int R = 64 * RPL;
    const int ax_off = NPF * 64 * 16;
    const size_t lds = (size_t)ax_off + (size_t)(RPL + 1) * 64 * sizeof(T);
    auto kern = gbmv_n_wave<T, RPL, NPF, NS>;
    static int occ = -1, cus = 0;
    if (occ < 0) {
        int dev = 0, b = 0;
        BH_HIP(hipGetDevice(&dev));
        BH_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        BH_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, (const void *)kern, 64, lds));
        occ = b;
    }
    if (occ < 1) return -1;
    const int ntiles = (int)cdiv(m, R);
    i64 grid = (i64)cus * occ;
    const i64 cap = 8 * cdiv(ntiles, 8);
    if (grid > cap) grid = cap;
    grid = (grid + 7) & ~(i64)7;
    hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(64), lds, st, m, n, kl, ku, alpha, A, x, y, ntiles, ax_off);
    BH_LAUNCH_CHECK();
    return BH_OK;
}

// Register batch (16-byte chunks per lane) that holds one tile's slab, or -1.
template <typename T, int RPL, int NS>
int wave_launch(int m, int n, int kl, int ku, T alpha, const T *A, const T *x, T *y, hipStream_t st) {
    constexpr int R = 64 * RPL;
    if (kl + ku > 64) return -1;
    const int h = kl + ku + 1;
    const i64 span = (i64)(R + kl + ku) * h * (i64)sizeof(T) + 16;  // bytes incl. the 16-byte alignment slack
    const i64 need = cdiv(span, 16 * 64);
#define BH_NPF(P)                                                                           \
    if (need <= P) {                                                                        \
        if constexpr (NS * P <= 48)                                                         \
            return wave_launch_npf<T, RPL, P, NS>(m, n, kl, ku, alpha, A, x, y, st);        \
        else                                                                                \
            return -1;                                                                      \
    }
    BH_NPF(4)
    BH_NPF(6)
    BH_NPF(8)
    BH_NPF(11)
    BH_NPF(14)
    BH_NPF(18)
#undef BH_NPF
    return -1;
}

}  // namespace

// 'N' gbmv through the LDS slab, or -1 when it does not apply (strided
// vectors, padded storage, bands too wide for the slab).  The workgroup
// shape (threads x rows per thread) can be forced with BANDEDHIP_GBMV_SLAB
// = "256x2" etc.
template <typename T>
int slab_gbmv_n(i64 m, i64 n, i64 kl, i64 ku, i64 nrhs, T alpha, const T *A, i64 lda, const T *X, i64 incx, i64 ldx,
                T beta, T *Y, i64 incy, i64 ldy, hipStream_t st) {
    if (incx != 1 || incy != 1 || lda != kl + ku + 1 || m >= (1ll << 31) || n >= (1ll << 31) || nrhs >= (1ll << 31))
        return -1;
    int nt = 256, rpt = 1;
    if (const char *e = getenv("BANDEDHIP_GBMV_SLAB")) {
        int a = 0, b = 0;
        if (sscanf(e, "%dx%d", &a, &b) == 2) {
            nt = a;
            rpt = b;
        }
    }
    const int mi = (int)m, ni = (int)n, li = (int)kl, ui = (int)ku, ri = (int)nrhs;
    const i64 ae = lda * n;
    const bool aligned16 = (reinterpret_cast<unsigned long long>(A) & 15) == 0 && ((ae * (i64)sizeof(T)) & 15) == 0;
    int wave_rpl = 1;
    if (const char *e = getenv("BANDEDHIP_GBMV_WAVE")) wave_rpl = atoi(e);
    if (nrhs == 1 && beta == T(0) && aligned16 && m <= n && wave_rpl > 0) {
        int r = -1;
        if (wave_rpl == 3) r = wave_launch<T, 1, 3>(mi, ni, li, ui, alpha, A, X, Y, st);
        if (wave_rpl == 4) r = wave_launch<T, 1, 4>(mi, ni, li, ui, alpha, A, X, Y, st);
        if (r == -1) r = wave_launch<T, 1, 2>(mi, ni, li, ui, alpha, A, X, Y, st);
        if (r != -1) return r;
    }
#define BH_SLAB(NTv, RPTv)                                                                                         \
    if (nt == NTv && rpt == RPTv) {                                                                                \
        const int r = slab_launch<T, NTv, RPTv>(mi, ni, li, ui, ri, alpha, A, ae, X, ldx, beta, Y, ldy, st);     \
        if (r != -1) return r;                                                                                     \
    }
    BH_SLAB(128, 1)
    BH_SLAB(128, 2)
    BH_SLAB(256, 1)
    BH_SLAB(256, 2)
    BH_SLAB(512, 1)
#undef BH_SLAB
    int r = slab_launch<T, 256, 1>(mi, ni, li, ui, ri, alpha, A, ae, X, ldx, beta, Y, ldy, st);
    if (r == -1) r = slab_launch<T, 512, 1>(mi, ni, li, ui, ri, alpha, A, ae, X, ldx, beta, Y, ldy, st);
    return r;
}

}  // namespace bh
