// bh_gbmm.hip -- banded x banded product C <- alpha*A*B + beta*C on gfx950.
//
// Restates gbmm! (src/banded/gbmm.jl:207-340).  The reference computes every
// column of C with one BLAS gbmv over a zero-copy band window of A
// (gbmm.jl:18-127), so each entry of C is
//     c_ij = (beta == 0 ? 0 : beta*c_ij);  c_ij = fma(alpha*b_vj, a_iv, c_ij), v ascending
// over the v where both a_iv and b_vj are in band.  Here all columns are one
// launch: a workgroup owns a run of consecutive C columns, stages the A and B
// band columns it needs in LDS (both are contiguous slabs in band storage)
// and one thread computes one stored entry of C, so C is written with
// coalesced stores.  Bands of C outside the product's bands are beta-scaled
// (gbmm.jl:286-288) and columns the product never reaches are beta-scaled
// (gbmm.jl:339), exactly as the reference does.
#include <stdlib.h>

#include <vector>

#include "bh_common.hpp"

namespace bh {

constexpr int kMmThreads = 256;
// entries per thread computed before their stores issue (measured on the
// general path at n = 2^20, (8,8) x (4,4): 105 us FP64 at 4, 115 us at 8,
// 108 us at 2, 135 us storing each entry at once)
constexpr int kMmBatch = 4;

template <typename T>
struct GbmmArgs {
    i64 n, k, m;
    T alpha, beta;
    const T *A;
    i64 lda, Al, Au;
    const T *B;
    i64 ldb, Bl, Bu;
    T *C;
    i64 ldc, Ctl, Ctu;  // C's own bands
    i64 Cl, Cu;         // bands the product writes (clipped)
};

__host__ __device__ inline i64 odd_stride(i64 h) { return (h % 2 == 0) ? h + 1 : h; }
template <typename I>
__device__ __forceinline__ I mm_min(I a, I b) {
    return a < b ? a : b;
}
template <typename I>
__device__ __forceinline__ I mm_max(I a, I b) {
    return a > b ? a : b;
}

// One workgroup per NC consecutive columns of C.
template <typename T, typename I, int NE>
__global__ void __launch_bounds__(kMmThreads) gbmm_kernel(GbmmArgs<T> p, i64 NC) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const i64 hA = p.Al + p.Au + 1, hB = p.Bl + p.Bu + 1, hC = p.Ctl + p.Ctu + 1;
    const i64 SA = odd_stride(hA), SB = odd_stride(hB);
    const i64 j0 = (i64)blockIdx.x * NC;
    const i64 j1 = imin(p.m, j0 + NC);
    // contraction range needed by columns [j0, j1): v in [j0-Bu, j1-1+Bl] ∩ [0,k)
    const i64 v0 = imax(0, j0 - p.Bu), v1 = imin(p.k, j1 + p.Bl);
    T *sB = reinterpret_cast<T *>(smem_raw);             // NC columns of B
    T *sA = sB + NC * SB;                                // A columns [v0, v1)
    // Staging: when the band storage has no padding (ld == h) and the LDS
    // stride equals h, the columns are one contiguous run on both sides and
    // are copied flat; otherwise element (column, row) indices come from a
    // 32-bit division (the slab sizes are small).  Each thread issues a whole
    // batch of loads (UB of B, UA of A; indices past the slab are clamped to
    // its last element, so every load is in bounds and none sits in a branch)
    // before the batch's LDS stores, so the workgroup waits for global memory
    // once per batch rather than once per element.
    const int nB = (int)((j1 - j0) * hB), nA = (int)imax(0, (v1 - v0) * hA);
    const int hBi = (int)hB, hAi = (int)hA, SBi = (int)SB, SAi = (int)SA;
    const bool flatB = p.ldb == hB && SB == hB, flatA = p.lda == hA && SA == hA;
    const T *srcB = p.B + j0 * p.ldb, *srcA = p.A + v0 * p.lda;
    auto offB = [&](int e) { return flatB ? (i64)e : (i64)(e / hBi) * p.ldb + (e % hBi); };
    auto offA = [&](int e) { return flatA ? (i64)e : (i64)(e / hAi) * p.lda + (e % hAi); };
    auto dstB = [&](int e) { return flatB ? e : (e / hBi) * SBi + (e % hBi); };
    auto dstA = [&](int e) { return flatA ? e : (e / hAi) * SAi + (e % hAi); };
    constexpr int UB = 4, UA = 8;
    const int bt = (int)blockDim.x, tid = (int)threadIdx.x;
    const int iters = (int)imax(cdiv(nB, UB * bt), cdiv(nA, UA * bt));
    for (int it = 0; it < iters; ++it) {
        T vb[UB], va[UA];
#pragma unroll
        for (int q = 0; q < UB; ++q) vb[q] = srcB[offB((int)imin((it * UB + q) * bt + tid, nB - 1))];
        if (nA > 0) {
#pragma unroll
            for (int q = 0; q < UA; ++q) va[q] = srcA[offA((int)imin((it * UA + q) * bt + tid, nA - 1))];
        }
#pragma unroll
        for (int q = 0; q < UB; ++q) {
            const int e = (it * UB + q) * bt + tid;
            if (e < nB) sB[dstB(e)] = p.alpha * vb[q];  // alpha*b, as the reference's gbmv
        }
#pragma unroll
        for (int q = 0; q < UA; ++q) {
            const int e = (it * UA + q) * bt + tid;
            if (e < nA) sA[dstA(e)] = va[q];
        }
    }
    __syncthreads();

    // The entries: thread tid takes e = tid, tid + bt, ... of the NC x hC
    // block (e = column * hC + row, so the stores are contiguous), stepping
    // (column, row) without a division; sizes in I (int when every dimension
    // is below 2^30, which keeps the index arithmetic 32-bit).  A batch of NE
    // entries is computed into registers before the batch's stores issue:
    // a store counts against the same wait counter as a load, and storing
    // each entry before computing the next made every entry wait for the
    // previous entry's store to complete.
    const I shift = (I)(p.Ctu - p.Cu);  // first stored row of the product bands
    const I Ctu = (I)p.Ctu, Cl = (I)p.Cl, Cu = (I)p.Cu, Al = (I)p.Al, Au = (I)p.Au, Bl = (I)p.Bl, Bu = (I)p.Bu;
    const I n = (I)p.n, k = (I)p.k, m = (I)p.m, jb = (I)j0, vb0 = (I)v0;
    const I rowlast = Ctu + Cl, jprod = k + Bu, jscale = mm_min(m, n + Cu);
    const bool b0 = p.beta == T(0);
    const int hCi = (int)hC, nC = (int)((j1 - j0) * hC);
    const int sj = bt / hCi, sr = bt - sj * hCi, sa1 = SAi - 1;
    // Interior workgroups (every column's product band lies inside B's
    // columns, inside A's rows and inside the matrix): thread (g, r) keeps
    // stored row r and takes columns g, g + G, ...  Row r's band offsets,
    // trip count and slab steps are then the same for all its columns:
    // v runs over j + dlo .. j + dhi, b_vj at sB[jj*SB + Bu + dlo + t],
    // a_iv at sA[(jj + Bu + dlo + t)*SA + Au + r - Ctu - dlo - t].
    const bool inner_wg = j0 >= p.Bu && j0 >= p.Cu && j1 - 1 + p.Bl <= p.k - 1 && j1 - 1 + p.Cl <= p.n - 1 &&
                          (i64)hC * 2 <= bt;
    if (inner_wg) {
        const int G = bt / hCi, g = tid / hCi, rr = tid - g * hCi;
        if (g >= G) return;  // no LDS use or barrier follows
        const int ncol = (int)(j1 - j0);
        const int rc = rr - (int)Ctu;  // row r's offset from the diagonal
        const bool prod = !(rr < (int)shift || rr > (int)rowlast);
        const int dlo = rc - (int)Al > -(int)Bu ? rc - (int)Al : -(int)Bu;
        const int dhi = rc + (int)Au < (int)Bl ? rc + (int)Au : (int)Bl;
        const int cnt = dhi - dlo + 1;
        const int boff = (int)Bu + dlo, aoff = ((int)Bu + dlo) * SAi + (int)Au + rc - dlo;
        T *cbase = p.C + j0 * p.ldc + rr;
        for (int c0 = g; c0 < ncol; c0 += NE * G) {
            T val[NE];
#pragma unroll
            for (int q = 0; q < NE; ++q) {
                const int jj2 = c0 + q * G;
                if (jj2 < ncol) {
                    T *cp = cbase + (i64)jj2 * p.ldc;
                    T acc = b0 ? T(0) : p.beta * *cp;
                    if (prod) {
                        const T *bp = sB + jj2 * SBi + boff;
                        const T *ap = sA + jj2 * SAi + aoff;
                        int v = 0;
                        for (; v + 4 <= cnt; v += 4) {  // four steps' LDS reads issued before their fmas
                            const T x0 = bp[0], x1 = bp[1], x2 = bp[2], x3 = bp[3];
                            const T a0 = ap[0], a1 = ap[sa1], a2 = ap[2 * sa1], a3 = ap[3 * sa1];
                            acc = dfma(x0, a0, acc);
                            acc = dfma(x1, a1, acc);
                            acc = dfma(x2, a2, acc);
                            acc = dfma(x3, a3, acc);
                            bp += 4;
                            ap += 4 * sa1;
                        }
                        for (; v < cnt; ++v) {
                            acc = dfma(*bp, *ap, acc);
                            ++bp;
                            ap += sa1;
                        }
                    }
                    val[q] = acc;
                }
            }
#pragma unroll
            for (int q = 0; q < NE; ++q) {
                const int jj2 = c0 + q * G;
                if (jj2 < ncol) cbase[(i64)jj2 * p.ldc] = val[q];
            }
        }
        return;
    }
    int jj = tid / hCi, r = tid - jj * hCi;
    for (int e0 = tid; e0 < nC; e0 += NE * bt) {
        T val[NE];
        T *dst[NE];
#pragma unroll
        for (int q = 0; q < NE; ++q) {
            const int e = e0 + q * bt;
            const I j = jb + jj;
            T *cp = p.C + (i64)j * p.ldc + r;
            dst[q] = nullptr;
            if (e < nC) {
                if (r < shift || r > rowlast) {
                    val[q] = b0 ? T(0) : p.beta * *cp;
                    dst[q] = cp;
                } else if (j >= jprod) {  // no product contribution in this column
                    if (j < jscale) {
                        val[q] = b0 ? T(0) : p.beta * *cp;
                        dst[q] = cp;
                    }
                } else {
                    const I i = j - Ctu + r;
                    if (i >= 0 && i < n) {  // rows outside the matrix are never touched
                        T acc = b0 ? T(0) : p.beta * *cp;
                        const I vlo = mm_max(mm_max((I)0, j - Bu), i - Al);
                        const I vhi = mm_min(mm_min(k - 1, j + Bl), i + Au);
                        // v ascending: b_vj at sB[jj*SB + (Bu + v - j)], a_iv at sA[(v - v0)*SA + (Au + i - v)]
                        const T *bp = sB + jj * SBi + (int)(Bu + vlo - j);
                        const T *ap = sA + (int)(vlo - vb0) * SAi + (int)(Au + i - vlo);
                        const int cnt = (int)(vhi - vlo) + 1;
                        int v = 0;
                        for (; v + 4 <= cnt; v += 4) {  // four steps' LDS reads issued before their fmas
                            const T x0 = bp[0], x1 = bp[1], x2 = bp[2], x3 = bp[3];
                            const T a0 = ap[0], a1 = ap[sa1], a2 = ap[2 * sa1], a3 = ap[3 * sa1];
                            acc = dfma(x0, a0, acc);
                            acc = dfma(x1, a1, acc);
                            acc = dfma(x2, a2, acc);
                            acc = dfma(x3, a3, acc);
                            bp += 4;
                            ap += 4 * sa1;
                        }
                        for (; v < cnt; ++v) {
                            acc = dfma(*bp, *ap, acc);
                            ++bp;
                            ap += sa1;
                        }
                        val[q] = acc;
                        dst[q] = cp;
                    }
                }
            }
            r += sr;
            jj += sj;
            if (r >= hCi) {
                r -= hCi;
                ++jj;
            }
        }
#pragma unroll
        for (int q = 0; q < NE; ++q)
            if (dst[q]) *dst[q] = val[q];
    }
}

template <typename T>
__global__ void scale_band(i64 rows, i64 cols, i64 ld, T beta, T *C) {
    i64 e = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= rows * cols) return;
    i64 j = e / rows, r = e - j * rows;
    T *p = C + j * ld + r;
    *p = (beta == T(0)) ? T(0) : beta * *p;
}

// flags[b] != 0 iff stored band row b (0 = top super-diagonal) of a
// rows x cols matrix with bands (l,u) holds a nonzero inside the matrix.
template <typename T>
__global__ void band_nonzero(i64 rows, i64 cols, i64 l, i64 u, const T *A, i64 lda, int *flags) {
    i64 e = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    const i64 h = l + u + 1;
    if (e >= h * cols) return;
    i64 j = e / h, t = e - j * h;
    i64 i = j - u + t;
    if (i < 0 || i >= rows) return;
    if (A[j * lda + t] != T(0)) atomicOr(&flags[t], 1);
}

template <typename T>
static int band_flags(i64 rows, i64 cols, i64 l, i64 u, const T *A, i64 lda, int *hflags) {
    const i64 h = l + u + 1;
    int *d = nullptr;
    {
        const int cs = scratch_buffer((size_t)h * sizeof(int), (void **)&d);
        if (cs) return cs;
    }
    BH_HIP(hipMemsetAsync(d, 0, h * sizeof(int), stream()));
    if (h * cols > 0)
        hipLaunchKernelGGL(band_nonzero<T>, dim3((unsigned)cdiv(h * cols, 256)), dim3(256), 0, stream(), rows, cols,
                           l, u, A, lda, d);
    BH_LAUNCH_CHECK();
    BH_HIP(hipMemcpyAsync(hflags, d, h * sizeof(int), hipMemcpyDeviceToHost, stream()));
    BH_HIP(hipStreamSynchronize(stream()));
    return BH_OK;
}

template <typename T>
static int scale_all(i64 rows, i64 cols, i64 ld, T beta, T *C) {
    if (rows * cols == 0) return BH_OK;
    hipLaunchKernelGGL(scale_band<T>, dim3((unsigned)cdiv(rows * cols, 256)), dim3(256), 0, stream(), rows, cols, ld,
                       beta, C);
    BH_LAUNCH_CHECK();
    return BH_OK;
}

template <typename T>
static int gbmm_impl(i64 n, i64 k, i64 m, T alpha, const T *A, i64 lda, i64 Al, i64 Au, const T *B, i64 ldb, i64 Bl,
                     i64 Bu, T beta, T *C, i64 ldc, i64 Ctl, i64 Ctu) {
    if (n == 0 || m == 0) return BH_OK;
    const i64 hC = Ctl + Ctu + 1;
    const i64 Cl = imin(n - 1, Al + Bl), Cu = imin(m - 1, Au + Bu);
    // Strip exactly-zero outer bands when C is too narrow (gbmm.jl:250-279).
    if (Ctu < Cu || Ctl < Cl) {
        const bool upper = Ctu < Cu;
        std::vector<int> fa(Al + Au + 1), fb(Bl + Bu + 1);
        int st = band_flags<T>(n, k, Al, Au, A, lda, fa.data());
        if (st) return st;
        st = band_flags<T>(k, m, Bl, Bu, B, ldb, fb.data());
        if (st) return st;
        auto lead = [](const std::vector<int> &f) {
            i64 b = 0;
            while (b < (i64)f.size() && !f[b]) ++b;
            return b;
        };
        auto trail = [](const std::vector<int> &f) {
            i64 b = 0;
            while (b < (i64)f.size() && !f[f.size() - 1 - b]) ++b;
            return b;
        };
        if (upper) {
            i64 Aur = lead(fa), Bur = lead(fb);
            if (Ctu < Cu - Aur - Bur) {
                set_error("BandError: C has %lld super-diagonals, the product needs %lld", (long long)Ctu,
                          (long long)(Cu - Aur - Bur));
                return BH_ERR_BANDERROR;
            }
            if (Au - Aur < -Al || Bu - Bur < -Bl) return scale_all<T>(hC, m, ldc, beta, C);
            return gbmm_impl<T>(n, k, m, alpha, A + Aur, lda, Al, Au - Aur, B + Bur, ldb, Bl, Bu - Bur, beta, C, ldc,
                                Ctl, Ctu);
        } else {
            i64 Alr = trail(fa), Blr = trail(fb);
            if (Ctl < Cl - Alr - Blr) {
                set_error("BandError: C has %lld sub-diagonals, the product needs %lld", (long long)Ctl,
                          (long long)(Cl - Alr - Blr));
                return BH_ERR_BANDERROR;
            }
            if (Al - Alr < -Au || Bl - Blr < -Bu) return scale_all<T>(hC, m, ldc, beta, C);
            return gbmm_impl<T>(n, k, m, alpha, A, lda, Al - Alr, Au, B, ldb, Bl - Blr, Bu, beta, C, ldc, Ctl, Ctu);
        }
    }
    if (Al < 0 || Au < 0 || Bl < 0 || Bu < 0) {
        set_error("gbmm: negative bandwidths after pruning are handled by the host layer");
        return -6;
    }
    GbmmArgs<T> p{n, k, m, alpha, beta, A, lda, Al, Au, B, ldb, Bl, Bu, C, ldc, Ctl, Ctu, Cl, Cu};
    const i64 hA = Al + Au + 1, hB = Bl + Bu + 1;
    // columns per workgroup: 64 FP64 / 128 FP32 (measured at n = 2^20,
    // (8,8)x(4,4), warm: FP64 92 us at 64 against 101 us at 128, 124 us at 32
    // and 155 us at 256; FP32 65 us at 128, 72 us at 64), halved while the
    // launch would have fewer than 512 workgroups, and kept within 64 KiB of
    // LDS.  The column count changes no result.  BANDEDHIP_GBMM_NC overrides
    // the start (a power of two <= 256).
    i64 NC = sizeof(T) == 4 ? 128 : 64;
    while (NC > 8 && cdiv(m, NC) < 512) NC /= 2;
    if (const char *e = getenv("BANDEDHIP_GBMM_NC")) {
        const i64 v = atoll(e);
        if (v >= 1 && v <= 256 && (v & (v - 1)) == 0) NC = v;
    }
    auto lds_for = [&](i64 nc) { return (size_t)(nc * odd_stride(hB) + (nc + Bl + Bu) * odd_stride(hA)) * sizeof(T); };
    while (NC > 1 && lds_for(NC) > 64 * 1024) NC /= 2;
    const size_t lds = lds_for(NC);
    if (lds > 64 * 1024) {
        set_error("gbmm: bandwidths too large for the LDS-staged kernel");
        return -6;
    }
    // 32-bit index arithmetic when every size (and the sums formed from them) fits
    const i64 big = imax(imax(n, k), imax(m, hC + hA + hB));
    if (big < (i64(1) << 30))
        hipLaunchKernelGGL((gbmm_kernel<T, int, kMmBatch>), dim3((unsigned)cdiv(m, NC)), dim3(kMmThreads), lds,
                           stream(), p, NC);
    else
        hipLaunchKernelGGL((gbmm_kernel<T, i64, kMmBatch>), dim3((unsigned)cdiv(m, NC)), dim3(kMmThreads), lds,
                           stream(), p, NC);
    BH_LAUNCH_CHECK();
    return BH_OK;
}

template <typename T>
static int gbmm_entry(i64 n, i64 k, i64 m, T alpha, const T *A, i64 lda, i64 Al, i64 Au, const T *B, i64 ldb, i64 Bl,
                      i64 Bu, T beta, T *C, i64 ldc, i64 Cl, i64 Cu) {
    BH_ARG(n >= 0, 1);
    BH_ARG(k >= 0, 2);
    BH_ARG(m >= 0, 3);
    BH_ARG(Al >= 0, 7);
    BH_ARG(Au >= 0, 8);
    BH_ARG(lda >= Al + Au + 1, 6);
    BH_ARG(Bl >= 0, 11);
    BH_ARG(Bu >= 0, 12);
    BH_ARG(ldb >= Bl + Bu + 1, 10);
    BH_ARG(Cl >= 0, 16);
    BH_ARG(Cu >= 0, 17);
    BH_ARG(ldc >= Cl + Cu + 1, 15);
    return gbmm_impl<T>(n, k, m, alpha, A, lda, Al, Au, B, ldb, Bl, Bu, beta, C, ldc, Cl, Cu);
}

}  // namespace bh

using namespace bh;

extern "C" {

int bh_dgbmm(bh_int n, bh_int k, bh_int m, double alpha, const double *A, bh_int lda, bh_int Al, bh_int Au,
             const double *B, bh_int ldb, bh_int Bl, bh_int Bu, double beta, double *C, bh_int ldc, bh_int Cl,
             bh_int Cu) {
    return gbmm_entry<double>(n, k, m, alpha, A, lda, Al, Au, B, ldb, Bl, Bu, beta, C, ldc, Cl, Cu);
}

int bh_sgbmm(bh_int n, bh_int k, bh_int m, float alpha, const float *A, bh_int lda, bh_int Al, bh_int Au,
             const float *B, bh_int ldb, bh_int Bl, bh_int Bu, float beta, float *C, bh_int ldc, bh_int Cl,
             bh_int Cu) {
    return gbmm_entry<float>(n, k, m, alpha, A, lda, Al, Au, B, ldb, Bl, Bu, beta, C, ldc, Cl, Cu);
}

}  // extern "C"
