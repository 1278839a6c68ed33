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

// One workgroup per NC consecutive columns of C.
template <typename T>
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
    // 32-bit division (the slab sizes are small).
    const int nB = (int)((j1 - j0) * hB), nA = (int)((v1 - v0) * hA);
    const int hBi = (int)hB, hAi = (int)hA, SBi = (int)SB, SAi = (int)SA;
    if (p.ldb == hB && SB == hB) {
        const T *src = p.B + j0 * p.ldb;
        for (int e = threadIdx.x; e < nB; e += blockDim.x) sB[e] = p.alpha * src[e];  // alpha*b, as the reference's gbmv
    } else {
        for (int e = threadIdx.x; e < nB; e += blockDim.x) {
            const int c = e / hBi, t = e - c * hBi;
            sB[c * SBi + t] = p.alpha * p.B[(j0 + c) * p.ldb + t];
        }
    }
    if (p.lda == hA && SA == hA) {
        const T *src = p.A + v0 * p.lda;
        for (int e = threadIdx.x; e < nA; e += blockDim.x) sA[e] = src[e];
    } else {
        for (int e = threadIdx.x; e < nA; e += blockDim.x) {
            const int c = e / hAi, t = e - c * hAi;
            sA[c * SAi + t] = p.A[(v0 + c) * p.lda + t];
        }
    }
    __syncthreads();

    const i64 shift = p.Ctu - p.Cu;  // first stored row of the product bands
    const int hCi = (int)hC, nC = (int)((j1 - j0) * hC);
    for (int e = threadIdx.x; e < nC; e += blockDim.x) {
        const int jj = e / hCi;
        const i64 r = e - jj * hCi, j = j0 + jj;
        T *cp = p.C + j * p.ldc + r;
        const bool extra = (r < shift) || (r > p.Ctu + p.Cl);
        if (extra) {
            *cp = (p.beta == T(0)) ? T(0) : p.beta * *cp;
            continue;
        }
        if (j >= p.k + p.Bu) {  // no product contribution in this column
            if (j < imin(p.m, p.n + p.Cu)) *cp = (p.beta == T(0)) ? T(0) : p.beta * *cp;
            continue;
        }
        const i64 i = j - p.Ctu + r;
        if (i < 0 || i >= p.n) continue;  // outside the matrix: never touched
        T acc = (p.beta == T(0)) ? T(0) : p.beta * *cp;
        const i64 vlo = imax(imax(0, j - p.Bu), i - p.Al), vhi = imin(imin(p.k - 1, j + p.Bl), i + p.Au);
        // v ascending: b_vj at sB[jj*SB + (Bu + v - j)], a_iv at sA[(v - v0)*SA + (Au + i - v)]
        const T *bp = sB + jj * SBi + (int)(p.Bu + vlo - j);
        const T *ap = sA + (int)(vlo - v0) * SAi + (int)(p.Au + i - vlo);
        const int cnt = (int)(vhi - vlo) + 1, sa1 = SAi - 1;
        int v = 0;
        for (; v + 4 <= cnt; v += 4) {  // four steps' LDS reads issued before their fmas
            const T b0 = bp[0], b1 = bp[1], b2 = bp[2], b3 = bp[3];
            const T a0 = ap[0], a1 = ap[sa1], a2 = ap[2 * sa1], a3 = ap[3 * sa1];
            acc = dfma(b0, a0, acc);
            acc = dfma(b1, a1, acc);
            acc = dfma(b2, a2, acc);
            acc = dfma(b3, a3, acc);
            bp += 4;
            ap += 4 * sa1;
        }
        for (; v < cnt; ++v) {
            acc = dfma(*bp, *ap, acc);
            ++bp;
            ap += sa1;
        }
        *cp = acc;
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
    // columns per workgroup: 64 (measured at n = 2^20, (8,8)x(4,4): 3.1 TB/s
    // against 1.5 TB/s at 256, where 54 KiB of LDS per workgroup left two
    // workgroups per CU), halved while the launch would have fewer than 512
    // workgroups, and kept within 64 KiB of LDS.  The column count changes no
    // result.  BANDEDHIP_GBMM_NC overrides the start (a power of two <= 256).
    i64 NC = 64;
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
    hipLaunchKernelGGL(gbmm_kernel<T>, dim3((unsigned)cdiv(m, NC)), dim3(kMmThreads), lds, stream(), p, NC);
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
