// bh_gbtrf.hip -- partial-pivot band LU on gfx950.
//
// Restates LAPACK xGBTF2, the routine the reference's LAPACK.gbtrf!
// (src/banded/BandedLU.jl:98) runs for ku <= 64, step for step:
//   for each column j: zero the fill-in rows of column j+kv; pick the pivot
//   as the first row of max |a| among rows j..j+km (IDAMAX); record ipiv;
//   swap rows across columns j..ju; scale the multipliers by the rounded
//   reciprocal of the pivot; rank-1 update a_ic = fma(l_i, -u_jc, a_ic).
// The rounding of every step matches the CPU routine, so factors and pivots
// are bit-identical to it.
//
// Kernel: one workgroup runs the column chain of one system; the working
// window lives in global memory (L2-resident: at most (2kl+ku+1) x (kl+ku+1)
// entries are live) and the phases of a step are separated by workgroup
// barriers.  Each phase is spread over the workgroup's lanes.
#include "bh_common.hpp"

#include <cstdlib>
#include <cstring>

namespace bh {

constexpr int kLuThreads = 256;

template <typename T>
struct LuSys {
    T *ab;
    i64 *ipiv;
    i64 m, n, kl, ku, ldab;
};

// Argmax over |col[kv+t]|, t in [0, km], first index on ties; whole block.
template <typename T>
__device__ __forceinline__ i64 block_pivot(const T *col, i64 km, T *sval, int *sidx) {
    T best = T(-1);
    int bidx = 0x7fffffff;
    for (i64 t = threadIdx.x; t <= km; t += blockDim.x) {
        T v = dabs(col[t]);
        if (v > best) {  // strictly greater keeps the first index per thread
            best = v;
            bidx = (int)t;
        }
    }
    // wave reduction: larger value wins, equal values keep the smaller index
    for (int off = 32; off > 0; off >>= 1) {
        T ov = __shfl_xor(best, off, 64);
        int oi = __shfl_xor(bidx, off, 64);
        if (ov > best || (ov == best && oi < bidx)) {
            best = ov;
            bidx = oi;
        }
    }
    const int wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        sval[wid] = best;
        sidx[wid] = bidx;
    }
    __syncthreads();
    T b = sval[0];
    int bi = sidx[0];
    for (int w = 1; w < nw; ++w) {
        if (sval[w] > b || (sval[w] == b && sidx[w] < bi)) {
            b = sval[w];
            bi = sidx[w];
        }
    }
    __syncthreads();
    return bi == 0x7fffffff ? 0 : bi;
}

template <typename T>
__device__ void gbtf2_block(const LuSys<T> s, i64 *info_out) {
    __shared__ T sval[16];
    __shared__ int sidx[16];
    const i64 kl = s.kl, ku = s.ku, kv = kl + ku, ld = s.ldab, m = s.m, n = s.n;
    T *ab = s.ab;
    i64 info = 0;
    if (m == 0 || n == 0) {
        if (threadIdx.x == 0 && info_out) *info_out = 0;
        return;
    }
    // fill-in rows of columns ku+1 .. min(kv,n)-1 start at zero
    for (i64 c = ku + 1; c < imin(kv, n); ++c)
        for (i64 r = kv - c + threadIdx.x; r < kl; r += blockDim.x) ab[c * ld + r] = T(0);
    __syncthreads();
    i64 ju = 0;
    const i64 steps = imin(m, n);
    for (i64 j = 0; j < steps; ++j) {
        if (j + kv < n)
            for (i64 r = threadIdx.x; r < kl; r += blockDim.x) ab[(j + kv) * ld + r] = T(0);
        const i64 km = imin(kl, m - 1 - j);
        T *colj = ab + j * ld + kv;  // colj[t] = A(j+t, j)
        const i64 jp = block_pivot(colj, km, sval, sidx);
        if (threadIdx.x == 0) s.ipiv[j] = jp + j + 1;
        const T piv = colj[jp];
        if (piv != T(0)) {
            ju = imax(ju, imin(j + ku + jp, n - 1));
            if (jp != 0) {
                for (i64 c = j + threadIdx.x; c <= ju; c += blockDim.x) {
                    T *p = ab + c * ld + kv + jp + j - c;
                    T *q = ab + c * ld + kv + j - c;
                    T tmp = *p;
                    *p = *q;
                    *q = tmp;
                }
                __syncthreads();
            }
            if (km > 0) {
                const T r = T(1) / colj[0];
                for (i64 t = 1 + threadIdx.x; t <= km; t += blockDim.x) colj[t] = colj[t] * r;
                __syncthreads();
                const i64 ncol = ju - j;
                const i64 tot = km * ncol;
                for (i64 e = threadIdx.x; e < tot; e += blockDim.x) {
                    const i64 cc = e / km, t = e - cc * km + 1, c = j + 1 + cc;
                    T *colc = ab + c * ld;
                    const T negu = -colc[kv + j - c];
                    colc[kv + j + t - c] = dfma(colj[t], negu, colc[kv + j + t - c]);
                }
            }
        } else if (info == 0) {
            info = j + 1;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0 && info_out) *info_out = info;
}

template <typename T>
__global__ void __launch_bounds__(kLuThreads) gbtf2_single(LuSys<T> s, i64 *info_out) {
    gbtf2_block<T>(s, info_out);
}

template <typename T>
int reg_gbtrf_launch(i64 m, i64 n, i64 kl, i64 ku, T *ab, i64 ldab, i64 *ipiv, i64 *dinfo, hipStream_t st);
template <typename T>
int blk_gbtrf_launch(i64 m, i64 n, i64 kl, i64 ku, T *ab, i64 ldab, i64 *ipiv, i64 *dinfo, hipStream_t st);
template <typename T>
int small_gbtrf_launch(i64 m, i64 n, i64 kl, i64 ku, T *ab, i64 ldab, i64 *ipiv, i64 *dinfo, hipStream_t st);
template <typename T>
int mfma_gbtrf_launch(i64 m, i64 n, i64 kl, i64 ku, T *ab, i64 ldab, i64 *ipiv, i64 *dinfo, hipStream_t st);

// Kernel choice: BANDEDHIP_LU_KERNEL = "small" | "blk" | "reg" | "global" |
// "mfma" forces one (where it applies).  By default, as LAPACK xGBTRF
// decides between its unblocked and blocked code (ILAENV block size 32 for
// ku > 64, blocked when the block size is at most kl): the MFMA-blocked
// kernel for ku >= 65 and kl >= 32, whose pivots are xGBTRF's and whose
// factors agree with it to rounding.  Otherwise the xGBTF2-exact kernels:
// the one-wave kernel for kl <= 16, the register-window blocked kernel for
// kl >= 16 where that did not apply, the per-step register kernel otherwise.
constexpr i64 kLongLu = 16384;  // systems at least this long take the MFMA factorisation (kl, ku >= 8)

static int lu_kernel_override() {
    const char *e = getenv("BANDEDHIP_LU_KERNEL");
    if (!e) return 0;
    if (!strcmp(e, "blk")) return 1;
    if (!strcmp(e, "reg")) return 2;
    if (!strcmp(e, "global")) return 3;
    if (!strcmp(e, "small")) return 4;
    if (!strcmp(e, "mfma")) return 5;
    return 0;
}

template <typename T>
static int gbtrf_impl(i64 m, i64 n, i64 kl, i64 ku, T *ab, i64 ldab, i64 *ipiv, i64 *info_host) {
    BH_ARG(m >= 0, 1);
    BH_ARG(n >= 0, 2);
    BH_ARG(kl >= 0, 3);
    BH_ARG(ku >= 0, 4);
    BH_ARG(ldab >= 2 * kl + ku + 1, 6);
    BH_ARG(info_host != nullptr, 8);
    *info_host = 0;
    if (m == 0 || n == 0) return BH_OK;
    i64 *dinfo = nullptr;
    {
        const int cs = ctrl_words(&dinfo);
        if (cs) return cs;
    }
    BH_HIP(hipMemsetAsync(dinfo, 0, 2 * sizeof(i64), stream()));
    // register-resident kernel when the window fits one of its shapes,
    // otherwise the barrier-stepped kernel working in global memory
    const int ov = lu_kernel_override();
    int st = -1;
    // The MFMA-blocked factorisation (pivots equal xGBTF2's, factors to
    // rounding): where the reference itself runs blocked xGBTRF (ku >= 65,
    // kl >= 32), and for long systems with kl, ku >= 8, where the one-wave
    // chain of the narrow kernels is the longer path (n = 262144, (16,16):
    // 142 ms against 278 ms).  Shorter narrow systems keep the bit-exact
    // xGBTF2 kernels.
    if (ov == 5 || (ov == 0 && ((ku >= 65 && kl >= 32) || (kl >= 8 && ku >= 8 && imin(m, n) >= kLongLu))))
        st = mfma_gbtrf_launch<T>(m, n, kl, ku, ab, ldab, ipiv, dinfo, stream());
    // The one-wave lane kernel exists for kl <= 8 and kl = 16; above 16 the
    // blocked kernel beats the generic one-wave kernel (at n = 65536, (24,24):
    // 113.5 ms against 278.6 ms, identical factors), so it goes first there.
    if (st == -1 && (ov == 4 || (ov == 0 && kl <= 16))) st = small_gbtrf_launch<T>(m, n, kl, ku, ab, ldab, ipiv, dinfo, stream());
    if (st == -1 && (ov == 1 || (ov == 0 && kl >= 16)))
        st = blk_gbtrf_launch<T>(m, n, kl, ku, ab, ldab, ipiv, dinfo, stream());
    if (st == -1 && ov != 3) st = reg_gbtrf_launch<T>(m, n, kl, ku, ab, ldab, ipiv, dinfo, stream());
    if (st == -1) {
        LuSys<T> s{ab, ipiv, m, n, kl, ku, ldab};
        hipLaunchKernelGGL(gbtf2_single<T>, dim3(1), dim3(kLuThreads), 0, stream(), s, dinfo);
        BH_LAUNCH_CHECK();
    } else if (st != BH_OK) {
        return st;
    }
    i64 hinfo[2] = {0, 0};
    BH_HIP(hipMemcpyAsync(hinfo, dinfo, 2 * sizeof(i64), hipMemcpyDeviceToHost, stream()));
    BH_HIP(hipStreamSynchronize(stream()));
    if (hinfo[1] != 0) {
        set_error("gbtrf: an internal wait of the blocked kernel gave up; the factors are not valid");
        return BH_ERR_INTERNAL;
    }
    *info_host = hinfo[0];
    return BH_OK;
}

}  // namespace bh

using namespace bh;

extern "C" {

int bh_dgbtrf(bh_int m, bh_int n, bh_int kl, bh_int ku, double *AB, bh_int ldab, bh_int *ipiv, bh_int *info_host) {
    return gbtrf_impl<double>(m, n, kl, ku, AB, ldab, ipiv, info_host);
}

int bh_sgbtrf(bh_int m, bh_int n, bh_int kl, bh_int ku, float *AB, bh_int ldab, bh_int *ipiv, bh_int *info_host) {
    return gbtrf_impl<float>(m, n, kl, ku, AB, ldab, ipiv, info_host);
}

}  // extern "C"
