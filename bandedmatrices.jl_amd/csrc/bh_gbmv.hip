// bh_gbmv.hip -- banded matrix-vector products on gfx950.
//
// y <- alpha*op(A)*x + beta*y with A in band storage.  The 'N' kernels give
// every y_i the exact rounding sequence of the column-ordered BLAS gbmv that
// BandedMatrices.jl calls (src/generic/matmul.jl:21-23):
//     y_i = (beta == 0 ? 0 : beta*y_i);  y_i = fma(alpha*x_j, a_ij, y_i), j ascending
// so results are bit-identical to that CPU path.  The kernels are HBM-bound
// (A is read once, 8*(kl+ku+1) bytes per column for FP64):
//   * narrow bands: a workgroup stages a contiguous slab of band columns plus
//     alpha*x in LDS with wide coalesced loads, then one thread per row walks
//     its band row from LDS (odd column stride => no LDS bank conflicts);
//   * wide bands (the slab would not fit in LDS): one wave per 64 rows walks
//     the columns j in order; lane i reads a_ij, and because a column's band is
//     contiguous the 64 lanes read one contiguous segment per step.
#include "bh_common.hpp"

namespace bh {

constexpr int kTileThreads = 256;
constexpr size_t kMaxTileLds = 64 * 1024;

// Column stride of the LDS slab: the band height rounded to an odd count so
// that walking a band row (stride S-1 in the slab) spreads over LDS banks.
__host__ __device__ inline i64 slab_stride(i64 kl, i64 ku) {
    i64 h = kl + ku + 1;
    return (h % 2 == 0) ? h + 1 : h;
}

template <typename T>
static size_t tile_lds_bytes(i64 R, i64 kl, i64 ku, int nvec) {
    i64 W = R + kl + ku;
    return (size_t)(W * slab_stride(kl, ku) + (i64)nvec * W) * sizeof(T);
}

// Stage band columns [c0,c1) (all kl+ku+1 stored rows) into the slab.  Each
// thread issues kStageBatch loads before it stores any (load indices clamped
// to the last element, so no load is conditional): a workgroup waits for
// memory a few times, not once per element.
constexpr int kStageBatch = 8;

template <typename T>
__device__ __forceinline__ void load_slab(T *__restrict__ slab, i64 S, const T *__restrict__ A, i64 lda,
                                          i64 c0, i64 c1, i64 h) {
    // The slab holds at most 64 KiB, so 32-bit indexing is exact here.
    const int total = (int)((c1 - c0) * h);
    if (total <= 0) return;
    const unsigned hh = (unsigned)h;
    const int nt = blockDim.x;
    const unsigned Su = (unsigned)S;
    if (lda == h) {
        const T *src = A + c0 * lda;
        for (int base = threadIdx.x; base < total; base += kStageBatch * nt) {
            T v[kStageBatch];
#pragma unroll
            for (int u = 0; u < kStageBatch; ++u) v[u] = src[min(base + u * nt, total - 1)];
#pragma unroll
            for (int u = 0; u < kStageBatch; ++u) {
                const int e = base + u * nt;
                if (e < total) {
                    const unsigned c = (unsigned)e / hh, t = (unsigned)e - c * hh;
                    slab[c * Su + t] = v[u];
                }
            }
        }
    } else {
        for (int base = threadIdx.x; base < total; base += kStageBatch * nt) {
            T v[kStageBatch];
#pragma unroll
            for (int u = 0; u < kStageBatch; ++u) {
                const int e = min(base + u * nt, total - 1);
                const unsigned c = (unsigned)e / hh, t = (unsigned)e - c * hh;
                v[u] = A[(c0 + c) * lda + t];
            }
#pragma unroll
            for (int u = 0; u < kStageBatch; ++u) {
                const int e = base + u * nt;
                if (e < total) {
                    const unsigned c = (unsigned)e / hh, t = (unsigned)e - c * hh;
                    slab[c * Su + t] = v[u];
                }
            }
        }
    }
}

// ---------------------------------------------------------------- 'N' tile
// One workgroup per R = blockDim.x consecutive rows; nrhs vectors handled in a
// loop so the slab (A) is read from HBM once for all right-hand sides.
template <typename T>
__global__ void __launch_bounds__(kTileThreads)
gbmv_n_tile(i64 m, i64 n, i64 kl, i64 ku, i64 nrhs, T alpha, const T *__restrict__ A, i64 lda,
            const T *__restrict__ X, i64 incx, i64 ldx, T beta, T *__restrict__ Y, i64 incy, i64 ldy) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T *slab = reinterpret_cast<T *>(smem_raw);
    const i64 R = blockDim.x;
    const i64 h = kl + ku + 1;
    const i64 S = slab_stride(kl, ku);
    const i64 r0 = (i64)blockIdx.x * R;
    const i64 c0 = imax(0, r0 - kl);
    const i64 c1 = imin(n, r0 + R + ku);
    const i64 W = R + kl + ku;
    T *ax = slab + W * S;

    load_slab(slab, S, A, lda, c0, c1, h);
    const i64 i = r0 + threadIdx.x;
    const i64 jlo = imax(0, i - kl), jhi = imin(n - 1, i + ku);
    for (i64 r = 0; r < nrhs; ++r) {
        const T *x = X + r * ldx;
        for (i64 c = threadIdx.x; c < c1 - c0; c += blockDim.x) ax[c] = alpha * x[(c0 + c) * incx];
        __syncthreads();
        if (i < m) {
            T *yp = Y + r * ldy + i * incy;
            T acc = (beta == T(0)) ? T(0) : beta * *yp;
            for (i64 j = jlo; j <= jhi; ++j) acc = dfma(ax[j - c0], slab[(j - c0) * S + (ku + i - j)], acc);
            *yp = acc;
        }
        __syncthreads();
    }
}

// -------------------------------------------------------------- 'N' direct
// One wave per 64 rows for bands too wide for the slab.  Step j reads the
// contiguous segment a[ku+r0-j .. ku+r0+63-j, j] of column j.
template <typename T>
__global__ void __launch_bounds__(256)
gbmv_n_direct(i64 m, i64 n, i64 kl, i64 ku, i64 nrhs, T alpha, const T *__restrict__ A, i64 lda,
              const T *__restrict__ X, i64 incx, i64 ldx, T beta, T *__restrict__ Y, i64 incy, i64 ldy) {
    const int lane = threadIdx.x & 63;
    const i64 wave = (i64)blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
    const i64 r0 = wave * 64;
    if (r0 >= m) return;
    const i64 i = r0 + lane;
    const i64 jlo = imax(0, r0 - kl), jhi = imin(n - 1, r0 + 63 + ku);
    for (i64 r = 0; r < nrhs; ++r) {
        const T *x = X + r * ldx;
        T *yp = Y + r * ldy + i * incy;
        T acc = T(0);
        if (i < m && beta != T(0)) acc = beta * *yp;
        for (i64 j = jlo; j <= jhi; ++j) {
            const T axj = alpha * x[j * incx];
            const i64 t = ku + i - j;
            if (t >= 0 && t < kl + ku + 1 && i < m) acc = dfma(axj, A[j * lda + t], acc);
        }
        if (i < m) *yp = acc;
    }
}

// -------------------------------------------------------------- 'N' wide
// Bands too wide for the slab, several right-hand sides: a workgroup owns 64
// rows x 16 right-hand sides and walks the band's columns in chunks of 16,
// staging the 64 x 16 block of A (each column's 64 rows are contiguous in
// band storage) and alpha*x for the chunk in LDS; a thread keeps 4
// right-hand sides of one row.  Every y_i still accumulates fma(alpha x_j,
// a_ij, acc) over ascending j, skipping entries outside the band, exactly
// as the direct kernel (and OpenBLAS's column sweep) does.
template <typename T>
__global__ void __launch_bounds__(256)
gbmv_n_wide(i64 m, i64 n, i64 kl, i64 ku, i64 nrhs, T alpha, const T *__restrict__ A, i64 lda,
            const T *__restrict__ X, i64 incx, i64 ldx, T beta, T *__restrict__ Y, i64 incy, i64 ldy) {
    __shared__ T As[16][65];  // [column in chunk][row]
    __shared__ T Xs[16][17];  // [column in chunk][right-hand side]: alpha * x
    const int tid = threadIdx.x, rr = tid & 63, rg = tid >> 6;  // row, group of 4 right-hand sides
    const i64 r0 = (i64)blockIdx.x * 64, c0 = (i64)blockIdx.y * 16;
    const i64 i = r0 + rr;
    const i64 h = kl + ku;
    T acc[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const i64 r = c0 + 4 * rg + q;
        acc[q] = (i < m && r < nrhs && beta != T(0)) ? beta * Y[r * ldy + i * incy] : T(0);
    }
    const i64 jlo = imax(0, r0 - kl), jhi = imin(n - 1, r0 + 63 + ku);
    for (i64 jb = jlo; jb <= jhi; jb += 16) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {  // 64 x 16 block of A, one column per 64 threads
            const int jj = rg + 4 * q;
            const i64 j = jb + jj, t = ku + i - j;
            As[jj][rr] = (j <= jhi && t >= 0 && t <= h && i < m) ? A[j * lda + t] : T(0);
        }
        {
            const int jj = tid >> 4, r = tid & 15;
            const i64 j = jb + jj;
            Xs[jj][r] = (j <= jhi && c0 + r < nrhs) ? alpha * X[(c0 + r) * ldx + j * incx] : T(0);
        }
        __syncthreads();
#pragma unroll
        for (int jj = 0; jj < 16; ++jj) {
            const i64 j = jb + jj, t = ku + i - j;
            if (j <= jhi && t >= 0 && t <= h) {
                const T a = As[jj][rr];
#pragma unroll
                for (int q = 0; q < 4; ++q) acc[q] = dfma(Xs[jj][4 * rg + q], a, acc[q]);
            }
        }
        __syncthreads();
    }
    if (i < m) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const i64 r = c0 + 4 * rg + q;
            if (r < nrhs) Y[r * ldy + i * incy] = acc[q];
        }
    }
}

// ----------------------------------------------------------------- 'T' tile
// y_j = beta*y_j + alpha * sum_i a_ij x_i: one thread per column j, the dot
// accumulated in ascending i from the LDS slab (columns are contiguous).
template <typename T>
__global__ void __launch_bounds__(kTileThreads)
gbmv_t_tile(i64 m, i64 n, i64 kl, i64 ku, i64 nrhs, T alpha, const T *__restrict__ A, i64 lda,
            const T *__restrict__ X, i64 incx, i64 ldx, T beta, T *__restrict__ Y, i64 incy, i64 ldy) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T *slab = reinterpret_cast<T *>(smem_raw);
    const i64 R = blockDim.x;
    const i64 h = kl + ku + 1;
    const i64 S = slab_stride(kl, ku);
    const i64 j0 = (i64)blockIdx.x * R;
    const i64 c1 = imin(n, j0 + R);
    const i64 W = R + kl + ku;
    T *xs = slab + W * S;  // x rows [j0-ku, j0+R+kl)
    const i64 x0 = j0 - ku;

    // the first right-hand side's x (the thread's first two entries) is
    // loaded before the slab, so both wait for memory together
    T xpre[2];
#pragma unroll
    for (int u = 0; u < 2; ++u) {
        const i64 e = threadIdx.x + u * R, row = x0 + e;
        xpre[u] = (e < W && row >= 0 && row < m) ? X[row * incx] : T(0);
    }
    load_slab(slab, S, A, lda, j0, c1, h);
    const i64 j = j0 + threadIdx.x;
    // the whole tile is full columns: j0 - ku >= 0, the last column's band ends inside the rows
    const bool interior = j0 - ku >= 0 && j0 + R <= n && j0 + R - 1 + kl <= m - 1;
    for (i64 r = 0; r < nrhs; ++r) {
        const T *x = X + r * ldx;
        i64 e = threadIdx.x;
        if (r == 0) {
#pragma unroll
            for (int u = 0; u < 2; ++u, e += blockDim.x)
                if (e < W) xs[e] = xpre[u];
        }
        for (; e < W; e += blockDim.x) {
            i64 row = x0 + e;
            xs[e] = (row >= 0 && row < m) ? x[row * incx] : T(0);
        }
        __syncthreads();
        if (interior) {
            // every column of the tile has its whole band inside the matrix:
            // 32-bit indices, eight LDS reads issued before their fmas, the
            // same ascending order
            const int jj = threadIdx.x, hi = (int)h, Si = (int)S;
            const T *ap = slab + jj * Si;  // a_(j-ku+t, j) at ap[t]
            const T *xp = xs + jj;         // x_(j-ku+t) at xp[t]
            T acc = T(0);
            int t = 0;
            for (; t + 8 <= hi; t += 8) {
                T av[8], xv[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    av[u] = ap[t + u];
                    xv[u] = xp[t + u];
                }
#pragma unroll
                for (int u = 0; u < 8; ++u) acc = dfma(av[u], xv[u], acc);
            }
            for (; t < hi; ++t) acc = dfma(ap[t], xp[t], acc);
            T *yp = Y + r * ldy + j * incy;
            T yv = (beta == T(0)) ? T(0) : beta * *yp;
            *yp = dfma(alpha, acc, yv);
        } else if (j < n) {
            const i64 ilo = imax(0, j - ku), ihi = imin(m - 1, j + kl);
            T acc = T(0);
            for (i64 ii = ilo; ii <= ihi; ++ii)
                acc = dfma(slab[(j - j0) * S + (ku + ii - j)], xs[ii - x0], acc);
            T *yp = Y + r * ldy + j * incy;
            T yv = (beta == T(0)) ? T(0) : beta * *yp;
            *yp = dfma(alpha, acc, yv);
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------- 'T' direct
// Wide bands: one wave per column, lanes stride the band, then a wave sum.
template <typename T>
__global__ void __launch_bounds__(256)
gbmv_t_direct(i64 m, i64 n, i64 kl, i64 ku, i64 nrhs, T alpha, const T *__restrict__ A, i64 lda,
              const T *__restrict__ X, i64 incx, i64 ldx, T beta, T *__restrict__ Y, i64 incy, i64 ldy) {
    const int lane = threadIdx.x & 63;
    const i64 j = (i64)blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
    if (j >= n) return;
    const i64 ilo = imax(0, j - ku), ihi = imin(m - 1, j + kl);
    for (i64 r = 0; r < nrhs; ++r) {
        const T *x = X + r * ldx;
        T acc = T(0);
        for (i64 ii = ilo + lane; ii <= ihi; ii += 64) acc = dfma(A[j * lda + ku + ii - j], x[ii * incx], acc);
        for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
        if (lane == 0) {
            T *yp = Y + r * ldy + j * incy;
            T yv = (beta == T(0)) ? T(0) : beta * *yp;
            *yp = dfma(alpha, acc, yv);
        }
    }
}

// -------------------------------------------------------------- beta only
template <typename T>
__global__ void scale_vec(i64 len, i64 nrhs, T beta, T *__restrict__ Y, i64 incy, i64 ldy) {
    i64 e = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= len * nrhs) return;
    i64 r = e / len, i = e - r * len;
    T *p = Y + r * ldy + i * incy;
    *p = (beta == T(0)) ? T(0) : beta * *p;
}

template <typename T>
int slab_gbmv_n(i64 m, i64 n, i64 kl, i64 ku, i64 nrhs, T alpha, const T *A, i64 lda, const T *X, i64 incx, i64 ldx,
                T beta, T *Y, i64 incy, i64 ldy, hipStream_t st);

template <typename T>
static int gbmv_impl(char trans, i64 m, i64 n, i64 kl, i64 ku, i64 nrhs, T alpha, const T *A, i64 lda,
                     const T *X, i64 incx, i64 ldx, T beta, T *Y, i64 incy, i64 ldy) {
    const bool nt = (trans == 'N' || trans == 'n');
    const i64 leny = nt ? m : n, lenx = nt ? n : m;
    if (leny == 0 || nrhs == 0) return BH_OK;
    hipStream_t s = stream();
    if (lenx == 0) {
        i64 tot = leny * nrhs;
        hipLaunchKernelGGL(scale_vec<T>, dim3((unsigned)cdiv(tot, 256)), dim3(256), 0, s, leny, nrhs, beta, Y, incy, ldy);
        BH_LAUNCH_CHECK();
        return BH_OK;
    }
    if (nt) {
        const int st = slab_gbmv_n<T>(m, n, kl, ku, nrhs, alpha, A, lda, X, incx, ldx, beta, Y, incy, ldy, s);
        if (st != -1) return st;
    }
    const i64 R = kTileThreads;
    const size_t lds = tile_lds_bytes<T>(R, kl, ku, 1);
    if (lds <= kMaxTileLds) {
        dim3 grid((unsigned)cdiv(leny, R));
        if (nt)
            hipLaunchKernelGGL(gbmv_n_tile<T>, grid, dim3(kTileThreads), lds, s, m, n, kl, ku, nrhs, alpha, A, lda,
                               X, incx, ldx, beta, Y, incy, ldy);
        else
            hipLaunchKernelGGL(gbmv_t_tile<T>, grid, dim3(kTileThreads), lds, s, m, n, kl, ku, nrhs, alpha, A, lda,
                               X, incx, ldx, beta, Y, incy, ldy);
    } else if (nt && nrhs > 1) {
        dim3 grid((unsigned)cdiv(m, 64), (unsigned)cdiv(nrhs, 16));
        hipLaunchKernelGGL(gbmv_n_wide<T>, grid, dim3(256), 0, s, m, n, kl, ku, nrhs, alpha, A, lda, X, incx, ldx,
                           beta, Y, incy, ldy);
    } else if (nt) {
        dim3 grid((unsigned)cdiv(cdiv(m, 64), 4));
        hipLaunchKernelGGL(gbmv_n_direct<T>, grid, dim3(256), 0, s, m, n, kl, ku, nrhs, alpha, A, lda, X, incx, ldx,
                           beta, Y, incy, ldy);
    } else {
        dim3 grid((unsigned)cdiv(n, 4));
        hipLaunchKernelGGL(gbmv_t_direct<T>, grid, dim3(256), 0, s, m, n, kl, ku, nrhs, alpha, A, lda, X, incx, ldx,
                           beta, Y, incy, ldy);
    }
    BH_LAUNCH_CHECK();
    return BH_OK;
}

template <typename T>
static int gbmv_check(char trans, i64 m, i64 n, i64 kl, i64 ku, i64 nrhs, const T *A, i64 lda, i64 incx,
                      i64 incy, bool multi) {
    BH_ARG(trans == 'N' || trans == 'n' || trans == 'T' || trans == 't' || trans == 'C' || trans == 'c', 1);
    BH_ARG(m >= 0, 2);
    BH_ARG(n >= 0, 3);
    BH_ARG(kl >= 0, 4);
    BH_ARG(ku >= 0, 5);
    if (multi) BH_ARG(nrhs >= 0, 6);
    BH_ARG(lda >= kl + ku + 1, multi ? 9 : 8);
    if (!multi) {
        BH_ARG(incx > 0, 10);
        BH_ARG(incy > 0, 13);
    }
    return BH_OK;
}

}  // namespace bh

using namespace bh;

extern "C" {

int bh_dgbmv(char trans, bh_int m, bh_int n, bh_int kl, bh_int ku, double alpha, const double *A, bh_int lda,
             const double *x, bh_int incx, double beta, double *y, bh_int incy) {
    int st = gbmv_check<double>(trans, m, n, kl, ku, 1, A, lda, incx, incy, false);
    if (st) return st;
    return gbmv_impl<double>(trans, m, n, kl, ku, 1, alpha, A, lda, x, incx, 0, beta, y, incy, 0);
}

int bh_sgbmv(char trans, bh_int m, bh_int n, bh_int kl, bh_int ku, float alpha, const float *A, bh_int lda,
             const float *x, bh_int incx, float beta, float *y, bh_int incy) {
    int st = gbmv_check<float>(trans, m, n, kl, ku, 1, A, lda, incx, incy, false);
    if (st) return st;
    return gbmv_impl<float>(trans, m, n, kl, ku, 1, alpha, A, lda, x, incx, 0, beta, y, incy, 0);
}

int bh_dgbmv_multi(char trans, bh_int m, bh_int n, bh_int kl, bh_int ku, bh_int nrhs, double alpha, const double *A,
                   bh_int lda, const double *X, bh_int ldx, double beta, double *Y, bh_int ldy) {
    int st = gbmv_check<double>(trans, m, n, kl, ku, nrhs, A, lda, 1, 1, true);
    if (st) return st;
    const bool nt = (trans == 'N' || trans == 'n');
    BH_ARG(ldx >= imax(1, nt ? n : m), 11);
    BH_ARG(ldy >= imax(1, nt ? m : n), 14);
    return gbmv_impl<double>(trans, m, n, kl, ku, nrhs, alpha, A, lda, X, 1, ldx, beta, Y, 1, ldy);
}

int bh_sgbmv_multi(char trans, bh_int m, bh_int n, bh_int kl, bh_int ku, bh_int nrhs, float alpha, const float *A,
                   bh_int lda, const float *X, bh_int ldx, float beta, float *Y, bh_int ldy) {
    int st = gbmv_check<float>(trans, m, n, kl, ku, nrhs, A, lda, 1, 1, true);
    if (st) return st;
    const bool nt = (trans == 'N' || trans == 'n');
    BH_ARG(ldx >= imax(1, nt ? n : m), 11);
    BH_ARG(ldy >= imax(1, nt ? m : n), 14);
    return gbmv_impl<float>(trans, m, n, kl, ku, nrhs, alpha, A, lda, X, 1, ldx, beta, Y, 1, ldy);
}

}  // extern "C"
