/*
 * bandoracle.c -- CPU oracle for the banded hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: it is compiled into
 * oracle/liboracle.so and loaded by tests/, by __graft_entry__.smoke() and by
 * the cpu_baseline leg of bench.py.  Nothing in the product
 * (bandedmatrices.jl_amd/) links, loads or calls it.
 *
 * It restates, in plain C, the arithmetic that BandedMatrices.jl reaches on its
 * hot path:
 *   - OpenBLAS ?gbmv 'N'/'T'        (src/generic/matmul.jl:21-23, src/blas.jl:16-28)
 *   - gbmm! column decomposition    (src/banded/gbmm.jl:18-158,207-340)
 *   - LAPACK ?gbtf2 / ?gbtrf        (src/banded/BandedLU.jl:90-103)
 *   - LAPACK ?gbtrs 'N'/'T'         (src/banded/linalg.jl:24-47)
 *   - BLAS ?tbsv upper              (src/tribanded.jl:75-83, src/blas.jl:123-141)
 *   - BLAS ?tbsv / ?tbmv, all cases (src/tribanded.jl:47-54,75-96, src/blas.jl:71-147)
 *   - LAPACK ?pbtf2 and BLAS ?sbmv  (src/symbanded/BandedCholesky.jl:1-11, symbanded.jl:72-91)
 *   - banded LDL' and its solve     (src/symbanded/ldlt.jl:8-69)
 *   - banded Householder QR         (src/banded/bandedqr.jl:8-35) with the
 *     reflector!/reflectorApply! pair it imports from ArrayLayouts
 *   - Q / Q' application and solve  (src/banded/bandedqr.jl:38-94,175-206)
 *
 * Band storage (docs/src/index.md:236-252, BandedMatrix.jl:414-415): entry
 * (i,j), 0-based, of a matrix with bandwidths (kl,ku) lives at
 * ab[(ku + i - j) + j*ldab] when -kl <= j-i <= ku.
 *
 * Rounding recipes (which operations are fused) follow black-box probes of
 * OpenBLAS 0.3.28: gbmv 'N' is y_i = fma(alpha*x_j, a_ij, y_i) with j
 * ascending; gbtf2 scales by a rounded reciprocal and updates with
 * fma(l_i, -u_jc, a_ic); gbtrs 'N' forward is fma(l, -b_j, b) and the back
 * substitution divides.  This file is built with -ffp-contract=off so that
 * only the explicit fma() calls are fused.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef int64_t i64;

#define IDX(ld, r, c) ((r) + (c) * (ld))
static i64 imax(i64 a, i64 b) { return a > b ? a : b; }
static i64 imin(i64 a, i64 b) { return a < b ? a : b; }

/* ------------------------------------------------------------------ gbmv */
/* y <- alpha*op(A)*x + beta*y, A m-by-n with bands (kl,ku) in band storage.
 * beta == 0 overwrites y (NaN-safe), as matmul.jl:26-39 / utils.jl:29-31
 * require.  'N' follows the column-ordered OpenBLAS kernel; 'T' is a
 * per-column dot product accumulated in ascending row order. */
#define GBMV_BODY(T, FMA)                                                        \
    if (m <= 0 || n <= 0) return;                                                \
    i64 leny = (trans == 'N' || trans == 'n') ? m : n;                           \
    i64 lenx = (trans == 'N' || trans == 'n') ? n : m;                           \
    for (i64 i = 0; i < leny; ++i) {                                             \
        T *yi = y + i * incy;                                                    \
        *yi = (beta == (T)0) ? (T)0 : beta * *yi;                                \
    }                                                                            \
    if (trans == 'N' || trans == 'n') {                                          \
        for (i64 j = 0; j < lenx; ++j) {                                         \
            T t = alpha * x[j * incx];                                           \
            i64 lo = imax(0, j - ku), hi = imin(m - 1, j + kl);                  \
            for (i64 i = lo; i <= hi; ++i)                                       \
                y[i * incy] = FMA(t, a[IDX(lda, ku + i - j, j)], y[i * incy]);   \
        }                                                                        \
    } else {                                                                     \
        for (i64 j = 0; j < leny; ++j) {                                         \
            T acc = 0;                                                           \
            i64 lo = imax(0, j - ku), hi = imin(m - 1, j + kl);                  \
            for (i64 i = lo; i <= hi; ++i)                                       \
                acc = FMA(a[IDX(lda, ku + i - j, j)], x[i * incx], acc);         \
            y[j * incy] = FMA(alpha, acc, y[j * incy]);                          \
        }                                                                        \
    }

void bo_dgbmv(char trans, i64 m, i64 n, i64 kl, i64 ku, double alpha,
              const double *a, i64 lda, const double *x, i64 incx, double beta,
              double *y, i64 incy) {
    GBMV_BODY(double, fma)
}
void bo_sgbmv(char trans, i64 m, i64 n, i64 kl, i64 ku, float alpha,
              const float *a, i64 lda, const float *x, i64 incx, float beta,
              float *y, i64 incy) {
    GBMV_BODY(float, fmaf)
}

/* --------------------------------------------------------------- gbmm! */
/* C <- alpha*A*B + beta*C for banded A (n x nu, bands Al,Au >= 0),
 * B (nu x m, Bl,Bu >= 0), C (n x m, Cl,Cu) -- the positive-band core of
 * gbmm! (gbmm.jl:282-340).  Every column j of C is one gbmv over the
 * contraction index nu ascending (gbmm.jl:306-336), so an entry is
 *   c_ij = fma(alpha*b_vj, a_iv, c_ij)   for v ascending,
 * seeded with beta*c_ij (or 0 when beta == 0).  Rows of C's band outside the
 * product's band and columns the product does not reach are beta-scaled
 * (gbmm.jl:286-289,339).  Returns 0, or -1 for BandError (C too narrow and
 * the outer bands of A,B are not exactly zero, gbmm.jl:250-279). */
static i64 zeroband_u(const double *a, i64 lda, i64 rows, i64 cols, i64 l, i64 u) {
    /* _num_zeroband_u (gbmm.jl:191-197): leading all-zero super bands */
    for (i64 b = 0; b <= l + u; ++b) {
        i64 k = u - b; /* band index: j - i = k */
        for (i64 j = imax(0, k); j < cols && j - k < rows; ++j)
            if (a[IDX(lda, u - k, j)] != 0.0) return b;
    }
    return l + u + 1;
}
static i64 zeroband_l(const double *a, i64 lda, i64 rows, i64 cols, i64 l, i64 u) {
    /* _num_zeroband_l (gbmm.jl:199-205): trailing all-zero sub bands */
    for (i64 b = 0; b <= l + u; ++b) {
        i64 k = -l + b;
        for (i64 j = imax(0, k); j < cols && j - k < rows; ++j)
            if (a[IDX(lda, u - k, j)] != 0.0) return b;
    }
    return l + u + 1;
}

static void scale_or_zero(double beta, double *p) { *p = (beta == 0.0) ? 0.0 : beta * *p; }

int bo_dgbmm(i64 n, i64 nu, i64 m, double alpha,
             const double *A, i64 lda, i64 Al, i64 Au,
             const double *B, i64 ldb, i64 Bl, i64 Bu,
             double beta, double *C, i64 ldc, i64 Ctl, i64 Ctu) {
    if (n == 0 || m == 0) return 0;
    /* product bands clipped to the size of C (gbmm.jl:228, broadcast.jl:895) */
    i64 Cl = imin(n - 1, Al + Bl), Cu = imin(m - 1, Au + Bu);
    if (Ctu < Cu) {
        i64 Aur = zeroband_u(A, lda, n, nu, Al, Au), Bur = zeroband_u(B, ldb, nu, m, Bl, Bu);
        if (Ctu < Cu - Aur - Bur) return -1;
        if (Au - Aur < -Al || Bu - Bur < -Bl) {
            for (i64 j = 0; j < m; ++j)
                for (i64 r = 0; r < Ctl + Ctu + 1; ++r) scale_or_zero(beta, &C[IDX(ldc, r, j)]);
            return 0;
        }
        return bo_dgbmm(n, nu, m, alpha, A + Aur, lda, Al, Au - Aur, B + Bur, ldb, Bl, Bu - Bur,
                        beta, C, ldc, Ctl, Ctu);
    }
    if (Ctl < Cl) {
        i64 Alr = zeroband_l(A, lda, n, nu, Al, Au), Blr = zeroband_l(B, ldb, nu, m, Bl, Bu);
        if (Ctl < Cl - Alr - Blr) return -1;
        if (Al - Alr < -Au || Bl - Blr < -Bu) {
            for (i64 j = 0; j < m; ++j)
                for (i64 r = 0; r < Ctl + Ctu + 1; ++r) scale_or_zero(beta, &C[IDX(ldc, r, j)]);
            return 0;
        }
        return bo_dgbmm(n, nu, m, alpha, A, lda, Al - Alr, Au, B, ldb, Bl - Blr, Bu,
                        beta, C, ldc, Ctl, Ctu);
    }
    /* extra bands of C outside (Cl,Cu) are beta-scaled (gbmm.jl:286-288) */
    for (i64 j = 0; j < m; ++j) {
        for (i64 r = 0; r < Ctu - Cu; ++r) scale_or_zero(beta, &C[IDX(ldc, r, j)]);
        for (i64 r = Ctu + Cl + 1; r < Ctl + Ctu + 1; ++r) scale_or_zero(beta, &C[IDX(ldc, r, j)]);
    }
    double *Cd = C + (Ctu - Cu); /* shifted view with bands (Cl,Cu) */
    for (i64 j = 0; j < m; ++j) {
        /* rows of column j that the product (A*B)[:,j] covers */
        if (j >= nu + Bu) { /* untouched columns (gbmm.jl:339) */
            if (j < imin(m, n + Cu))
                for (i64 r = 0; r < Cl + Cu + 1; ++r) scale_or_zero(beta, &Cd[IDX(ldc, r, j)]);
            continue;
        }
        i64 ilo = imax(0, j - Cu), ihi = imin(n - 1, j + Cl);
        for (i64 i = ilo; i <= ihi; ++i) {
            double *c = &Cd[IDX(ldc, Cu + i - j, j)];
            double acc = (beta == 0.0) ? 0.0 : beta * *c;
            i64 vlo = imax(imax(0, j - Bu), i - Al), vhi = imin(imin(nu - 1, j + Bl), i + Au);
            for (i64 v = vlo; v <= vhi; ++v) {
                double t = alpha * B[IDX(ldb, Bu + v - j, j)];
                acc = fma(t, A[IDX(lda, Au + i - v, v)], acc);
            }
            *c = acc;
        }
    }
    return 0;
}

/* ------------------------------------------------------------- gbtf2 */
/* Unblocked partial-pivot band LU, restating LAPACK xGBTF2 (the routine
 * xGBTRF runs for ku <= 64).  ab is (2kl+ku+1) x n, diagonal at row kl+ku.
 * ipiv is 1-based (LAPACK convention, BandedLU.jl:12).  Returns info. */
#define GBTF2_BODY(T, FMA, ABS)                                                  \
    i64 kv = ku + kl, info = 0;                                                  \
    if (m == 0 || n == 0) return 0;                                              \
    for (i64 j = ku + 1; j < imin(kv, n); ++j)                                   \
        for (i64 i = kv - j; i < kl; ++i) ab[IDX(ldab, i, j)] = 0;               \
    i64 ju = 0;                                                                  \
    for (i64 j = 0; j < imin(m, n); ++j) {                                       \
        if (j + kv < n)                                                          \
            for (i64 i = 0; i < kl; ++i) ab[IDX(ldab, i, j + kv)] = 0;           \
        i64 km = imin(kl, m - 1 - j);                                            \
        i64 jp = 0;                                                              \
        T best = ABS(ab[IDX(ldab, kv, j)]);                                      \
        for (i64 t = 1; t <= km; ++t) {                                          \
            T v = ABS(ab[IDX(ldab, kv + t, j)]);                                 \
            if (v > best) { best = v; jp = t; }                                  \
        }                                                                        \
        ipiv[j] = jp + j + 1;                                                    \
        if (ab[IDX(ldab, kv + jp, j)] != 0) {                                    \
            ju = imax(ju, imin(j + ku + jp, n - 1));                             \
            if (jp != 0)                                                         \
                for (i64 c = j; c <= ju; ++c) {                                  \
                    T *p = &ab[IDX(ldab, kv + jp + j - c, c)];                   \
                    T *q = &ab[IDX(ldab, kv + j - c, c)];                        \
                    T s = *p; *p = *q; *q = s;                                   \
                }                                                                \
            if (km > 0) {                                                        \
                T r = (T)1 / ab[IDX(ldab, kv, j)];                               \
                for (i64 t = 1; t <= km; ++t) {                                  \
                    if (variant == 1) ab[IDX(ldab, kv + t, j)] /= ab[IDX(ldab, kv, j)]; \
                    else ab[IDX(ldab, kv + t, j)] *= r;                          \
                }                                                                \
                for (i64 c = j + 1; c <= ju; ++c) {                              \
                    T nu_ = -ab[IDX(ldab, kv + j - c, c)];                       \
                    for (i64 t = 1; t <= km; ++t) {                              \
                        T *q_ = &ab[IDX(ldab, kv + j + t - c, c)];               \
                        if (variant == 2) *q_ = *q_ + ab[IDX(ldab, kv + t, j)] * nu_; \
                        else *q_ = FMA(ab[IDX(ldab, kv + t, j)], nu_, *q_);      \
                    }                                                            \
                }                                                                \
            }                                                                    \
        } else if (info == 0) {                                                  \
            info = j + 1;                                                        \
        }                                                                        \
    }                                                                            \
    return info;

i64 bo_dgbtf2(i64 m, i64 n, i64 kl, i64 ku, double *ab, i64 ldab, i64 *ipiv) {
    const int variant = 0;
    GBTF2_BODY(double, fma, fabs)
}
i64 bo_sgbtf2(i64 m, i64 n, i64 kl, i64 ku, float *ab, i64 ldab, i64 *ipiv) {
    const int variant = 0;
    GBTF2_BODY(float, fmaf, fabsf)
}
/* Planted mistakes for the pin tests only: variant 1 divides by the pivot
 * instead of scaling by its reciprocal, variant 2 does not fuse the update.
 * Each must fail the bit pin against LAPACK. */
i64 bo_dgbtf2_planted(int variant, i64 m, i64 n, i64 kl, i64 ku, double *ab, i64 ldab, i64 *ipiv) {
    GBTF2_BODY(double, fma, fabs)
}

/* ------------------------------------------------------------- tbsv */
/* Upper-triangular band solve, x <- U^{-1} x ('N') or U^{-T} x ('T'),
 * U n x n with k super-diagonals stored with the diagonal at row k of a.
 * 'N' restates the column-oriented BLAS xTBSV: x_j = b_j / u_jj, then
 * b_i = fma(u_ij, -x_j, b_i).  'T' accumulates each dot in ascending order. */
#define TBSV_BODY(T, FMA)                                                        \
    if (trans == 'N' || trans == 'n') {                                          \
        for (i64 j = n - 1; j >= 0; --j) {                                       \
            T xj = x[j * incx] / a[IDX(lda, k, j)];                              \
            x[j * incx] = xj;                                                    \
            T nx = -xj;                                                          \
            for (i64 i = imax(0, j - k); i < j; ++i)                             \
                x[i * incx] = FMA(a[IDX(lda, k + i - j, j)], nx, x[i * incx]);   \
        }                                                                        \
    } else {                                                                     \
        for (i64 j = 0; j < n; ++j) {                                            \
            T acc = x[j * incx];                                                 \
            for (i64 i = imax(0, j - k); i < j; ++i)                             \
                acc = FMA(-a[IDX(lda, k + i - j, j)], x[i * incx], acc);         \
            x[j * incx] = acc / a[IDX(lda, k, j)];                               \
        }                                                                        \
    }

void bo_dtbsv_u(char trans, i64 n, i64 k, const double *a, i64 lda, double *x, i64 incx) {
    TBSV_BODY(double, fma)
}
void bo_stbsv_u(char trans, i64 n, i64 k, const float *a, i64 lda, float *x, i64 incx) {
    TBSV_BODY(float, fmaf)
}

/* General triangular band solve and product, x <- op(A)^{-1} x (xTBSV) and
 * x <- op(A) x (xTBMV), as tbsv!/tbmv! call them (src/blas.jl:71-147) from
 * triangular \ and * on banded data (src/tribanded.jl:47-54,75-96).
 * uplo 'U': A(i,j) at a[k + i - j + j*lda] (diagonal in row k); 'L': A(i,j)
 * at a[i - j + j*lda] (diagonal in row 0).  diag 'U' takes a unit diagonal
 * that is never read.
 * Solve 'N' restates the column-oriented BLAS sweep (upper: j descending,
 * lower: j ascending): x_j = b_j / a_jj, then b_i = fma(a_ij, -x_j, b_i).
 * Solve 'T' accumulates each dot over the off-diagonal entries in order of
 * their distance from the diagonal, then divides.
 * Product: y_i = d_i, d_i = a_ii x_i (x_i for a unit diagonal), then
 * y_i = fma(a, x, y_i) over the off-diagonal terms of row i ('N') or column i
 * ('T') in order of their distance t = 1..k from the diagonal. */
#define TB_AT(t_, j_) (up ? a[IDX(lda, k - (t_), j_)] : a[IDX(lda, (t_), j_)])
#define TBSV_GEN_BODY(T, FMA)                                                          \
    const int up = (uplo == 'U' || uplo == 'u');                                         \
    const int tr = !(trans == 'N' || trans == 'n');                                      \
    const int unit = (diag == 'U' || diag == 'u');                                       \
    const i64 dr = up ? k : 0;                                                           \
    for (i64 q = 0; q < n; ++q) {                                                        \
        /* 'N' upper and 'T' lower run j descending */                                   \
        const i64 j = (up != tr) ? n - 1 - q : q;                                        \
        /* the step touches entries after j (lower 'N', lower 'T') or before j */        \
        const int fut = !up;                                                             \
        const i64 tm = fut ? imin(k, n - 1 - j) : imin(k, j);                            \
        if (!tr) {                                                                       \
            T xj = x[j * incx];                                                          \
            if (!unit) xj = xj / a[IDX(lda, dr, j)];                                     \
            x[j * incx] = xj;                                                            \
            const T nx = -xj;                                                            \
            for (i64 t = 1; t <= tm; ++t) {                                              \
                const i64 i = fut ? j + t : j - t;                                       \
                x[i * incx] = FMA(TB_AT(t, j), nx, x[i * incx]);                         \
            }                                                                            \
        } else {                                                                         \
            T part = 0;                                                                  \
            for (i64 t = 1; t <= tm; ++t) {                                              \
                const i64 i = fut ? j + t : j - t;                                       \
                part = FMA(TB_AT(t, j), x[i * incx], part);                              \
            }                                                                            \
            const T v = x[j * incx] - part;                                              \
            x[j * incx] = unit ? v : v / a[IDX(lda, dr, j)];                             \
        }                                                                                \
    }

#define TBMV_GEN_BODY(T, FMA)                                                          \
    const int up = (uplo == 'U' || uplo == 'u');                                         \
    const int tr = !(trans == 'N' || trans == 'n');                                      \
    const int unit = (diag == 'U' || diag == 'u');                                       \
    const i64 dr = up ? k : 0;                                                           \
    T *y = (T *)malloc(sizeof(T) * (size_t)(n > 0 ? n : 1));                             \
    for (i64 i = 0; i < n; ++i) {                                                        \
        T acc = unit ? x[i * incx] : a[IDX(lda, dr, i)] * x[i * incx];                   \
        if (!tr && up) { /* row i of U: A(i, i+t) */                                     \
            for (i64 t = 1; t <= imin(k, n - 1 - i); ++t)                                \
                acc = FMA(a[IDX(lda, k - t, i + t)], x[(i + t) * incx], acc);            \
        } else if (!tr) { /* row i of L: A(i, i-t) */                                    \
            for (i64 t = 1; t <= imin(k, i); ++t)                                        \
                acc = FMA(a[IDX(lda, t, i - t)], x[(i - t) * incx], acc);                \
        } else if (up) { /* column i of U: A(i-t, i) */                                  \
            for (i64 t = 1; t <= imin(k, i); ++t)                                        \
                acc = FMA(a[IDX(lda, k - t, i)], x[(i - t) * incx], acc);                \
        } else { /* column i of L: A(i+t, i) */                                          \
            for (i64 t = 1; t <= imin(k, n - 1 - i); ++t)                                \
                acc = FMA(a[IDX(lda, t, i)], x[(i + t) * incx], acc);                    \
        }                                                                                \
        y[i] = acc;                                                                      \
    }                                                                                    \
    for (i64 i = 0; i < n; ++i) x[i * incx] = y[i];                                      \
    free(y);

void bo_dtbsv(char uplo, char trans, char diag, i64 n, i64 k, const double *a, i64 lda, double *x, i64 incx) {
    TBSV_GEN_BODY(double, fma)
}
void bo_stbsv(char uplo, char trans, char diag, i64 n, i64 k, const float *a, i64 lda, float *x, i64 incx) {
    TBSV_GEN_BODY(float, fmaf)
}
void bo_dtbmv(char uplo, char trans, char diag, i64 n, i64 k, const double *a, i64 lda, double *x, i64 incx) {
    TBMV_GEN_BODY(double, fma)
}
void bo_stbmv(char uplo, char trans, char diag, i64 n, i64 k, const float *a, i64 lda, float *x, i64 incx) {
    TBMV_GEN_BODY(float, fmaf)
}

/* ------------------------------------------------------- symmetric */
/* Band Cholesky A = U'U ('U') or L L' ('L') in LAPACK's symmetric band storage
 * (uplo 'U': A(i,j), i <= j, at ab[kd + i - j + j*ldab]; 'L': A(i,j), i >= j,
 * at ab[i - j + j*ldab]), restating xPBTF2, which xPBTRF runs for kd < 32
 * (src/symbanded/BandedCholesky.jl:1-11, src/lapack.jl:268-292): for each
 * column j, a_jj <- sqrt(a_jj) (info = j+1 when a_jj <= 0; NaN goes on), the kn
 * off-diagonal entries of its row (U) or column (L) times the rounded
 * reciprocal 1/a_jj, then the trailing kn x kn triangle updated as xSYR does
 * with alpha = -1: column q skipped when v_q == 0, else
 * a_pq = fma(-v_q, v_p, a_pq). */
#define PBTF2_BODY(T, SQRT, FMA)                                                        \
    const int up = (uplo == 'U' || uplo == 'u');                                        \
    for (i64 j = 0; j < n; ++j) {                                                       \
        T *dj = &ab[IDX(ldab, up ? kd : 0, j)];                                         \
        T ajj = *dj;                                                                    \
        if (ajj <= 0) return j + 1; /* NaN passes, as in LAPACK */                      \
        ajj = SQRT(ajj);                                                                \
        *dj = ajj;                                                                      \
        const i64 kn = imin(kd, n - 1 - j);                                             \
        if (kn == 0) continue;                                                          \
        const T r = (T)1 / ajj;                                                         \
        /* v_c: U(j, j+c) at row kd-c of column j+c, or L(j+c, j) at row c of column j */ \
        for (i64 c = 1; c <= kn; ++c) {                                                 \
            T *v = up ? &ab[IDX(ldab, kd - c, j + c)] : &ab[IDX(ldab, c, j)];           \
            *v = *v * r;                                                                \
        }                                                                               \
        for (i64 q = 1; q <= kn; ++q) {                                                 \
            const T vq = up ? ab[IDX(ldab, kd - q, j + q)] : ab[IDX(ldab, q, j)];       \
            if (vq == 0) continue;                                                      \
            const T tq = -vq;                                                           \
            if (up) { /* column j+q, rows j+1 .. j+q: A(j+p, j+q) at row kd + p - q */  \
                for (i64 p = 1; p <= q; ++p)                                            \
                    ab[IDX(ldab, kd + p - q, j + q)] =                                  \
                        FMA(tq, ab[IDX(ldab, kd - p, j + p)], ab[IDX(ldab, kd + p - q, j + q)]); \
            } else { /* column j+q, rows j+q .. j+kn: A(j+p, j+q) at row p - q */       \
                for (i64 p = q; p <= kn; ++p)                                           \
                    ab[IDX(ldab, p - q, j + q)] =                                       \
                        FMA(tq, ab[IDX(ldab, p, j)], ab[IDX(ldab, p - q, j + q)]);      \
            }                                                                           \
        }                                                                               \
    }                                                                                   \
    return 0;

i64 bo_dpbtf2(char uplo, i64 n, i64 kd, double *ab, i64 ldab) { PBTF2_BODY(double, sqrt, fma) }
i64 bo_spbtf2(char uplo, i64 n, i64 kd, float *ab, i64 ldab) { PBTF2_BODY(float, sqrtf, fmaf) }

/* Banded LDL' without pivoting, restating ldlt! (src/symbanded/ldlt.jl:8-37)
 * on the stored triangle (storage as for pbtf2), left-looking in j with each
 * sum taken in ascending k (the reference's @simd loops leave the order to
 * the compiler, so results are defined to rounding):
 *   d_j = a_jj - sum_k l_jk^2 d_k,  l_ij = (a_ij - sum_k l_ik l_jk d_k) / d_j.
 * The solve restates ldiv!(::LDLt, B) (ldlt.jl:47-69): forward with unit L,
 * divide by d, backward with L'. */
#define LDL_AT(r, c) (up ? ab[IDX(ldab, kd + (r) - (c), (c))] : ab[IDX(ldab, (c) - (r), (r))]) /* A(max, min) */
#define SBLDLT_BODY(T)                                                                  \
    const int up = (uplo == 'U' || uplo == 'u');                                        \
    /* element (i, j), i >= j, of the lower factor: L(i,j) lives where A(i,j) does */     \
    for (i64 j = 0; j < n; ++j) {                                                       \
        T *djj = up ? &ab[IDX(ldab, kd, j)] : &ab[IDX(ldab, 0, j)];                     \
        for (i64 k = imax(0, j - kd); k < j; ++k) {                                     \
            const T ljk = up ? ab[IDX(ldab, kd + k - j, j)] : ab[IDX(ldab, j - k, k)];  \
            const T dk = up ? ab[IDX(ldab, kd, k)] : ab[IDX(ldab, 0, k)];               \
            *djj -= ljk * ljk * dk;                                                     \
        }                                                                               \
        for (i64 i = j + 1; i <= imin(n - 1, j + kd); ++i) {                            \
            T *aij = up ? &ab[IDX(ldab, kd + j - i, i)] : &ab[IDX(ldab, i - j, j)];     \
            for (i64 k = imax(0, i - kd); k < j; ++k) {                                 \
                const T lik = up ? ab[IDX(ldab, kd + k - i, i)] : ab[IDX(ldab, i - k, k)]; \
                const T ljk = up ? ab[IDX(ldab, kd + k - j, j)] : ab[IDX(ldab, j - k, k)]; \
                const T dk = up ? ab[IDX(ldab, kd, k)] : ab[IDX(ldab, 0, k)];           \
                *aij -= lik * ljk * dk;                                                 \
            }                                                                           \
            *aij /= *djj;                                                               \
        }                                                                               \
    }

void bo_dsbldlt(char uplo, i64 n, i64 kd, double *ab, i64 ldab) { SBLDLT_BODY(double) }
void bo_ssbldlt(char uplo, i64 n, i64 kd, float *ab, i64 ldab) { SBLDLT_BODY(float) }

#define SBLDLTS_BODY(T)                                                                 \
    const int up = (uplo == 'U' || uplo == 'u');                                        \
    for (i64 j = 0; j < n; ++j)                                                         \
        for (i64 k = imax(0, j - kd); k < j; ++k)                                       \
            b[j] -= (up ? ab[IDX(ldab, kd + k - j, j)] : ab[IDX(ldab, j - k, k)]) * b[k]; \
    for (i64 j = 0; j < n; ++j) b[j] /= (up ? ab[IDX(ldab, kd, j)] : ab[IDX(ldab, 0, j)]); \
    for (i64 j = n - 1; j >= 0; --j)                                                    \
        for (i64 k = j + 1; k <= imin(n - 1, j + kd); ++k)                              \
            b[j] -= (up ? ab[IDX(ldab, kd + j - k, k)] : ab[IDX(ldab, k - j, j)]) * b[k];

void bo_dsbldlts(char uplo, i64 n, i64 kd, const double *ab, i64 ldab, double *b) { SBLDLTS_BODY(double) }

/* Symmetric band product y <- alpha*A*x + beta*y (BLAS xSBMV, as
 * src/symbanded/symbanded.jl:72-91 calls it), A stored as one triangle like
 * the Cholesky input.  Row i: t = sum over j = i-k .. i+k ascending of
 * fma(A(i,j), x_j, t) from t = 0; y_i = alpha*t (beta == 0) or
 * alpha*t + beta*y_i. */
#define SBMV_BODY(T, FMA)                                                               \
    const int up = (uplo == 'U' || uplo == 'u');                                        \
    for (i64 i = 0; i < n; ++i) {                                                       \
        T t = 0;                                                                        \
        for (i64 j = imax(0, i - k); j <= imin(n - 1, i + k); ++j) {                    \
            const i64 r = imin(i, j), c = imax(i, j); /* stored entry (r, c), r <= c */ \
            const T a = up ? ab[IDX(ldab, k + r - c, c)] : ab[IDX(ldab, c - r, r)];     \
            t = FMA(a, x[j], t);                                                        \
        }                                                                               \
        y[i] = (beta == 0) ? alpha * t : alpha * t + beta * y[i];                       \
    }

void bo_dsbmv(char uplo, i64 n, i64 k, double alpha, const double *ab, i64 ldab, const double *x, double beta,
              double *y) {
    SBMV_BODY(double, fma)
}
void bo_ssbmv(char uplo, i64 n, i64 k, float alpha, const float *ab, i64 ldab, const float *x, float beta,
              float *y) {
    SBMV_BODY(float, fmaf)
}

/* ------------------------------------------------------------- gbtrs */
/* Solve with the factors of gbtf2/gbtrf, restating LAPACK xGBTRS.
 * 'N': forward sweep with the stored interchanges and multipliers, then the
 * upper band solve with kl+ku super-diagonals.  'T': transposed upper solve,
 * then the transposed unit-lower sweep in reverse with the interchanges. */
#define GBTRS_BODY(T, FMA, TBSV)                                                 \
    i64 kd = ku + kl;                                                            \
    if (n == 0 || nrhs == 0) return;                                             \
    if (trans == 'N' || trans == 'n') {                                          \
        if (kl > 0)                                                              \
            for (i64 j = 0; j < n - 1; ++j) {                                    \
                i64 lm = imin(kl, n - 1 - j), p = ipiv[j] - 1;                   \
                for (i64 r = 0; r < nrhs; ++r) {                                 \
                    T *bc = b + r * ldb;                                         \
                    if (p != j) { T s = bc[p]; bc[p] = bc[j]; bc[j] = s; }       \
                    T nb = -bc[j];                                               \
                    for (i64 t = 1; t <= lm; ++t)                                \
                        bc[j + t] = FMA(ab[IDX(ldab, kd + t, j)], nb, bc[j + t]);\
                }                                                                \
            }                                                                    \
        for (i64 r = 0; r < nrhs; ++r) TBSV('N', n, kd, ab, ldab, b + r * ldb, 1);\
    } else {                                                                     \
        for (i64 r = 0; r < nrhs; ++r) TBSV('T', n, kd, ab, ldab, b + r * ldb, 1);\
        if (kl > 0)                                                              \
            for (i64 j = n - 2; j >= 0; --j) {                                   \
                i64 lm = imin(kl, n - 1 - j), p = ipiv[j] - 1;                   \
                for (i64 r = 0; r < nrhs; ++r) {                                 \
                    T *bc = b + r * ldb;                                         \
                    T acc = 0;                                                   \
                    for (i64 t = 1; t <= lm; ++t)                                \
                        acc = FMA(bc[j + t], ab[IDX(ldab, kd + t, j)], acc);     \
                    bc[j] = bc[j] - acc;                                         \
                    if (p != j) { T s = bc[p]; bc[p] = bc[j]; bc[j] = s; }       \
                }                                                                \
            }                                                                    \
    }

void bo_dgbtrs(char trans, i64 n, i64 kl, i64 ku, i64 nrhs, const double *ab, i64 ldab,
               const i64 *ipiv, double *b, i64 ldb) {
    GBTRS_BODY(double, fma, bo_dtbsv_u)
}
void bo_sgbtrs(char trans, i64 n, i64 kl, i64 ku, i64 nrhs, const float *ab, i64 ldab,
               const i64 *ipiv, float *b, i64 ldb) {
    GBTRS_BODY(float, fmaf, bo_stbsv_u)
}

/* -------------------------------------------------------- banded QR */
/* reflector! (ArrayLayouts, imported at src/BandedMatrices.jl:28-35):
 * x <- [beta; v(2:end)], returns tau with H = I - tau*[1;v][1;v]'.
 * beta = -copysign(||x||, x1); tau = (x1 + copysign(||x||,x1)) / copysign(||x||,x1).
 * ||x|| == 0 returns tau = 0 and leaves x.  When x(2:end) == 0 and x1 != 0
 * this gives tau = 2 (LAPACK's dlarfg would return 0 there). */
static double reflector(double *x, i64 len) {
    if (len == 0) return 0.0;
    double xi1 = x[0];
    double nrm = xi1 * xi1;
    for (i64 i = 1; i < len; ++i) nrm += x[i] * x[i];
    if (nrm == 0.0) return 0.0;
    nrm = sqrt(nrm);
    double nu = copysign(nrm, xi1);
    xi1 += nu;
    x[0] = -nu;
    for (i64 i = 1; i < len; ++i) x[i] /= xi1;
    return xi1 / nu;
}
/* reflectorApply!: y <- (I - tau [1;v][1;v]') y for one column slice y. */
static void reflector_apply(const double *x, i64 len, double tau, double *y) {
    if (len == 0) return;
    double v = y[0];
    for (i64 i = 1; i < len; ++i) v += x[i] * y[i];
    v = tau * v;
    y[0] -= v;
    for (i64 i = 1; i < len; ++i) y[i] -= x[i] * v;
}

/* _banded_qr! (bandedqr.jl:8-28) on the padded copy: R is m x n with bands
 * (l, u) where u already includes the l fill-in super-diagonals
 * (bandedqr.jl:2).  tau has min(m,n) entries; entries past ncols stay 0. */
void bo_dgbqrf(i64 m, i64 n, i64 l, i64 u, double *D, i64 ldd, double *tau) {
    i64 nu = l + u + 1;
    i64 ncols = imin(m - 1, n);
    for (i64 k = 0; k < imin(m, n); ++k) tau[k] = 0.0;
    for (i64 k = 0; k < ncols; ++k) {
        i64 len = imin(nu, m - k + u) - u; /* rows u..min(nu,m-k+u)-1 */
        double *x = &D[IDX(ldd, u, k)];
        double tk = reflector(x, len);
        tau[k] = tk;
        for (i64 j = 1; j <= imin(u, n - 1 - k); ++j) {
            double *y = &D[IDX(ldd, u - j, k + j)];
            reflector_apply(x, len, tk, y);
        }
    }
}

/* banded_qr_lmul! for Q' (bandedqr.jl:66-94): B <- Q' B, B is m x nrhs dense. */
void bo_dgbqr_qtmul(i64 m, i64 n, i64 l, i64 u, const double *D, i64 ldd, const double *tau,
                    i64 ntau, i64 nrhs, double *B, i64 ldb) {
    for (i64 j = 0; j < nrhs; ++j) {
        double *bj = B + j * ldb;
        i64 kmax = imin(imin(m, n), ntau);
        for (i64 k = 0; k < kmax; ++k) {
            double v = bj[k];
            i64 ilo = k + imax(1, -u), ihi = imin(k + l, m - 1);
            for (i64 i = ilo; i <= ihi; ++i) v += D[IDX(ldd, i - k + u, k)] * bj[i];
            v = tau[k] * v;
            bj[k] -= v;
            for (i64 i = ilo; i <= ihi; ++i) bj[i] -= D[IDX(ldd, i - k + u, k)] * v;
        }
    }
}

/* banded_qr_lmul! for Q (bandedqr.jl:38-64): B <- Q B. */
void bo_dgbqr_qmul(i64 m, i64 n, i64 l, i64 u, const double *D, i64 ldd, const double *tau,
                   i64 ntau, i64 nrhs, double *B, i64 ldb) {
    i64 kmax = imin(imin(m, n), ntau);
    for (i64 k = kmax - 1; k >= 0; --k) {
        for (i64 j = 0; j < nrhs; ++j) {
            double *bj = B + j * ldb;
            double v = bj[k];
            i64 ilo = k + imax(1, -u), ihi = imin(k + l, m - 1);
            for (i64 i = ilo; i <= ihi; ++i) v += D[IDX(ldd, i - k + u, k)] * bj[i];
            v = tau[k] * v;
            bj[k] -= v;
            for (i64 i = ilo; i <= ihi; ++i) bj[i] -= D[IDX(ldd, i - k + u, k)] * v;
        }
    }
}

/* banded_qr_rmul! (bandedqr.jl:96-148): A <- A Q (trans 'N', reflectors in
 * ascending order) or A Q' ('T', descending); A is mA x mQ dense, Q comes from
 * the mQ x nQ factors D.  Per row: v = A(i,k) + sum_j A(i,j) v_j, v *= tau_k,
 * A(i,k) -= v, A(i,j) -= v v_j for j = k+1 .. min(k+l, mQ-1). */
void bo_dgbqr_rmul(char trans, i64 mA, i64 mQ, i64 nQ, i64 l, i64 u, const double *D, i64 ldd,
                   const double *tau, double *A, i64 lda) {
    i64 kmax = imin(mQ, nQ);
    for (i64 t = 0; t < kmax; ++t) {
        i64 k = (trans == 'N' || trans == 'n') ? t : kmax - 1 - t;
        i64 jhi = imin(k + l, mQ - 1);
        for (i64 i = 0; i < mA; ++i) {
            double v = A[IDX(lda, i, k)];
            for (i64 j = k + 1; j <= jhi; ++j) v += A[IDX(lda, i, j)] * D[IDX(ldd, j - k + u, k)];
            v = v * tau[k];
            A[IDX(lda, i, k)] -= v;
            for (i64 j = k + 1; j <= jhi; ++j) A[IDX(lda, i, j)] -= v * D[IDX(ldd, j - k + u, k)];
        }
    }
}

/* ldiv!(QR, B) for m >= n (bandedqr.jl:178-191): B <- Q'B, then
 * R[1:n,1:n] \ B[1:n,:] with the tbsv upper solve of bandwidth u. */
void bo_dgbqrs(i64 m, i64 n, i64 l, i64 u, const double *D, i64 ldd, const double *tau,
               i64 nrhs, double *B, i64 ldb) {
    bo_dgbqr_qtmul(m, n, l, u, D, ldd, tau, imin(m, n), nrhs, B, ldb);
    for (i64 r = 0; r < nrhs; ++r) bo_dtbsv_u('N', n, u, D, ldd, B + r * ldb, 1);
}

/* ------------------------------------------------ adjoint of A\b + 1/2|Ax0|^2 */
/* Gradient of L = w' A^{-1} b + 0.5*||A x0||^2 w.r.t. the band entries of A
 * and b, given x = A\b, lambda = A^{-T} w and y0 = A x0:
 *   dL/db = lambda,   dL/dA_ij = -lambda_i x_j + y0_i x0_j   (in band). */
void bo_dgb_band_grad(i64 n, i64 kl, i64 ku, const double *lam, const double *x,
                      const double *y0, const double *x0, double *G, i64 ldg) {
    for (i64 j = 0; j < n; ++j)
        for (i64 i = imax(0, j - ku); i <= imin(n - 1, j + kl); ++i)
            G[IDX(ldg, ku + i - j, j)] = fma(y0[i], x0[j], -lam[i] * x[j]);
}
