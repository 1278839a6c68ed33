/*
 * bandedhip.h -- C ABI of libbandedhip.so, the MI355X (gfx950) banded hot path.
 *
 * Each entry point replaces one native call (or one Julia loop of native
 * calls) that BandedMatrices.jl makes on its hot path; the citation next to
 * each names the reference interface it stands in for.  Conventions follow
 * the Fortran BLAS/LAPACK interfaces that the reference binds with ccall
 * (src/blas.jl:16-28, stdlib LAPACK.gbtrf!/gbtrs!):
 *   - character flags 'N' / 'T' (real types: 'C' is treated as 'T'),
 *   - 64-bit integer sizes and leading dimensions (BlasInt = Int64),
 *   - column-major band storage: A[i,j] (0-based) at A[(ku+i-j) + j*lda],
 *   - every pointer named d* / A / B / C / x / y / ipiv is DEVICE memory
 *     owned by the caller; host pointers are named *_host,
 *   - ipiv is 1-based (LAPACK convention) and int64.
 * Return value: 0 on success; -i when argument i is invalid (LAPACK xerbla
 * numbering, 1-based); BH_ERR_* (<= -1000) for runtime failures.  Singular
 * factorisations are not errors: they report info > 0 like LAPACK.
 * Work is stream-ordered on the library stream (bh_set_stream); calls that
 * hand a value back to the host (info, iteration counts) synchronise it.
 */
#ifndef BANDEDHIP_H
#define BANDEDHIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int64_t bh_int;

#define BH_OK 0
#define BH_ERR_HIP (-1000)        /* a HIP runtime call failed (see bh_last_error) */
#define BH_ERR_ALLOC (-1001)      /* device allocation failed */
#define BH_ERR_NODEVICE (-1002)   /* no usable gfx950 device */
#define BH_ERR_BANDERROR (-1003)  /* C lacks bands for A*B (BandError, Band.jl:75-88) */
#define BH_ERR_WORKSPACE (-1004)  /* workspace too small */
#define BH_ERR_INTERNAL (-1005)   /* a kernel's bounded internal wait gave up (results invalid) */
#define BH_ERR_BUSY (-1006)       /* a batch plan still has captured graphs (schedule fixed, plan kept) */

/* ---------------------------------------------------------------- runtime */
/* No reference counterpart: device/memory plumbing the Julia extension uses to
 * place BandedMatrix data in HBM (the reference's examples do the same with a
 * GPU array container, examples/clarrays.jl:15). */
int bh_init(int device);
int bh_device_count(int *count);
const char *bh_last_error(void);
int bh_malloc(void **dptr, size_t bytes);
int bh_free(void *dptr);
int bh_memcpy_htod(void *dst, const void *src_host, size_t bytes);
int bh_memcpy_dtoh(void *dst_host, const void *src, size_t bytes);
int bh_memcpy_dtod(void *dst, const void *src, size_t bytes);
int bh_memset(void *dptr, int value, size_t bytes);
int bh_sync(void);
int bh_set_stream(void *hip_stream); /* NULL selects the library's own stream */
void *bh_get_stream(void);
int bh_event_create(void **ev);
int bh_event_record(void *ev);        /* on the library stream */
int bh_event_elapsed_ms(void *start, void *stop, float *ms); /* synchronises stop */
int bh_event_destroy(void *ev);

/* ------------------------------------------------------------------ gbmv */
/* y <- alpha*op(A)*x + beta*y, A m x n with bands (kl,ku).  beta == 0
 * overwrites y (NaN-safe, src/generic/utils.jl:29-31).  'N' reproduces the
 * rounding of the column-ordered BLAS kernel bit for bit.
 * Replaces BLAS.gbmv! via banded_gbmv! (src/generic/matmul.jl:21-23) and the
 * pointer wrapper gbmv! (src/blas.jl:16-28). */
int bh_dgbmv(char trans, bh_int m, bh_int n, bh_int kl, bh_int ku, double alpha,
             const double *A, bh_int lda, const double *x, bh_int incx, double beta,
             double *y, bh_int incy);
int bh_sgbmv(char trans, bh_int m, bh_int n, bh_int kl, bh_int ku, float alpha,
             const float *A, bh_int lda, const float *x, bh_int incx, float beta,
             float *y, bh_int incy);

/* Y <- alpha*op(A)*X + beta*Y for a dense column-major X with nrhs columns,
 * A read from HBM once.  Each column equals a bh_?gbmv call bit for bit.
 * Replaces the per-column mul! loop of
 * materialize!(::MatMulMatAdd{<:BandedColumns,...}) (src/generic/matmul.jl:243-256). */
int bh_dgbmv_multi(char trans, bh_int m, bh_int n, bh_int kl, bh_int ku, bh_int nrhs,
                   double alpha, const double *A, bh_int lda, const double *X, bh_int ldx,
                   double beta, double *Y, bh_int ldy);
int bh_sgbmv_multi(char trans, bh_int m, bh_int n, bh_int kl, bh_int ku, bh_int nrhs,
                   float alpha, const float *A, bh_int lda, const float *X, bh_int ldx,
                   float beta, float *Y, bh_int ldy);

/* -------------------------------------------------------------- gbmm */
/* C <- alpha*A*B + beta*C, all three banded (A n x k (Al,Au), B k x m (Bl,Bu),
 * C n x m (Cl,Cu)), non-negative bandwidths.  Each entry follows the rounding
 * of gbmm!'s per-column gbmv bit for bit.  When C is narrower than the
 * product, all-zero outer bands of A and B are stripped first, else
 * BH_ERR_BANDERROR.  Replaces gbmm! (src/banded/gbmm.jl:207-340). */
int bh_dgbmm(bh_int n, bh_int k, bh_int m, double alpha,
             const double *A, bh_int lda, bh_int Al, bh_int Au,
             const double *B, bh_int ldb, bh_int Bl, bh_int Bu,
             double beta, double *C, bh_int ldc, bh_int Cl, bh_int Cu);
int bh_sgbmm(bh_int n, bh_int k, bh_int m, float alpha,
             const float *A, bh_int lda, bh_int Al, bh_int Au,
             const float *B, bh_int ldb, bh_int Bl, bh_int Bu,
             float beta, float *C, bh_int ldc, bh_int Cl, bh_int Cu);

/* ------------------------------------------------------------- padding */
/* AB <- the padded copy of A (m x n, bands (kl,ku)) with kl extra zero
 * super-diagonals on top: LAPACK's AB layout, ldab >= 2kl+ku+1.  Entries
 * outside A's band are written as zero.  bh_dsgbpad converts FP64 -> FP32.
 * Replaces BandedMatrix{T}(A, (l, l+u)) (src/banded/BandedMatrix.jl:222-232)
 * as used by _lu (src/banded/BandedLU.jl:108-111) and _banded_qr
 * (src/banded/bandedqr.jl:2). */
int bh_dgbpad(bh_int m, bh_int n, bh_int kl, bh_int ku, const double *A, bh_int lda, double *AB,
              bh_int ldab);
int bh_sgbpad(bh_int m, bh_int n, bh_int kl, bh_int ku, const float *A, bh_int lda, float *AB,
              bh_int ldab);
int bh_dsgbpad(bh_int m, bh_int n, bh_int kl, bh_int ku, const double *A, bh_int lda, float *AB,
               bh_int ldab);

/* ------------------------------------------------------------- gbtrf */
/* Partial-pivot band LU in place on the padded array AB (ldab >= 2kl+ku+1,
 * diagonal at row kl+ku), pivots into ipiv (device, 1-based, min(m,n)
 * entries), LAPACK info into *info_host (0 ok, j>0 zero pivot in column j).
 * Pivots equal LAPACK dgbtf2's.  Factors equal dgbtf2's bit for bit except
 * where an MFMA-blocked factorisation runs: ku >= 65 with kl >= 32 (where
 * the reference's LAPACK.gbtrf! itself runs blocked dgbtrf) and
 * min(m, n) >= 16384 with kl, ku >= 8; there the factors agree to rounding
 * (the updates of a 16-column panel are summed at once).  Bands with kl >= 16
 * below those bounds use a blocked factorisation that keeps dgbtf2's
 * per-entry operation order (on inputs containing -0.0 or non-finite values
 * it may differ in zero signs and in how far NaN spreads).
 * Replaces LAPACK.gbtrf! in lu! (src/banded/BandedLU.jl:90-103). */
int bh_dgbtrf(bh_int m, bh_int n, bh_int kl, bh_int ku, double *AB, bh_int ldab,
              bh_int *ipiv, bh_int *info_host);
int bh_sgbtrf(bh_int m, bh_int n, bh_int kl, bh_int ku, float *AB, bh_int ldab,
              bh_int *ipiv, bh_int *info_host);

/* ------------------------------------------------------------- gbtrs */
/* Solve op(A) X = B with the factors of bh_?gbtrf; B is n x nrhs, overwritten.
 * 'N' equals LAPACK dgbtrs bit for bit.
 * Replaces LAPACK.gbtrs! in ldiv!(::BandedLU, B) (src/banded/linalg.jl:24-30)
 * and its transposed form (src/banded/linalg.jl:41-47,55-56). */
int bh_dgbtrs(char trans, bh_int n, bh_int kl, bh_int ku, bh_int nrhs, const double *AB,
              bh_int ldab, const bh_int *ipiv, double *B, bh_int ldb);
int bh_sgbtrs(char trans, bh_int n, bh_int kl, bh_int ku, bh_int nrhs, const float *AB,
              bh_int ldab, const bh_int *ipiv, float *B, bh_int ldb);

/* Blocked solve A X = B ('N') with the factors of bh_?gbtrf: panels of 32
 * rows on the matrix cores, 16 right-hand sides per workgroup.  Same result
 * as bh_?gbtrs to rounding (the diagonal blocks through their inverses, a
 * panel's updates summed at once), not bit for bit; for many right-hand
 * sides with wide bands.  work: device workspace of
 * bh_gbtrs_blocked_workspace(prec 'd'/'s', ...) bytes (0: bands outside the
 * supported range -- FP32 1 <= kl <= 384, kl+ku <= 512; FP64 1 <= kl <= 128,
 * kl+ku <= 256 -- then BH_ERR_WORKSPACE).
 * Same job as LAPACK.gbtrs! in ldiv!(::BandedLU, B) (src/banded/linalg.jl:24-30). */
int bh_gbtrs_blocked_workspace(char prec, bh_int n, bh_int kl, bh_int ku, size_t *bytes);
int bh_dgbtrs_blocked(bh_int n, bh_int kl, bh_int ku, bh_int nrhs, const double *AB, bh_int ldab,
                      const bh_int *ipiv, double *B, bh_int ldb, void *work, size_t lwork);
int bh_sgbtrs_blocked(bh_int n, bh_int kl, bh_int ku, bh_int nrhs, const float *AB, bh_int ldab,
                      const bh_int *ipiv, float *B, bh_int ldb, void *work, size_t lwork);

/* Partitioned solve A X = B ('N') or A^T X = B ('T'/'C') with the factors of
 * bh_?gbtrf: the rows split into blocks of about sqrt(n) solved at once
 * across the GPU, joined through the linear map each block applies to the
 * state it takes over (a window of KW > bandwidth values).  Same result as
 * bh_?gbtrs to rounding, not bit for bit (bh_?gbtrs keeps the bit-exact
 * sequential chain); for long systems with narrow bands: kl <= 31 and
 * kl + ku <= 63 (bh_gbtrs_part_workspace gives 0 bytes outside that range,
 * and the solve returns BH_ERR_WORKSPACE).  work: device workspace of
 * bh_gbtrs_part_workspace(prec 'd'/'s', n, kl, ku, nrhs) bytes.  The
 * partitioned result's rounding error grows with the blocks' transfer
 * matrices; when an entry of one exceeds 1e8 (FP64) / 1e3 (FP32) in
 * magnitude the call runs bh_?gbtrs instead (B is written once, by whichever
 * ran).  One host synchronisation per call (that check).
 * Same job as LAPACK.gbtrs! in ldiv!(::BandedLU, B) and
 * ldiv!(transpose(::BandedLU), B) (src/banded/linalg.jl:24-47). */
int bh_gbtrs_part_workspace(char prec, bh_int n, bh_int kl, bh_int ku, bh_int nrhs, size_t *bytes);
int bh_dgbtrs_part(char trans, bh_int n, bh_int kl, bh_int ku, bh_int nrhs, const double *AB, bh_int ldab,
                   const bh_int *ipiv, double *B, bh_int ldb, void *work, size_t lwork);
int bh_sgbtrs_part(char trans, bh_int n, bh_int kl, bh_int ku, bh_int nrhs, const float *AB, bh_int ldab,
                   const bh_int *ipiv, float *B, bh_int ldb, void *work, size_t lwork);

/* ------------------------------------------------------------- banded QR */
/* Householder QR in band storage: D is the padded copy (bh_?gbpad with
 * kl = l, ku = u_orig) viewed as bands (l, u) with u = l + u_orig, ldd >= l+u+1,
 * diagonal at row u.  On return R is in the upper u bands and the reflectors'
 * v (v1 = 1 implicit) in the lower l band; tau (device, min(m,n) entries) is
 * zero past the last reflector (min(m-1, n) of them).  Reflector convention:
 * ArrayLayouts.reflector! (tau = 2 when x(2:end) = 0).
 * Replaces _banded_qr! via qr/qr! (src/banded/bandedqr.jl:1-35). */
int bh_dgbqrf(bh_int m, bh_int n, bh_int l, bh_int u, double *D, bh_int ldd, double *tau);
int bh_sgbqrf(bh_int m, bh_int n, bh_int l, bh_int u, float *D, bh_int ldd, float *tau);
/* B <- Q' B (trans 'T' or 'C') or Q B ('N'); B is m x nrhs (ldb >= m).
 * Replaces banded_qr_lmul! (src/banded/bandedqr.jl:38-94) behind lmul!(Q', B)
 * and lmul!(Q, B) (bandedqr.jl:150-172). */
int bh_dgbqrm(char trans, bh_int m, bh_int n, bh_int l, bh_int u, const double *D, bh_int ldd,
              const double *tau, bh_int nrhs, double *B, bh_int ldb);
int bh_sgbqrm(char trans, bh_int m, bh_int n, bh_int l, bh_int u, const float *D, bh_int ldd,
              const float *tau, bh_int nrhs, float *B, bh_int ldb);
/* A <- A Q (trans 'N') or A Q' ('T' / 'C'); A is mA x mQ (lda >= mA), Q the
 * orthogonal factor of the mQ x nQ band QR (D, tau as bh_?gbqrf returns them).
 * Replaces banded_qr_rmul! (src/banded/bandedqr.jl:96-148) behind rmul!(A, Q)
 * and rmul!(A, Q') (bandedqr.jl:210-214). */
int bh_dgbqrmr(char trans, bh_int mA, bh_int mQ, bh_int nQ, bh_int l, bh_int u, const double *D,
               bh_int ldd, const double *tau, double *A, bh_int lda);
int bh_sgbqrmr(char trans, bh_int mA, bh_int mQ, bh_int nQ, bh_int l, bh_int u, const float *D,
               bh_int ldd, const float *tau, float *A, bh_int lda);
/* Least-squares / square solve with the factors: B (m x nrhs, m >= n) <- Q'B,
 * then the first n rows <- R \ (Q'B)[1:n] (upper band solve, u super-diagonals).
 * m < n is an argument error (-1), as the reference raises for wide QR.
 * Replaces ldiv!(::QR{<:Any,<:AbstractBandedMatrix}, B) (src/banded/bandedqr.jl:175-206)
 * and its tbsv back substitution (src/banded/tribanded.jl:75-83). */
int bh_dgbqrs(bh_int m, bh_int n, bh_int l, bh_int u, const double *D, bh_int ldd, const double *tau,
              bh_int nrhs, double *B, bh_int ldb);
int bh_sgbqrs(bh_int m, bh_int n, bh_int l, bh_int u, const float *D, bh_int ldd, const float *tau,
              bh_int nrhs, float *B, bh_int ldb);

/* ----------------------------------------------------------- symmetric */
/* Symmetric positive definite band Cholesky A = U'U (uplo 'U') or L L' ('L')
 * in LAPACK symmetric band storage (U: A(i,j), i <= j, at AB[kd+i-j + j*ldab];
 * L: A(i,j), i >= j, at AB[i-j + j*ldab]), in place, with LAPACK's info
 * (> 0: the leading minor of that order is not positive definite; the
 * columns up to it are factored and the rest left as the update left them).
 * Bit-identical to xPBTF2.  kd up to 137 (FP64) / 195 (FP32): the live
 * columns sit in LDS.  Replaces pbtrf! (src/lapack.jl:268-292) as
 * cholesky(Symmetric(A)) calls it (src/symbanded/BandedCholesky.jl:1-11). */
int bh_dpbtrf(char uplo, bh_int n, bh_int kd, double *AB, bh_int ldab, bh_int *info_host);
int bh_spbtrf(char uplo, bh_int n, bh_int kd, float *AB, bh_int ldab, bh_int *info_host);
/* Solve A X = B with the factor of bh_?pbtrf: two triangular band solves per
 * column (U' then U, or L then L'), all columns in one launch each.
 * Replaces pbtrs! (src/lapack.jl:297-333) as ldiv!(::Cholesky, B) calls it
 * (src/symbanded/BandedCholesky.jl:66-74). */
int bh_dpbtrs(char uplo, bh_int n, bh_int kd, bh_int nrhs, const double *AB, bh_int ldab, double *B,
              bh_int ldb);
int bh_spbtrs(char uplo, bh_int n, bh_int kd, bh_int nrhs, const float *AB, bh_int ldab, float *B,
              bh_int ldb);
/* Banded LDL' without pivoting of the stored triangle (storage as for
 * bh_?pbtrf), in place: L's entries where A's were, d on the diagonal.  No
 * positivity check (a zero pivot gives Inf/NaN, as in the reference).  The
 * right-looking form of the reference's loops: equal up to rounding.
 * Replaces ldlt!(::Symmetric{<:Any,<:BandedMatrix}) (src/symbanded/ldlt.jl:8-44). */
int bh_dsbldlt(char uplo, bh_int n, bh_int kd, double *AB, bh_int ldab);
int bh_ssbldlt(char uplo, bh_int n, bh_int kd, float *AB, bh_int ldab);
/* Solve with that factor: unit-L sweep, division by d, unit-L' sweep, all
 * columns of B in one launch each.  Replaces ldiv!(::LDLt, B) (ldlt.jl:47-69). */
int bh_dsbldlts(char uplo, bh_int n, bh_int kd, bh_int nrhs, const double *AB, bh_int ldab, double *B,
                bh_int ldb);
int bh_ssbldlts(char uplo, bh_int n, bh_int kd, bh_int nrhs, const float *AB, bh_int ldab, float *B,
                bh_int ldb);
/* y <- alpha*A*x + beta*y for the symmetric band A stored as one triangle (as
 * above, k off-diagonals); y not read when beta == 0.  Row i sums its terms
 * in column order with one fma each.  Replaces BLAS.sbmv! as
 * materialize!(::MulAdd{SymmetricLayout{BandedColumns}}) calls it
 * (src/symbanded/symbanded.jl:72-91). */
int bh_dsbmv(char uplo, bh_int n, bh_int k, double alpha, const double *A, bh_int lda,
             const double *x, bh_int incx, double beta, double *y, bh_int incy);
int bh_ssbmv(char uplo, bh_int n, bh_int k, float alpha, const float *A, bh_int lda,
             const float *x, bh_int incx, float beta, float *y, bh_int incy);

/* -------------------------------------------------------- element-wise */
/* Scale the stored entries A(i,j), r0 <= i < r1, c0 <= j < c1 (0-based,
 * half-open), of an m x n band (kl, ku, either may be negative with
 * kl + ku >= -1): A(i,j) <- alpha*A(i,j); with strong_zero != 0 and alpha == 0
 * they are set to 0 (NaN and Inf included).  Slots outside the matrix are not
 * touched.  Replaces lmul!(α, A) / rmul!(A, α) on a BandedMatrix
 * (src/generic/broadcast.jl:942-958; strong_zero = 0) and the partial
 * _fill_lmul!(β, view(C, ...)) of the band-pruned gbmm! (src/banded/gbmm.jl:231-279). */
int bh_dgbscal(bh_int m, bh_int n, bh_int kl, bh_int ku, double alpha, double *A, bh_int lda,
               bh_int r0, bh_int r1, bh_int c0, bh_int c1, int strong_zero);
int bh_sgbscal(bh_int m, bh_int n, bh_int kl, bh_int ku, float alpha, float *A, bh_int lda,
               bh_int r0, bh_int r1, bh_int c0, bh_int c1, int strong_zero);
/* Y <- alpha*X + Y for m x n bands X (xl, xu) and Y (yl, yu) over the bands
 * they share, each entry alpha*x rounded, then added.  A nonzero of X outside
 * Y's bands returns BH_ERR_BANDERROR before Y is modified.
 * Replaces banded_generic_axpy! (src/generic/broadcast.jl:978-1020). */
int bh_dgbaxpy(bh_int m, bh_int n, double alpha, const double *X, bh_int ldx, bh_int xl, bh_int xu,
               double *Y, bh_int ldy, bh_int yl, bh_int yu);
int bh_sgbaxpy(bh_int m, bh_int n, float alpha, const float *X, bh_int ldx, bh_int xl, bh_int xu,
               float *Y, bh_int ldy, bh_int yl, bh_int yu);
/* B <- A between m x n bands of different bandwidths: B's entries on A's bands
 * copied bit for bit, the rest of B's band set to 0; a nonzero of A outside
 * B's bands returns BH_ERR_BANDERROR before B is modified.  Replaces
 * copyto!(B, A) / BandedMatrix(A, (l, u)) on banded data
 * (src/banded/BandedMatrix.jl:124-165). */
int bh_dgbcopy(bh_int m, bh_int n, const double *A, bh_int lda, bh_int al, bh_int au, double *B,
               bh_int ldb, bh_int bl, bh_int bu);
int bh_sgbcopy(bh_int m, bh_int n, const float *A, bh_int lda, bh_int al, bh_int au, float *B,
               bh_int ldb, bh_int bl, bh_int bu);

/* ---------------------------------------------------------- triangular */
/* Triangular band solve x <- op(A)^{-1} x and product x <- op(A) x in BLAS
 * xTBSV / xTBMV form: n x n triangle with k off-diagonals, uplo 'U' (A(i,j)
 * at A[k+i-j + j*lda]) or 'L' (A(i,j) at A[i-j + j*lda]), diag 'U' for a unit
 * diagonal that is never read, lda >= k+1, x with stride incx >= 1.
 * A may point into a BandedMatrix's data (a row offset of its band) with the
 * parent's leading dimension; only rows 0..k of each column are read.
 * Solve 'N' equals OpenBLAS xtbsv bit for bit (division, then one fma per
 * entry); 'T' sums the dots across a wavefront (agrees to rounding).
 * The product sums each row's terms in order of distance from the diagonal;
 * it takes n temporary elements from the stream-ordered allocator.
 * Replace tbsv! / tbmv! (src/blas.jl:71-147) as triangular \ and * on banded
 * data call them (src/tribanded.jl:47-54,75-96). */
int bh_dtbsv(char uplo, char trans, char diag, bh_int n, bh_int k, const double *A, bh_int lda,
             double *x, bh_int incx);
int bh_stbsv(char uplo, char trans, char diag, bh_int n, bh_int k, const float *A, bh_int lda,
             float *x, bh_int incx);
/* The same solve for the nrhs columns of X (n x nrhs, leading dimension ldx),
 * one wavefront per column in one launch; each column's result equals the
 * single-vector call's bit for bit.  Replaces the column loop of tbsv! calls
 * that ldiv!(Tri(A), B) makes for a matrix B (src/tribanded.jl:75-96). */
int bh_dtbsv_multi(char uplo, char trans, char diag, bh_int n, bh_int k, const double *A, bh_int lda,
                   bh_int nrhs, double *X, bh_int ldx);
int bh_stbsv_multi(char uplo, char trans, char diag, bh_int n, bh_int k, const float *A, bh_int lda,
                   bh_int nrhs, float *X, bh_int ldx);
int bh_dtbmv(char uplo, char trans, char diag, bh_int n, bh_int k, const double *A, bh_int lda,
             double *x, bh_int incx);
int bh_stbmv(char uplo, char trans, char diag, bh_int n, bh_int k, const float *A, bh_int lda,
             float *x, bh_int incx);

/* ------------------------------------------------------------- batched */
/* Many independent square systems A_s x_s = b_s (kl_s <= 8, kl_s + ku_s <= 63)
 * factored and solved in one persistent launch: each of the W resident waves
 * walks its share of the schedule (systems w, w+W, ...; by default ordered by
 * decreasing work and dealt out in snake rounds), running the band LU with
 * the forward substitution fused into the factorisation sweep, then the back
 * substitution.  Per system the results equal LAPACK dgbtf2 +
 * dgbtrs 'N' bit for bit.
 * Packed layout (bh_batch_plan_layout): system s's padded AB (ldab = 2kl+ku+1,
 * as bh_?gbpad writes it) at AB + ab_off[s], its b / ipiv (1-based) at
 * B + vec_off[s] / ipiv + vec_off[s]; info[s] is LAPACK's info.
 * No reference counterpart (every reference call factors one matrix,
 * src/banded/BandedLU.jl:90-111); callers are loops of lu(A_s) \ b_s. */
typedef struct bh_batch_plan bh_batch_plan;
typedef struct bh_batch_graph bh_batch_graph;
int bh_batch_plan_create(bh_int nsys, const bh_int *n, const bh_int *kl, const bh_int *ku,
                         bh_batch_plan **plan);
int bh_batch_plan_layout(const bh_batch_plan *plan, bh_int *ab_off, bh_int *vec_off,
                         bh_int *ab_elems, bh_int *vec_elems);
/* Schedule: order 0 = decreasing work (snake rounds), 1 = as given
 * (round-robin), 2 = band classes interleaved; waves per workgroup (1..4);
 * resident workgroups. */
int bh_batch_plan_schedule(bh_batch_plan *plan, int order_kind, int waves_per_block, int blocks);
int bh_batch_plan_get_schedule(const bh_batch_plan *plan, int *order_kind, int *waves_per_block,
                               int *blocks);
int bh_batch_plan_destroy(bh_batch_plan *plan);
int bh_dgbsv_batched(bh_batch_plan *plan, double *AB, double *B, bh_int *ipiv, bh_int *info);
int bh_sgbsv_batched(bh_batch_plan *plan, float *AB, float *B, bh_int *ipiv, bh_int *info);
/* Time the candidate schedules on scratch copies of AB / B (the inputs are
 * not modified) and keep the fastest in the plan; *best_ms its time. */
int bh_dgbsv_batched_autotune(bh_batch_plan *plan, const double *AB, const double *B, double *best_ms);
int bh_sgbsv_batched_autotune(bh_batch_plan *plan, const float *AB, const float *B, double *best_ms);
/* Capture the batched solve (one kernel node, no memset node) into a graph
 * bound to these buffers; replay with bh_batch_graph_launch on the library
 * stream. */
int bh_dgbsv_batched_graph(bh_batch_plan *plan, double *AB, double *B, bh_int *ipiv, bh_int *info,
                           bh_batch_graph **graph);
int bh_sgbsv_batched_graph(bh_batch_plan *plan, float *AB, float *B, bh_int *ipiv, bh_int *info,
                           bh_batch_graph **graph);
int bh_batch_graph_launch(bh_batch_graph *graph);
int bh_batch_graph_destroy(bh_batch_graph *graph);

/* ------------------------------------------------------------- adjoint */
/* Gradient of L = w' A^{-1} b + 1/2 ||A x0||^2 with respect to A's band, given
 * lam = A^{-T} w (bh_?gbtrs 'T' with the same factors), x = A \ b, y0 = A x0:
 * G_ij = fma(y0_i, x0_j, -(lam_i x_j)) for every (i,j) of the m x n band
 * (G is band storage, ldg >= kl+ku+1; slots outside the matrix untouched).
 * No reference counterpart (the reference has no rrule); the adjoint solve
 * reuses the transposed LU solve of src/banded/linalg.jl:41-47. */
int bh_dgb_band_grad(bh_int m, bh_int n, bh_int kl, bh_int ku, const double *lam, const double *x,
                     const double *y0, const double *x0, double *G, bh_int ldg);
int bh_sgb_band_grad(bh_int m, bh_int n, bh_int kl, bh_int ku, const float *lam, const float *x,
                     const float *y0, const float *x0, float *G, bh_int ldg);
/* G <- beta G + P(alpha1 u1 v1' + alpha2 u2 v2'), P = restriction to the band,
 * each term one fma(alpha u_i, v_j, g) in that order; a NULL (u, v) pair is
 * skipped.  The pullbacks of A \ b (alpha = -1, u = A^{-T} xbar, v = x) and of
 * A * x0 (alpha = 1, u = ybar, v = x0) in band storage. */
int bh_dgb_band_outer(bh_int m, bh_int n, bh_int kl, bh_int ku, double alpha1, const double *u1,
                      const double *v1, double alpha2, const double *u2, const double *v2, double beta,
                      double *G, bh_int ldg);
int bh_sgb_band_outer(bh_int m, bh_int n, bh_int kl, bh_int ku, float alpha1, const float *u1,
                      const float *v1, float alpha2, const float *u2, const float *v2, float beta,
                      float *G, bh_int ldg);

/* ------------------------------------------------------- mixed precision */
/* X = A \ B for an FP64 band matrix A ((kl+ku+1) x n, lda) via an FP32 band LU
 * and FP64 iterative refinement: residuals B - A X in FP64, corrections from
 * FP32 solves, stop when every column meets LAPACK dsgesv's test
 * ||r||_inf <= sqrt(n) eps64 ||A||_inf ||x||_inf (at most 30 iterations), else
 * fall back to an FP64 LU (then *iters_host < 0).  B is not modified.
 * work: device workspace of bh_dsgbsv_workspace() bytes.  ipiv32 (device,
 * optional) receives the FP32 factorisation's pivots.
 * No reference counterpart (the reference factors in the input precision,
 * src/banded/BandedLU.jl:108-111); it is a drop-in for A\b (linalg.jl:5-9). */
int bh_dsgbsv_workspace(bh_int n, bh_int kl, bh_int ku, bh_int nrhs, size_t *bytes);
int bh_dsgbsv(bh_int n, bh_int kl, bh_int ku, bh_int nrhs, const double *A, bh_int lda,
              const double *B, bh_int ldb, double *X, bh_int ldx, void *work, size_t lwork,
              bh_int *ipiv32, bh_int *iters_host, bh_int *info_host);

/* bh_dsgbsv without a workspace argument and without any device allocation:
 * no reference counterpart; it is for callers that cannot allocate device
 * memory (serving loops, harnesses that forbid hipMalloc).  The workspace and
 * the FP64 fallback's factors live in one statically allocated 64 MiB device
 * array per GPU (held from library load until process exit), so
 * max(bh_dsgbsv_workspace bytes, 256-aligned (2kl+ku+1) n 8 + 8 n) must not
 * exceed 64 MiB, else BH_ERR_WORKSPACE.  bh_init(device) must have run on the
 * current device (else BH_ERR_NODEVICE).  Every call on a GPU shares that
 * array: calls that overlap on one GPU (other host threads or streams) race. */
int bh_dsgbsv_noalloc(bh_int n, bh_int kl, bh_int ku, bh_int nrhs, const double *A, bh_int lda,
                      const double *B, bh_int ldb, double *X, bh_int ldx, bh_int *ipiv32,
                      bh_int *iters_host, bh_int *info_host);

#ifdef __cplusplus
}
#endif

#endif /* BANDEDHIP_H */
