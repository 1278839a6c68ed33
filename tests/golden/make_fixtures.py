"""Write the golden fixtures in this directory.

The reference stores no numeric vectors for the banded hot path (its tests
compare against dense products and factorizations at run time), so these are
generated here:

* from scipy's OpenBLAS 0.3.28 -- the library whose dgbmv / dgbtrf / dgbtrs /
  dgeqrf the reference's BLAS.gbmv! / LAPACK.gbtrf! / LAPACK.gbtrs! calls
  reach -- on small seeded inputs, with NaN in the corner slots of the band
  storage (never read), ties in pivot columns and a singular column;
* for the triangular and symmetric rows, from scipy's dtbsv / dtbmv / dpbtrf /
  dpbtrs / dsbmv (the routines tbsv!, tbmv!, pbtrf!, pbtrs! and BLAS.sbmv!
  call, src/blas.jl:71-147, src/lapack.jl:268-333, src/symbanded/symbanded.jl:72-91);
* from the reference tests' deterministic integer cases, whose products are
  exact: the symmetric band of test/test_linalg.jl:54 and the
  (-1, 1, 2)-banded matrix of test/test_linalg.jl:99, and the right-hand side
  b = 1:5 of test/test_bandedlu.jl:13-14.

Run from the repository root:  python tests/golden/make_fixtures.py
Each fixture is an .npz of plain arrays (no pickles).
"""
import os

import numpy as np
import scipy.linalg.blas as sblas
import scipy.linalg.lapack as slap

HERE = os.path.dirname(os.path.abspath(__file__))


def band_from_dense(M, kl, ku):
    m, n = M.shape
    D = np.full((kl + ku + 1, n), np.nan)
    for j in range(n):
        for i in range(max(0, j - ku), min(m, j + kl + 1)):
            D[ku + i - j, j] = M[i, j]
    return np.asfortranarray(D)


def pad_for_lu(D, kl, ku):
    n = D.shape[1]
    AB = np.zeros((2 * kl + ku + 1, n), D.dtype, order="F")
    AB[kl:, :] = np.nan_to_num(D, nan=0.0)
    return AB


def random_band(rng, m, n, kl, ku, dtype, dist="n"):
    D = (rng.standard_normal((kl + ku + 1, n)) if dist == "n" else rng.random((kl + ku + 1, n))).astype(dtype)
    D = np.asfortranarray(D)
    for j in range(n):
        for t in range(kl + ku + 1):
            i = j - ku + t
            if i < 0 or i >= m:
                D[t, j] = np.nan
    return D


def gbmv_fixtures():
    rng = np.random.default_rng(20240101)
    out = {}
    for tag, (m, n, kl, ku, alpha, beta, dt) in {
        "sq_d": (37, 37, 3, 2, 1.5, 0.5, np.float64),
        "wide_d": (23, 41, 4, 6, -2.0, 0.0, np.float64),
        "tall_d": (41, 23, 6, 1, 0.75, 2.0, np.float64),
        "sq_s": (37, 37, 3, 2, 1.5, 0.5, np.float32),
    }.items():
        D = random_band(rng, m, n, kl, ku, dt)
        x = rng.standard_normal(n).astype(dt)
        y = rng.standard_normal(m).astype(dt)
        f = sblas.dgbmv if dt == np.float64 else sblas.sgbmv
        ref = f(m, n, kl, ku, alpha, np.nan_to_num(D, nan=0.0), x, beta=beta, y=y.copy())
        out.update({f"{tag}_dims": np.array([m, n, kl, ku]), f"{tag}_ab": np.array([alpha, beta]), f"{tag}_D": D,
                    f"{tag}_x": x, f"{tag}_y": y, f"{tag}_ref": ref})
    np.savez(os.path.join(HERE, "gbmv_openblas.npz"), **out)


def lu_fixtures():
    rng = np.random.default_rng(7)
    out = {}
    cases = {}
    cases["piv_d"] = (random_band(rng, 60, 60, 4, 3, np.float64, "u"), 60, 4, 3)
    cases["piv_s"] = (random_band(rng, 60, 60, 4, 3, np.float32, "u"), 60, 4, 3)
    cases["wide_d"] = (random_band(rng, 50, 50, 8, 20, np.float64, "n"), 50, 8, 20)
    # ties: entries from {-1, 1, 0.5}, equal magnitudes in every pivot column
    T = np.asfortranarray(rng.choice([-1.0, 1.0, 0.5], size=(3 + 2 + 1, 30)))
    for j in range(30):
        for t in range(6):
            i = j - 2 + t
            if i < 0 or i >= 30:
                T[t, j] = np.nan
    cases["ties_d"] = (T, 30, 3, 2)
    # a zero column: info > 0, LAPACK continues past it
    Z = random_band(rng, 25, 25, 2, 2, np.float64, "n")
    for t in range(5):
        if not np.isnan(Z[t, 11]):
            Z[t, 11] = 0.0
    cases["singular_d"] = (Z, 25, 2, 2)
    for tag, (D, n, kl, ku) in cases.items():
        AB = pad_for_lu(D, kl, ku)
        f = slap.dgbtrf if D.dtype == np.float64 else slap.sgbtrf
        lub, piv, info = f(AB.copy(order="F"), kl, ku)
        B = np.asfortranarray(rng.standard_normal((n, 2)).astype(D.dtype))
        out.update({f"{tag}_dims": np.array([n, kl, ku]), f"{tag}_D": D, f"{tag}_lu": np.asfortranarray(lub),
                    f"{tag}_ipiv": (piv + 1).astype(np.int64), f"{tag}_info": np.array([info]), f"{tag}_B": B})
        if info == 0:
            g = slap.dgbtrs if D.dtype == np.float64 else slap.sgbtrs
            xN, _ = g(lub, kl, ku, B.copy(order="F"), piv, trans=0)
            xT, _ = g(lub, kl, ku, B.copy(order="F"), piv, trans=1)
            out[f"{tag}_xN"] = np.asfortranarray(xN)
            out[f"{tag}_xT"] = np.asfortranarray(xT)
    # test_bandedlu.jl:13-14: b = 1:5 with a 5x5 (1,1) band (fixed entries here)
    M = np.array([[4.0, 1, 0, 0, 0], [2, 5, -1, 0, 0], [0, 1, 3, 2, 0], [0, 0, -2, 6, 1], [0, 0, 0, 1, 2]])
    D5 = band_from_dense(M, 1, 1)
    AB = pad_for_lu(D5, 1, 1)
    lub, piv, info = slap.dgbtrf(AB.copy(order="F"), 1, 1)
    b = np.arange(1.0, 6.0).reshape(5, 1, order="F")
    x, _ = slap.dgbtrs(lub, 1, 1, b.copy(order="F"), piv)
    out.update({"b15_dims": np.array([5, 1, 1]), "b15_D": D5, "b15_lu": np.asfortranarray(lub),
                "b15_ipiv": (piv + 1).astype(np.int64), "b15_info": np.array([info]), "b15_B": b,
                "b15_xN": np.asfortranarray(x), "b15_dense_x": np.linalg.solve(M, np.arange(1.0, 6.0))})
    np.savez(os.path.join(HERE, "lu_openblas.npz"), **out)


def integer_fixtures():
    out = {}
    # test_linalg.jl:54: Symmetric(BandedMatrix(0=>1:10, 1=>11:19, 2=>21:28)), v = 1:10
    M = np.diag(np.arange(1.0, 11.0)) + np.diag(np.arange(11.0, 20.0), 1) + np.diag(np.arange(21.0, 29.0), 2)
    M = np.triu(M) + np.triu(M, 1).T
    out.update({"sym_D": band_from_dense(M, 2, 2), "sym_v": np.arange(1.0, 11.0), "sym_Mv": M @ np.arange(1.0, 11.0),
                "sym_MM": band_from_dense(M @ M, 4, 4)})
    # test_linalg.jl:99: BandedMatrix(-1=>1:n, 1=>11:n+10, 2=>21:n+19), n = 100 -> 101 x 101, bands (1,2)
    n = 100
    N = n + 1
    B = np.zeros((N, N))
    B[np.arange(1, N), np.arange(0, N - 1)] = np.arange(1.0, n + 1)
    B[np.arange(0, N - 1), np.arange(1, N)] = np.arange(11.0, n + 11)
    B[np.arange(0, N - 2), np.arange(2, N)] = np.arange(21.0, n + 20)
    x = np.arange(1.0, N + 1)
    out.update({"strided_D": band_from_dense(B, 1, 2), "strided_x": x, "strided_Bx": B @ x,
                "strided_BB": band_from_dense(B @ B, 2, 4)})
    np.savez(os.path.join(HERE, "integer_closed_form.npz"), **out)


def qr_fixtures():
    rng = np.random.default_rng(11)
    out = {}
    for tag, (n, l, u) in {"a": (12, 2, 1), "b": (20, 3, 3)}.items():
        D = random_band(rng, n, n, l, u, np.float64)
        M = np.zeros((n, n))
        for j in range(n):
            for t in range(l + u + 1):
                i = j - u + t
                if 0 <= i < n:
                    M[i, j] = D[t, j]
        qr, tau, work, info = slap.dgeqrf(np.asfortranarray(M))
        out.update({f"{tag}_dims": np.array([n, l, u]), f"{tag}_D": D, f"{tag}_R": np.triu(qr), f"{tag}_tau": tau})
    np.savez(os.path.join(HERE, "qr_lapack.npz"), **out)


def triangular_fixtures():
    rng = np.random.default_rng(21)
    out = {}
    for uplo in ("U", "L"):
        for diag in ("N", "U"):
            n, k = 90, 5
            a = rng.uniform(-1.0, 1.0, size=(k + 1, n)) * 0.1
            dr = k if uplo == "U" else 0
            a[dr, :] = 1.0 + rng.random(n)
            a = np.asfortranarray(a)
            x = rng.standard_normal(n)
            tag = f"{uplo}{diag}"
            out[f"{tag}_dims"] = np.array([n, k])
            out[f"{tag}_a"] = a
            out[f"{tag}_x"] = x
            for trans in ("N", "T"):
                kw = dict(lower=int(uplo == "L"), trans=int(trans == "T"), diag=int(diag == "U"))
                out[f"{tag}_sv{trans}"] = sblas.dtbsv(k, a, x.copy(), **kw)
                out[f"{tag}_mv{trans}"] = sblas.dtbmv(k, a, x.copy(), **kw)
    np.savez(os.path.join(HERE, "tri_openblas.npz"), **out)


def symmetric_fixtures():
    rng = np.random.default_rng(22)
    out = {}
    for uplo in ("U", "L"):
        for n, kd in ((70, 4), (120, 20)):
            a = rng.uniform(-1.0, 1.0, size=(kd + 1, n))
            dr = kd if uplo == "U" else 0
            a[dr, :] = kd + 1 + rng.random(n)
            for j in range(n):  # corner slots outside the matrix hold 0
                for t in range(1, kd + 1):
                    if uplo == "U" and j - t < 0:
                        a[kd - t, j] = 0
                    if uplo == "L" and j + t >= n:
                        a[t, j] = 0
            a = np.asfortranarray(a)
            c, info = slap.dpbtrf(a.copy(order="F"), lower=int(uplo == "L"))
            b = rng.standard_normal(n)
            xs, _ = slap.dpbtrs(c, b, lower=int(uplo == "L"))
            x = rng.standard_normal(n)
            y = sblas.dsbmv(kd, 1.25, a, x, lower=int(uplo == "L"))
            tag = f"{uplo}{kd}"
            out.update({f"{tag}_dims": np.array([n, kd]), f"{tag}_a": a, f"{tag}_chol": np.asfortranarray(c),
                        f"{tag}_info": np.array([info]), f"{tag}_b": b, f"{tag}_x": xs, f"{tag}_sx": x,
                        f"{tag}_sy": y})
    np.savez(os.path.join(HERE, "sym_lapack.npz"), **out)


if __name__ == "__main__":
    import sys

    only = sys.argv[1:]  # e.g. "triangular symmetric": regenerate only those
    for name, fn in (("gbmv", gbmv_fixtures), ("lu", lu_fixtures), ("integer", integer_fixtures),
                     ("qr", qr_fixtures), ("triangular", triangular_fixtures), ("symmetric", symmetric_fixtures)):
        if not only or name in only:
            fn()
    for f in sorted(os.listdir(HERE)):
        if f.endswith(".npz"):
            print(f, os.path.getsize(os.path.join(HERE, f)), "bytes")
