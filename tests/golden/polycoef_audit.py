"""Roundoff audit of the reference's polynomial fits (mod_intrastuff.f90:127-175).

polycoef fits each axis polynomial of a quartet by a Vandermonde solve
M C = b with M_ij = y_i^(L-j), y_i = sqe (x_i + P_ik - P_jl), x_i = i - 1/2,
and b_i the Gauss-Hermite sum of the product of four linear factors; it
zeroes the fit when LAPACK reports a singular matrix or any |C_k| > 1e3.
Where M is ill-conditioned, the fitted C are dominated by rounding and the
zeroing decision may depend on the arithmetic (libm pow or repeated
multiplication, FMA contraction, the BLAS).  This module decides, fit by
fit, whether the mathematics fixes the decision:

  C*  the exact solution (mpmath, 60 digits) of the system whose entries
      are formed exactly from the fit's FP64 inputs (the y_i are the FP64
      values every faithful implementation computes: (x_i + P_ik) - P_jl,
      then times sqe);
  d   a first-order bound on |C - C*| for ANY faithful FP64 implementation:
        2 |M^-1| ( e_M |M| |C| + gamma(3 np) P^T |L||U| |C| + db )
      with e_M = k u for a power y^k (k >= 2), the Gauss-elimination
      backward error gamma(3n) |L||U| (Higham, Accuracy and Stability of
      Numerical Algorithms, Thm 9.4) and db = (10 L + n + 30) u
      sum_k |w_k| prod_a F_a^t_a, F_a the sum of the magnitudes of the
      terms of each linear factor (u = 2^-53, n nodes);
  class  0 "keep": every |C*_k| + d_k <= 1e3 -- no faithful implementation
           zeroes the fit;
         1 "zero": some |C*_k| - d_k > 1e3 -- every one zeroes it;
         2 "roundoff": neither -- the decision belongs to the rounding;
         3 "unbounded": the perturbations can move C by more than its size
           (|M^-1| |dM| 1 >= 0.1 in some row): roundoff, no bound.

Well-conditioned fits (condition number below 1e8) are classified in FP64
(numpy; the inverse is then accurate to 1e-8 relative) with the same bound
evaluated on the FP64 solution; every other fit is solved at 60 digits.
A seeded sample of the FP64-classified fits is checked at 60 digits too.

Test infrastructure: used by gen_polycoef_audit.py (fixture generation) and
tests/test_polycoef_audit.py.  Gauss-Hermite roots and weights are the
oracle's (orc_gauherm, mode 0) FP64 values.
"""
from __future__ import annotations

import mpmath as mp
import numpy as np

U = 2.0 ** -53
LIMIT = 1.0e3
KEEP, ZERO, ROUNDOFF, UNBOUNDED = 0, 1, 2, 3
COND_FP64 = 1.0e8


def gauherm(L):
    """Oracle (mode 0) Gauss-Hermite nodes and weights for degree L."""
    from oracle import oracle as O
    rh = np.zeros(11)
    wh = np.zeros(11)
    n = O.lib().orc_gauherm(int(L), rh, wh)
    return rh[:n].copy(), wh[:n].copy()


def _unpack(row):
    Ri, Rj, Rk, Rl, Pik, Pjl, P = (float(v) for v in row[2:9])
    t = [int(v) for v in row[9:13]]
    npt = int(row[13])
    sqe, invz, alf = (float(v) for v in row[14:17])
    return (Ri, Rj, Rk, Rl), Pik, Pjl, P, t, npt, sqe, invz, alf


def _sign(a):
    """alfijkl - 1/2 for the factors of i and k, + 1/2 for j and l."""
    return -0.5 if a in (0, 2) else 0.5


def _y_values(Pik, Pjl, sqe, npt):
    # the FP64 arithmetic of every faithful implementation (Python floats
    # are IEEE doubles without contraction)
    return [sqe * ((-0.5 + (i + 1) + Pik) - Pjl) for i in range(npt)]


def _rhs_terms(Rs, P, t, invz, alf, xs, rh, wh, exact):
    """b_i (exact, mpmath) or the magnitude sums sum_k |w_k| prod F_a^t_a."""
    out = []
    for x in xs:
        acc = mp.mpf(0) if exact else 0.0
        for r, w in zip(rh, wh):
            prod = mp.mpf(1) if exact else 1.0
            for a in range(4):
                if t[a] == 0:
                    continue
                s = _sign(a)
                if exact:
                    f = mp.mpf(invz) * mp.mpf(r) + (mp.mpf(alf) + s) * mp.mpf(x) + \
                        (mp.mpf(P) - mp.mpf(Rs[a]))
                else:
                    f = abs(invz * r) + abs((alf + s) * x) + abs(P) + abs(Rs[a])
                prod = prod * f ** t[a]
            acc = acc + (mp.mpf(w) * prod if exact else abs(w) * prod)
        out.append(acc)
    return out


def _lu(A):
    """Partial-pivot LU (first maximal |pivot|, as LAPACK) of a square list
    matrix; returns (perm, L, U) with A[perm] = L U."""
    n = len(A)
    A = [row[:] for row in A]
    perm = list(range(n))
    for k in range(n):
        p = max(range(k, n), key=lambda i: (abs(A[i][k]), -i))
        A[k], A[p] = A[p], A[k]
        perm[k], perm[p] = perm[p], perm[k]
        piv = A[k][k]
        for i in range(k + 1, n):
            A[i][k] = A[i][k] / piv
            for j in range(k + 1, n):
                A[i][j] = A[i][j] - A[i][k] * A[k][j]
    L = [[(A[i][j] if j < i else (1 if i == j else 0)) for j in range(n)] for i in range(n)]
    Uu = [[(A[i][j] if j >= i else 0) for j in range(n)] for i in range(n)]
    return perm, L, Uu


def _lu_solve(perm, L, Uu, b):
    n = len(b)
    x = [b[perm[i]] for i in range(n)]
    for i in range(n):
        for k in range(i):
            x[i] = x[i] - L[i][k] * x[k]
    for i in range(n - 1, -1, -1):
        for k in range(i + 1, n):
            x[i] = x[i] - Uu[i][k] * x[k]
        x[i] = x[i] / Uu[i][i]
    return x


def _classify(C, d, valid):
    if not valid:
        return UNBOUNDED
    aC = np.abs(C)
    if np.all(aC + d <= LIMIT):
        return KEEP
    if np.any(aC - d > LIMIT):
        return ZERO
    return ROUNDOFF


def exact_fit(row, dps=60):
    """(C* as float64 array, bound d, valid, class) of one fit at `dps` digits."""
    Rs, Pik, Pjl, P, t, npt, sqe, invz, alf = _unpack(row)
    L = npt - 1
    rh, wh = gauherm(L)
    xs = [-0.5 + (i + 1) for i in range(npt)]
    with mp.workdps(dps):
        ys = _y_values(Pik, Pjl, sqe, npt)
        M = [[mp.mpf(y) ** (L - j) for j in range(npt)] for y in ys]
        b = _rhs_terms(Rs, P, t, invz, alf, xs, rh, wh, exact=True)
        perm, Lf, Uf = _lu(M)
        if any(Uf[i][i] == 0 for i in range(npt)):
            return np.full(npt, np.inf), np.full(npt, np.inf), False, UNBOUNDED
        Cs = _lu_solve(perm, Lf, Uf, b)
        Minv_cols = []
        for j in range(npt):
            e = [mp.mpf(1) if i == j else mp.mpf(0) for i in range(npt)]
            Minv_cols.append(_lu_solve(perm, Lf, Uf, e))
        Minv = np.array([[float(abs(Minv_cols[j][i])) for j in range(npt)] for i in range(npt)])
        Cabs = np.array([float(abs(c)) for c in Cs])
        Cf = np.array([float(c) for c in Cs])
        Mabs = np.array([[float(abs(v)) for v in r] for r in M])
        LU = np.abs(np.array([[float(v) for v in r] for r in Lf])) @ \
            np.abs(np.array([[float(v) for v in r] for r in Uf]))
    LU_orig = np.zeros_like(LU)
    LU_orig[perm] = LU  # rows of P^T |L||U| in the original order
    return _bound(Cf, Cabs, Minv, Mabs, LU_orig, Rs, P, t, invz, alf, xs, rh, wh, L, npt)


def _bound(Cf, Cabs, Minv, Mabs, LU, Rs, P, t, invz, alf, xs, rh, wh, L, npt):
    pw = np.array([max(L - j, 0) for j in range(npt)], dtype=float)
    eM = np.where(pw >= 2, pw * U, 0.0)[None, :] * Mabs
    gam = 3 * npt * U / (1 - 3 * npt * U)
    dM = eM + gam * LU
    n = len(rh)
    db = (10 * L + n + 30) * U * np.array(_rhs_terms(Rs, P, t, invz, alf, xs, rh, wh, exact=False))
    d = 2.0 * Minv @ (dM @ Cabs + db)
    valid = bool(np.max(Minv @ dM @ np.ones(npt)) < 0.1)
    return Cf, d, valid, _classify(Cf, d, valid)


def fp64_classify(rows, C, status):
    """FP64 classification of well-conditioned fits.  rows (n, 17), C (n, 22)
    the FP64 fits (mode 0) and status their zeroing (a zeroed fit's C is 0,
    so those always go to the 60-digit solve).  Returns (class, d_max, cond)
    with class -1 for the fits that need the 60-digit solve (condition >=
    COND_FP64, zeroed, or the decision not settled by the FP64 bound)."""
    n = rows.shape[0]
    cls = np.full(n, -1, np.int8)
    dmax = np.full(n, np.inf)
    cond = np.full(n, np.inf)
    for npt in np.unique(rows[:, 13].astype(int)):
        sel = np.nonzero(rows[:, 13].astype(int) == npt)[0]
        L = npt - 1
        rh, wh = gauherm(L)
        R = rows[sel]
        xs = -0.5 + np.arange(1, npt + 1)
        ys = R[:, 14:15] * ((xs[None, :] + R[:, 6:7]) - R[:, 7:8])
        pw = (L - np.arange(npt)).astype(float)
        M = ys[:, :, None] ** pw[None, None, :]
        with np.errstate(all="ignore"):
            Minv = np.linalg.inv(M)
            cnd = np.abs(M).sum(1).max(1) * np.abs(Minv).sum(1).max(1)
        ok = np.isfinite(cnd) & (cnd < COND_FP64) & (np.asarray(status)[sel] == 0)
        cond[sel] = cnd
        idx = sel[ok]
        if idx.size == 0:
            continue
        Mo, Mi = M[ok], np.abs(Minv[ok])
        Cabs = np.abs(C[idx, :npt])
        # vectorised partial-pivot LU for |L||U|
        A = Mo.copy()
        k_rows = np.arange(A.shape[0])
        perm = np.tile(np.arange(npt), (A.shape[0], 1))
        for k in range(npt):
            p = k + np.argmax(np.abs(A[:, k:, k]), axis=1)
            rk = A[k_rows, k].copy()
            A[k_rows, k] = A[k_rows, p]
            A[k_rows, p] = rk
            pk = perm[k_rows, k].copy()
            perm[k_rows, k] = perm[k_rows, p]
            perm[k_rows, p] = pk
            A[:, k + 1:, k] /= A[:, k:k + 1, k]
            A[:, k + 1:, k + 1:] -= A[:, k + 1:, k:k + 1] * A[:, k:k + 1, k + 1:]
        Lf = np.tril(A, -1) + np.eye(npt)[None]
        Uf = np.triu(A)
        LUp = np.abs(Lf) @ np.abs(Uf)
        LU = np.empty_like(LUp)
        LU[k_rows[:, None], perm] = LUp
        eM = np.where(pw >= 2, pw * U, 0.0)[None, None, :] * np.abs(Mo)
        gam = 3 * npt * U / (1 - 3 * npt * U)
        dM = eM + gam * LU
        # |b| magnitude sums
        F = np.ones((idx.size, npt))
        db = np.zeros((idx.size, npt))
        for r, w in zip(rh, wh):
            prod = np.ones((idx.size, npt))
            for a in range(4):
                ta = R[ok, 9 + a]
                s = _sign(a)
                f = np.abs(R[ok, 15:16] * r) + np.abs((R[ok, 16:17] + s) * xs[None, :]) + \
                    np.abs(R[ok, 8:9]) + np.abs(R[ok, 2 + a:3 + a])
                prod = prod * f ** ta[:, None]
            db += abs(w) * prod
        del F
        db *= (10 * L + len(rh) + 30) * U
        d = 2.0 * np.einsum("nij,nj->ni", Mi, np.einsum("nij,nj->ni", dM, Cabs) + db)
        # the FP64 solution is within d/2 of C*; settle only what 2d settles
        valid = np.einsum("nij,nj->ni", Mi, np.einsum("nij,nj->ni", dM,
                                                       np.ones((idx.size, npt)))).max(1) < 0.1
        keep = valid & np.all(Cabs + 2 * d <= LIMIT, axis=1)
        zero = valid & np.any(Cabs - 2 * d > LIMIT, axis=1)
        c = np.full(idx.size, -1, np.int8)
        c[keep] = KEEP
        c[zero] = ZERO
        cls[idx] = c
        dmax[idx] = d.max(1)
    return cls, dmax, cond
