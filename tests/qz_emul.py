"""numpy statement of the GPU QZ iteration (paraht_amd/csrc/qz.hip), used by the
CPU tests to check the algorithm against LAPACK (scipy) before it runs on the
GPU, and by the GPU tests as a second opinion.

The reference has no QZ step: its driver stops at the HT form and compares
with DGGHD3 (src/profilereduction.f90:152-162).  The LAPACK routine that would
come next is DHGEQZ; this is the Moler-Stewart implicit double-shift QZ
iteration (C. B. Moler, G. W. Stewart, SIAM J. Numer. Anal. 10 (1973);
Golub & Van Loan, Matrix Computations, 4th ed., Alg. 7.7.2 and section 7.7.5
for infinite eigenvalues), stated with the same transformations, in the same
order, as the kernel:

  * every transformation is a 2- or 3-element Householder reflector
    P = I - tau v v^T (v general, so sign flips are reflectors with v = e_i,
    tau = 2), applied from the left to rows or from the right to columns;
  * negligible subdiagonals of H: |h(j,j-1)| <= ulp (|h(j,j)| + |h(j-1,j-1)|),
    or <= ulp ||H||_F where both diagonal entries are zero; negligible
    diagonal entries of T: |t(j,j)| < ulp ||T||_F (these become exact zeros,
    i.e. infinite eigenvalues);
  * a zero on T's diagonal is chased to the bottom of its block (left
    reflector on rows j, j+1 zeroing t(j+1,j+1), right reflector on columns
    j-1, j removing the fill h(j+1,j-1)), then deflated with beta = 0;
  * 2x2 blocks: real pairs are split with a right reflector onto the null
    vector of beta A - alpha B and a left reflector that restores T; complex
    pairs have their T block made diagonal and positive (one-sided Jacobi);
  * double-shift sweep over an unreduced block [lo, m] of order >= 3: shifts
    from the trailing 2x2 pencil (exceptional shifts every 10th iteration
    without deflation), then per bulge position k a left reflector on rows
    k..k+2 and right reflectors on columns k..k+2 and k..k+1 restoring T.

Eigenvalues only (schur=False): transformations reach the active block only,
as DHGEQZ with JOB='E'.  Schur form (schur=True): whole rows and columns, and
optionally Q <- Q P, Z <- Z P.  Returns (alphar, alphai, beta, info) in
DHGEQZ's convention; info = 0, or the 1-based index of the last unconverged
eigenvalue when 30 n iterations did not suffice.
"""
import math

import numpy as np

ULP = 2.0 ** -52
SAFMIN = np.finfo(float).tiny


def house(x):
    """v with v[0] = 1, tau, beta such that (I - tau v v^T) x = beta e1."""
    x = np.asarray(x, float)
    alpha = x[0]
    xn = float(np.linalg.norm(x[1:]))
    v = np.zeros(len(x))
    v[0] = 1.0
    if xn == 0.0:
        return v, 0.0, alpha
    beta = -math.copysign(math.hypot(alpha, xn), alpha)
    tau = (beta - alpha) / beta
    v[1:] = x[1:] / (alpha - beta)
    return v, tau, beta


def house_rev(w):
    """v with v[-1] = 1, tau such that w (I - tau v v^T) = (0, ..., 0, beta)."""
    v, tau, beta = house(np.asarray(w, float)[::-1])
    return v[::-1].copy(), tau, beta


def flip(nv, i):
    """The reflector that negates entry i of an nv-vector."""
    v = np.zeros(nv)
    v[i] = 1.0
    return v, 2.0


class QZ:
    def __init__(self, H, T, Q=None, Z=None, schur=False, itmax_factor=30):
        self.H, self.T, self.Q, self.Z = H, T, Q, Z
        self.n = H.shape[0]
        self.schur = schur
        self.itmax = itmax_factor * max(self.n, 1)
        self.ar = np.zeros(self.n)
        self.ai = np.zeros(self.n)
        self.be = np.zeros(self.n)

    # ---- transformations --------------------------------------------
    def left(self, r0, v, tau, lo, m, hc0=None):
        """Rows r0..r0+len(v)-1 of H (columns hc0..last) and T (columns
        r0..last), columns of Q."""
        if tau == 0.0:
            return
        k = len(v)
        last = self.n - 1 if self.schur else m
        hc0 = r0 if hc0 is None else hc0
        for M, c0 in ((self.H, hc0), (self.T, r0)):
            blk = M[r0:r0 + k, c0:last + 1]
            blk -= tau * np.outer(v, v @ blk)
        if self.schur and self.Q is not None:
            blk = self.Q[:, r0:r0 + k]
            blk -= tau * np.outer(blk @ v, v)

    def right(self, c0, v, tau, lo, m, hr1, tr1):
        """Columns c0..c0+len(v)-1 of H (rows first..hr1) and T (rows
        first..tr1), columns of Z."""
        if tau == 0.0:
            return
        k = len(v)
        first = 0 if self.schur else lo
        for M, r1 in ((self.H, hr1), (self.T, tr1)):
            blk = M[first:r1 + 1, c0:c0 + k]
            blk -= tau * np.outer(blk @ v, v)
        if self.schur and self.Z is not None:
            blk = self.Z[:, c0:c0 + k]
            blk -= tau * np.outer(blk @ v, v)

    # ---- deflation ----------------------------------------------------
    def one(self, j, lo, m):
        """1x1 block at j: beta >= 0."""
        H, T = self.H, self.T
        if T[j, j] < 0.0:
            v, tau = flip(1, 0)
            if self.schur:
                self.right(j, v, tau, lo, m, j, j)
            else:
                H[j, j], T[j, j] = -H[j, j], -T[j, j]
        self.ar[j], self.ai[j], self.be[j] = H[j, j], 0.0, T[j, j]

    def two(self, lo, m):
        """2x2 block at rows/columns m-1, m."""
        H, T = self.H, self.T
        a = H[m - 1:m + 1, m - 1:m + 1].copy()
        b = T[m - 1:m + 1, m - 1:m + 1].copy()
        s = a[0, 0] / b[0, 0] + a[1, 1] / b[1, 1] - a[1, 0] * b[0, 1] / (b[0, 0] * b[1, 1])
        p = (a[0, 0] * a[1, 1] - a[0, 1] * a[1, 0]) / (b[0, 0] * b[1, 1])
        disc = 0.25 * s * s - p
        if disc >= 0.0:
            # real pair: split on the null vector of beta A - alpha B
            l1 = 0.5 * s + math.copysign(math.sqrt(disc), s)
            lam = l1
            if abs(lam) > 1.0:
                al, bt = 1.0, 1.0 / lam
            else:
                al, bt = lam, 1.0
            an = max(np.abs(a).max(), SAFMIN)
            bn = max(np.abs(b).max(), SAFMIN)
            M = bt * a - al * b
            r = 0 if abs(M[0, 0]) + abs(M[0, 1]) >= abs(M[1, 0]) + abs(M[1, 1]) else 1
            # M is singular: a right reflector taking its larger row to
            # (0, *) maps e1 onto its null vector
            w = np.array([M[r, 0], M[r, 1]])
            v, tau, _ = house_rev(w)
            self.right(m - 1, v, tau, lo, m, m, m)
            # left reflector restoring T (or H where T's column is smaller)
            ca = H[m - 1:m + 1, m - 1]
            cb = T[m - 1:m + 1, m - 1]
            x = cb if np.linalg.norm(cb) / bn >= np.linalg.norm(ca) / an else ca
            v, tau, _ = house(x.copy())
            self.left(m - 1, v, tau, lo, m, m - 1)
            H[m, m - 1] = 0.0
            T[m, m - 1] = 0.0
            self.one(m - 1, lo, m)
            self.one(m, lo, m)
            return
        # complex pair: make T's block diagonal (one-sided Jacobi), positive
        g11 = b[0, 0] * b[0, 0]
        g12 = b[0, 0] * b[0, 1]
        g22 = b[0, 1] * b[0, 1] + b[1, 1] * b[1, 1]
        if g12 != 0.0:
            zeta = (g22 - g11) / (2.0 * g12)
            t = math.copysign(1.0, zeta) / (abs(zeta) + math.sqrt(1.0 + zeta * zeta))
            c = 1.0 / math.sqrt(1.0 + t * t)
            sn = t * c
            # the rotation [[c, sn], [-sn, c]] as a reflector followed by a
            # column sign flip: columns (c, -sn) and (sn, c)
            v, tau, _ = house_rev(np.array([sn, c]))
            self.right(m - 1, v, tau, lo, m, m, m)
        x = T[m - 1:m + 1, m - 1].copy()
        v, tau, _ = house(x)
        self.left(m - 1, v, tau, lo, m, m - 1)
        T[m, m - 1] = 0.0
        T[m - 1, m] = 0.0
        for j in (m - 1, m):
            if T[j, j] < 0.0:
                v, tau = flip(2, j - (m - 1))
                self.left(m - 1, v, tau, lo, m, m - 1)
        b1, b2 = T[m - 1, m - 1], T[m, m]
        a = H[m - 1:m + 1, m - 1:m + 1]
        s = a[0, 0] / b1 + a[1, 1] / b2
        p = (a[0, 0] * a[1, 1] - a[0, 1] * a[1, 0]) / (b1 * b2)
        re = 0.5 * s
        im = math.sqrt(max(p - re * re, 0.0))
        self.ar[m - 1], self.ai[m - 1], self.be[m - 1] = re * b1, im * b1, b1
        self.ar[m], self.ai[m], self.be[m] = re * b2, -im * b2, b2

    def chase_zero(self, kz, lo, m):
        """T(kz, kz) = 0: move the zero to T(m, m), then deflate m."""
        H, T = self.H, self.T
        for j in range(kz, m):
            v, tau, _ = house(np.array([T[j, j + 1], T[j + 1, j + 1]]))
            self.left(j, v, tau, lo, m, max(j - 1, lo))
            T[j + 1, j + 1] = 0.0
            if j > lo:
                v, tau, _ = house_rev(np.array([H[j + 1, j - 1], H[j + 1, j]]))
                self.right(j - 1, v, tau, lo, m, j + 1, j)
                H[j + 1, j - 1] = 0.0
                T[j, j - 1] = 0.0
        if m > lo:
            v, tau, _ = house_rev(np.array([H[m, m - 1], H[m, m]]))
            self.right(m - 1, v, tau, lo, m, m, m)
            H[m, m - 1] = 0.0
            T[m, m - 1] = 0.0
        self.ar[m], self.ai[m], self.be[m] = H[m, m], 0.0, 0.0

    # ---- the double-shift sweep --------------------------------------
    def shifts(self, lo, m, exceptional):
        H, T = self.H, self.T
        if exceptional:
            sg = H[m, m] / T[m, m] + 0.75 * abs(H[m, m - 1] / T[m - 1, m - 1])
            return 2.0 * sg, sg * sg
        a, b, d = T[m - 1, m - 1], T[m - 1, m], T[m, m]
        e, f, g, h = H[m - 1, m - 1], H[m - 1, m], H[m, m - 1], H[m, m]
        s = e / a + h / d - g * b / (a * d)
        p = (e * h - f * g) / (a * d)
        return s, p

    def first_column(self, lo, s, p):
        H, T = self.H, self.T
        y1 = H[lo, lo] / T[lo, lo]
        y2 = H[lo + 1, lo] / T[lo, lo]
        w2 = y2 / T[lo + 1, lo + 1]
        w1 = (y1 - T[lo, lo + 1] * w2) / T[lo, lo]
        x0 = H[lo, lo] * w1 + H[lo, lo + 1] * w2 - s * y1 + p
        x1 = H[lo + 1, lo] * w1 + H[lo + 1, lo + 1] * w2 - s * y2
        x2 = H[lo + 2, lo + 1] * w2
        return np.array([x0, x1, x2])

    def sweep(self, lo, m, exceptional):
        H, T = self.H, self.T
        s, p = self.shifts(lo, m, exceptional)
        for k in range(lo, m):
            nr = 3 if k <= m - 2 else 2
            x = self.first_column(lo, s, p)[:nr] if k == lo else H[k:k + nr, k - 1].copy()
            v, tau, _ = house(x)
            self.left(k, v, tau, lo, m, k if k == lo else k - 1)
            if k > lo:
                H[k + 1:k + nr, k - 1] = 0.0
            hr1 = min(k + 3, m)
            if nr == 3:
                v, tau, _ = house_rev(T[k + 2, k:k + 3].copy())
                self.right(k, v, tau, lo, m, hr1, k + 2)
                T[k + 2, k:k + 2] = 0.0
            v, tau, _ = house_rev(T[k + 1, k:k + 2].copy())
            self.right(k, v, tau, lo, m, hr1, k + 1)
            T[k + 1, k] = 0.0

    # ---- driver -------------------------------------------------------
    def run(self):
        H, T, n = self.H, self.T, self.n
        if n == 0:
            return self.ar, self.ai, self.be, 0
        anorm = float(np.linalg.norm(H))
        bnorm = float(np.linalg.norm(T))
        atol = max(SAFMIN, ULP * anorm)
        btol = max(SAFMIN, ULP * bnorm)
        m = n - 1
        it = 0
        since = 0
        while m >= 0:
            lo = 0
            for j in range(m, 0, -1):
                d = abs(H[j, j]) + abs(H[j - 1, j - 1])
                tol = max(SAFMIN, ULP * d) if d > 0.0 else atol
                if abs(H[j, j - 1]) <= tol:
                    H[j, j - 1] = 0.0
                    lo = j
                    break
            kz = -1
            for j in range(m, lo - 1, -1):
                if abs(T[j, j]) < btol:
                    T[j, j] = 0.0
                    kz = j
                    break
            if lo == m:
                if kz == m:
                    self.ar[m], self.ai[m], self.be[m] = H[m, m], 0.0, 0.0
                else:
                    self.one(m, lo, m)
                m -= 1
                since = 0
                continue
            if kz >= 0:
                self.chase_zero(kz, lo, m)
                m -= 1
                since = 0
                continue
            if lo == m - 1:
                self.two(lo, m)
                m -= 2
                since = 0
                continue
            it += 1
            since += 1
            if it > self.itmax:
                return self.ar, self.ai, self.be, m + 1
            self.sweep(lo, m, since % 10 == 0)
        return self.ar, self.ai, self.be, 0


def hgeqz(H, T, Q=None, Z=None, schur=False):
    """In place on H, T (and Q, Z with schur=True); (alphar, alphai, beta, info)."""
    return QZ(H, T, Q, Z, schur).run()


def chordal(a1, b1, a2, b2):
    """Chordal distance between eigenvalues a1/b1 and a2/b2 (complex)."""
    num = abs(a1 * b2 - b1 * a2)
    return num / (math.hypot(abs(a1), abs(b1)) * math.hypot(abs(a2), abs(b2)))


def match_chordal(alpha1, beta1, alpha2, beta2):
    """Largest chordal distance under the best one-to-one matching."""
    from scipy.optimize import linear_sum_assignment
    a1, b1 = np.asarray(alpha1, complex), np.asarray(beta1, complex)
    a2, b2 = np.asarray(alpha2, complex), np.asarray(beta2, complex)
    n1 = np.hypot(np.abs(a1), np.abs(b1))
    n2 = np.hypot(np.abs(a2), np.abs(b2))
    D = np.abs(np.outer(a1, b2) - np.outer(b1, a2)) / np.outer(n1, n2)
    r, c = linear_sum_assignment(D)
    return float(D[r, c].max()) if len(r) else 0.0
