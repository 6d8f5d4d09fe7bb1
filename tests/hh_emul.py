"""CPU emulation of the GPU schedule of the Householder stage 2 -- test helper.

The GPU runs one sweep of rht2ht_householder (src/rht2ht.f90:1346-2051) as:
  * generation: one leader workgroup per column ic of the sweep chasing its
    bulge block by block, step (ic, ib) starting once column ic-1 has finished
    step ib+2 (leader and helpers) or all of its steps; helper workgroups apply
    each step's right reflector to the rows above the window
    (rht2ht.f90:1595-1596 of the reference's step) after the leader publishes
    the step;
  * T factors of every block (dlarft, rht2ht.f90:1599-1610);
  * the remaining right reflectors on rows of A and B as compact-WY products
    over per-row ranges of a block's sequences, the remaining left reflectors
    on columns right of the window, and Q, Z (rht2ht.f90:1630-2024).

This module restates the reference's sweep in numpy with the generation split
into those tasks, so the tests can check on the CPU that
  (1) any order of the tasks that respects the GPU's waits gives the same
      pencil bit for bit as the reference's column-by-column order;
  (2) the per-row masked compact-WY form of the row updates equals the
      reference's mix of dlarf (edges) and dlarfb (bulk) to rounding;
  (3) generating sweep k+1 before sweep k's far (non-lookahead) updates gives
      the same pencil bit for bit (the far updates and the next generation
      touch disjoint entries).
Indices follow the reference (1-based) through the helpers E/S.
"""
from __future__ import annotations

import math

import numpy as np


# ----------------------------------------------------------------- LAPACK bits
def dlarfg(alpha: float, x: np.ndarray):
    """(beta, tau, v) with DLARFG's conventions; x == 0 gives tau = 0, v = x."""
    if x.size == 0:
        return alpha, 0.0, x.copy()
    xnorm = float(np.sqrt(np.dot(x, x)))
    if xnorm == 0.0:
        return alpha, 0.0, x.copy()
    beta = -math.copysign(math.hypot(alpha, xnorm), alpha)
    tau = (beta - alpha) / beta
    return beta, tau, x * (1.0 / (alpha - beta))


def larf_left(C: np.ndarray, v: np.ndarray, tau: float) -> None:
    if tau != 0.0:
        C -= tau * np.outer(v, v @ C)


def larf_right(C: np.ndarray, v: np.ndarray, tau: float) -> None:
    if tau != 0.0:
        C -= tau * np.outer(C @ v, v)


def dgerq2_q1(Bc: np.ndarray) -> np.ndarray:
    """First row of Q of the RQ factorization Bc = R Q by DGERQ2's reflectors
    (rows nb..2, bottom-up), formed as e1^T H(2) ... H(nb)."""
    nb = Bc.shape[0]
    M = Bc.copy()
    refl = []
    for i in range(nb, 1, -1):  # row i (1-based); H(1) is the identity
        beta, tau, v = dlarfg(M[i - 1, i - 1], M[i - 1, : i - 1])
        vv = np.concatenate([v, [1.0]])
        larf_right(M[: i - 1, :i], vv, tau)
        M[i - 1, i - 1] = beta
        M[i - 1, : i - 1] = 0.0
        refl.append((i, vv, tau))
    x = np.zeros(nb)
    x[0] = 1.0
    for i, vv, tau in reversed(refl):  # H(2) first
        x[:i] -= tau * (x[:i] @ vv) * vv
    return x


def dlarft_fc(V: np.ndarray, tau: np.ndarray) -> np.ndarray:
    """T of H(1)...H(k) = I - V T V^T (DLARFT 'F','C')."""
    k = V.shape[1]
    T = np.zeros((k, k))
    for i in range(k):
        if tau[i] == 0.0:
            continue
        T[:i, i] = -tau[i] * (V[:, :i].T @ V[:, i])
        T[:i, i] = T[:i, :i] @ T[:i, i]
        T[i, i] = tau[i]
    return T


# ------------------------------------------------------------------ the sweep
class Sweep:
    """One sweep (columns s..s+nq-1) of the Householder stage 2 on A, B, Q, Z
    (numpy arrays, updated in place)."""

    def __init__(self, A, B, Q, Z, r, nq, s):
        self.A, self.B, self.Q, self.Z = A, B, Q, Z
        self.n = A.shape[0]
        self.r, self.nq, self.s = r, nq, s
        n = self.n
        self.nblocks = (n - s) // r + ((n - s) % r != 0)
        self.ldu = r + nq
        nbl = self.nblocks + 2
        self.U = np.zeros((self.ldu, nq, nbl))
        self.V = np.zeros((self.ldu, nq, nbl))
        self.tauU = np.zeros((nq, nbl))
        self.tauV = np.zeros((nq, nbl))
        self.TU = None
        self.TV = None
        # helper workgroups run beside their column's next steps only when
        # nq <= r+1: then no later step of the column reads the rows above
        # its earlier bulges (the GPU kernel runs the rows inside the step
        # otherwise)
        self.async_helpers = nq <= r + 1

    # -- geometry of step (ic, ib), 1-based like rht2ht.f90:1450-1596
    def step_geom(self, ic, ib):
        """(row0, col, nb, istart) of the step, or None past the matrix."""
        n, r, nq, s = self.n, self.r, self.nq, self.s
        jc = s + ic - 1
        if ib == 1:
            if jc > n:
                return None
            return jc + 1, jc, min(r, n - jc), s + 1
        jb = jc + 1 + (ib - 2) * r
        if jb > n:
            return None
        istart = max(s + 1, s + 1 + (ib - 1) * r - (nq - ic) * r)
        return jb + r, jb, min(r, n - jb - r + 1), istart

    def nsteps(self, ic):
        ib = 1
        while self.step_geom(ic, ib) is not None:
            ib += 1
        return ib - 1

    def leader(self, ic, ib):
        """Step (ic, ib) without the rows above the window."""
        A, B, U, V = self.A, self.B, self.U, self.V
        n, r, s = self.n, self.r, self.s
        row0, col, nb, _ = self.step_geom(ic, ib)
        g0 = s + (ib - 1) * r  # global row of local index 0 of block ib
        bcol = row0 + r - 1
        if bcol <= n:
            for i in range(1, ic):
                a = g0 + i  # first row (1-based) of reflector i
                larf_left(B[a - 1:a - 1 + r, bcol - 1:bcol], U[i - 1:i - 1 + r, i - 1, ib - 1],
                          self.tauU[i - 1, ib - 1])
        for i in range(1, ic):
            a = g0 + i
            if a <= n:
                ln = min(r, n - a + 1)
                larf_left(A[a - 1:a - 1 + ln, col - 1:col], U[i - 1:i - 1 + ln, i - 1, ib - 1],
                          self.tauU[i - 1, ib - 1])
        if nb < 2:
            return
        lo, hi = row0 - 1, row0 - 1 + nb  # 0-based rows/cols of the block
        beta, tau, v = dlarfg(A[lo, col - 1], A[lo + 1:hi, col - 1])
        self.tauU[ic - 1, ib - 1] = tau
        U[ic - 1, ic - 1, ib - 1] = 1.0
        U[ic:ic + nb - 1, ic - 1, ib - 1] = v
        A[lo, col - 1] = beta
        A[lo + 1:hi, col - 1] = 0.0
        u = U[ic - 1:ic - 1 + nb, ic - 1, ib - 1]
        larf_left(B[lo:hi, lo:hi], u, tau)
        q1 = dgerq2_q1(B[lo:hi, lo:hi])
        beta, tau, v = dlarfg(q1[0], q1[1:])
        self.tauV[ic - 1, ib - 1] = tau
        V[ic - 1, ic - 1, ib - 1] = 1.0
        V[ic:ic + nb - 1, ic - 1, ib - 1] = v
        w = V[ic - 1:ic - 1 + nb, ic - 1, ib - 1]
        last = min(n, row0 + r + nb - 1)
        larf_right(A[lo:last, lo:hi], w, tau)
        larf_right(B[lo:hi, lo:hi], w, tau)
        B[lo + 1:hi, lo] = 0.0

    def helper(self, ic, ib):
        """Rows istart..row0-1 of A and B of step (ic, ib)."""
        row0, col, nb, istart = self.step_geom(ic, ib)
        if nb < 2 or istart > row0 - 1:
            return
        w = self.V[ic - 1:ic - 1 + nb, ic - 1, ib - 1]
        tau = self.tauV[ic - 1, ib - 1]
        lo, hi = row0 - 1, row0 - 1 + nb
        larf_right(self.A[istart - 1:row0 - 1, lo:hi], w, tau)
        larf_right(self.B[istart - 1:row0 - 1, lo:hi], w, tau)

    def generate(self, rng=None, lag=2):
        """All steps; rng=None: the reference's order (column by column, each
        step with its helper rows); else a random order that respects the GPU
        kernel's waits."""
        nq = self.nq
        steps = [self.nsteps(ic) for ic in range(1, nq + 1)]
        if rng is None:
            for ic in range(1, nq + 1):
                for ib in range(1, steps[ic - 1] + 1):
                    self.leader(ic, ib)
                    self.helper(ic, ib)
            return
        led = [0] * (nq + 1)  # leader steps done per column (1-based ic)
        hel = [0] * (nq + 1)  # helper steps done per column
        while True:
            ready = []
            for ic in range(1, nq + 1):
                tot = steps[ic - 1]
                if led[ic] < tot:
                    ib = led[ic] + 1
                    if ic == 1:
                        ok = True
                    else:
                        tp = steps[ic - 2]
                        need = min(ib + lag, tp)
                        ok = led[ic - 1] >= need and hel[ic - 1] >= need
                    if ok:
                        ready.append(("L", ic))
                if hel[ic] < led[ic]:
                    # the helper of step ib waits for the leader's step ib
                    ready.append(("H", ic))
            if not ready:
                break
            kind, ic = ready[rng.integers(len(ready))]
            if kind == "L":
                led[ic] += 1
                self.leader(ic, led[ic])
                if not self.async_helpers:  # the step's rows above, in the step
                    hel[ic] += 1
                    self.helper(ic, hel[ic])
            else:
                hel[ic] += 1
                self.helper(ic, hel[ic])
        assert all(led[ic] == steps[ic - 1] and hel[ic] == steps[ic - 1]
                   for ic in range(1, nq + 1))

    # -- T factors, rht2ht.f90:1599-1610
    def factors(self):
        n, r, nq, s = self.n, self.r, self.nq, self.s
        self.TU = np.zeros((nq, nq, self.nblocks + 2))
        self.TV = np.zeros((nq, nq, self.nblocks + 2))
        for ib in range(self.nblocks, 0, -1):
            jb = s + (ib - 1) * r
            nb = min(n - jb, r + nq)
            self.TU[:, :, ib - 1] = dlarft_fc(self.U[:nb, :, ib - 1], self.tauU[:, ib - 1])
            self.TV[:, :, ib - 1] = dlarft_fc(self.V[:nb, :, ib - 1], self.tauV[:, ib - 1])

    # -- the reference's update phases (rht2ht.f90:1630-2024)
    def _wy_right(self, C, ib, rows, c0, nb, use_v=True):
        """C[rows, c0:c0+nb] <- C * (I - V T V^T), V = tape(:nb, :k)."""
        if nb <= 0 or len(rows) == 0:
            return
        k = min(self.nq, nb)
        W = self.V if use_v else self.U
        T = (self.TV if use_v else self.TU)[:k, :k, ib - 1]
        Vb = W[:nb, :k, ib - 1]
        Cs = C[rows[0]:rows[-1] + 1, c0:c0 + nb]
        Cs -= ((Cs @ Vb) @ T) @ Vb.T

    def _wy_left(self, C, ib, r0, nb, cols):
        """C[r0:r0+nb, cols] <- (I - U T^T U^T) C."""
        if nb <= 0 or cols[1] <= cols[0]:
            return
        k = min(self.nq, nb)
        T = self.TU[:k, :k, ib - 1]
        Ub = self.U[:nb, :k, ib - 1]
        Cs = C[r0:r0 + nb, cols[0]:cols[1]]
        Cs -= Ub @ (T.T @ (Ub.T @ Cs))

    def lookahead(self, M):
        """rht2ht.f90:1644-1765 on M (A or B): the rows within nq*r above each
        bulge, then the columns within r*nq right of it."""
        n, r, nq, s = self.n, self.r, self.nq, self.s
        for ib in range(self.nblocks, 0, -1):
            jb = s + (ib - 1) * r
            for ic in range(1, nq + 1):
                ist = max(s + 1, s + 1 + (ib - 1) * r - nq * r - (nq - ic) * r)
                ist2 = max(s + 1, s + 1 + (ib - 1) * r - (nq - ic) * r)
                if ist2 - 1 >= ist and jb + ic <= n:
                    ln = min(r, n - (jb + ic) + 1)
                    larf_right(M[ist - 1:ist2 - 1, jb + ic - 1:jb + ic - 1 + ln],
                               self.V[ic - 1:ic - 1 + ln, ic - 1, ib - 1],
                               self.tauV[ic - 1, ib - 1])
        for ib in range(self.nblocks, 0, -1):
            r0, nb, istop = self._left_geom(M, ib)
            istop2 = min(n, s + (ib - 1) * r + r + nq + r * nq - 1)
            if istop2 >= istop + 1:
                self._wy_left(M, ib, r0, nb, (istop, istop2))

    def _left_geom(self, M, ib):
        """(first row 0-based, rows, istop) of block ib's left update of M."""
        n, r, nq, s = self.n, self.r, self.nq, self.s
        if M is self.B:
            jb = s + (ib - 1) * r
            return jb, min(n - jb, r + nq), min(n, s + (ib - 1) * r + r + nq - 1)
        if ib == 1:
            jb = s
            return jb, min(n - jb, r + nq), min(n, jb + nq - 1)
        jb = s + 1 + (ib - 2) * r
        return jb + r - 1, min(n - jb - r + 1, r + nq), min(n, jb + nq - 1)

    def far(self, M):
        """rht2ht.f90:1818-1951 on M: rows above the lookahead rows (edge
        reflectors by dlarf, the bulk by dlarfb), columns beyond."""
        n, r, nq, s = self.n, self.r, self.nq, self.s
        for ib in range(self.nblocks, 0, -1):
            jb = s + (ib - 1) * r
            nb = min(n - jb, r + nq)
            ist = max(s + 1, s + 1 + (ib - 1) * r - (2 * nq - 1) * r)
            for ic in range(2, nq + 1):
                ist2 = max(s + 1, s + 1 + (ib - 1) * r - (2 * nq - ic) * r)
                if ist2 - 1 >= ist and jb + ic <= n:
                    ln = min(r, n - (jb + ic) + 1)
                    larf_right(M[ist - 1:ist2 - 1, jb + ic - 1:jb + ic - 1 + ln],
                               self.V[ic - 1:ic - 1 + ln, ic - 1, ib - 1],
                               self.tauV[ic - 1, ib - 1])
            self._wy_right(M, ib, list(range(0, ist - 1)), jb, nb)
        for ib in range(self.nblocks, 0, -1):
            r0, nb, _ = self._left_geom(M, ib)
            istop = min(n, s + (ib - 1) * r + r + nq + r * nq - 1)
            if n >= istop + 1:
                self._wy_left(M, ib, r0, nb, (istop, n))

    def qz(self):
        """rht2ht.f90:1976-2018."""
        n, r, nq, s = self.n, self.r, self.nq, self.s
        rows = list(range(n))
        for ib in range(self.nblocks, 0, -1):
            jb = s + (ib - 1) * r
            self._wy_right(self.Q, ib, rows, jb, min(n - jb, r + nq - 1), use_v=False)
        for ib in range(self.nblocks, 0, -1):
            jb = s + (ib - 1) * r
            self._wy_right(self.Z, ib, rows, jb, min(n - jb, r + nq))

    # -- the GPU's form of the row updates: per-row ranges of sequences
    def row_range(self, i, ib, phase):
        """Sequences [lo, hi] (1-based, empty if lo > hi) of block ib that row
        i (1-based) receives in phase 'look', 'far' or 'all' (both)."""
        r, nq, s = self.r, self.nq, self.s
        if i <= s:
            return (1, 0) if phase == "look" else (1, nq)
        base = s + 1 + (ib - 1) * r
        if i >= base:
            return 1, 0
        d = (i - base) // r  # floor
        if phase == "all":
            return max(1, d + nq + 1), nq
        if phase == "look":
            return max(1, d + nq + 1), min(nq, d + 2 * nq)
        if d <= -2 * nq:
            return 1, nq
        return max(1, d + 2 * nq + 1), nq

    def rows_masked(self, M, phase):
        """All blocks' right reflectors for `phase` on the rows of M, as the
        GPU applies them: W = C V, W and W T masked to the row's range."""
        n, r, nq, s = self.n, self.r, self.nq, self.s
        for ib in range(self.nblocks, 0, -1):
            jb = s + (ib - 1) * r
            nb = min(n - jb, r + nq)
            k = min(nq, nb)
            base = s + 1 + (ib - 1) * r
            top = min(n, base - 1)  # rows above the block's first window row
            if top < 1:
                continue
            C = M[:top, jb:jb + nb]
            Vb = self.V[:nb, :k, ib - 1]
            T = self.TV[:k, :k, ib - 1]
            mask = np.zeros((top, k))
            for i in range(1, top + 1):
                lo, hi = self.row_range(i, ib, phase)
                lo, hi = max(lo, 1), min(hi, k)
                if lo <= hi:
                    mask[i - 1, lo - 1:hi] = 1.0
            W = (C @ Vb) * mask
            W2 = (W @ T) * mask
            C -= W2 @ Vb.T


def reduce_sweeps(A, B, Q, Z, r, nq, *, rng=None, masked=False, next_first=False,
                  one_pass=False):
    """Stage 2 by this module's sweeps (in place).  rng: random generation
    order; masked: the GPU form of the row updates; next_first: each sweep's
    far updates after the next sweep's generation (and lookahead); one_pass:
    the GPU's single-stream form, all row updates of a matrix in one masked
    pass, then all its column updates."""
    n = A.shape[0]
    pending = None
    s = 1
    while s <= n - 2:
        sw = Sweep(A, B, Q, Z, r, nq, s)
        sw.generate(rng)
        sw.factors()
        if one_pass:
            for M in (B, A):
                sw.rows_masked(M, "all")
                for ib in range(sw.nblocks, 0, -1):
                    r0, nb, istop = sw._left_geom(M, ib)
                    if n >= istop + 1:
                        sw._wy_left(M, ib, r0, nb, (istop, n))
            sw.qz()
            s += nq
            continue
        if pending is not None:  # the previous sweep's far updates
            _far_all(pending, masked)
            pending = None
        if masked:
            sw.rows_masked(B, "look")
            _cols_only_look(sw, B)
            sw.rows_masked(A, "look")
            _cols_only_look(sw, A)
        else:
            sw.lookahead(B)
            sw.lookahead(A)
        if next_first:
            pending = sw
        else:
            _far_all(sw, masked)
        s += nq
    if pending is not None:
        _far_all(pending, masked)


def _cols_only_look(sw, M):
    n, r, nq, s = sw.n, sw.r, sw.nq, sw.s
    for ib in range(sw.nblocks, 0, -1):
        r0, nb, istop = sw._left_geom(M, ib)
        istop2 = min(n, s + (ib - 1) * r + r + nq + r * nq - 1)
        if istop2 >= istop + 1:
            sw._wy_left(M, ib, r0, nb, (istop, istop2))


def _far_all(sw, masked):
    if masked:
        for M in (sw.B, sw.A):
            sw.rows_masked(M, "far")
            n, r, nq, s = sw.n, sw.r, sw.nq, sw.s
            for ib in range(sw.nblocks, 0, -1):
                r0, nb, _ = sw._left_geom(M, ib)
                istop = min(n, s + (ib - 1) * r + r + nq + r * nq - 1)
                if n >= istop + 1:
                    sw._wy_left(M, ib, r0, nb, (istop, n))
    else:
        sw.far(sw.B)
        sw.far(sw.A)
    sw.qz()
