"""CPU oracle of the reverse mode of the HT reduction -- TEST INFRASTRUCTURE ONLY.

Only the test-suite may import this module, as the checker of the GPU adjoint
(``paraht_amd.adjoint``).  The reference has no adjoint; what pins this
restatement is (tests/test_adjoint_oracle.py):
  * central finite differences of the oracle's own reduction
    (``oracle.reduce``, the C restatement of src/ab2rht.f90 + src/rht2ht.f90);
  * the transpose (dot-product) identity between ``tangent`` and
    ``transpose_recurrence``, two separately written recurrences;
  * torch FP64 autograd through ``autograd_ref.reduce_unblocked``, a
    restatement of src/rht2ht.f90:2053-2180.

Mathematics.  For H = Q^T A Z, T = Q^T B Z with Q e1 = e1 the HT form is
unique up to diagonal signs, which are locally constant, so the reduction is
a smooth map.  With E_A = Q^T dA Z, E_B = Q^T dB Z, Om_Q = Q^T dQ and
Om_Z = Z^T dZ (skew, Om_Q e1 = 0):
    dH = E_A - Om_Q H + H Om_Z,     dT = E_B - Om_Q T + T Om_Z,
and dH being zero below the subdiagonal, dT below the diagonal, fixes the
strictly lower parts w_Q, w_Z of Om_Q, Om_Z column by column (``tangent``).
"""
from __future__ import annotations

import numpy as np


def tangent(H, T, EA, EB):
    """(w_Q, w_Z): strictly lower parts of Q^T dQ and Z^T dZ for the
    transformed perturbations EA = Q^T dA Z, EB = Q^T dB Z.  For j = 0..n-2:
      T(j+1:,j+1:) w_Z(j+1:,j) = -E_B(j+1:,j) + w_Q(j+1:,:j+1) T(:j+1,j)
      w_Q(j+2:,j+1) = [E_A(j+2:,j) - w_Q(j+2:,:j+1) H(:j+1,j)
                       + H(j+2:,j+1:) w_Z(j+1:,j)] / H(j+1,j)."""
    n = H.shape[0]
    wQ = np.zeros((n, n))
    wZ = np.zeros((n, n))
    for j in range(n - 1):
        rhs = -EB[j + 1:, j] + wQ[j + 1:, :j + 1] @ T[:j + 1, j]
        wZ[j + 1:, j] = np.linalg.solve(T[j + 1:, j + 1:], rhs)
        if j + 2 < n:
            num = (EA[j + 2:, j] - wQ[j + 2:, :j + 1] @ H[:j + 1, j]
                   + H[j + 2:, j + 1:] @ wZ[j + 1:, j])
            wQ[j + 2:, j + 1] = num / H[j + 1, j]
    return wQ, wZ


def jvp(H, T, Q, Z, dA, dB):
    """Forward-mode derivative (dH, dT, dQ, dZ) of the reduction."""
    EA = Q.T @ dA @ Z
    EB = Q.T @ dB @ Z
    wQ, wZ = tangent(H, T, EA, EB)
    OQ = wQ - wQ.T
    OZ = wZ - wZ.T
    dH = EA - OQ @ H + H @ OZ
    dT = EB - OQ @ T + T @ OZ
    return np.triu(dH, -1), np.triu(dT), Q @ OQ, Z @ OZ


def transpose_recurrence(H, T, gQ, gZ):
    """(X_A, X_B) with <gQ, w_Q> + <gZ, w_Z> = <X_A, E_A> + <X_B, E_B> for
    (w_Q, w_Z) = tangent(H, T, E_A, E_B): the recurrence of ``tangent`` run
    backwards, j = n-2..0.  The contributions of later steps to column j+1
    of w_Q's cotangent are kept as the columns of two matrices (Aa, Rb)
    instead of being pushed forward by rank-2 updates."""
    n = H.shape[0]
    Tinv = np.linalg.inv(T)
    Aa = np.zeros((n, n))
    Rb = np.zeros((n, n))
    XA = np.zeros((n, n))
    XB = np.zeros((n, n))
    for j in range(n - 2, -1, -1):
        cq = gQ[:, j + 1] - Aa[:, j + 1:] @ H[j + 1, j + 1:] + Rb[:, j + 1:] @ T[j + 1, j + 1:]
        a = np.zeros(n)
        if j + 2 < n:
            a[j + 2:] = cq[j + 2:] / H[j + 1, j]
        Aa[:, j] = a
        XA[j + 2:, j] = a[j + 2:]
        cz = gZ[j + 1:, j] + H[j + 2:, j + 1:].T @ a[j + 2:]
        rb = Tinv[j + 1:, j + 1:].T @ cz
        Rb[j + 1:, j] = rb
        XB[j + 1:, j] = -rb
    return XA, XB


def omega_cotangents(H, T, Q, Z, Hb, Tb, Qb, Zb):
    """(g_Q, g_Z): cotangents of w_Q, w_Z from those of (H, T, Q, Z)."""
    Hb = np.triu(Hb, -1)
    Tb = np.triu(Tb)
    GQ = Q.T @ Qb - Hb @ H.T - Tb @ T.T
    GZ = Z.T @ Zb + H.T @ Hb + T.T @ Tb
    gQ = np.tril(GQ - GQ.T, -1)
    gQ[:, 0] = 0.0
    gZ = np.tril(GZ - GZ.T, -1)
    return gQ, gZ


def vjp(H, T, Q, Z, Hb, Tb, Qb, Zb):
    """Reverse-mode derivative (Abar, Bbar); Bbar on B's upper triangle."""
    gQ, gZ = omega_cotangents(H, T, Q, Z, Hb, Tb, Qb, Zb)
    XA, XB = transpose_recurrence(H, T, gQ, gZ)
    EA = np.triu(Hb, -1) + XA
    EB = np.triu(Tb) + XB
    return Q @ EA @ Z.T, np.triu(Q @ EB @ Z.T)


def weighted_loss(outs, weights):
    """sum W*H + V*T + P*Q + S*Z over the structured part of H and T."""
    H, T, Q, Z = outs
    W, V, P, S = weights
    return (np.sum(np.triu(W, -1) * H) + np.sum(np.triu(V) * T)
            + np.sum(P * Q) + np.sum(S * Z))


def align_signs(outs, base):
    """outs carried to the column signs of base: (D_Q H D_Z, D_Q T D_Z, Q D_Q,
    Z D_Z) with D_Q = sign(diag(Q_base^T Q)), D_Z likewise."""
    H, T, Q, Z = outs
    dq = np.sign(np.sum(base[2] * Q, axis=0))
    dz = np.sign(np.sum(base[3] * Z, axis=0))
    return (dq[:, None] * H * dz[None, :], dq[:, None] * T * dz[None, :],
            Q * dq[None, :], Z * dz[None, :])


def central_difference(forward, A0, B0, dA, dB, weights, h=1e-6, base=None):
    """(L(A0 + h dA, B0 + h dB) - L(A0 - h dA, B0 - h dB)) / 2h of the
    weighted loss of ``forward(A, B) -> (H, T, Q, Z)``.

    The reduction is smooth only between the surfaces where one of its
    rotations or reflectors changes branch (DLARTG's c >= 0 and DLARFG's
    beta = -sign(alpha) conventions: a kept entry crossing zero flips the
    signs of two columns of Q or Z).  With ``base`` (the unperturbed
    decomposition) both perturbed decompositions are carried to its signs,
    which gives the difference quotient of the branch the derivative is
    taken on even when a step of size h crosses such a surface."""
    plus = forward(A0 + h * dA, B0 + h * dB)
    minus = forward(A0 - h * dA, B0 - h * dB)
    if base is not None:
        plus, minus = align_signs(plus, base), align_signs(minus, base)
    return (weighted_loss(plus, weights) - weighted_loss(minus, weights)) / (2.0 * h)
