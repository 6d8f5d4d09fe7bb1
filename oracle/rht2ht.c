/*
 * Stage 2 of the reference, restated for the CPU oracle (test infrastructure
 * only): the three variants of /src/rht2ht.f90 that reduce an r-Hessenberg-
 * triangular pencil to Hessenberg-triangular form.
 *
 *   oracle_rht2ht_householder           rht2ht.f90:1346-2051 (driver's variant)
 *   oracle_rht2ht_givens                rht2ht.f90:696-1344
 *   oracle_rht2ht_householder_unblocked rht2ht.f90:2053-2180
 *
 * OpenMP tasks are executed in program order; the reference's tape copies
 * (U2/V2/...) hold the same values as U/V in program order and are not
 * duplicated.  Rotations and reflector entries that lie past row/column n are
 * identities/zeros in the reference and are skipped here (same arithmetic).
 */
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define E(a, ld, i, j) (a)[((size_t)(j)-1) * (size_t)(ld) + ((size_t)(i)-1)]
#define MIN(a, b) ((a) < (b) ? (a) : (b))
#define MAX(a, b) ((a) > (b) ? (a) : (b))

static int check_args(int n, int lda, int ldb, int ldq, int ldz, int r,
                      int nq) {
  if (n < 0) return -1;
  if (lda < MAX(1, n)) return -3;
  if (ldb < MAX(1, n)) return -5;
  if (ldq < MAX(1, n)) return -7;
  if (ldz < MAX(1, n)) return -9;
  if (r < 1) return -10;
  if (nq < 1) return -11;
  return 0;
}

/* C(1:m,1:k) <- C(1:m,1:k) * M(1:k,1:k)   (dgemm 'N','N' + dlacpy) */
static void gemm_right(int m, int k, double* C, int ldc, const double* M,
                       int ldm) {
  if (m <= 0 || k <= 0) return;
  double* W = (double*)malloc(sizeof(double) * (size_t)m * k);
#pragma omp parallel for schedule(static) if ((long)m * k * k > 2000000L)
  for (int j = 1; j <= k; ++j) {
    double* w = W + (size_t)(j - 1) * m;
    for (int i = 0; i < m; ++i) w[i] = 0.0;
    for (int l = 1; l <= k; ++l) {
      double t = E(M, ldm, l, j);
      if (t == 0.0) continue;
      const double* cl = &E(C, ldc, 1, l);
      for (int i = 0; i < m; ++i) w[i] += cl[i] * t;
    }
  }
  for (int j = 1; j <= k; ++j)
    memcpy(&E(C, ldc, 1, j), W + (size_t)(j - 1) * m, sizeof(double) * m);
  free(W);
}

/* C(1:k,1:m) <- M(1:k,1:k)^T * C(1:k,1:m)   (dgemm 'T','N' + dlacpy) */
static void gemm_left_t(int k, int m, double* C, int ldc, const double* M,
                        int ldm) {
  if (m <= 0 || k <= 0) return;
#pragma omp parallel for schedule(static) if ((long)m * k * k > 2000000L)
  for (int j = 1; j <= m; ++j) {
    double w[512];
    double* wp = k <= 512 ? w : (double*)malloc(sizeof(double) * k);
    for (int i = 1; i <= k; ++i) {
      double s = 0.0;
      for (int l = 1; l <= k; ++l) s += E(M, ldm, l, i) * E(C, ldc, l, j);
      wp[i - 1] = s;
    }
    for (int i = 1; i <= k; ++i) E(C, ldc, i, j) = wp[i - 1];
    if (wp != w) free(wp);
  }
}

/* ------------------------------------------------------------------------ */
/* rht2ht.f90:2053-2180 — every reflector applied at once to all of A,B,Q,Z. */
int oracle_rht2ht_householder_unblocked(int n, double* A, int lda, double* B,
                                        int ldb, double* Q, int ldq, double* Z,
                                        int ldz, int r) {
  int info = check_args(n, lda, ldb, ldq, ldz, r, 1);
  if (info) return info;
  double* Bc = (double*)calloc((size_t)r * r, sizeof(double));
  double* tau = (double*)calloc((size_t)r, sizeof(double));
  double t, temp;
  for (int jc = 1; jc <= n - 2; ++jc) {
    int nb = MIN(r, n - jc);
    /* rht2ht.f90:2085-2100 */
    o_dlarfg(nb, &E(A, lda, jc + 1, jc), &E(A, lda, jc + 2, jc), 1, &t);
    temp = E(A, lda, jc + 1, jc);
    E(A, lda, jc + 1, jc) = 1.0;
    o_dlarf('L', nb, n - jc, &E(A, lda, jc + 1, jc), 1, t,
            &E(A, lda, jc + 1, jc + 1), lda);
    o_dlarf('L', nb, n - jc, &E(A, lda, jc + 1, jc), 1, t,
            &E(B, ldb, jc + 1, jc + 1), ldb);
    o_dlarf('R', n, nb, &E(A, lda, jc + 1, jc), 1, t, &E(Q, ldq, 1, jc + 1),
            ldq);
    E(A, lda, jc + 1, jc) = temp;
    for (int i = 2; i <= nb; ++i) E(A, lda, jc + i, jc) = 0.0;
    /* rht2ht.f90:2102-2122 */
    for (int j = 1; j <= nb; ++j)
      for (int i = 1; i <= nb; ++i)
        E(Bc, r, i, j) = E(B, ldb, jc + i, jc + j);
    o_dgerq2(nb, nb, Bc, r, tau);
    o_dorgr2(nb, nb, nb, Bc, r, tau);
    o_dlarfg(nb, &E(Bc, r, 1, 1), &E(Bc, r, 1, 2), r, &t);
    E(Bc, r, 1, 1) = 1.0;
    o_dlarf('R', n, nb, &E(Bc, r, 1, 1), r, t, &E(A, lda, 1, jc + 1), lda);
    o_dlarf('R', jc + nb, nb, &E(Bc, r, 1, 1), r, t, &E(B, ldb, 1, jc + 1),
            ldb);
    o_dlarf('R', n, nb, &E(Bc, r, 1, 1), r, t, &E(Z, ldz, 1, jc + 1), ldz);
    for (int i = 2; i <= nb; ++i) E(B, ldb, jc + i, jc + 1) = 0.0;
    /* chase, rht2ht.f90:2129-2173 */
    for (int jb = jc + 1; jb <= n - r; jb += r) {
      nb = MIN(r, n - jb - r + 1);
      o_dlarfg(nb, &E(A, lda, jb + r, jb), &E(A, lda, jb + r + 1, jb), 1, &t);
      temp = E(A, lda, jb + r, jb);
      E(A, lda, jb + r, jb) = 1.0;
      o_dlarf('L', nb, n - jb, &E(A, lda, jb + r, jb), 1, t,
              &E(A, lda, jb + r, jb + 1), lda);
      o_dlarf('L', nb, n - jb - r + 1, &E(A, lda, jb + r, jb), 1, t,
              &E(B, ldb, jb + r, jb + r), ldb);
      o_dlarf('R', n, nb, &E(A, lda, jb + r, jb), 1, t, &E(Q, ldq, 1, jb + r),
              ldq);
      E(A, lda, jb + r, jb) = temp;
      for (int i = 1; i <= nb - 1; ++i) E(A, lda, jb + r + i, jb) = 0.0;
      for (int j = 1; j <= nb; ++j)
        for (int i = 1; i <= nb; ++i)
          E(Bc, r, i, j) = E(B, ldb, jb + r + i - 1, jb + r + j - 1);
      o_dgerq2(nb, nb, Bc, r, tau);
      o_dorgr2(nb, nb, nb, Bc, r, tau);
      o_dlarfg(nb, &E(Bc, r, 1, 1), &E(Bc, r, 1, 2), r, &t);
      E(Bc, r, 1, 1) = 1.0;
      o_dlarf('R', n, nb, &E(Bc, r, 1, 1), r, t, &E(A, lda, 1, jb + r), lda);
      o_dlarf('R', jb + r + nb - 1, nb, &E(Bc, r, 1, 1), r, t,
              &E(B, ldb, 1, jb + r), ldb);
      o_dlarf('R', n, nb, &E(Bc, r, 1, 1), r, t, &E(Z, ldz, 1, jb + r), ldz);
      for (int i = 1; i <= nb - 1; ++i) E(B, ldb, jb + r + i, jb + r) = 0.0;
    }
  }
  free(Bc);
  free(tau);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* rht2ht.f90:1346-2051 — sweeps of nq columns; reflectors of bulge block ib
 * stored as U(:,:,ib), V(:,:,ib) ((r+nq) x nq) with compact-WY factors. */
int oracle_rht2ht_householder(int n, double* A, int lda, double* B, int ldb,
                              double* Q, int ldq, double* Z, int ldz, int r,
                              int nq) {
  int info = check_args(n, lda, ldb, ldq, ldz, r, nq);
  if (info) return info;
  const int ldu = r + nq;
  int nbl_all = (n - 1) / r + (((n - 1) % r) != 0);
  if (nbl_all < 1) nbl_all = 1;
  size_t usz = (size_t)ldu * nq * nbl_all, tsz = (size_t)nq * nq * nbl_all;
  double* U = (double*)calloc(usz, sizeof(double));
  double* V = (double*)calloc(usz, sizeof(double));
  double* TU = (double*)calloc(tsz, sizeof(double));
  double* TV = (double*)calloc(tsz, sizeof(double));
  double* tauU = (double*)calloc((size_t)nq * nbl_all, sizeof(double));
  double* tauV = (double*)calloc((size_t)nq * nbl_all, sizeof(double));
  double* Bc = (double*)calloc((size_t)r * r, sizeof(double));
  double* tau = (double*)calloc((size_t)r, sizeof(double));
#define U3(i, j, b) U[((size_t)(b)-1) * ldu * nq + ((size_t)(j)-1) * ldu + ((i)-1)]
#define V3(i, j, b) V[((size_t)(b)-1) * ldu * nq + ((size_t)(j)-1) * ldu + ((i)-1)]
#define TAUU(i, b) tauU[((size_t)(b)-1) * nq + ((i)-1)]
#define TAUV(i, b) tauV[((size_t)(b)-1) * nq + ((i)-1)]
#define UB(b) (&U3(1, 1, b))
#define VB(b) (&V3(1, 1, b))
#define TUB(b) (TU + ((size_t)(b)-1) * nq * nq)
#define TVB(b) (TV + ((size_t)(b)-1) * nq * nq)

  for (int jblock = 1; jblock <= n - 2; jblock += nq) {
    int nblocks = (n - jblock) / r + (((n - jblock) % r) != 0);
    memset(U, 0, usz * sizeof(double));
    memset(V, 0, usz * sizeof(double));
    memset(TU, 0, tsz * sizeof(double));
    memset(TV, 0, tsz * sizeof(double));
    memset(tauU, 0, (size_t)nq * nbl_all * sizeof(double));
    memset(tauV, 0, (size_t)nq * nbl_all * sizeof(double));

    /* generation, rht2ht.f90:1450-1596 */
    for (int ic = 1; ic <= nq; ++ic) {
      int jc = jblock + ic - 1;
      int nb = MIN(r, n - jc);
      int istart = jblock + 1;
      if (jc + r <= n)
        for (int i = 1; i <= ic - 1; ++i)
          o_dlarf('L', r, 1, &U3(i, i, 1), 1, TAUU(i, 1),
                  &E(B, ldb, jblock + i, jc + r), ldb);
      for (int i = 1; i <= ic - 1; ++i)
        if (jc <= n)
          o_dlarf('L', MIN(r, n - jblock - i + 1), 1, &U3(i, i, 1), 1,
                  TAUU(i, 1), &E(A, lda, jblock + i, jc), lda);
      if (nb < 2) continue;
      o_dlarfg(nb, &E(A, lda, jc + 1, jc), &E(A, lda, jc + 2, jc), 1,
               &TAUU(ic, 1));
      for (int i = 2; i <= nb; ++i) {
        U3(ic + i - 1, ic, 1) = E(A, lda, jc + i, jc);
        E(A, lda, jc + i, jc) = 0.0;
      }
      U3(ic, ic, 1) = 1.0;
      o_dlarf('L', nb, nb, &U3(ic, ic, 1), 1, TAUU(ic, 1),
              &E(B, ldb, jc + 1, jc + 1), ldb);
      for (int j = 1; j <= nb; ++j)
        for (int i = 1; i <= nb; ++i)
          E(Bc, r, i, j) = E(B, ldb, jc + i, jc + j);
      o_dgerq2(nb, nb, Bc, r, tau);
      o_dorgr2(nb, nb, nb, Bc, r, tau);
      o_dlarfg(nb, &E(Bc, r, 1, 1), &E(Bc, r, 1, 2), r, &TAUV(ic, 1));
      for (int i = 2; i <= nb; ++i) V3(ic + i - 1, ic, 1) = E(Bc, r, 1, i);
      V3(ic, ic, 1) = 1.0;
      o_dlarf('R', MIN(n, jc + r + nb) - jc, nb, &V3(ic, ic, 1), 1,
              TAUV(ic, 1), &E(A, lda, jc + 1, jc + 1), lda);
      o_dlarf('R', nb, nb, &V3(ic, ic, 1), 1, TAUV(ic, 1),
              &E(B, ldb, jc + 1, jc + 1), ldb);
      for (int i = 2; i <= nb; ++i) E(B, ldb, jc + i, jc + 1) = 0.0;
      o_dlarf('R', jc - istart + 1, nb, &V3(ic, ic, 1), 1, TAUV(ic, 1),
              &E(A, lda, istart, jc + 1), lda);
      o_dlarf('R', jc - istart + 1, nb, &V3(ic, ic, 1), 1, TAUV(ic, 1),
              &E(B, ldb, istart, jc + 1), ldb);

      for (int jb = jc + 1; jb <= n; jb += r) {
        int ib = (jb - (jc + 1)) / r + 2;
        nb = MIN(r, n - jb - r + 1);
        istart = MAX(jblock + 1, jblock + 1 + (ib - 1) * r - (nq - ic) * r);
        if (jb + r + r - 1 <= n)
          for (int i = 1; i <= ic - 1; ++i)
            o_dlarf('L', r, 1, &U3(i, i, ib), 1, TAUU(i, ib),
                    &E(B, ldb, jb + r - ic + i, jb + r + r - 1), ldb);
        for (int i = 1; i <= ic - 1; ++i)
          if (jb + r - ic + i <= n)
            o_dlarf('L', MIN(r, n - (jb + r - ic + i) + 1), 1, &U3(i, i, ib),
                    1, TAUU(i, ib), &E(A, lda, jb + r - ic + i, jb), lda);
        if (nb < 2) continue;
        o_dlarfg(nb, &E(A, lda, jb + r, jb), &E(A, lda, jb + r + 1, jb), 1,
                 &TAUU(ic, ib));
        for (int i = 1; i <= nb - 1; ++i) {
          U3(ic + i, ic, ib) = E(A, lda, jb + r + i, jb);
          E(A, lda, jb + r + i, jb) = 0.0;
        }
        U3(ic, ic, ib) = 1.0;
        o_dlarf('L', nb, nb, &U3(ic, ic, ib), 1, TAUU(ic, ib),
                &E(B, ldb, jb + r, jb + r), ldb);
        for (int j = 1; j <= nb; ++j)
          for (int i = 1; i <= nb; ++i)
            E(Bc, r, i, j) = E(B, ldb, jb + r + i - 1, jb + r + j - 1);
        o_dgerq2(nb, nb, Bc, r, tau);
        o_dorgr2(nb, nb, nb, Bc, r, tau);
        o_dlarfg(nb, &E(Bc, r, 1, 1), &E(Bc, r, 1, 2), r, &TAUV(ic, ib));
        for (int i = 2; i <= nb; ++i) V3(ic + i - 1, ic, ib) = E(Bc, r, 1, i);
        V3(ic, ic, ib) = 1.0;
        o_dlarf('R', MIN(n, jb + 2 * r + nb - 1) - jb - r + 1, nb,
                &V3(ic, ic, ib), 1, TAUV(ic, ib), &E(A, lda, jb + r, jb + r),
                lda);
        o_dlarf('R', nb, nb, &V3(ic, ic, ib), 1, TAUV(ic, ib),
                &E(B, ldb, jb + r, jb + r), ldb);
        for (int i = 1; i <= nb - 1; ++i) E(B, ldb, jb + r + i, jb + r) = 0.0;
        o_dlarf('R', jb + r - istart, nb, &V3(ic, ic, ib), 1, TAUV(ic, ib),
                &E(A, lda, istart, jb + r), lda);
        o_dlarf('R', jb + r - istart, nb, &V3(ic, ic, ib), 1, TAUV(ic, ib),
                &E(B, ldb, istart, jb + r), ldb);
      }
    }
    /* rht2ht.f90:1599-1610 */
    for (int ib = nblocks; ib >= 1; --ib) {
      int jb = jblock + (ib - 1) * r;
      int nb = MIN(n - jb, r + nq);
      o_dlarft('F', 'C', nb, nq, UB(ib), ldu, &TAUU(1, ib), TUB(ib), nq);
      o_dlarft('F', 'C', nb, nq, VB(ib), ldu, &TAUV(1, ib), TVB(ib), nq);
    }

    /* lookahead on B, rht2ht.f90:1644-1684 */
    for (int ib = nblocks; ib >= 1; --ib) {
      int jb = jblock + (ib - 1) * r;
      for (int ic = 1; ic <= nq; ++ic) {
        int istart = MAX(jblock + 1, jblock + 1 + (ib - 1) * r - nq * r - (nq - ic) * r);
        int istart2 = MAX(jblock + 1, jblock + 1 + (ib - 1) * r - (nq - ic) * r);
        if (istart2 - 1 >= istart && jb + ic <= n)
          o_dlarf('R', istart2 - istart, MIN(r, n - (jb + ic) + 1),
                  &V3(ic, ic, ib), 1, TAUV(ic, ib), &E(B, ldb, istart, jb + ic),
                  ldb);
      }
    }
    for (int ib = nblocks; ib >= 1; --ib) {
      int jb = jblock + (ib - 1) * r;
      int nb = MIN(n - jb, r + nq);
      int istop = MIN(n, jblock + (ib - 1) * r + r + nq - 1);
      int istop2 = MIN(n, jblock + (ib - 1) * r + r + nq + r * nq - 1);
      if (istop2 >= istop + 1)
        o_dlarfb('L', 'T', 'F', 'C', nb, istop2 - istop, MIN(nq, nb), UB(ib),
                 ldu, TUB(ib), nq, &E(B, ldb, jb + 1, istop + 1), ldb);
    }
    /* lookahead on A, rht2ht.f90:1710-1765 */
    for (int ib = nblocks; ib >= 1; --ib) {
      int jb = jblock + (ib - 1) * r;
      for (int ic = 1; ic <= nq; ++ic) {
        int istart = MAX(jblock + 1, jblock + 1 + (ib - 1) * r - nq * r - (nq - ic) * r);
        int istart2 = MAX(jblock + 1, jblock + 1 + (ib - 1) * r - (nq - ic) * r);
        if (istart2 - 1 >= istart && jb + ic <= n)
          o_dlarf('R', istart2 - istart, MIN(r, n - (jb + ic) + 1),
                  &V3(ic, ic, ib), 1, TAUV(ic, ib), &E(A, lda, istart, jb + ic),
                  lda);
      }
    }
    for (int ib = nblocks; ib >= 1; --ib) {
      int jb, nb, row0;
      if (ib == 1) {
        jb = jblock;
        nb = MIN(n - jb, r + nq);
        row0 = jb + 1;
      } else {
        jb = jblock + 1 + (ib - 2) * r;
        nb = MIN(n - jb - r + 1, r + nq);
        row0 = jb + r;
      }
      int istop = MIN(n, jb + nq - 1);
      int istop2 = MIN(n, jblock + (ib - 1) * r + r + nq + r * nq - 1);
      if (istop2 >= istop + 1)
        o_dlarfb('L', 'T', 'F', 'C', nb, istop2 - istop, MIN(nq, nb), UB(ib),
                 ldu, TUB(ib), nq, &E(A, lda, row0, istop + 1), lda);
    }

    /* far updates of B, rht2ht.f90:1818-1866 */
    for (int ib = nblocks; ib >= 1; --ib) {
      int jb = jblock + (ib - 1) * r;
      int nb = MIN(n - jb, r + nq);
      int istart = MAX(jblock + 1, jblock + 1 + (ib - 1) * r - (2 * nq - 1) * r);
      for (int ic = 2; ic <= nq; ++ic) {
        int istart2 = MAX(jblock + 1, jblock + 1 + (ib - 1) * r - nq * r - (nq - ic) * r);
        if (istart2 - 1 >= istart && jb + ic <= n)
          o_dlarf('R', istart2 - istart, MIN(r, n - (jb + ic) + 1),
                  &V3(ic, ic, ib), 1, TAUV(ic, ib), &E(B, ldb, istart, jb + ic),
                  ldb);
      }
      o_dlarfb('R', 'N', 'F', 'C', istart - 1, nb, MIN(nq, nb), VB(ib), ldu,
               TVB(ib), nq, &E(B, ldb, 1, jb + 1), ldb);
    }
    for (int ib = nblocks; ib >= 1; --ib) {
      int jb = jblock + (ib - 1) * r;
      int nb = MIN(n - jb, r + nq);
      int istop = MIN(n, jblock + (ib - 1) * r + r + nq + r * nq - 1);
      if (n >= istop + 1)
        o_dlarfb('L', 'T', 'F', 'C', nb, n - istop, MIN(nq, nb), UB(ib), ldu,
                 TUB(ib), nq, &E(B, ldb, jb + 1, istop + 1), ldb);
    }
    /* far updates of A, rht2ht.f90:1894-1951 */
    for (int ib = nblocks; ib >= 1; --ib) {
      int jb = jblock + (ib - 1) * r;
      int nb = MIN(n - jb, r + nq);
      int istart = MAX(jblock + 1, jblock + 1 + (ib - 1) * r - (2 * nq - 1) * r);
      for (int ic = 2; ic <= nq; ++ic) {
        int istart2 = MAX(jblock + 1, jblock + 1 + (ib - 1) * r - (2 * nq - ic) * r);
        if (istart2 - 1 >= istart && jb + ic <= n)
          o_dlarf('R', istart2 - istart, MIN(r, n - (jb + ic) + 1),
                  &V3(ic, ic, ib), 1, TAUV(ic, ib), &E(A, lda, istart, jb + ic),
                  lda);
      }
      o_dlarfb('R', 'N', 'F', 'C', istart - 1, nb, MIN(nq, nb), VB(ib), ldu,
               TVB(ib), nq, &E(A, lda, 1, jb + 1), lda);
    }
    for (int ib = nblocks; ib >= 1; --ib) {
      int istop = MIN(n, jblock + (ib - 1) * r + r + nq + r * nq - 1);
      if (n < istop + 1) continue;
      int jb, nb, row0;
      if (ib == 1) {
        jb = jblock;
        nb = MIN(n - jb, r + nq);
        row0 = jb + 1;
      } else {
        jb = jblock + 1 + (ib - 2) * r;
        nb = MIN(n - jb - r + 1, r + nq);
        row0 = jb + r;
      }
      o_dlarfb('L', 'T', 'F', 'C', nb, n - istop, MIN(nq, nb), UB(ib), ldu,
               TUB(ib), nq, &E(A, lda, row0, istop + 1), lda);
    }
    /* Q and Z, rht2ht.f90:1976-2018 */
    for (int ib = nblocks; ib >= 1; --ib) {
      int jb = jblock + (ib - 1) * r;
      int nb = MIN(n - jb, r + nq - 1);
      o_dlarfb('R', 'N', 'F', 'C', n, nb, MIN(nq, nb), UB(ib), ldu, TUB(ib),
               nq, &E(Q, ldq, 1, jb + 1), ldq);
    }
    for (int ib = nblocks; ib >= 1; --ib) {
      int jb = jblock + (ib - 1) * r;
      int nb = MIN(n - jb, r + nq);
      o_dlarfb('R', 'N', 'F', 'C', n, nb, MIN(nq, nb), VB(ib), ldu, TVB(ib),
               nq, &E(Z, ldz, 1, jb + 1), ldz);
    }
  }
  free(U);
  free(V);
  free(TU);
  free(TV);
  free(tauU);
  free(tauV);
  free(Bc);
  free(tau);
  return 0;
#undef U3
#undef V3
#undef TAUU
#undef TAUV
#undef UB
#undef VB
#undef TUB
#undef TVB
}

/* ------------------------------------------------------------------------ */
/* rht2ht.f90:696-1344 — the same sweeps with Givens rotations; rotations of
 * bulge block ib are kept as CU,SU,CV,SV(r+nq,nq,ib) and accumulated into
 * dense (r+nq)x(r+nq) U(:,:,ib), V(:,:,ib) for the far updates (mb = 1). */
int oracle_rht2ht_givens(int n, double* A, int lda, double* B, int ldb,
                         double* Q, int ldq, double* Z, int ldz, int r,
                         int nq) {
  int info = check_args(n, lda, ldb, ldq, ldz, r, nq);
  if (info) return info;
  const int ld = r + nq;
  int nbl_all = (n - 1) / r + (((n - 1) % r) != 0);
  if (nbl_all < 1) nbl_all = 1;
  size_t rsz = (size_t)ld * nq * nbl_all, msz = (size_t)ld * ld * nbl_all;
  double* CU = (double*)malloc(rsz * sizeof(double));
  double* SU = (double*)malloc(rsz * sizeof(double));
  double* CV = (double*)malloc(rsz * sizeof(double));
  double* SV = (double*)malloc(rsz * sizeof(double));
  double* U = (double*)malloc(msz * sizeof(double));
  double* V = (double*)malloc(msz * sizeof(double));
#define R3(a, i, j, b) a[((size_t)(b)-1) * ld * nq + ((size_t)(j)-1) * ld + ((i)-1)]
#define MB(a, b) (a + ((size_t)(b)-1) * ld * ld)
  double temp;

  for (int jblock = 1; jblock <= n - 2; jblock += nq) {
    int nblocks = (n - jblock) / r + (((n - jblock) % r) != 0);
    for (size_t i = 0; i < rsz; ++i) {
      CU[i] = 1.0;
      CV[i] = 1.0;
      SU[i] = 0.0;
      SV[i] = 0.0;
    }
    /* generation, rht2ht.f90:787-957 */
    for (int ic = 1; ic <= nq; ++ic) {
      int jc = jblock + ic - 1;
      int nb = MIN(r, n - jc);
      int istart = jblock + 1;
      if (jc + r <= n)
        for (int i = 1; i <= ic - 1; ++i)
          for (int ig = r - 1; ig >= 1; --ig)
            o_drot(1, &E(B, ldb, jblock + i - 1 + ig, jc + r), ldb,
                   &E(B, ldb, jblock + i + ig, jc + r), ldb,
                   R3(CU, i + ig - 1, i, 1), R3(SU, i + ig - 1, i, 1));
      for (int i = 1; i <= ic - 1; ++i)
        if (jc <= n)
          for (int ig = MIN(r - 1, n - (jblock + i)); ig >= 1; --ig)
            o_drot(1, &E(A, lda, jblock + i - 1 + ig, jc), lda,
                   &E(A, lda, jblock + i + ig, jc), lda,
                   R3(CU, i + ig - 1, i, 1), R3(SU, i + ig - 1, i, 1));
      if (nb < 2) continue;
      for (int ig = nb - 1; ig >= 1; --ig) {
        o_dlartg(E(A, lda, jc + ig, jc), E(A, lda, jc + ig + 1, jc),
                 &R3(CU, ic + ig - 1, ic, 1), &R3(SU, ic + ig - 1, ic, 1),
                 &temp);
        E(A, lda, jc + ig, jc) = temp;
        E(A, lda, jc + ig + 1, jc) = 0.0;
      }
      for (int ig = nb - 1; ig >= 1; --ig)
        o_drot(nb, &E(B, ldb, jc + ig, jc + 1), ldb,
               &E(B, ldb, jc + ig + 1, jc + 1), ldb,
               R3(CU, ic + ig - 1, ic, 1), R3(SU, ic + ig - 1, ic, 1));
      for (int ig = nb - 1; ig >= 1; --ig) {
        double* c = &R3(CV, ic + ig - 1, ic, 1);
        double* s = &R3(SV, ic + ig - 1, ic, 1);
        o_dlartg(E(B, ldb, jc + ig + 1, jc + ig + 1),
                 E(B, ldb, jc + ig + 1, jc + ig), c, s, &temp);
        E(B, ldb, jc + ig + 1, jc + ig + 1) = temp;
        E(B, ldb, jc + ig + 1, jc + ig) = 0.0;
        o_drot(ig, &E(B, ldb, jc + 1, jc + ig + 1), 1,
               &E(B, ldb, jc + 1, jc + ig), 1, *c, *s);
      }
      for (int ig = nb - 1; ig >= 1; --ig)
        o_drot(MIN(n, jc + r + nb) - jc, &E(A, lda, jc + 1, jc + ig + 1), 1,
               &E(A, lda, jc + 1, jc + ig), 1, R3(CV, ic + ig - 1, ic, 1),
               R3(SV, ic + ig - 1, ic, 1));
      for (int ig = nb - 1; ig >= 1; --ig)
        o_drot(jc - istart + 1, &E(B, ldb, istart, jc + ig + 1), 1,
               &E(B, ldb, istart, jc + ig), 1, R3(CV, ic + ig - 1, ic, 1),
               R3(SV, ic + ig - 1, ic, 1));
      for (int ig = nb - 1; ig >= 1; --ig)
        o_drot(jc - istart + 1, &E(A, lda, istart, jc + ig + 1), 1,
               &E(A, lda, istart, jc + ig), 1, R3(CV, ic + ig - 1, ic, 1),
               R3(SV, ic + ig - 1, ic, 1));

      for (int jb = jc + 1; jb <= n; jb += r) {
        int ib = (jb - (jc + 1)) / r + 2;
        nb = MIN(r, n - jb - r + 1);
        istart = MAX(jblock + 1, jblock + 1 + (ib - 1) * r - (nq - ic) * r);
        if (jb + r + r - 1 <= n)
          for (int i = 1; i <= ic - 1; ++i)
            for (int ig = r - 1; ig >= 1; --ig)
              o_drot(1, &E(B, ldb, jb + r - ic + i + ig - 1, jb + r + r - 1),
                     ldb, &E(B, ldb, jb + r - ic + i + ig, jb + r + r - 1),
                     ldb, R3(CU, i + ig - 1, i, ib), R3(SU, i + ig - 1, i, ib));
        for (int i = 1; i <= ic - 1; ++i)
          if (jb + r - ic + i <= n)
            for (int ig = MIN(r, n - (jb + r - ic + i) + 1) - 1; ig >= 1; --ig)
              o_drot(1, &E(A, lda, jb + r - ic + i + ig - 1, jb), lda,
                     &E(A, lda, jb + r - ic + i + ig, jb), lda,
                     R3(CU, i + ig - 1, i, ib), R3(SU, i + ig - 1, i, ib));
        if (nb < 2) continue;
        for (int ig = nb - 1; ig >= 1; --ig) {
          o_dlartg(E(A, lda, jb + r - 1 + ig, jb), E(A, lda, jb + r + ig, jb),
                   &R3(CU, ic + ig - 1, ic, ib), &R3(SU, ic + ig - 1, ic, ib),
                   &temp);
          E(A, lda, jb + ig + r - 1, jb) = temp;
          E(A, lda, jb + ig + r, jb) = 0.0;
        }
        for (int ig = nb - 1; ig >= 1; --ig)
          o_drot(nb, &E(B, ldb, jb + r - 1 + ig, jb + r), ldb,
                 &E(B, ldb, jb + ig + r, jb + r), ldb,
                 R3(CU, ic + ig - 1, ic, ib), R3(SU, ic + ig - 1, ic, ib));
        for (int ig = nb - 1; ig >= 1; --ig) {
          double* c = &R3(CV, ic + ig - 1, ic, ib);
          double* s = &R3(SV, ic + ig - 1, ic, ib);
          o_dlartg(E(B, ldb, jb + r + ig, jb + r + ig),
                   E(B, ldb, jb + r + ig, jb + r + ig - 1), c, s, &temp);
          E(B, ldb, jb + r + ig, jb + r + ig) = temp;
          E(B, ldb, jb + r + ig, jb + r + ig - 1) = 0.0;
          o_drot(ig, &E(B, ldb, jb + r, jb + r + ig), 1,
                 &E(B, ldb, jb + r, jb + r + ig - 1), 1, *c, *s);
        }
        for (int ig = nb - 1; ig >= 1; --ig)
          o_drot(MIN(n, jb + 2 * r + nb - 1) - (jb + r) + 1,
                 &E(A, lda, jb + r, jb + r + ig), 1,
                 &E(A, lda, jb + r, jb + r + ig - 1), 1,
                 R3(CV, ic + ig - 1, ic, ib), R3(SV, ic + ig - 1, ic, ib));
        for (int ig = nb - 1; ig >= 1; --ig)
          o_drot(jb + r - istart, &E(B, ldb, istart, jb + r + ig), 1,
                 &E(B, ldb, istart, jb + r + ig - 1), 1,
                 R3(CV, ic + ig - 1, ic, ib), R3(SV, ic + ig - 1, ic, ib));
        for (int ig = nb - 1; ig >= 1; --ig)
          o_drot(jb + r - istart, &E(A, lda, istart, jb + r + ig), 1,
                 &E(A, lda, istart, jb + r + ig - 1), 1,
                 R3(CV, ic + ig - 1, ic, ib), R3(SV, ic + ig - 1, ic, ib));
      }
    }
    /* accumulate, rht2ht.f90:975-1002 */
    for (int ib = 1; ib <= nblocks; ++ib) {
      double* Ub = MB(U, ib);
      double* Vb = MB(V, ib);
      memset(Ub, 0, sizeof(double) * ld * ld);
      memset(Vb, 0, sizeof(double) * ld * ld);
      for (int i = 1; i <= ld; ++i) {
        E(Ub, ld, i, i) = 1.0;
        E(Vb, ld, i, i) = 1.0;
      }
      for (int ic = 1; ic <= nq; ++ic)
        for (int ig = r - 1; ig >= 1; --ig) {
          o_drot(ld, &E(Ub, ld, 1, ic - 1 + ig), 1, &E(Ub, ld, 1, ic + ig), 1,
                 R3(CU, ic + ig - 1, ic, ib), R3(SU, ic + ig - 1, ic, ib));
          o_drot(ld, &E(Vb, ld, 1, ic + ig), 1, &E(Vb, ld, 1, ic + ig - 1), 1,
                 R3(CV, ic + ig - 1, ic, ib), R3(SV, ic + ig - 1, ic, ib));
        }
    }
    /* lookahead on B, rht2ht.f90:1019-1049 */
    for (int ib = nblocks; ib >= 1; --ib) {
      int jb = jblock + (ib - 1) * r;
      for (int ic = 1; ic <= nq; ++ic) {
        int istart = MAX(jblock + 1, jblock + 1 + (ib - 1) * r - nq * r - (nq - ic) * r);
        int istart2 = MAX(jblock + 1, jblock + 1 + (ib - 1) * r - (nq - ic) * r);
        if (istart2 > istart)
          for (int ig = r - 1; ig >= 1; --ig) {
            if (jb + ic + ig > n) continue;
            o_drot(istart2 - istart, &E(B, ldb, istart, jb + ic + ig), 1,
                   &E(B, ldb, istart, jb + ic + ig - 1), 1,
                   R3(CV, ic + ig - 1, ic, ib), R3(SV, ic + ig - 1, ic, ib));
          }
      }
    }
    for (int ib = nblocks; ib >= 1; --ib) {
      int jb = jblock + (ib - 1) * r;
      int nb = MIN(n - jb, r + nq);
      int istop = MIN(n, jblock + (ib - 1) * r + r + nq - 1);
      int istop2 = MIN(n, jblock + (ib - 1) * r + r + nq + r * nq - 1);
      if (istop2 > istop)
        gemm_left_t(nb, istop2 - istop, &E(B, ldb, jb + 1, istop + 1), ldb,
                    MB(U, ib), ld);
    }
    /* lookahead on A, rht2ht.f90:1072-1115 */
    for (int ib = nblocks; ib >= 1; --ib) {
      int jb = jblock + (ib - 1) * r;
      for (int ic = 1; ic <= nq; ++ic) {
        int istart = MAX(jblock + 1, jblock + 1 + (ib - 1) * r - nq * r - (nq - ic) * r);
        int istart2 = MAX(jblock + 1, jblock + 1 + (ib - 1) * r - (nq - ic) * r);
        if (istart2 > istart)
          for (int ig = MIN(r - 1, n - jb - ic); ig >= 1; --ig)
            o_drot(istart2 - istart, &E(A, lda, istart, jb + ic + ig), 1,
                   &E(A, lda, istart, jb + ic + ig - 1), 1,
                   R3(CV, ic + ig - 1, ic, ib), R3(SV, ic + ig - 1, ic, ib));
      }
    }
    for (int ib = nblocks; ib >= 1; --ib) {
      int jb, nb, row0;
      if (ib == 1) {
        jb = jblock;
        nb = MIN(n - jb, r + nq);
        row0 = jb + 1;
      } else {
        jb = jblock + 1 + (ib - 2) * r;
        nb = MIN(n - jb - r + 1, r + nq);
        row0 = jb + r;
      }
      int istop = MIN(n, jb + nq - 1);
      int istop2 = MIN(n, jblock + (ib - 1) * r + r + nq + r * nq - 1);
      if (istop2 > istop)
        gemm_left_t(nb, istop2 - istop, &E(A, lda, row0, istop + 1), lda,
                    MB(U, ib), ld);
    }
    /* far updates of B, rht2ht.f90:1195-1229 */
    for (int ib = nblocks; ib >= 1; --ib) {
      int jb = jblock + (ib - 1) * r;
      int nb = MIN(n - jb, r + nq);
      int istart = MAX(jblock + 1, jblock + 1 + (ib - 1) * r - (2 * nq - 1) * r);
      for (int ic = 1; ic <= nq; ++ic) {
        int istart2 = MAX(jblock + 1, jblock + 1 + (ib - 1) * r - nq * r - (nq - ic) * r);
        if (istart2 > istart)
          for (int ig = r - 1; ig >= 1; --ig) {
            if (jb + ic + ig > n) continue;
            o_drot(istart2 - istart, &E(B, ldb, istart, jb + ic + ig), 1,
                   &E(B, ldb, istart, jb + ic + ig - 1), 1,
                   R3(CV, ic + ig - 1, ic, ib), R3(SV, ic + ig - 1, ic, ib));
          }
      }
      gemm_right(istart - 1, nb, &E(B, ldb, 1, jb + 1), ldb, MB(V, ib), ld);
    }
    for (int ib = nblocks; ib >= 1; --ib) {
      int jb = jblock + (ib - 1) * r;
      int nb = MIN(n - jb, r + nq);
      int istop = MIN(n, jblock + (ib - 1) * r + r + nq + r * nq - 1);
      if (n > istop)
        gemm_left_t(nb, n - istop, &E(B, ldb, jb + 1, istop + 1), ldb,
                    MB(U, ib), ld);
    }
    /* far updates of A, rht2ht.f90:1250-1291 */
    for (int ib = nblocks; ib >= 1; --ib) {
      int jb = jblock + (ib - 1) * r;
      int nb = MIN(n - jb, r + nq);
      int istart = MAX(jblock + 1, jblock + 1 + (ib - 1) * r - (2 * nq - 1) * r);
      for (int ic = 1; ic <= nq; ++ic) {
        int istart2 = MAX(jblock + 1, jblock + 1 + (ib - 1) * r - (2 * nq - ic) * r);
        for (int ig = MIN(r - 1, n - jb - ic); ig >= 1; --ig)
          o_drot(istart2 - istart, &E(A, lda, istart, jb + ic + ig), 1,
                 &E(A, lda, istart, jb + ic + ig - 1), 1,
                 R3(CV, ic + ig - 1, ic, ib), R3(SV, ic + ig - 1, ic, ib));
      }
      gemm_right(istart - 1, nb, &E(A, lda, 1, jb + 1), lda, MB(V, ib), ld);
    }
    for (int ib = nblocks; ib >= 1; --ib) {
      int istop = MIN(n, jblock + (ib - 1) * r + r + nq + r * nq - 1);
      if (n <= istop) continue;
      int jb, nb, row0;
      if (ib == 1) {
        jb = jblock;
        nb = MIN(n - jb, r + nq);
        row0 = jb + 1;
      } else {
        jb = jblock + 1 + (ib - 2) * r;
        nb = MIN(n - jb - r + 1, r + nq);
        row0 = jb + r;
      }
      gemm_left_t(nb, n - istop, &E(A, lda, row0, istop + 1), lda, MB(U, ib),
                  ld);
    }
    /* Q and Z, rht2ht.f90:1304-1325 */
    for (int ib = nblocks; ib >= 1; --ib) {
      int jb = jblock + (ib - 1) * r;
      int nb = MIN(n - jb, r + nq);
      gemm_right(n, nb, &E(Q, ldq, 1, jb + 1), ldq, MB(U, ib), ld);
    }
    for (int ib = nblocks; ib >= 1; --ib) {
      int jb = jblock + (ib - 1) * r;
      int nb = MIN(n - jb, r + nq);
      gemm_right(n, nb, &E(Z, ldz, 1, jb + 1), ldz, MB(V, ib), ld);
    }
  }
  free(CU);
  free(SU);
  free(CV);
  free(SV);
  free(U);
  free(V);
  return 0;
#undef R3
#undef MB
}
