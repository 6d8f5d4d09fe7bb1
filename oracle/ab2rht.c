/*
 * Stage 1 of the reference, restated for the CPU oracle (test infrastructure
 * only): ab2rht_simple, /src/ab2rht.f90:8-495.
 *
 * The reference runs its steps as OpenMP tasks; with OpenMP tasks executed in
 * program order (a valid topological order of the task graph) the arithmetic
 * is the same, so this restatement follows program order.  Column/row chunking
 * of the reference's taskloops (blocksize) only splits independent columns or
 * rows and is dropped.
 */
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define E(a, ld, i, j) (a)[((size_t)(j)-1) * (size_t)(ld) + ((size_t)(i)-1)]
#define MIN(a, b) ((a) < (b) ? (a) : (b))
#define MAX(a, b) ((a) > (b) ? (a) : (b))

/* ab2rht.f90:8 — reduce (A,B), B upper triangular, to r-Hessenberg-triangular
 * form with nb = r; Q <- Q*Q1, Z <- Z*Z1. */
int oracle_ab2rht_simple(int n, double* A, int lda, double* B, int ldb,
                         double* Q, int ldq, double* Z, int ldz, int nb,
                         int p) {
  if (n < 0) return -1;
  if (lda < MAX(1, n)) return -3;
  if (ldb < MAX(1, n)) return -5;
  if (ldq < MAX(1, n)) return -7;
  if (ldz < MAX(1, n)) return -9;
  if (nb < 1) return -10;
  if (p < 2) return -11;
  const int P = p * nb;
  /* ab2rht.f90:34-35 */
  int nbmax = (n - nb) / ((p - 1) * nb);
  if ((n - nb) % ((p - 1) * nb) != 0) nbmax += 1;
  if (nbmax < 1) nbmax = 1;
  double* VL = (double*)calloc((size_t)P * nb * nbmax, sizeof(double));
  double* VR = (double*)calloc((size_t)nb * P * nbmax, sizeof(double));
  double* TL = (double*)calloc((size_t)nb * nb * nbmax, sizeof(double));
  double* TR = (double*)calloc((size_t)nb * nb * nbmax, sizeof(double));
  double* Bc = (double*)calloc((size_t)P * P, sizeof(double));
  double* tau = (double*)calloc((size_t)P + nb, sizeof(double));
  double* tau2 = (double*)calloc((size_t)P + nb, sizeof(double));
#define VLb(bl) (VL + (size_t)((bl)-1) * P * nb)
#define VRb(bl) (VR + (size_t)((bl)-1) * nb * P)
#define TLb(bl) (TL + (size_t)((bl)-1) * nb * nb)
#define TRb(bl) (TR + (size_t)((bl)-1) * nb * nb)

  /* ab2rht.f90:71 panel loop */
  for (int jb = 1; jb <= n - nb - 1; jb += nb) {
    /* ab2rht.f90:74-75 */
    int nblocks = (n - nb - jb + 1) / ((p - 1) * nb);
    if ((n - nb - jb + 1) % ((p - 1) * nb) != 0) nblocks += 1;
    const int ibtop = jb + nb, ibbot = jb + nb + (nblocks - 1) * (p - 1) * nb;

    /* left generation, ab2rht.f90:93-118 (blocks bottom-up) */
    int bl = 0;
    for (int ib = ibbot; ib >= ibtop; ib -= nb * (p - 1)) {
      ++bl;
      int ib2 = MIN(n, ib + P - 1), nb2 = ib2 - ib + 1, k = MIN(nb, nb2);
      o_dgeqr2(nb2, nb, &E(A, lda, ib, jb), lda, tau);
      for (int j = 1; j <= nb; ++j)
        for (int i = 1; i <= nb2; ++i)
          E(VLb(bl), P, i, j) = E(A, lda, ib + i - 1, jb + j - 1);
      for (int j = jb; j <= jb + nb - 1; ++j)
        for (int i = ib + j - jb + 1; i <= ib2; ++i) E(A, lda, i, j) = 0.0;
      o_dlarft('F', 'C', nb2, k, VLb(bl), P, tau, TLb(bl), nb);
    }

    /* left application to B, ab2rht.f90:149-166 */
    bl = 0;
    for (int ib = ibbot; ib >= ibtop; ib -= nb * (p - 1)) {
      ++bl;
      int ib2 = MIN(n, ib + P - 1), nb2 = ib2 - ib + 1, k = MIN(nb, nb2);
      o_dlarfb('L', 'T', 'F', 'C', nb2, n - ib + 1, k, VLb(bl), P, TLb(bl), nb,
               &E(B, ldb, ib, ib), ldb);
    }
    /* left application to A, ab2rht.f90:189-207 */
    bl = 0;
    for (int ib = ibbot; ib >= ibtop; ib -= nb * (p - 1)) {
      ++bl;
      int ib2 = MIN(n, ib + P - 1), nb2 = ib2 - ib + 1, k = MIN(nb, nb2);
      int c1 = jb + nb;
      if (n >= c1)
        o_dlarfb('L', 'T', 'F', 'C', nb2, n - c1 + 1, k, VLb(bl), P, TLb(bl),
                 nb, &E(A, lda, ib, c1), lda);
    }
    /* application to Q, ab2rht.f90:230-245 */
    bl = 0;
    for (int ib = ibbot; ib >= ibtop; ib -= nb * (p - 1)) {
      ++bl;
      int ib2 = MIN(n, ib + P - 1), nb2 = ib2 - ib + 1, k = MIN(nb, nb2);
      o_dlarfb('R', 'N', 'F', 'C', n, nb2, k, VLb(bl), P, TLb(bl), nb,
               &E(Q, ldq, 1, ib), ldq);
    }

    /* right generation with lookahead and rest of B, ab2rht.f90:271-369 */
    bl = 0;
    for (int ib = ibbot; ib >= ibtop; ib -= nb * (p - 1)) {
      ++bl;
      int ib2 = MIN(n, ib + P - 1), nb2 = ib2 - ib + 1;
      /* ab2rht.f90:290-311 */
      for (int j = 1; j <= nb2; ++j)
        for (int i = 1; i <= nb2; ++i)
          E(Bc, P, i, j) = E(B, ldb, ib + i - 1, ib + j - 1);
      o_dgerq2(nb2, nb2, Bc, P, tau);
      if (nb2 > nb) {
        o_dorgr2(nb2, nb2, nb2, Bc, P, tau);
        o_dgelq2(nb, nb2, Bc, P, tau2);
        for (int j = 1; j <= nb2; ++j)
          for (int i = 1; i <= nb; ++i) E(VRb(bl), nb, i, j) = E(Bc, P, i, j);
        o_dlarft('F', 'R', nb2, nb, VRb(bl), nb, tau2, TRb(bl), nb);
      } else {
        o_dlarft('B', 'R', nb2, nb2, Bc, P, tau, TRb(bl), nb);
        for (int j = 1; j <= nb2; ++j)
          for (int i = 1; i <= nb2; ++i) E(VRb(bl), nb, i, j) = E(Bc, P, i, j);
      }
      /* lookahead, ab2rht.f90:324-337 */
      int c1 = MAX(1, ib - nb * (p - 1)), c2 = ib2;
      if (nb2 > nb) {
        o_dlarfb('R', 'N', 'F', 'R', c2 - c1 + 1, nb2, nb, VRb(bl), nb,
                 TRb(bl), nb, &E(B, ldb, c1, ib), ldb);
        for (int j = ib; j <= ib + nb - 1; ++j)
          for (int i = j + 1; i <= ib2; ++i) E(B, ldb, i, j) = 0.0;
      } else {
        o_dlarfb('R', 'N', 'B', 'R', c2 - c1 + 1, nb2, nb2, VRb(bl), nb,
                 TRb(bl), nb, &E(B, ldb, c1, ib), ldb);
      }
      /* rest of B, ab2rht.f90:354-366 */
      int rlast = MAX(1, ib - nb * (p - 1)) - 1;
      if (rlast >= 1) {
        if (nb2 > nb)
          o_dlarfb('R', 'N', 'F', 'R', rlast, nb2, nb, VRb(bl), nb, TRb(bl),
                   nb, &E(B, ldb, 1, ib), ldb);
        else
          o_dlarfb('R', 'N', 'B', 'R', rlast, nb2, nb2, VRb(bl), nb, TRb(bl),
                   nb, &E(B, ldb, 1, ib), ldb);
      }
    }

    /* right application to A and Z, ab2rht.f90:400-461 */
    bl = 0;
    for (int ib = ibbot; ib >= ibtop; ib -= nb * (p - 1)) {
      ++bl;
      int ib2 = MIN(n, ib + P - 1), nb2 = ib2 - ib + 1;
      if (nb2 > nb)
        o_dlarfb('R', 'N', 'F', 'R', n, nb2, nb, VRb(bl), nb, TRb(bl), nb,
                 &E(A, lda, 1, ib), lda);
      else
        o_dlarfb('R', 'N', 'B', 'R', n, nb2, nb2, VRb(bl), nb, TRb(bl), nb,
                 &E(A, lda, 1, ib), lda);
    }
    bl = 0;
    for (int ib = ibbot; ib >= ibtop; ib -= nb * (p - 1)) {
      ++bl;
      int ib2 = MIN(n, ib + P - 1), nb2 = ib2 - ib + 1;
      if (nb2 > nb)
        o_dlarfb('R', 'N', 'F', 'R', n, nb2, nb, VRb(bl), nb, TRb(bl), nb,
                 &E(Z, ldz, 1, ib), ldz);
      else
        o_dlarfb('R', 'N', 'B', 'R', n, nb2, nb2, VRb(bl), nb, TRb(bl), nb,
                 &E(Z, ldz, 1, ib), ldz);
    }
  }
  free(VL);
  free(VR);
  free(TL);
  free(TR);
  free(Bc);
  free(tau);
  free(tau2);
  return 0;
}
