/*
 * CPU oracle for the two-stage Hessenberg-triangular (HT) reduction.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker, never the product: only the
 * test-suite, the smoke check and the CPU-baseline leg of the benchmark load
 * it.  The product path (libparaht_hip.so) never links or calls it.
 *
 * It restates, in plain C, the reference's algorithm (ParaHT, Fortran) and
 * the LAPACK kernels the reference calls (which are not part of the
 * reference).  Every routine cites the reference file:line it follows.
 * All matrices are column-major FP64 with leading dimensions, as in the
 * reference; indices in comments are 1-based like the reference.
 */
#ifndef PARAHT_ORACLE_H
#define PARAHT_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- LAPACK-equivalent kernels, written from the published mathematics ---- */
double o_dnrm2(int n, const double* x, int incx);
double o_dlapy2(double x, double y);
void o_dlarfg(int n, double* alpha, double* x, int incx, double* tau);
void o_dlartg(double f, double g, double* c, double* s, double* r);
void o_drot(int n, double* x, int incx, double* y, int incy, double c, double s);
void o_dlarf(char side, int m, int n, const double* v, int incv, double tau,
             double* C, int ldc);
void o_dgeqr2(int m, int n, double* A, int lda, double* tau);
void o_dgerq2(int m, int n, double* A, int lda, double* tau);
void o_dorgr2(int m, int n, int k, double* A, int lda, const double* tau);
void o_dgelq2(int m, int n, double* A, int lda, double* tau);
void o_dlarft(char direct, char storev, int n, int k, const double* V, int ldv,
              const double* tau, double* T, int ldt);
void o_dlarfb(char side, char trans, char direct, char storev, int m, int n,
              int k, const double* V, int ldv, const double* T, int ldt,
              double* C, int ldc);

/* ---- The reference's stages ---- */
/* ab2rht.f90:8-495 */
int oracle_ab2rht_simple(int n, double* A, int lda, double* B, int ldb,
                         double* Q, int ldq, double* Z, int ldz, int nb, int p);
/* rht2ht.f90:1346-2051 */
int oracle_rht2ht_householder(int n, double* A, int lda, double* B, int ldb,
                              double* Q, int ldq, double* Z, int ldz, int r,
                              int nq);
/* rht2ht.f90:696-1344 */
int oracle_rht2ht_givens(int n, double* A, int lda, double* B, int ldb,
                         double* Q, int ldq, double* Z, int ldz, int r, int nq);
/* rht2ht.f90:2053-2180 */
int oracle_rht2ht_householder_unblocked(int n, double* A, int lda, double* B,
                                        int ldb, double* Q, int ldq, double* Z,
                                        int ldz, int r);

/* Number of OpenMP threads the oracle's large updates use (1 without OpenMP). */
int oracle_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif
