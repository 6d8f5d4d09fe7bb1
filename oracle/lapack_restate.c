/*
 * LAPACK/BLAS-equivalent kernels for the CPU oracle (test infrastructure only).
 *
 * The reference (ParaHT) does every flop through LAPACK/BLAS (MKL), which is
 * not part of the reference.  These are restatements of the published
 * algorithms, not of any library's code:
 *   - Householder reflectors, QR/RQ/LQ and compact-WY block reflectors as in
 *     Golub & Van Loan, "Matrix Computations" (4th ed.), sections 5.1-5.2, and
 *     Schreiber & Van Loan (1989) for the triangular factor T;
 *   - the LAPACK argument conventions (DLARFG: beta = -sign(alpha)*||x||, DLARF,
 *     DGEQR2, DGERQ2, DORGR2, DGELQ2, DLARFT, DLARFB), from the LAPACK Users'
 *     Guide and the routines' documented interfaces;
 *   - Givens rotations with safe scaling after E. Anderson, "Algorithm 978:
 *     Safe scaling in the Level 1 BLAS", ACM TOMS 44 (2017).
 * Column-major; the macro E(a,ld,i,j) takes 1-based (i,j).
 */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define E(a, ld, i, j) (a)[((size_t)(j)-1) * (size_t)(ld) + ((size_t)(i)-1)]

int oracle_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

double o_dlapy2(double x, double y) {
  double xa = fabs(x), ya = fabs(y);
  double w = xa > ya ? xa : ya;
  double z = xa > ya ? ya : xa;
  if (z == 0.0 || w > DBL_MAX) return w;
  double q = z / w;
  return w * sqrt(1.0 + q * q);
}

/* Scaled 2-norm: max|x| * sqrt(sum (x/max|x|)^2). */
double o_dnrm2(int n, const double* x, int incx) {
  if (n < 1) return 0.0;
  double scale = 0.0;
  for (int i = 0; i < n; ++i) {
    double a = fabs(x[(size_t)i * incx]);
    if (a > scale) scale = a;
  }
  if (scale == 0.0 || scale > DBL_MAX) return scale;
  double ssq = 0.0;
  for (int i = 0; i < n; ++i) {
    double t = x[(size_t)i * incx] / scale;
    ssq += t * t;
  }
  return scale * sqrt(ssq);
}

/* Householder reflector H = I - tau*[1;v]*[1;v]^T with H*[alpha;x] = [beta;0],
 * beta = -sign(alpha)*||[alpha;x]||; x is overwritten by v.  tau = 0 when x = 0.
 * Tiny beta is rescaled by 1/safmin (up to 20 times) as in the DLARFG contract. */
void o_dlarfg(int n, double* alpha, double* x, int incx, double* tau) {
  if (n <= 1) {
    *tau = 0.0;
    return;
  }
  double xnorm = o_dnrm2(n - 1, x, incx);
  if (xnorm == 0.0) {
    *tau = 0.0;
    return;
  }
  double beta = -copysign(o_dlapy2(*alpha, xnorm), *alpha);
  const double safmin = DBL_MIN / (DBL_EPSILON * 0.5);
  int knt = 0;
  if (fabs(beta) < safmin) {
    const double rsafmn = 1.0 / safmin;
    do {
      ++knt;
      for (int i = 0; i < n - 1; ++i) x[(size_t)i * incx] *= rsafmn;
      beta *= rsafmn;
      *alpha *= rsafmn;
    } while (fabs(beta) < safmin && knt < 20);
    xnorm = o_dnrm2(n - 1, x, incx);
    beta = -copysign(o_dlapy2(*alpha, xnorm), *alpha);
  }
  *tau = (beta - *alpha) / beta;
  double scal = 1.0 / (*alpha - beta);
  for (int i = 0; i < n - 1; ++i) x[(size_t)i * incx] *= scal;
  for (int j = 0; j < knt; ++j) beta *= safmin;
  *alpha = beta;
}

/* Plane rotation [c s; -s c]*[f; g] = [r; 0] (Anderson 2017, Algorithm 978):
 * c >= 0, r carries the sign of f; f == 0 gives c = 0, s = sign(g), r = |g|. */
void o_dlartg(double f, double g, double* c, double* s, double* r) {
  const double safmin = DBL_MIN;
  const double safmax = 1.0 / safmin;
  const double rtmin = sqrt(safmin);
  const double rtmax = sqrt(safmax / 2.0);
  double f1 = fabs(f), g1 = fabs(g);
  if (g == 0.0) {
    *c = 1.0;
    *s = 0.0;
    *r = f;
  } else if (f == 0.0) {
    *c = 0.0;
    *s = copysign(1.0, g);
    *r = g1;
  } else if (f1 > rtmin && f1 < rtmax && g1 > rtmin && g1 < rtmax) {
    double d = sqrt(f * f + g * g);
    *c = f1 / d;
    *r = copysign(d, f);
    *s = g / *r;
  } else {
    double u = f1 > g1 ? f1 : g1;
    if (u < safmin) u = safmin;
    if (u > safmax) u = safmax;
    double fs = f / u, gs = g / u;
    double d = sqrt(fs * fs + gs * gs);
    *c = fabs(fs) / d;
    double rr = copysign(d, f);
    *s = gs / rr;
    *r = rr * u;
  }
}

/* x <- c*x + s*y ; y <- c*y - s*x */
void o_drot(int n, double* x, int incx, double* y, int incy, double c,
            double s) {
  for (int i = 0; i < n; ++i) {
    double xv = x[(size_t)i * incx], yv = y[(size_t)i * incy];
    x[(size_t)i * incx] = c * xv + s * yv;
    y[(size_t)i * incy] = c * yv - s * xv;
  }
}

/* C <- H*C (side 'L', C m x n, v length m) or C <- C*H (side 'R', v length n),
 * H = I - tau*v*v^T. */
void o_dlarf(char side, int m, int n, const double* v, int incv, double tau,
             double* C, int ldc) {
  if (tau == 0.0 || m <= 0 || n <= 0) return;
  if (side == 'L') {
    for (int j = 1; j <= n; ++j) {
      double w = 0.0;
      for (int i = 1; i <= m; ++i) w += E(C, ldc, i, j) * v[(size_t)(i - 1) * incv];
      w *= tau;
      for (int i = 1; i <= m; ++i) E(C, ldc, i, j) -= v[(size_t)(i - 1) * incv] * w;
    }
  } else {
    double* w = (double*)calloc((size_t)m, sizeof(double));
    for (int j = 1; j <= n; ++j) {
      double vj = v[(size_t)(j - 1) * incv];
      for (int i = 1; i <= m; ++i) w[i - 1] += E(C, ldc, i, j) * vj;
    }
    for (int j = 1; j <= n; ++j) {
      double vj = tau * v[(size_t)(j - 1) * incv];
      for (int i = 1; i <= m; ++i) E(C, ldc, i, j) -= w[i - 1] * vj;
    }
    free(w);
  }
}

/* Unblocked Householder QR, A = Q*R, reflectors below the diagonal. */
void o_dgeqr2(int m, int n, double* A, int lda, double* tau) {
  int k = m < n ? m : n;
  for (int i = 1; i <= k; ++i) {
    int ip = i + 1 < m ? i + 1 : m;
    o_dlarfg(m - i + 1, &E(A, lda, i, i), &E(A, lda, ip, i), 1, &tau[i - 1]);
    if (i < n) {
      double aii = E(A, lda, i, i);
      E(A, lda, i, i) = 1.0;
      o_dlarf('L', m - i + 1, n - i, &E(A, lda, i, i), 1, tau[i - 1],
              &E(A, lda, i, i + 1), lda);
      E(A, lda, i, i) = aii;
    }
  }
}

/* Unblocked Householder RQ, A = R*Q with Q = H(1)...H(k); reflector i has its
 * unit entry at column n-k+i and is stored in row m-k+i left of it. */
void o_dgerq2(int m, int n, double* A, int lda, double* tau) {
  int k = m < n ? m : n;
  for (int i = k; i >= 1; --i) {
    int row = m - k + i, col = n - k + i;
    o_dlarfg(col, &E(A, lda, row, col), &E(A, lda, row, 1), lda, &tau[i - 1]);
    double aii = E(A, lda, row, col);
    E(A, lda, row, col) = 1.0;
    o_dlarf('R', row - 1, col, &E(A, lda, row, 1), lda, tau[i - 1], A, lda);
    E(A, lda, row, col) = aii;
  }
}

/* Forms the m x n matrix Q with orthonormal rows, the last m rows of
 * H(1)...H(k) from o_dgerq2. */
void o_dorgr2(int m, int n, int k, double* A, int lda, const double* tau) {
  if (m <= 0) return;
  if (k < m) {
    for (int j = 1; j <= n; ++j) {
      for (int l = 1; l <= m - k; ++l) E(A, lda, l, j) = 0.0;
      if (j > n - m && j <= n - k) E(A, lda, m - n + j, j) = 1.0;
    }
  }
  for (int i = 1; i <= k; ++i) {
    int ii = m - k + i;
    int col = n - m + ii;
    E(A, lda, ii, col) = 1.0;
    o_dlarf('R', ii - 1, col, &E(A, lda, ii, 1), lda, tau[i - 1], A, lda);
    for (int l = 1; l <= col - 1; ++l) E(A, lda, ii, l) *= -tau[i - 1];
    E(A, lda, ii, col) = 1.0 - tau[i - 1];
    for (int l = col + 1; l <= n; ++l) E(A, lda, ii, l) = 0.0;
  }
}

/* Unblocked Householder LQ, A = L*Q, reflectors right of the diagonal. */
void o_dgelq2(int m, int n, double* A, int lda, double* tau) {
  int k = m < n ? m : n;
  for (int i = 1; i <= k; ++i) {
    int ip = i + 1 < n ? i + 1 : n;
    o_dlarfg(n - i + 1, &E(A, lda, i, i), &E(A, lda, i, ip), lda, &tau[i - 1]);
    if (i < m) {
      double aii = E(A, lda, i, i);
      E(A, lda, i, i) = 1.0;
      o_dlarf('R', m - i, n - i + 1, &E(A, lda, i, i), lda, tau[i - 1],
              &E(A, lda, i + 1, i), lda);
      E(A, lda, i, i) = aii;
    }
  }
}

/* Explicit reflector i (1-based) of a stored block, as a full vector of length
 * nn: column storage 'C' forward, row storage 'R' forward or backward. */
static double vfull(char direct, char storev, int nn, int k, const double* V,
                    int ldv, int l, int i) {
  if (direct == 'F') {
    if (l < i) return 0.0;
    if (l == i) return 1.0;
    return storev == 'C' ? E(V, ldv, l, i) : E(V, ldv, i, l);
  } else { /* backward: unit entry at nn-k+i */
    int u = nn - k + i;
    if (l > u) return 0.0;
    if (l == u) return 1.0;
    return storev == 'C' ? E(V, ldv, l, i) : E(V, ldv, i, l);
  }
}

/* Triangular factor T of a block reflector (Schreiber-Van Loan):
 * forward: H = H(1)...H(k) = I - V*T*V^T, T upper;
 * backward: H = H(k)...H(1) = I - V*T*V^T, T lower (V explicit as columns). */
void o_dlarft(char direct, char storev, int n, int k, const double* V, int ldv,
              const double* tau, double* T, int ldt) {
  if (n <= 0) return;
  if (direct == 'F') {
    for (int i = 1; i <= k; ++i) {
      if (tau[i - 1] == 0.0) {
        for (int j = 1; j <= i; ++j) E(T, ldt, j, i) = 0.0;
        continue;
      }
      for (int j = 1; j < i; ++j) {
        double s = 0.0;
        for (int l = i; l <= n; ++l)
          s += vfull(direct, storev, n, k, V, ldv, l, j) *
               vfull(direct, storev, n, k, V, ldv, l, i);
        E(T, ldt, j, i) = -tau[i - 1] * s;
      }
      /* T(1:i-1,i) = T(1:i-1,1:i-1) * T(1:i-1,i), T upper triangular */
      for (int j = 1; j < i; ++j) {
        double s = 0.0;
        for (int l = j; l < i; ++l) s += E(T, ldt, j, l) * E(T, ldt, l, i);
        E(T, ldt, j, i) = s;
      }
      E(T, ldt, i, i) = tau[i - 1];
    }
  } else {
    for (int i = k; i >= 1; --i) {
      if (tau[i - 1] == 0.0) {
        for (int j = i; j <= k; ++j) E(T, ldt, j, i) = 0.0;
        continue;
      }
      int u = n - k + i; /* reflector i is zero beyond u */
      for (int j = i + 1; j <= k; ++j) {
        double s = 0.0;
        for (int l = 1; l <= u; ++l)
          s += vfull(direct, storev, n, k, V, ldv, l, j) *
               vfull(direct, storev, n, k, V, ldv, l, i);
        E(T, ldt, j, i) = -tau[i - 1] * s;
      }
      /* T(i+1:k,i) = T(i+1:k,i+1:k) * T(i+1:k,i), T lower triangular */
      for (int j = k; j > i; --j) {
        double s = 0.0;
        for (int l = i + 1; l <= j; ++l) s += E(T, ldt, j, l) * E(T, ldt, l, i);
        E(T, ldt, j, i) = s;
      }
      E(T, ldt, i, i) = tau[i - 1];
    }
  }
}

/* Applies H = I - V*T*V^T (or H^T for trans 'T') from the left (C <- op(H)*C,
 * C m x n) or right (C <- C*op(H)).  V is given in LAPACK storage and made
 * explicit first (unit entries and structural zeros implied by direct/storev).
 * Large updates run their independent columns (left) or rows (right) in
 * OpenMP threads. */
void o_dlarfb(char side, char trans, char direct, char storev, int m, int n,
              int k, const double* V, int ldv, const double* T, int ldt,
              double* C, int ldc) {
  if (m <= 0 || n <= 0 || k <= 0) return;
  int nn = side == 'L' ? m : n;
  double* Vf = (double*)malloc(sizeof(double) * (size_t)nn * k);
  for (int i = 1; i <= k; ++i)
    for (int l = 1; l <= nn; ++l)
      Vf[(size_t)(i - 1) * nn + (l - 1)] =
          vfull(direct, storev, nn, k, V, ldv, l, i);
  /* op(T) explicit, k x k, zero outside its triangle */
  double* Tf = (double*)calloc((size_t)k * k, sizeof(double));
  for (int j = 1; j <= k; ++j)
    for (int i = 1; i <= k; ++i) {
      int inside = direct == 'F' ? (i <= j) : (i >= j);
      if (!inside) continue;
      double t = E(T, ldt, i, j);
      if (trans == 'N')
        Tf[(size_t)(j - 1) * k + (i - 1)] = t;
      else
        Tf[(size_t)(i - 1) * k + (j - 1)] = t;
    }
  if (side == 'L') {
    /* C <- C - Vf * op(T) * (Vf^T C), column by column */
#pragma omp parallel for schedule(static) if ((long)m * n * k > 2000000L)
    for (int j = 1; j <= n; ++j) {
      double w[1024];
      double* wp = k <= 512 ? w : (double*)malloc(sizeof(double) * 2 * k);
      double* w2 = wp + k;
      for (int i = 0; i < k; ++i) {
        double s = 0.0;
        const double* vi = Vf + (size_t)i * nn;
        for (int l = 1; l <= m; ++l) s += vi[l - 1] * E(C, ldc, l, j);
        wp[i] = s;
      }
      for (int i = 0; i < k; ++i) {
        double s = 0.0;
        for (int l = 0; l < k; ++l) s += Tf[(size_t)l * k + i] * wp[l];
        w2[i] = s;
      }
      for (int i = 0; i < k; ++i) {
        const double* vi = Vf + (size_t)i * nn;
        double wi = w2[i];
        for (int l = 1; l <= m; ++l) E(C, ldc, l, j) -= vi[l - 1] * wi;
      }
      if (wp != w) free(wp);
    }
  } else {
    /* C <- C - (C Vf) * op(T) * Vf^T, row block by row block */
    const int RB = 64;
    int nrb = (m + RB - 1) / RB;
#pragma omp parallel for schedule(static) if ((long)m * n * k > 2000000L)
    for (int b = 0; b < nrb; ++b) {
      int r0 = b * RB + 1, r1 = r0 + RB - 1 < m ? r0 + RB - 1 : m;
      int rows = r1 - r0 + 1;
      double* W = (double*)calloc((size_t)rows * k, sizeof(double));
      double* W2 = (double*)calloc((size_t)rows * k, sizeof(double));
      for (int i = 0; i < k; ++i) {
        const double* vi = Vf + (size_t)i * nn;
        double* wi = W + (size_t)i * rows;
        for (int l = 1; l <= n; ++l) {
          double v = vi[l - 1];
          if (v == 0.0) continue;
          const double* cl = &E(C, ldc, r0, l);
          for (int q = 0; q < rows; ++q) wi[q] += cl[q] * v;
        }
      }
      for (int i = 0; i < k; ++i) {
        double* w2i = W2 + (size_t)i * rows;
        for (int l = 0; l < k; ++l) {
          double t = Tf[(size_t)i * k + l];
          if (t == 0.0) continue;
          const double* wl = W + (size_t)l * rows;
          for (int q = 0; q < rows; ++q) w2i[q] += wl[q] * t;
        }
      }
      for (int i = 0; i < k; ++i) {
        const double* vi = Vf + (size_t)i * nn;
        const double* w2i = W2 + (size_t)i * rows;
        for (int l = 1; l <= n; ++l) {
          double v = vi[l - 1];
          if (v == 0.0) continue;
          double* cl = &E(C, ldc, r0, l);
          for (int q = 0; q < rows; ++q) cl[q] -= w2i[q] * v;
        }
      }
      free(W);
      free(W2);
    }
  }
  free(Vf);
  free(Tf);
}
