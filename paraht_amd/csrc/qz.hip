// QZ iteration on Hessenberg-triangular pencils: generalized eigenvalues, and
// optionally the generalized real Schur form with Q and Z accumulated.
//
// No counterpart in the reference: its driver stops at the HT form
// (src/profilereduction.f90:152-162).  The LAPACK routine that comes next is
// DHGEQZ; this is the Moler-Stewart implicit double-shift iteration (Moler &
// Stewart 1973; Golub & Van Loan, Matrix Computations 4th ed., Alg. 7.7.2,
// section 7.7.5 for infinite eigenvalues) with every transformation a 1-, 2- or
// 3-element Householder reflector.  tests/qz_emul.py states the same
// transformations in the same order in numpy.
//
// One workgroup per pencil.  Thread 0 makes every decision and generates
// every reflector; the workgroup applies it (one thread per column for a
// left reflector, per row for a right one) between barriers.  In a bulge
// sweep, thread 0 also applies the step's three reflectors to the corner the
// next step reads (rows k..k+3, columns k-1..k+2 of H; rows and columns
// k..k+2 of T), so each step needs one serial phase and two barriers.  The
// corner stays in LDS from step to step; its new column comes from the
// far-column update thread 0 takes itself and from entries no step has
// touched (prefetched a step ahead), and its stores go out after the first
// barrier, beside the other threads' work, so the serial phase waits on no
// global memory.
#include "device_util.hpp"
#include "kernels.hpp"

namespace pht {
namespace {

constexpr double kQzUlp = 2.220446049250313e-16;       // 2^-52
constexpr double kQzSafmin = 2.2250738585072014e-308;  // smallest normal

// P = I - tau v v^T on nv (1..3) entries
struct QzRefl {
  double v0, v1, v2, tau;
  int nv;
};

// v = (1, v1, v2), tau, beta with P x = beta e1 (x2 unused for nv = 2):
// beta = -sign(x0) ||x||, tau = (beta - x0) / beta, v = x / (x0 - beta).
// Without division or square root: x is scaled by the power of two that
// brings its largest magnitude into [0.5, 1) (exact), 1/||x|| comes from
// v_rsq_f64 and 1/(|x0| + ||x||) from v_rcp_f64, each refined by two Newton
// steps (relative error below 2^-52), and tau = 1 + |x0| / ||x||.
__device__ inline QzRefl qz_house(double x0, double x1, double x2, int nv, double* beta) {
  QzRefl r;
  r.nv = nv;
  r.v0 = 1.0;
  r.v1 = 0.0;
  r.v2 = 0.0;
  if (nv != 3) x2 = 0.0;
  if (x1 == 0.0 && x2 == 0.0) {
    r.tau = 0.0;
    if (beta) *beta = x0;
    return r;
  }
  const double mx = fmax(fabs(x0), fmax(fabs(x1), fabs(x2)));
  const int e = __builtin_amdgcn_frexp_exp(mx);
  const double s0 = ldexp(x0, -e), s1 = ldexp(x1, -e), s2 = ldexp(x2, -e);
  const double d2 = fma(s0, s0, fma(s1, s1, s2 * s2));  // in [0.25, 3]
  double y = __builtin_amdgcn_rsq(d2);
  double q = fma(-d2 * y, y, 1.0);
  y = fma(0.5 * y, q, y);
  q = fma(-d2 * y, y, 1.0);
  y = fma(0.5 * y, q, y);
  const double d = d2 * y;  // ||x|| scaled
  const double a0 = fabs(s0);
  const double den = a0 + d;  // |x0 - beta| scaled, in [0.5, 3.5]
  double w = __builtin_amdgcn_rcp(den);
  w = fma(w, fma(-den, w, 1.0), w);
  w = fma(w, fma(-den, w, 1.0), w);
  const double sc = x0 < 0.0 ? -w : w;  // 1 / (x0 - beta), scaled
  r.tau = fma(a0, y, 1.0);
  r.v1 = s1 * sc;
  r.v2 = s2 * sc;
  if (beta) *beta = ldexp(x0 < 0.0 ? d : -d, e);
  return r;
}

// the reflector with w P = (0, .., 0, beta) for a row w of nv entries
__device__ inline QzRefl qz_house_rev(double w0, double w1, double w2, int nv) {
  QzRefl r = nv == 3 ? qz_house(w2, w1, w0, 3, nullptr) : qz_house(w1, w0, 0.0, 2, nullptr);
  if (nv == 3) {
    const double t = r.v0;
    r.v0 = r.v2;
    r.v2 = t;
  } else {
    const double t = r.v0;
    r.v0 = r.v1;
    r.v1 = t;
  }
  return r;
}

__device__ inline QzRefl qz_flip(int i, int nv) {
  QzRefl r;
  r.nv = nv;
  r.v0 = i == 0 ? 1.0 : 0.0;
  r.v1 = i == 1 ? 1.0 : 0.0;
  r.v2 = 0.0;
  r.tau = 2.0;
  return r;
}

// x <- P x on nv entries x[0], x[s], x[2s]
__device__ inline void qz_apply(const QzRefl& R, double* x, size_t s) {
  const double x0 = x[0];
  const double x1 = R.nv > 1 ? x[s] : 0.0;
  const double x2 = R.nv > 2 ? x[2 * s] : 0.0;
  const double w = R.tau * (R.v0 * x0 + R.v1 * x1 + R.v2 * x2);
  x[0] = x0 - w * R.v0;
  if (R.nv > 1) x[s] = x1 - w * R.v1;
  if (R.nv > 2) x[2 * s] = x2 - w * R.v2;
}

// the same on three registers
__device__ __forceinline__ void qz_apply3(const QzRefl& R, double& x0, double& x1, double& x2) {
  const double w = R.tau * (R.v0 * x0 + R.v1 * x1 + R.v2 * x2);
  x0 -= w * R.v0;
  x1 -= w * R.v1;
  x2 -= w * R.v2;
}

struct QzMats {
  double *H, *T, *Q, *Z;
  int ldh, ldt, ldq, ldz, n;
  bool vq, vz;
  __device__ double& h(int i, int j) const { return H[PHT_IDX(i, j, ldh)]; }
  __device__ double& t(int i, int j) const { return T[PHT_IDX(i, j, ldt)]; }
};

struct QzCorner {
  double h[4][4];  // h[i][j] = H(k+i, k-1+j)
  double t[3][3];  // t[i][j] = T(k+i, k+j)
};

struct QzShared {
  QzRefl L, R1, R2;
  double x0, x1, x2;  // first column of the shift polynomial
  double red[32];
  int ired[2];
  int act;
  int flipa, flipb;
  QzCorner corner;    // thread 0's corner between bulge steps
  double next[7];     // H(k+1..k+2, k+3), T(k+1..k+2, k+3), H(k+3, k+3), H(k+4, k+3), T(k+3, k+3)
};

// left reflector on rows r0.. of H (columns hc0..last), T (tc0..last) and the
// columns r0.. of Q (Q <- Q P)
__device__ void qz_left(const QzMats& M, int r0, const QzRefl& R, int hc0, int tc0, int last) {
  if (R.tau == 0.0) return;
  const int nh = max(last - hc0 + 1, 0), nt = max(last - tc0 + 1, 0);
  const int nqr = M.vq ? M.n : 0;
  const int tot = nh + nt + nqr;
  for (int x = threadIdx.x; x < tot; x += blockDim.x) {
    if (x < nh)
      qz_apply(R, &M.h(r0, hc0 + x), 1);
    else if (x < nh + nt)
      qz_apply(R, &M.t(r0, tc0 + x - nh), 1);
    else
      qz_apply(R, &M.Q[PHT_IDX(x - nh - nt, r0, M.ldq)], M.ldq);
  }
}

// right reflector on columns c0.. of H (rows first..hr1), T (first..tr1) and Z
__device__ void qz_right(const QzMats& M, int c0, const QzRefl& R, int first, int hr1, int tr1) {
  if (R.tau == 0.0) return;
  const int nh = max(hr1 - first + 1, 0), nt = max(tr1 - first + 1, 0);
  const int nzr = M.vz ? M.n : 0;
  const int tot = nh + nt + nzr;
  for (int x = threadIdx.x; x < tot; x += blockDim.x) {
    if (x < nh)
      qz_apply(R, &M.h(first + x, c0), M.ldh);
    else if (x < nh + nt)
      qz_apply(R, &M.t(first + x - nh, c0), M.ldt);
    else
      qz_apply(R, &M.Z[PHT_IDX(x - nh - nt, c0, M.ldz)], M.ldz);
  }
}

// ---- one bulge step: the corner on thread 0 ------------------------------
// Rows k..k+nr-1 get the left reflector L, columns k..k+nr-1 the right
// reflectors R1 (nr = 3: zeroes T(k+2, k:k+1)) and R2 (zeroes T(k+1, k)).
// The corner is H(k..k+3, k-1..k+2) and T(k..k+2, k..k+2); thread 0 keeps it
// in LDS from one step to the next (entries outside the block [lo, m] or
// below the first step's column are 0).

// which corner entries exist at step k: rows k..k+2 from column k-1 (k-1 >= lo
// only below the top) to min(k+2, m); row k+3 (if <= m) from column k
__device__ __forceinline__ bool qz_hin(int i, int j, int k, int lo, int m) {
  return k + i <= m && (i < 3 ? (j > 0 || k > lo) && k - 1 + j <= m : j > 0);
}

// the first step of a sweep: the corner from memory (row k+3 holds only
// H(k+3, k+2) before the step)
__device__ __forceinline__ void qz_corner_load(const QzMats& M, QzCorner& C, int k, int lo,
                                               int m) {
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j)
      C.h[i][j] = (qz_hin(i, j, k, lo, m) && (i < 3 || j == 3)) ? M.h(k + i, k - 1 + j) : 0.0;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) C.t[i][j] = (k + i <= m && k + j <= m) ? M.t(k + i, k + j) : 0.0;
}

// step k's corner from step k-1's: the overlap moves up and left; column
// k+2 comes from the far-column update of step k-1 (rows k, k+1: hk, tk) and
// from entries no step has touched yet (H(k+2, k+2), H(k+3, k+2), T(k+2, k+2))
__device__ __forceinline__ void qz_corner_shift(QzCorner& C, int k, int m, double hk0, double hk1,
                                                double tk0, double tk1, double h22, double h32,
                                                double t22) {
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) C.h[i][j] = C.h[i + 1][j + 1];
  const bool c2 = k + 2 <= m, r3 = k + 3 <= m;
  C.h[0][3] = c2 ? hk0 : 0.0;
  C.h[1][3] = c2 ? hk1 : 0.0;
  C.h[2][3] = c2 ? h22 : 0.0;
  C.h[3][0] = 0.0;
  C.h[3][1] = 0.0;
  C.h[3][2] = 0.0;
  C.h[3][3] = r3 ? h32 : 0.0;
  C.t[0][0] = C.t[1][1];
  C.t[0][1] = C.t[1][2];
  C.t[1][0] = C.t[2][1];
  C.t[1][1] = C.t[2][2];
  C.t[0][2] = c2 ? tk0 : 0.0;
  C.t[1][2] = c2 ? tk1 : 0.0;
  C.t[2][2] = c2 ? t22 : 0.0;
  C.t[2][0] = 0.0;
  C.t[2][1] = 0.0;
}

// the step's reflectors, applied to the corner; published in sh
__device__ __forceinline__ void qz_corner_step(QzCorner& C, QzShared& sh, int k, int lo, int m) {
  const int nr = k <= m - 2 ? 3 : 2;
  const bool top = k == lo;
  double beta = 0.0;
  QzRefl L;
  if (top)
    L = qz_house(sh.x0, sh.x1, sh.x2, nr, &beta);
  else
    L = qz_house(C.h[0][0], C.h[1][0], C.h[2][0], nr, &beta);
  if (!top) {
    C.h[0][0] = beta;
    C.h[1][0] = 0.0;
    C.h[2][0] = 0.0;
  }
#pragma unroll
  for (int j = 1; j < 4; ++j) qz_apply3(L, C.h[0][j], C.h[1][j], C.h[2][j]);
#pragma unroll
  for (int j = 0; j < 3; ++j) qz_apply3(L, C.t[0][j], C.t[1][j], C.t[2][j]);
  QzRefl R1;
  R1.tau = 0.0;
  R1.nv = 3;
  R1.v0 = R1.v1 = R1.v2 = 0.0;
  if (nr == 3) {
    R1 = qz_house_rev(C.t[2][0], C.t[2][1], C.t[2][2], 3);
#pragma unroll
    for (int i = 0; i < 4; ++i) qz_apply3(R1, C.h[i][1], C.h[i][2], C.h[i][3]);
#pragma unroll
    for (int i = 0; i < 3; ++i) qz_apply3(R1, C.t[i][0], C.t[i][1], C.t[i][2]);
    C.t[2][0] = 0.0;
    C.t[2][1] = 0.0;
  }
  QzRefl R2 = qz_house_rev(C.t[1][0], C.t[1][1], 0.0, 2);
  R2.v2 = 0.0;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    double z = 0.0;
    qz_apply3(R2, C.h[i][1], C.h[i][2], z);
  }
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    double z = 0.0;
    qz_apply3(R2, C.t[i][0], C.t[i][1], z);
  }
  C.t[1][0] = 0.0;
  if (k + 3 > m) {  // keep entries outside the block at 0 for the next shift
#pragma unroll
    for (int j = 0; j < 4; ++j) C.h[3][j] = 0.0;
  }
  sh.L = L;
  sh.R1 = R1;
  sh.R2 = R2;
}

__device__ __forceinline__ void qz_corner_store(const QzMats& M, const QzCorner& C, int k, int lo,
                                                int m) {
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j)
      if (qz_hin(i, j, k, lo, m)) M.h(k + i, k - 1 + j) = C.h[i][j];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j)
      if (k + i <= m && k + j <= m) M.t(k + i, k + j) = C.t[i][j];
}

// ---- the kernel ----------------------------------------------------------
__global__ __launch_bounds__(1024) void k_qz(const QzTask* __restrict__ tasks, double* Hp,
                                             double* Tp, double* Qp, double* Zp,
                                             double* __restrict__ ar, double* __restrict__ ai,
                                             double* __restrict__ be, int* __restrict__ info,
                                             int flags) {
  __shared__ QzShared sh;
  const QzTask tk = tasks[blockIdx.x];
  const bool schur = (flags & kQzSchur) != 0;
  QzMats M;
  M.n = tk.n;
  M.H = Hp + tk.off;
  M.T = Tp + tk.off;
  M.Q = Qp ? Qp + tk.off : nullptr;
  M.Z = Zp ? Zp + tk.off : nullptr;
  M.ldh = tk.ldh;
  M.ldt = tk.ldt;
  M.ldq = tk.ldq;
  M.ldz = tk.ldz;
  M.vq = schur && Qp != nullptr;
  M.vz = schur && Zp != nullptr;
  const int n = tk.n;
  const int tid = threadIdx.x;
  double* are = ar + tk.eoff;
  double* aie = ai + tk.eoff;
  double* bee = be + tk.eoff;
  for (int j = tid; j < n; j += blockDim.x) {
    are[j] = 0.0;
    aie[j] = 0.0;
    bee[j] = 0.0;
  }
  // Frobenius norms of the Hessenberg part of H and the triangle of T
  double sa = 0.0, sb = 0.0;
  for (int j = tid; j < n; j += blockDim.x) {
    const int ie = min(j + 1, n - 1);
    for (int i = 0; i <= ie; ++i) {
      const double a = M.h(i, j);
      sa += a * a;
    }
    for (int i = 0; i <= j; ++i) {
      const double b = M.t(i, j);
      sb += b * b;
    }
  }
  const double anorm = sqrt(block_sum(sa, sh.red));
  const double bnorm = sqrt(block_sum(sb, sh.red + 16));
  const double atol = fmax(kQzSafmin, kQzUlp * anorm);
  const double btol = fmax(kQzSafmin, kQzUlp * bnorm);
  const long long itmax = 30LL * (long long)max(n, 1);
  long long it = 0;
  int since = 0;
  int status = 0;
  int m = n - 1;
  // NaN or Inf entries never deflate: report the whole pencil unconverged
  // rather than sweep 30 n times
  if (!isfinite(anorm) || !isfinite(bnorm)) {
    status = n;
    m = -1;
  }
  if (tid == 0) {
    sh.ired[0] = 0;
    sh.ired[1] = -1;
  }
  __syncthreads();
  while (m >= 0) {
    // the block [lo, m]: the largest j with a negligible H(j, j-1)
    for (int j = m - tid; j >= 1; j -= blockDim.x) {
      const double d = fabs(M.h(j, j)) + fabs(M.h(j - 1, j - 1));
      const double tol = d > 0.0 ? fmax(kQzSafmin, kQzUlp * d) : atol;
      if (fabs(M.h(j, j - 1)) <= tol) {
        M.h(j, j - 1) = 0.0;
        atomicMax(&sh.ired[0], j);
        break;
      }
    }
    __syncthreads();
    const int lo = sh.ired[0];
    // the largest j in [lo, m] with a negligible T(j, j)
    for (int j = m - tid; j >= lo; j -= blockDim.x) {
      if (fabs(M.t(j, j)) < btol) {
        M.t(j, j) = 0.0;
        atomicMax(&sh.ired[1], j);
        break;
      }
    }
    __syncthreads();
    const int kz = sh.ired[1];
    __syncthreads();  // everyone has read ired before it is reset
    if (tid == 0) {
      sh.ired[0] = 0;
      sh.ired[1] = -1;
    }
    const int first = schur ? 0 : lo;
    const int last = schur ? n - 1 : m;

    if (lo == m) {
      // 1x1 block: beta >= 0 (a column sign flip)
      if (tid == 0) sh.act = (kz != m && M.t(m, m) < 0.0) ? 1 : 0;
      __syncthreads();
      if (sh.act) qz_right(M, m, qz_flip(0, 1), first, m, m);
      __syncthreads();
      if (tid == 0) {
        are[m] = M.h(m, m);
        aie[m] = 0.0;
        bee[m] = kz == m ? 0.0 : M.t(m, m);
      }
      --m;
      since = 0;
      continue;
    }
    if (kz >= 0) {
      // chase the zero at T(kz, kz) down to T(m, m), then deflate m
      for (int j = kz; j < m; ++j) {
        if (tid == 0) sh.L = qz_house(M.t(j, j + 1), M.t(j + 1, j + 1), 0.0, 2, nullptr);
        __syncthreads();
        const QzRefl Lj = sh.L;
        qz_left(M, j, Lj, max(j - 1, lo), j, last);
        __syncthreads();
        if (tid == 0) {
          M.t(j + 1, j + 1) = 0.0;
          if (j > lo) sh.R1 = qz_house_rev(M.h(j + 1, j - 1), M.h(j + 1, j), 0.0, 2);
        }
        __syncthreads();
        if (j > lo) {
          const QzRefl Rj = sh.R1;
          qz_right(M, j - 1, Rj, first, j + 1, j);
          __syncthreads();
          if (tid == 0) {
            M.h(j + 1, j - 1) = 0.0;
            M.t(j, j - 1) = 0.0;
          }
        }
      }
      if (tid == 0) sh.R1 = qz_house_rev(M.h(m, m - 1), M.h(m, m), 0.0, 2);
      __syncthreads();
      {
        const QzRefl Rm = sh.R1;
        qz_right(M, m - 1, Rm, first, m, m);
      }
      __syncthreads();
      if (tid == 0) {
        M.h(m, m - 1) = 0.0;
        M.t(m, m - 1) = 0.0;
        are[m] = M.h(m, m);
        aie[m] = 0.0;
        bee[m] = 0.0;
      }
      __syncthreads();
      --m;
      since = 0;
      continue;
    }
    if (lo == m - 1) {
      // 2x2 block at m-1, m
      if (tid == 0) {
        const double a00 = M.h(m - 1, m - 1), a01 = M.h(m - 1, m), a10 = M.h(m, m - 1),
                     a11 = M.h(m, m);
        const double b00 = M.t(m - 1, m - 1), b01 = M.t(m - 1, m), b11 = M.t(m, m);
        const double s = a00 / b00 + a11 / b11 - a10 * b01 / (b00 * b11);
        const double p = (a00 * a11 - a01 * a10) / (b00 * b11);
        const double disc = 0.25 * s * s - p;
        if (disc >= 0.0) {
          sh.act = 1;
          const double lam = 0.5 * s + copysign(sqrt(disc), s);
          double al, bt;
          if (fabs(lam) > 1.0) {
            al = 1.0;
            bt = 1.0 / lam;
          } else {
            al = lam;
            bt = 1.0;
          }
          const double m00 = bt * a00 - al * b00, m01 = bt * a01 - al * b01;
          const double m10 = bt * a10, m11 = bt * a11 - al * b11;
          if (fabs(m00) + fabs(m01) >= fabs(m10) + fabs(m11))
            sh.R1 = qz_house_rev(m00, m01, 0.0, 2);
          else
            sh.R1 = qz_house_rev(m10, m11, 0.0, 2);
          sh.x0 = fmax(fmax(fabs(a00), fabs(a01)), fmax(fabs(a10), fabs(a11)));
          sh.x1 = fmax(fmax(fabs(b00), fabs(b01)), fabs(b11));
          sh.x0 = fmax(sh.x0, kQzSafmin);
          sh.x1 = fmax(sh.x1, kQzSafmin);
        } else {
          sh.act = 2;
          const double g11 = b00 * b00, g12 = b00 * b01, g22 = b01 * b01 + b11 * b11;
          if (g12 != 0.0) {
            const double zeta = (g22 - g11) / (2.0 * g12);
            const double tt = copysign(1.0, zeta) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
            const double c = 1.0 / sqrt(1.0 + tt * tt);
            sh.R1 = qz_house_rev(tt * c, c, 0.0, 2);
          } else {
            sh.R1 = qz_flip(0, 2);
            sh.R1.tau = 0.0;
          }
        }
      }
      __syncthreads();
      const int act = sh.act;
      {
        const QzRefl R = sh.R1;
        qz_right(M, m - 1, R, first, m, m);
      }
      __syncthreads();
      if (tid == 0) {
        const double cb0 = M.t(m - 1, m - 1), cb1 = M.t(m, m - 1);
        if (act == 1) {
          const double ca0 = M.h(m - 1, m - 1), ca1 = M.h(m, m - 1);
          const bool useb = hypot(cb0, cb1) / sh.x1 >= hypot(ca0, ca1) / sh.x0;
          sh.L = useb ? qz_house(cb0, cb1, 0.0, 2, nullptr) : qz_house(ca0, ca1, 0.0, 2, nullptr);
        } else {
          sh.L = qz_house(cb0, cb1, 0.0, 2, nullptr);
        }
      }
      __syncthreads();
      {
        const QzRefl L = sh.L;
        qz_left(M, m - 1, L, m - 1, m - 1, last);
      }
      __syncthreads();
      if (tid == 0) {
        M.t(m, m - 1) = 0.0;
        if (act == 1) {
          M.h(m, m - 1) = 0.0;
        } else {
          M.t(m - 1, m) = 0.0;
        }
        sh.flipa = M.t(m - 1, m - 1) < 0.0;
        sh.flipb = M.t(m, m) < 0.0;
      }
      __syncthreads();
      // signs: real pairs negate columns, complex pairs rows (T's diagonal > 0)
      if (act == 1) {
        if (sh.flipa) qz_right(M, m - 1, qz_flip(0, 1), first, m - 1, m - 1);
        if (sh.flipb) qz_right(M, m, qz_flip(0, 1), first, m, m);
      } else {
        if (sh.flipa) qz_left(M, m - 1, qz_flip(0, 1), m - 1, m - 1, last);
        if (sh.flipb) qz_left(M, m, qz_flip(0, 1), m - 1, m - 1, last);
      }
      __syncthreads();
      if (tid == 0) {
        if (act == 1) {
          are[m - 1] = M.h(m - 1, m - 1);
          bee[m - 1] = M.t(m - 1, m - 1);
          are[m] = M.h(m, m);
          bee[m] = M.t(m, m);
          aie[m - 1] = 0.0;
          aie[m] = 0.0;
        } else {
          const double b1 = M.t(m - 1, m - 1), b2 = M.t(m, m);
          const double a00 = M.h(m - 1, m - 1), a01 = M.h(m - 1, m), a10 = M.h(m, m - 1),
                       a11 = M.h(m, m);
          const double s = a00 / b1 + a11 / b2;
          const double p = (a00 * a11 - a01 * a10) / (b1 * b2);
          const double re = 0.5 * s;
          const double im = sqrt(fmax(p - re * re, 0.0));
          are[m - 1] = re * b1;
          aie[m - 1] = im * b1;
          bee[m - 1] = b1;
          are[m] = re * b2;
          aie[m] = -im * b2;
          bee[m] = b2;
        }
      }
      __syncthreads();
      m -= 2;
      since = 0;
      continue;
    }
    ++it;
    ++since;
    if (it > itmax) {
      status = m + 1;
      break;
    }
    // double-shift sweep over [lo, m]
    if (tid == 0) {
      double s, p;
      if (since % 10 == 0) {
        const double sg = M.h(m, m) / M.t(m, m) + 0.75 * fabs(M.h(m, m - 1) / M.t(m - 1, m - 1));
        s = 2.0 * sg;
        p = sg * sg;
      } else {
        const double a = M.t(m - 1, m - 1), b = M.t(m - 1, m), d = M.t(m, m);
        const double e = M.h(m - 1, m - 1), f = M.h(m - 1, m), g = M.h(m, m - 1),
                     hh = M.h(m, m);
        s = e / a + hh / d - g * b / (a * d);
        p = (e * hh - f * g) / (a * d);
      }
      const double t11 = M.t(lo, lo), t12 = M.t(lo, lo + 1), t22 = M.t(lo + 1, lo + 1);
      const double h11 = M.h(lo, lo), h21 = M.h(lo + 1, lo), h12 = M.h(lo, lo + 1),
                   h22 = M.h(lo + 1, lo + 1), h32 = M.h(lo + 2, lo + 1);
      const double y1 = h11 / t11, y2 = h21 / t11;
      const double w2 = y2 / t22;
      const double w1 = (y1 - t12 * w2) / t11;
      sh.x0 = h11 * w1 + h12 * w2 - s * y1 + p;
      sh.x1 = h21 * w1 + h22 * w2 - s * y2;
      sh.x2 = h32 * w2;
    }
    // Per step: thread 0 builds the corner (from the previous step's corner,
    // the far column it updated itself and prefetched entries), barrier, then
    // every thread takes far columns (H and T together: thread 0 takes
    // column k+nr, which the next corner needs), rows above the corner, and
    // rows of Q and Z, while thread 0 also stores the corner; barrier.
    for (int k = lo; k < m; ++k) {
      if (tid == 0) {
        QzCorner C;
        if (k == lo) {
          qz_corner_load(M, C, k, lo, m);
        } else {
          C = sh.corner;
          qz_corner_shift(C, k, m, sh.next[0], sh.next[1], sh.next[2], sh.next[3], sh.next[4],
                          sh.next[5], sh.next[6]);
        }
        qz_corner_step(C, sh, k, lo, m);
        sh.corner = C;
      }
      __syncthreads();
      const QzRefl L = sh.L, R1 = sh.R1, R2 = sh.R2;
      const int nr = L.nv;
      // untouched until the next step, which needs them: loaded now, kept in
      // LDS after thread 0's own updates (their latencies overlap)
      double pre0 = 0.0, pre1 = 0.0, pre2 = 0.0;
      if (tid == 0) {
        pre0 = k + 3 <= m ? M.h(k + 3, k + 3) : 0.0;
        pre1 = k + 4 <= m ? M.h(k + 4, k + 3) : 0.0;
        pre2 = k + 3 <= m ? M.t(k + 3, k + 3) : 0.0;
        qz_corner_store(M, sh.corner, k, lo, m);
      }
      const int c0 = k + nr;
      const int nc = max(last - c0 + 1, 0);
      const int nrw = max(k - first, 0);
      const int nqr = M.vq ? n : 0, nzr = M.vz ? n : 0;
      const int tot = nc + nrw + nqr + nzr;
      for (int x = tid; x < tot; x += blockDim.x) {
        if (x < nc) {
          double* ph = &M.h(k, c0 + x);
          double* pt = &M.t(k, c0 + x);
          double a0 = ph[0], a1 = ph[1], a2 = nr == 3 ? ph[2] : 0.0;
          double b0 = pt[0], b1 = pt[1], b2 = nr == 3 ? pt[2] : 0.0;
          qz_apply3(L, a0, a1, a2);
          qz_apply3(L, b0, b1, b2);
          ph[0] = a0;
          ph[1] = a1;
          pt[0] = b0;
          pt[1] = b1;
          if (nr == 3) {
            ph[2] = a2;
            pt[2] = b2;
          }
          if (x == 0) {  // rows k+1, k+2 of column k+3 for the next corner
            sh.next[0] = a1;
            sh.next[1] = a2;
            sh.next[2] = b1;
            sh.next[3] = b2;
          }
        } else if (x < nc + nrw) {
          const int i = first + x - nc;
          double* ph = &M.h(i, k);
          double* pt = &M.t(i, k);
          if (nr == 3) {
            qz_apply(R1, ph, M.ldh);
            qz_apply(R1, pt, M.ldt);
          }
          qz_apply(R2, ph, M.ldh);
          qz_apply(R2, pt, M.ldt);
        } else if (x < nc + nrw + nqr) {
          qz_apply(L, &M.Q[PHT_IDX(x - nc - nrw, k, M.ldq)], M.ldq);
        } else {
          double* p = &M.Z[PHT_IDX(x - nc - nrw - nqr, k, M.ldz)];
          if (nr == 3) qz_apply(R1, p, M.ldz);
          qz_apply(R2, p, M.ldz);
        }
      }
      if (tid == 0) {
        sh.next[4] = pre0;
        sh.next[5] = pre1;
        sh.next[6] = pre2;
      }
      __syncthreads();
    }
  }
  if (tid == 0) info[tk.idx] = status;
}

}  // namespace

int qz_threads(int n) { return n >= 384 ? 1024 : (n >= 192 ? 512 : 256); }

hipError_t qz_launch(hipStream_t st, const QzTask* dtasks, int count, int threads, double* H,
                     double* T, double* Q, double* Z, double* ar, double* ai, double* be,
                     int* dinfo, int flags) {
  if (count <= 0) return hipSuccess;
  if (threads < 64 || threads > 1024 || threads % 64) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_qz, dim3(count), dim3(threads), 0, st, dtasks, H, T, Q, Z, ar, ai, be,
                     dinfo, flags);
  return hipGetLastError();
}

}  // namespace pht
