// Register-resident stage-1 generators for blocks of at most 128 rows and
// 64 reflectors (P = p*nb <= 128, nb <= 64), the shapes of the benchmark
// configurations (nb = 64, p = 2; nb = 16, p = 8).  Same mathematics as the
// generic kernels in stage1.hip (DGEQR2 / DGERQ2 + DORGRQ + DGELQ2 / DLARFT
// semantics, src/ab2rht.f90:105-116,290-311); the difference is data
// placement and control flow:
//  * the block lives in registers, element j of a thread's slice being index
//    4*j + q (or 2*j + h) of its row/column, so the lanes of a wave that read
//    one LDS vector touch consecutive words (no bank conflicts);
//  * every reflector is stored zero-padded to 128 entries in its own LDS row,
//    which is both the broadcast buffer of its update and the explicit V of
//    the block (one barrier per reflector, no masks in the update loops);
//  * the work of a reflector is bounded by skipping whole 8-element chunks
//    on a workgroup-uniform test instead of per-element predicates;
//  * T comes from the DLARFT forward recurrence
//    T(0:j-1,j) = -tau_j T(0:j-1,0:j-1) G(0:j-1,j), G = V^T V,
//    one barrier per column with the sums spread over 4 lanes per row.
#include "device_util.hpp"
#include "kernels.hpp"

namespace pht {

namespace {
constexpr int kT = 256;
constexpr int kLd = 129;  // LDS row length of stored reflectors (odd: spreads banks)

__device__ __forceinline__ double grp2_sum(double v) { return v + __shfl_xor(v, 1, 64); }
__device__ __forceinline__ double grp4_sum(double v) {
  v += __shfl_xor(v, 1, 64);
  return v + __shfl_xor(v, 2, 64);
}

// the plain sum of squares is usable unless it over/underflowed
__device__ __forceinline__ bool ss_safe(double ss) {
  return ss == 0.0 || (ss > 1e-290 && ss < 1e290);
}

// G(x,y) = v_x . v_y for x < y < k (rows of Vs, kLd apart, zero-padded to
// 128 entries), then T (kfull x kfull, ld ldt, zeros below the diagonal and
// beyond k) of H = H(0) ... H(k-1) = I - V T V^T.  Called by all kT threads.
__device__ void gram_and_t(const double* Vs, int k, const double* tau, double* G,
                           double* Tl, double* Tout, int ldt, int kfull) {
  const int a = threadIdx.x >> 2, q = threadIdx.x & 3;
  double va[32];
#pragma unroll
  for (int j = 0; j < 32; ++j) va[j] = Vs[a * kLd + 4 * j + q];
  for (int y = 1; y < k; ++y) {
    const double* vy = Vs + y * kLd;
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
#pragma unroll
    for (int j = 0; j < 32; j += 4) {
      s0 += va[j] * vy[4 * j + q];
      s1 += va[j + 1] * vy[4 * j + 4 + q];
      s2 += va[j + 2] * vy[4 * j + 8 + q];
      s3 += va[j + 3] * vy[4 * j + 12 + q];
    }
    const double s = grp4_sum((s0 + s1) + (s2 + s3));
    if (q == 0 && a < y) G[a * 65 + y] = s;
  }
  __syncthreads();
  for (int j = 0; j < k; ++j) {
    double s = 0.0;
    if (a < j)
      for (int l = a + q; l < j; l += 4) s += Tl[a * 65 + l] * G[l * 65 + j];
    s = grp4_sum(s);
    if (q == 0) {
      if (a < j)
        Tl[a * 65 + j] = -tau[j] * s;
      else if (a == j)
        Tl[j * 65 + j] = tau[j];
    }
    __syncthreads();
  }
  for (int e = threadIdx.x; e < kfull * kfull; e += kT) {
    const int r = e % kfull, c = e / kfull;
    Tout[PHT_IDX(r, c, ldt)] = (r <= c && c < k) ? Tl[r * 65 + c] : 0.0;
  }
}

// ||x|| for a group of 4 lanes holding 32 values each, entries selected by
// `sel`; the scaled evaluation runs only when the plain sum of squares
// over/underflowed.
template <typename Sel>
__device__ __forceinline__ double grp4_norm(const double* X, Sel sel) {
  double ss = 0.0;
#pragma unroll
  for (int j = 0; j < 32; ++j) ss += sel(j) ? X[j] * X[j] : 0.0;
  ss = grp4_sum(ss);
  if (ss_safe(ss)) return sqrt(ss);
  double mx = 0.0;
#pragma unroll
  for (int j = 0; j < 32; ++j) mx = fmax(mx, sel(j) ? fabs(X[j]) : 0.0);
  mx = fmax(mx, __shfl_xor(mx, 1, 64));
  mx = fmax(mx, __shfl_xor(mx, 2, 64));
  if (mx == 0.0) return 0.0;
  double s2 = 0.0;
#pragma unroll
  for (int j = 0; j < 32; ++j) {
    const double t = sel(j) ? X[j] / mx : 0.0;
    s2 += t * t;
  }
  return mx * sqrt(grp4_sum(s2));
}
}  // namespace

// ---------------------------------------------------------------------------
// Left generation of one panel (ab2rht.f90:93-118), blocks bottom-up.
// Thread t owns column t/4 of the block, rows 4*j + t%4 (j < 32).
__global__ __launch_bounds__(kT) void k_s1_panel_qr_fast(
    double* __restrict__ A, int lda, int n, int jb, int nb, int p, int nblocks,
    double* __restrict__ Vt, int ldv, long long vstride, double* __restrict__ Tt,
    int ldt, long long tstride) {
  extern __shared__ double sm[];
  double* Vc = sm;               // [64][kLd]: reflector i in row i
  double* G = Vc + 64 * kLd;     // [64][65]
  double* Tl = G + 64 * 65;      // [64][65]
  double* tau = Tl + 64 * 65;    // [64]
  const int tid = threadIdx.x;
  const int col = tid >> 2, q = tid & 3;
  const int P = p * nb;

  for (int bl = 0; bl < nblocks; ++bl) {
    const int i0 = jb + nb + (nblocks - 1 - bl) * (p - 1) * nb;
    const int len = min(n - i0, P);
    const int k = min(nb, len);
    double X[32];
#pragma unroll
    for (int j = 0; j < 32; ++j) {
      const int r = 4 * j + q;
      X[j] = (col < nb && r < len) ? A[PHT_IDX(i0 + r, jb + col, lda)] : 0.0;
    }
    // DGEQR2: reflector i from column i (threads 4i..4i+3), applied to the
    // columns right of it; it is zero above row i.
    for (int i = 0; i < k; ++i) {
      double* vi = Vc + i * kLd;
      if (col == i) {
        double al = 0.0;
#pragma unroll
        for (int j = 0; j < 32; ++j) al += (4 * j + q == i) ? X[j] : 0.0;
        al = grp4_sum(al);
        const double xnorm = grp4_norm(X, [&](int j) { return 4 * j + q > i; });
        const Refl R = make_refl(al, xnorm);
        const double sc = (R.tau != 0.0) ? R.scal : 1.0;
#pragma unroll
        for (int j = 0; j < 32; ++j) {
          const int r = 4 * j + q;
          const double v = (r > i) ? X[j] * sc : (r == i ? 1.0 : 0.0);
          vi[r] = v;
          X[j] = (r > i) ? v : (r == i ? R.beta : X[j]);
        }
        if (q == 0) tau[i] = R.tau;
      }
      __syncthreads();
      const double t = tau[i];
      if (col > i && col < nb && t != 0.0) {
        double w = 0.0;
#pragma unroll
        for (int jc = 0; jc < 4; ++jc) {
          if (32 * jc + 31 < i) continue;  // chunk entirely above the reflector
#pragma unroll
          for (int j = 8 * jc; j < 8 * jc + 8; ++j) w += X[j] * vi[4 * j + q];
        }
        w = grp4_sum(w) * t;
#pragma unroll
        for (int jc = 0; jc < 4; ++jc) {
          if (32 * jc + 31 < i) continue;
#pragma unroll
          for (int j = 8 * jc; j < 8 * jc + 8; ++j) X[j] -= w * vi[4 * j + q];
        }
      }
    }
    // R (upper trapezoid) and zeros back into A
    if (col < nb) {
#pragma unroll
      for (int j = 0; j < 32; ++j) {
        const int r = 4 * j + q;
        if (r < len) A[PHT_IDX(i0 + r, jb + col, lda)] = (r <= col) ? X[j] : 0.0;
      }
    }
    // explicit V to the tape (coalesced along the rows)
    double* Vb = Vt + (long long)bl * vstride;
    for (int e = tid; e < len * nb; e += kT) {
      const int r = e % len, c = e / len;
      Vb[PHT_IDX(r, c, ldv)] = (c < k) ? Vc[c * kLd + r] : 0.0;
    }
    gram_and_t(Vc, k, tau, G, Tl, Tt + (long long)bl * tstride, ldt, nb);
    // the next block (above) re-reads the R rows just written to A, and
    // overwrites Vc / tau
    __threadfence();
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Right generator for one B block with nb < len <= 128 (ab2rht.f90:290-303).
// RQ phase: thread t owns row t/2 of the block, columns 2*j + t%2 (j < 64).
// Y / LQ phase: thread t owns row t/4 of Y = Q(0:nb-1,:), columns 4*j + t%4.
__global__ __launch_bounds__(kT) void k_s1_rq_fast(
    const double* __restrict__ B, int ldb, int i0, int len, int nb,
    double* __restrict__ Vout, int ldv, double* __restrict__ Tout, int ldt) {
  extern __shared__ double sm[];
  double* Vr = sm;                  // [128][kLd]: RQ reflector i in row i
  double* taus = Vr + 128 * kLd;    // [128]
  double* tau2 = taus + 128;        // [64]
  double* Ys = Vr;                  // [64][kLd]: LQ reflectors, after the Y phase
  double* G = Vr + 64 * kLd;        // [64][65]
  double* Tl = G + 64 * 65;         // [64][65] (ends inside taus, dead by then)
  const int tid = threadIdx.x;

  // ---- DGERQ2 on the len x len block: reflector i from row i, zero right of
  // column i, applied to the rows above it.
  {
    const int row = tid >> 1, h = tid & 1;
    double X[64];
#pragma unroll
    for (int j = 0; j < 64; ++j) {
      const int c = 2 * j + h;
      X[j] = (row < len && c < len) ? B[PHT_IDX(i0 + row, i0 + c, ldb)] : 0.0;
    }
    for (int i = len - 1; i >= 0; --i) {
      double* vi = Vr + i * kLd;
      if (row == i) {
        double ss = 0.0, al = 0.0;
#pragma unroll
        for (int j = 0; j < 64; ++j) {
          const int c = 2 * j + h;
          ss += (c < i) ? X[j] * X[j] : 0.0;
          al += (c == i) ? X[j] : 0.0;
        }
        ss = grp2_sum(ss);
        al = grp2_sum(al);
        double xnorm;
        if (ss_safe(ss)) {
          xnorm = sqrt(ss);
        } else {
          double mx = 0.0;
#pragma unroll
          for (int j = 0; j < 64; ++j) mx = fmax(mx, (2 * j + h < i) ? fabs(X[j]) : 0.0);
          mx = fmax(mx, __shfl_xor(mx, 1, 64));
          double s2 = 0.0;
          if (mx > 0.0) {
#pragma unroll
            for (int j = 0; j < 64; ++j) {
              const double t = (2 * j + h < i) ? X[j] / mx : 0.0;
              s2 += t * t;
            }
          }
          xnorm = mx * sqrt(grp2_sum(s2));
        }
        const Refl R = make_refl(al, xnorm);
        const double sc = (R.tau != 0.0) ? R.scal : 1.0;
#pragma unroll
        for (int j = 0; j < 64; ++j) {
          const int c = 2 * j + h;
          vi[c] = (c < i) ? X[j] * sc : (c == i ? 1.0 : 0.0);
        }
        if (h == 0) taus[i] = R.tau;
      }
      __syncthreads();
      const double t = taus[i];
      if (row < i && t != 0.0) {
        double w0 = 0.0, w1 = 0.0;
#pragma unroll
        for (int jc = 0; jc < 8; ++jc) {
          if (16 * jc > i) continue;  // chunk entirely right of the reflector
#pragma unroll
          for (int j = 8 * jc; j < 8 * jc + 8; j += 2) {
            w0 += X[j] * vi[2 * j + h];
            w1 += X[j + 1] * vi[2 * j + 2 + h];
          }
        }
        const double w = grp2_sum(w0 + w1) * t;
#pragma unroll
        for (int jc = 0; jc < 8; ++jc) {
          if (16 * jc > i) continue;
#pragma unroll
          for (int j = 8 * jc; j < 8 * jc + 8; ++j) X[j] -= w * vi[2 * j + h];
        }
      }
    }
  }
  __syncthreads();

  // ---- Y = E^T H(0) H(1) ... H(len-1): the first nb rows of the RQ's Q
  const int a = tid >> 2, q = tid & 3;
  double Y[32];
#pragma unroll
  for (int j = 0; j < 32; ++j) Y[j] = (a < nb && 4 * j + q == a) ? 1.0 : 0.0;
  if (a < nb) {
    for (int i = 1; i < len; ++i) {  // H(0) is the identity (length-1 reflector)
      const double t = taus[i];
      if (t == 0.0) continue;
      const double* vi = Vr + i * kLd;
      double w0 = 0.0, w1 = 0.0;
#pragma unroll
      for (int jc = 0; jc < 4; ++jc) {
        if (32 * jc > i) continue;
#pragma unroll
        for (int j = 8 * jc; j < 8 * jc + 8; j += 2) {
          w0 += Y[j] * vi[4 * j + q];
          w1 += Y[j + 1] * vi[4 * j + 4 + q];
        }
      }
      const double w = grp4_sum(w0 + w1) * t;
#pragma unroll
      for (int jc = 0; jc < 4; ++jc) {
        if (32 * jc > i) continue;
#pragma unroll
        for (int j = 8 * jc; j < 8 * jc + 8; ++j) Y[j] -= w * vi[4 * j + q];
      }
    }
  }
  __syncthreads();  // Ys overwrites Vr from here on

  // ---- DGELQ2 on Y (nb x len): reflector i from row i, zero left of column
  // i, applied to the rows below it.  Its LDS row is the explicit V row.
  for (int i = 0; i < nb; ++i) {
    double* vi = Ys + i * kLd;
    if (a == i) {
      double al = 0.0, ss = 0.0;
#pragma unroll
      for (int j = 0; j < 32; ++j) {
        const int c = 4 * j + q;
        al += (c == i) ? Y[j] : 0.0;
        ss += (c > i) ? Y[j] * Y[j] : 0.0;  // rows of an orthogonal factor: |entries| <= 1
      }
      al = grp4_sum(al);
      const double xnorm = sqrt(grp4_sum(ss));
      const Refl R = make_refl(al, xnorm);
      const double sc = (R.tau != 0.0) ? R.scal : 1.0;
#pragma unroll
      for (int j = 0; j < 32; ++j) {
        const int c = 4 * j + q;
        vi[c] = (c > i) ? Y[j] * sc : (c == i ? 1.0 : 0.0);
      }
      if (q == 0) tau2[i] = R.tau;
    }
    __syncthreads();
    const double t = tau2[i];
    if (a > i && a < nb && t != 0.0) {
      double w0 = 0.0, w1 = 0.0;
#pragma unroll
      for (int jc = 0; jc < 4; ++jc) {
        if (32 * jc + 31 < i) continue;  // chunk entirely left of the reflector
#pragma unroll
        for (int j = 8 * jc; j < 8 * jc + 8; j += 2) {
          w0 += Y[j] * vi[4 * j + q];
          w1 += Y[j + 1] * vi[4 * j + 4 + q];
        }
      }
      const double w = grp4_sum(w0 + w1) * t;
#pragma unroll
      for (int jc = 0; jc < 4; ++jc) {
        if (32 * jc + 31 < i) continue;
#pragma unroll
        for (int j = 8 * jc; j < 8 * jc + 8; ++j) Y[j] -= w * vi[4 * j + q];
      }
    }
  }
  // explicit V (len x nb), coalesced along the rows
  for (int e = tid; e < len * nb; e += kT) {
    const int c = e % len, r = e / len;
    Vout[PHT_IDX(c, r, ldv)] = Ys[r * kLd + c];
  }
  gram_and_t(Ys, nb, tau2, G, Tl, Tout, ldt, nb);
}

namespace {
constexpr size_t kPanelLds = (size_t)(64 * kLd + 2 * 64 * 65 + 64) * sizeof(double);
constexpr size_t kRqLds = (size_t)(128 * kLd + 128 + 64) * sizeof(double);
static_assert(kPanelLds <= 160 * 1024 && kRqLds <= 160 * 1024, "LDS budget");
static_assert(64 * kLd + 2 * 64 * 65 <= 128 * kLd + 128, "Gram/T alias in the RQ kernel");

hipError_t allow_lds(const void* fn, size_t bytes, bool& done) {
  if (done) return hipSuccess;
  hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (int)bytes);
  if (e == hipSuccess) done = true;
  return e;
}
}  // namespace

hipError_t s1_panel_qr_fast(hipStream_t st, double* A, int lda, int n, int jb,
                            int nb, int p, int nblocks, double* Vt, int ldv,
                            long long vstride, double* Tt, int ldt,
                            long long tstride) {
  if (nb > 64 || p * nb > 128) return hipErrorInvalidValue;
  static bool attr = false;
  if (hipError_t e = allow_lds((const void*)k_s1_panel_qr_fast, kPanelLds, attr)) return e;
  hipLaunchKernelGGL(k_s1_panel_qr_fast, dim3(1), dim3(kT), kPanelLds, st, A, lda, n, jb,
                     nb, p, nblocks, Vt, ldv, vstride, Tt, ldt, tstride);
  return hipGetLastError();
}

hipError_t s1_rq_fast(hipStream_t st, const double* B, int ldb, int i0, int len,
                      int nb, double* V, int ldv, double* T, int ldt) {
  if (nb > 64 || len > 128 || len <= nb) return hipErrorInvalidValue;
  static bool attr = false;
  if (hipError_t e = allow_lds((const void*)k_s1_rq_fast, kRqLds, attr)) return e;
  hipLaunchKernelGGL(k_s1_rq_fast, dim3(1), dim3(kT), kRqLds, st, B, ldb, i0, len, nb,
                     V, ldv, T, ldt);
  return hipGetLastError();
}

}  // namespace pht
