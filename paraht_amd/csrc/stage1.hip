// Stage-1 generation kernels (ab2rht_simple, src/ab2rht.f90:8-495).
//
//  k_s1_panel_qr : left generation, ab2rht.f90:93-118.  One workgroup walks
//                  the panel's blocks bottom-up; each P x nb block of
//                  A(:, jb:jb+nb-1) is QR-factored in LDS (DGEQR2 semantics,
//                  ab2rht.f90:105), its R and the explicit zeros below it are
//                  written back (ab2rht.f90:108-113), and the explicit
//                  reflector block V and its triangular factor T (DLARFT
//                  'F','C', ab2rht.f90:116) go to the tape.
//  k_s1_rq_gen   : right generation for one B block, ab2rht.f90:290-311.
//                  RQ of the nb2 x nb2 block (DGERQF); if nb2 > nb the first
//                  nb rows of its orthogonal factor are formed and
//                  LQ-factored (DORGRQ + DGELQ2) so that nb row reflectors
//                  make the block's first nb columns upper triangular;
//                  otherwise the RQ reflectors themselves are used (DLARFT
//                  'B','R').  Output is the explicit column-form V and T.
#include "device_util.hpp"
#include "kernels.hpp"

namespace pht {

namespace {

constexpr int kThreads = 256;

// ||x|| over a strided vector held by the workgroup (all threads return it).
__device__ double wg_norm(const double* x, int len, int inc, double* red) {
  const int tid = threadIdx.x;
  double ss = 0.0;
  for (int e = tid; e < len; e += blockDim.x) {
    double t = x[(size_t)e * inc];
    ss += t * t;
  }
  ss = block_sum(ss, red);
  if (ss > 1e-290 && ss < 1e290) return sqrt(ss);
  // scaled two-pass evaluation when the plain sum may have over/underflowed
  double mx = 0.0;
  for (int e = tid; e < len; e += blockDim.x) mx = fmax(mx, fabs(x[(size_t)e * inc]));
  mx = block_max(mx, red);
  if (mx == 0.0 || !(mx <= DBL_MAX)) return mx;
  double s2 = 0.0;
  for (int e = tid; e < len; e += blockDim.x) {
    double t = x[(size_t)e * inc] / mx;
    s2 += t * t;
  }
  s2 = block_sum(s2, red);
  return mx * sqrt(s2);
}

}  // namespace

// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kThreads) void k_s1_panel_qr(
    double* __restrict__ A, int lda, int n, int jb, int nb, int p, int nblocks,
    double* __restrict__ Vt, int ldv, long long vstride, double* __restrict__ Tt,
    int ldt, long long tstride) {
  extern __shared__ double sm[];
  const int P = p * nb;
  double* X = sm;           // P x nb, ld P; reused as G/T (nb x nb, ld nb)
  double* tau = X + (size_t)P * nb;
  double* gcol = tau + nb;  // nb
  double* red = gcol + nb;  // 32
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;

  for (int bl = 0; bl < nblocks; ++bl) {
    const int i0 = jb + nb + (nblocks - 1 - bl) * (p - 1) * nb;
    const int len = min(n - i0, P);
    const int k = min(nb, len);
    for (int e = tid; e < len * nb; e += kThreads) {
      const int r = e % len, c = e / len;
      X[r + c * P] = A[PHT_IDX(i0 + r, jb + c, lda)];
    }
    __syncthreads();
    // DGEQR2
    for (int i = 0; i < k; ++i) {
      const double alpha = X[i + i * P];
      const double xnorm = wg_norm(X + (i + 1) + i * P, len - i - 1, 1, red);
      const Refl R = make_refl(alpha, xnorm);
      if (R.tau != 0.0)
        for (int r = i + 1 + tid; r < len; r += kThreads) X[r + i * P] *= R.scal;
      __syncthreads();
      if (tid == 0) {
        X[i + i * P] = R.beta;
        tau[i] = R.tau;
      }
      if (R.tau != 0.0) {
        for (int j = i + 1 + wave; j < nb; j += kThreads / 64) {
          double w = 0.0;
          for (int r = i + 1 + lane; r < len; r += 64) w += X[r + i * P] * X[r + j * P];
          w = (wave_sum(w) + X[i + j * P]) * R.tau;
          __builtin_amdgcn_wave_barrier();
          if (lane == 0) X[i + j * P] -= w;
          for (int r = i + 1 + lane; r < len; r += 64) X[r + j * P] -= X[r + i * P] * w;
        }
      }
      __syncthreads();
    }
    // explicit V to the tape, R and zeros back into A
    double* Vb = Vt + (long long)bl * vstride;
    for (int e = tid; e < len * nb; e += kThreads) {
      const int r = e % len, c = e / len;
      const double x = X[r + c * P];
      Vb[PHT_IDX(r, c, ldv)] = (c >= k || r < c) ? 0.0 : (r == c ? 1.0 : x);
      A[PHT_IDX(i0 + r, jb + c, lda)] = (r <= c) ? x : 0.0;
    }
    // Gram entries G(a,b) = v_a . v_b (a < b < k), kept in registers
    const int npairs = k * (k - 1) / 2;
    double g[32];
#pragma unroll
    for (int t = 0; t < 32; ++t) {
      const int e = tid + t * kThreads;
      double s = 0.0;
      if (e < npairs) {
        // decode e -> (a, b), pairs ordered by b then a
        int b = (int)((1.0 + sqrt(1.0 + 8.0 * e)) * 0.5);
        while (b * (b - 1) / 2 > e) --b;
        while ((b + 1) * b / 2 <= e) ++b;
        const int a = e - b * (b - 1) / 2;
        s = X[b + a * P];
        for (int r = b + 1; r < len; ++r) s += X[r + a * P] * X[r + b * P];
      }
      g[t] = s;
    }
    __syncthreads();
    double* TG = X;  // nb x nb, ld nb
    for (int e = tid; e < nb * nb; e += kThreads) TG[e] = 0.0;
    __syncthreads();
#pragma unroll
    for (int t = 0; t < 32; ++t) {
      const int e = tid + t * kThreads;
      if (e < npairs) {
        int b = (int)((1.0 + sqrt(1.0 + 8.0 * e)) * 0.5);
        while (b * (b - 1) / 2 > e) --b;
        while ((b + 1) * b / 2 <= e) ++b;
        const int a = e - b * (b - 1) / 2;
        TG[a + b * nb] = g[t];
      }
    }
    __syncthreads();
    // DLARFT 'F','C': T(0:b-1,b) = -tau_b T(0:b-1,0:b-1) G(0:b-1,b)
    for (int b = 0; b < k; ++b) {
      if (tid < b) gcol[tid] = TG[tid + b * nb];
      __syncthreads();
      const double tb = tau[b];
      if (tid < b) {
        double s = 0.0;
        for (int c = tid; c < b; ++c) s += TG[tid + c * nb] * gcol[c];
        TG[tid + b * nb] = (tb == 0.0) ? 0.0 : -tb * s;
      }
      if (tid == 0) TG[b + b * nb] = tb;
      __syncthreads();
    }
    double* Tb = Tt + (long long)bl * tstride;
    for (int e = tid; e < nb * nb; e += kThreads) {
      const int r = e % nb, c = e / nb;
      Tb[PHT_IDX(r, c, ldt)] = (r <= c && c < k) ? TG[r + c * nb] : 0.0;
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// X: len x len (ld len), Y: nb x len (ld nb); both may be LDS or global.
__global__ __launch_bounds__(kThreads) void k_s1_rq_gen(
    const double* __restrict__ B, int ldb, int i0, int len, int nb,
    double* __restrict__ Vout, int ldv, double* __restrict__ Tout, int ldt,
    double* gX, double* gY, int x_in_lds, int y_in_lds) {
  extern __shared__ double sm[];
  double* red = sm;         // 32
  double* tau = red + 32;   // len
  double* tau2 = tau + len; // nb
  double* gcol = tau2 + nb; // max(len, nb)
  double* ldsfree = gcol + (len > nb ? len : nb);
  double* X = x_in_lds ? ldsfree : gX;
  double* Y = y_in_lds ? (x_in_lds ? ldsfree + (size_t)len * len : ldsfree) : gY;
  const int tid = threadIdx.x;
  const int ldx = len;

  for (int e = tid; e < len * len; e += kThreads) {
    const int r = e % len, c = e / len;
    X[r + c * ldx] = B[PHT_IDX(i0 + r, i0 + c, ldb)];
  }
  __syncthreads();
  // DGERQ2 (square): reflector i from row i, unit at column i
  for (int i = len - 1; i >= 0; --i) {
    const double alpha = X[i + i * ldx];
    const double xnorm = wg_norm(X + i, i, ldx, red);
    const Refl R = make_refl(alpha, xnorm);
    if (R.tau != 0.0)
      for (int c = tid; c < i; c += kThreads) X[i + c * ldx] *= R.scal;
    __syncthreads();
    if (tid == 0) {
      X[i + i * ldx] = R.beta;
      tau[i] = R.tau;
    }
    if (R.tau != 0.0) {
      for (int r = tid; r < i; r += kThreads) {
        double w = X[r + i * ldx];
        for (int c = 0; c < i; ++c) w += X[r + c * ldx] * X[i + c * ldx];
        w *= R.tau;
        X[r + i * ldx] -= w;
        for (int c = 0; c < i; ++c) X[r + c * ldx] -= w * X[i + c * ldx];
      }
    }
    __syncthreads();
  }

  if (len > nb) {
    // Y = first nb rows of Q = H(0) H(1) ... H(len-1)
    for (int e = tid; e < nb * len; e += kThreads) {
      const int a = e % nb, c = e / nb;
      Y[a + c * nb] = (a == c) ? 1.0 : 0.0;
    }
    __syncthreads();
    for (int i = 0; i < len; ++i) {
      const double t = tau[i];
      if (t != 0.0) {
        for (int a = tid; a < nb; a += kThreads) {
          double w = Y[a + i * nb];
          for (int c = 0; c < i; ++c) w += Y[a + c * nb] * X[i + c * ldx];
          w *= t;
          Y[a + i * nb] -= w;
          for (int c = 0; c < i; ++c) Y[a + c * nb] -= w * X[i + c * ldx];
        }
      }
      __syncthreads();
    }
    // DGELQ2 on Y (nb x len)
    for (int i = 0; i < nb; ++i) {
      const double alpha = Y[i + i * nb];
      const double xnorm = wg_norm(Y + i + (i + 1) * nb, len - i - 1, nb, red);
      const Refl R = make_refl(alpha, xnorm);
      if (R.tau != 0.0)
        for (int c = i + 1 + tid; c < len; c += kThreads) Y[i + c * nb] *= R.scal;
      __syncthreads();
      if (tid == 0) {
        Y[i + i * nb] = R.beta;
        tau2[i] = R.tau;
      }
      if (R.tau != 0.0) {
        for (int a = i + 1 + tid; a < nb; a += kThreads) {
          double w = Y[a + i * nb];
          for (int c = i + 1; c < len; ++c) w += Y[a + c * nb] * Y[i + c * nb];
          w *= R.tau;
          Y[a + i * nb] -= w;
          for (int c = i + 1; c < len; ++c) Y[a + c * nb] -= w * Y[i + c * nb];
        }
      }
      __syncthreads();
    }
    // explicit V (len x nb): v_i(c) = 0 (c<i), 1 (c=i), Y(i,c) (c>i)
    for (int e = tid; e < len * nb; e += kThreads) {
      const int c = e % len, i = e / len;
      Vout[PHT_IDX(c, i, ldv)] = c < i ? 0.0 : (c == i ? 1.0 : Y[i + c * nb]);
    }
    // DLARFT 'F','R': T upper, column b from G(a,b) = v_a . v_b
    for (int e = tid; e < nb * nb; e += kThreads) Tout[PHT_IDX(e % nb, e / nb, ldt)] = 0.0;
    __syncthreads();
    for (int b = 0; b < nb; ++b) {
      if (tid < b) {
        double s = Y[tid + b * nb];
        for (int c = b + 1; c < len; ++c) s += Y[tid + c * nb] * Y[b + c * nb];
        gcol[tid] = s;
      }
      __syncthreads();
      const double tb = tau2[b];
      double s = 0.0;
      if (tid < b)
        for (int c = tid; c < b; ++c) s += Tout[PHT_IDX(tid, c, ldt)] * gcol[c];
      __syncthreads();
      if (tid < b) Tout[PHT_IDX(tid, b, ldt)] = (tb == 0.0) ? 0.0 : -tb * s;
      if (tid == 0) Tout[PHT_IDX(b, b, ldt)] = tb;
      __syncthreads();
    }
  } else {
    const int k = len;
    // explicit V (backward rowwise): v_i(c) = X(i,c) (c<i), 1 (c=i), 0 (c>i)
    for (int e = tid; e < len * k; e += kThreads) {
      const int c = e % len, i = e / len;
      Vout[PHT_IDX(c, i, ldv)] = c < i ? X[i + c * ldx] : (c == i ? 1.0 : 0.0);
    }
    for (int e = tid; e < k * k; e += kThreads) Tout[PHT_IDX(e % k, e / k, ldt)] = 0.0;
    __syncthreads();
    // DLARFT 'B','R': T lower; T(i+1:k,i) = -tau_i T(i+1:k,i+1:k) G(i+1:k,i)
    for (int i = k - 1; i >= 0; --i) {
      const int l = i + 1 + tid;
      if (l < k) {
        double s = X[l + i * ldx];
        for (int c = 0; c < i; ++c) s += X[l + c * ldx] * X[i + c * ldx];
        gcol[l] = s;
      }
      __syncthreads();
      const double ti = tau[i];
      double s = 0.0;
      if (l < k)
        for (int m = i + 1; m <= l; ++m) s += Tout[PHT_IDX(l, m, ldt)] * gcol[m];
      __syncthreads();
      if (l < k) Tout[PHT_IDX(l, i, ldt)] = (ti == 0.0) ? 0.0 : -ti * s;
      if (tid == 0) Tout[PHT_IDX(i, i, ldt)] = ti;
      __syncthreads();
    }
  }
}

__global__ void k_zero_below(double* B, int ldb, int i0, int len, int ncols) {
  const int c = blockIdx.x;
  if (c >= ncols) return;
  const int j = i0 + c;
  for (int r = j + 1 + threadIdx.x; r < i0 + len; r += blockDim.x)
    B[PHT_IDX(r, j, ldb)] = 0.0;
}

// One workgroup per panel: the panel's blocks, after the blocks of the
// panels before it.
__global__ void k_s1_block_lists(ApplyBlock* __restrict__ lists, int nd, int n, int nb, int p) {
  const int panel = blockIdx.x, jb0 = panel * nb;
  int first = 0;
  for (int q = 0; q < panel; ++q) first += s1_nblocks(n, nb, p, q * nb);
  const int nblocks = s1_nblocks(n, nb, p, jb0);
  const int P = p * nb;
  for (int bl = threadIdx.x; bl < nblocks; bl += blockDim.x) {
    const S1Block b = s1_block(n, nb, p, jb0, nblocks, bl);
    const long long voff = (long long)bl * P * nb, toff = (long long)bl * nb * nb;
    const int e = first + bl;
    lists[e] = {b.i0, b.len, b.k, b.i0, n, 0, voff, toff};                   // B left
    lists[nd + e] = {b.i0, b.len, b.k, jb0 + nb, n, 0, voff, toff};          // A left
    lists[2 * nd + e] = {b.i0, b.len, b.k, 0, n, 0, voff, toff};             // Q
    lists[3 * nd + e] = {b.i0, b.len, b.kr, 0, n, 0, voff, toff};            // A, Z right
    lists[4 * nd + e] = {b.i0, b.len, b.kr, 0, b.i0 + b.len, 0, voff, toff}; // B right
  }
}

hipError_t s1_block_lists(hipStream_t st, ApplyBlock* lists, int nd, int n, int nb, int p) {
  const int np = s1_npanels(n, nb);
  if (np <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_s1_block_lists, dim3(np), dim3(128), 0, st, lists, nd, n, nb, p);
  return hipGetLastError();
}

// One workgroup per QR panel of B (qr_block_lists in kernels.hpp); the block
// rows are given in B's own coordinates (the panel kernels see them nb rows
// lower).
__global__ void k_qr_block_lists(ApplyBlock* __restrict__ lists, int nd, int n, int nb, int p) {
  const int panel = blockIdx.x;
  int first = 0;
  for (int q = 0; q < panel; ++q) first += s1_nblocks(n + nb, nb, p, qr_panel_col(n, nb, q));
  const int jb = qr_panel_col(n, nb, panel);
  const int nblocks = s1_nblocks(n + nb, nb, p, jb);
  const int P = p * nb;
  for (int bl = threadIdx.x; bl < nblocks; bl += blockDim.x) {
    const S1Block b = s1_block(n + nb, nb, p, jb, nblocks, bl);
    const long long voff = (long long)bl * P * nb, toff = (long long)bl * nb * nb;
    const int e = first + bl, i0 = b.i0 - nb;
    lists[e] = {i0, b.len, b.k, jb + nb, n, 0, voff, toff};  // B, trailing columns
    lists[nd + e] = {i0, b.len, b.k, 0, n, 0, voff, toff};   // A
    lists[2 * nd + e] = {i0, b.len, b.k, 0, n, 0, voff, toff};  // Q
  }
}

hipError_t qr_block_lists(hipStream_t st, ApplyBlock* lists, int nd, int n, int nb, int p) {
  const int np = qr_npanels(n, nb);
  if (np <= 0 || nd <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_qr_block_lists, dim3(np), dim3(128), 0, st, lists, nd, n, nb, p);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
static constexpr size_t kMaxLds = 160 * 1024;

bool s1_supported(int nb, int p) {
  if (nb <= 64 && (size_t)p * nb <= 128) return true;  // register-resident generators
  if (nb > 128) return false;
  // the generic panel QR keeps a P x nb block in LDS
  return ((size_t)p * nb * nb + 2 * (size_t)nb + 32) * sizeof(double) <= kMaxLds;
}

hipError_t s1_panel_qr(hipStream_t st, double* A, int lda, int n, int jb,
                       int nb, int p, int nblocks, double* Vt, int ldv,
                       long long vstride, double* Tt, int ldt,
                       long long tstride) {
  if (nb > 128) return hipErrorInvalidValue;
  const size_t P = (size_t)p * nb;
  const size_t lds = (P * nb + 2 * (size_t)nb + 32) * sizeof(double);
  if (lds > kMaxLds) return hipErrorInvalidValue;
  hipError_t e = hipFuncSetAttribute((const void*)k_s1_panel_qr,
                                     hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (int)kMaxLds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_s1_panel_qr, dim3(1), dim3(kThreads), lds, st, A, lda, n,
                     jb, nb, p, nblocks, Vt, ldv, vstride, Tt, ldt, tstride);
  return hipGetLastError();
}

hipError_t s1_rq_gen(hipStream_t st, const double* B, int ldb, int i0,
                     int len, int nb, double* V, int ldv, double* T, int ldt,
                     double* scratch) {
  const size_t base = (32 + (size_t)len + nb + (size_t)(len > nb ? len : nb)) * sizeof(double);
  const size_t xs = (size_t)len * len * sizeof(double);
  const size_t ys = (len > nb) ? (size_t)nb * len * sizeof(double) : 0;
  int x_lds = 0, y_lds = 0;
  size_t lds = base;
  if (base + xs + ys <= kMaxLds) {
    x_lds = y_lds = 1;
    lds = base + xs + ys;
  } else if (base + xs <= kMaxLds) {
    x_lds = 1;
    lds = base + xs;
  } else if (base + ys <= kMaxLds) {
    y_lds = 1;
    lds = base + ys;
  }
  hipError_t e = hipFuncSetAttribute((const void*)k_s1_rq_gen,
                                     hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (int)kMaxLds);
  if (e != hipSuccess) return e;
  double* gX = scratch;
  double* gY = scratch + (size_t)len * len;
  hipLaunchKernelGGL(k_s1_rq_gen, dim3(1), dim3(kThreads), lds, st, B, ldb, i0,
                     len, nb, V, ldv, T, ldt, gX, gY, x_lds, y_lds);
  return hipGetLastError();
}

hipError_t zero_below(hipStream_t st, double* B, int ldb, int i0, int len,
                      int ncols) {
  if (ncols <= 0 || len <= 1) return hipSuccess;
  hipLaunchKernelGGL(k_zero_below, dim3(ncols), dim3(128), 0, st, B, ldb, i0,
                     len, ncols);
  return hipGetLastError();
}

}  // namespace pht
