// Register-resident stage-1 generators for blocks of at most 128 rows and
// 64 reflectors (P = p*nb <= 128, nb <= 64), the shapes of the benchmark
// configurations (nb = 64, p = 2; nb = 16, p = 8).  Same mathematics as the
// generic kernels in stage1.hip (DGEQR2 / DGERQ2 + DORGRQ + DGELQ2 / DLARFT
// semantics, src/ab2rht.f90:105-116,290-311); the difference is data
// placement: the block lives in registers, the reflector being applied is
// broadcast through a double-buffered LDS vector (one barrier per reflector),
// and the triangular factor T is obtained column by column from
// T = D (I - triu(G,1) T), D = diag(tau), G = V^T V.
#include "device_util.hpp"
#include "kernels.hpp"

namespace pht {

namespace {
constexpr int kT = 256;

// sum over the 2 (xor 1) or 4 (xor 1, 2) adjacent lanes of a group
__device__ __forceinline__ double grp2_sum(double v) { return v + __shfl_xor(v, 1, 64); }
__device__ __forceinline__ double grp4_sum(double v) {
  v += __shfl_xor(v, 1, 64);
  return v + __shfl_xor(v, 2, 64);
}

// ||x|| from a sum of squares, falling back to a scaled evaluation by the
// owners of the vector when the plain sum over/underflowed.
__device__ __forceinline__ bool ss_safe(double ss) { return ss > 1e-290 && ss < 1e290; }

// T (k x k, ld ldt, upper triangular; zeros stored below) of H = H(0)...H(k-1)
// = I - V T V^T from tau[] and the strictly upper Gram matrix G (ld 65):
// column j: T(j,j) = tau_j, T(i,j) = -tau_i sum_{l=i+1..j} G(i,l) T(l,j).
__device__ void t_from_gram(const double* G, const double* tau, int k,
                            double* Tl, double* Tout, int ldt, int kfull) {
  const int j = threadIdx.x;
  if (j < k) {
    double* tj = Tl + j * 65;
    tj[j] = tau[j];
    for (int i = j - 1; i >= 0; --i) {
      double s = 0.0;
      for (int l = i + 1; l <= j; ++l) s += G[i * 65 + l] * tj[l];
      tj[i] = -tau[i] * s;
    }
  }
  __syncthreads();
  for (int e = threadIdx.x; e < kfull * kfull; e += kT) {
    const int r = e % kfull, c = e / kfull;
    Tout[PHT_IDX(r, c, ldt)] = (r <= c && c < k) ? Tl[c * 65 + r] : 0.0;
  }
}
}  // namespace

// ---------------------------------------------------------------------------
// Left generation of one panel (ab2rht.f90:93-118), blocks bottom-up.
// Thread t owns column t/4 of the block and rows 32*(t%4) .. +31.
__global__ __launch_bounds__(kT) void k_s1_panel_qr_fast(
    double* __restrict__ A, int lda, int n, int jb, int nb, int p, int nblocks,
    double* __restrict__ Vt, int ldv, long long vstride, double* __restrict__ Tt,
    int ldt, long long tstride) {
  __shared__ double vbuf[2][128];
  __shared__ double tau[64];
  __shared__ double G[64 * 65];
  __shared__ double Tl[64 * 65];
  __shared__ double Vs[128 * 65];
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
      const int r = q * 32 + j;
      X[j] = (col < nb && r < len) ? A[PHT_IDX(i0 + r, jb + col, lda)] : 0.0;
    }
    // DGEQR2: reflector i from column i (owned by threads 4i..4i+3)
    for (int i = 0; i < k; ++i) {
      double* buf = vbuf[i & 1];
      if (col == i) {
        double ss = 0.0, al = 0.0;
#pragma unroll
        for (int j = 0; j < 32; ++j) {
          const int r = q * 32 + j;
          if (r > i) ss += X[j] * X[j];
          if (r == i) al = X[j];
        }
        ss = grp4_sum(ss);
        al = grp4_sum(al);
        double xnorm;
        if (ss_safe(ss) || ss == 0.0) {
          xnorm = sqrt(ss);
        } else {
          double mx = 0.0;
#pragma unroll
          for (int j = 0; j < 32; ++j)
            if (q * 32 + j > i) mx = fmax(mx, fabs(X[j]));
          mx = fmax(mx, __shfl_xor(mx, 1, 64));
          mx = fmax(mx, __shfl_xor(mx, 2, 64));
          double s2 = 0.0;
#pragma unroll
          for (int j = 0; j < 32; ++j)
            if (q * 32 + j > i && mx > 0.0) {
              const double t = X[j] / mx;
              s2 += t * t;
            }
          s2 = grp4_sum(s2);
          xnorm = mx * sqrt(s2);
        }
        const Refl R = make_refl(al, xnorm);
#pragma unroll
        for (int j = 0; j < 32; ++j) {
          const int r = q * 32 + j;
          if (r > i) {
            if (R.tau != 0.0) X[j] *= R.scal;
            buf[r] = X[j];
          } else if (r == i) {
            X[j] = R.beta;
            buf[r] = 1.0;
          } else {
            buf[r] = 0.0;
          }
        }
        if (q == 0) tau[i] = R.tau;
      }
      __syncthreads();
      const double t = tau[i];
      if (col > i && col < nb && t != 0.0) {
        double w = 0.0;
#pragma unroll
        for (int j = 0; j < 32; ++j) {
          const int r = q * 32 + j;
          if (r >= i && r < len) w += buf[r] * X[j];
        }
        w = grp4_sum(w) * t;
#pragma unroll
        for (int j = 0; j < 32; ++j) {
          const int r = q * 32 + j;
          if (r >= i && r < len) X[j] -= w * buf[r];
        }
      }
    }
    // V (explicit, to LDS and the tape), R and zeros back into A
    double* Vb = Vt + (long long)bl * vstride;
    if (col < nb) {
#pragma unroll
      for (int j = 0; j < 32; ++j) {
        const int r = q * 32 + j;
        if (r < len) {
          const double v = (col >= k || r < col) ? 0.0 : (r == col ? 1.0 : X[j]);
          Vs[r * 65 + col] = v;
          Vb[PHT_IDX(r, col, ldv)] = v;
          A[PHT_IDX(i0 + r, jb + col, lda)] = (r <= col) ? X[j] : 0.0;
        }
      }
    }
    __syncthreads();
    // Gram G(a,b) = v_a . v_b, a < b < k
    for (int e = tid; e < k * k; e += kT) {
      const int a = e % k, b = e / k;
      if (a < b) {
        double s = 0.0;
        for (int r = b; r < len; ++r) s += Vs[r * 65 + a] * Vs[r * 65 + b];
        G[a * 65 + b] = s;
      }
    }
    __syncthreads();
    t_from_gram(G, tau, k, Tl, Tt + (long long)bl * tstride, ldt, nb);
    // the next block (above) re-reads the R rows just written to A
    __threadfence();
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Right generator for one B block with nb < len <= 128 (ab2rht.f90:290-303).
// RQ phase: thread t owns row t/2, columns 64*(t%2) .. +63.
// Y / LQ phase: thread t owns row t/4 of Y = Q(1:nb,:), columns 32*(t%4)..+31.
__global__ __launch_bounds__(kT) void k_s1_rq_fast(
    const double* __restrict__ B, int ldb, int i0, int len, int nb,
    double* __restrict__ Vout, int ldv, double* __restrict__ Tout, int ldt) {
  extern __shared__ double sm[];
  double* Vr = sm;                   // packed RQ reflectors: row i at i(i-1)/2
  double* Ys = Vr + 128 * 129 / 2;   // final LQ reflectors, [64][129]
  double* vbuf = Ys + 64 * 129;      // [2][128]
  double* taus = vbuf + 256;         // 128
  double* tau2 = taus + 128;         // 64
  double* G = Vr;                    // [64][65], reuses Vr after the Y phase
  double* Tl = Ys;                   // [64][65], reuses Ys once G is formed
  const int tid = threadIdx.x;

  // ---- DGERQ2 on the len x len block
  {
    const int row = tid >> 1, h = tid & 1;
    double X[64];
#pragma unroll
    for (int j = 0; j < 64; ++j) {
      const int c = h * 64 + j;
      X[j] = (row < len && c < len) ? B[PHT_IDX(i0 + row, i0 + c, ldb)] : 0.0;
    }
    for (int i = len - 1; i >= 0; --i) {
      double* buf = vbuf + (i & 1) * 128;
      if (row == i) {
        double ss = 0.0, al = 0.0;
#pragma unroll
        for (int j = 0; j < 64; ++j) {
          const int c = h * 64 + j;
          if (c < i) ss += X[j] * X[j];
          if (c == i) al = X[j];
        }
        ss = grp2_sum(ss);
        al = grp2_sum(al);
        double xnorm;
        if (ss_safe(ss) || ss == 0.0) {
          xnorm = sqrt(ss);
        } else {
          double mx = 0.0;
#pragma unroll
          for (int j = 0; j < 64; ++j)
            if (h * 64 + j < i) mx = fmax(mx, fabs(X[j]));
          mx = fmax(mx, __shfl_xor(mx, 1, 64));
          double s2 = 0.0;
#pragma unroll
          for (int j = 0; j < 64; ++j)
            if (h * 64 + j < i && mx > 0.0) {
              const double t = X[j] / mx;
              s2 += t * t;
            }
          s2 = grp2_sum(s2);
          xnorm = mx * sqrt(s2);
        }
        const Refl R = make_refl(al, xnorm);
        double* vi = Vr + i * (i - 1) / 2;
#pragma unroll
        for (int j = 0; j < 64; ++j) {
          const int c = h * 64 + j;
          if (c < i) {
            if (R.tau != 0.0) X[j] *= R.scal;
            buf[c] = X[j];
            vi[c] = X[j];
          } else if (c == i) {
            buf[c] = 1.0;
          }
        }
        if (h == 0) taus[i] = R.tau;
      }
      __syncthreads();
      const double t = taus[i];
      if (row < i && t != 0.0) {
        double w = 0.0;
#pragma unroll
        for (int j = 0; j < 64; ++j) {
          const int c = h * 64 + j;
          if (c <= i) w += X[j] * buf[c];
        }
        w = grp2_sum(w) * t;
#pragma unroll
        for (int j = 0; j < 64; ++j) {
          const int c = h * 64 + j;
          if (c <= i) X[j] -= w * buf[c];
        }
      }
    }
  }
  __syncthreads();

  // ---- Y = E^T H(0) H(1) ... H(len-1) (the first nb rows of the RQ's Q)
  const int a = tid >> 2, q = tid & 3;
  double Y[32];
#pragma unroll
  for (int j = 0; j < 32; ++j) {
    const int c = q * 32 + j;
    Y[j] = (a < nb && c == a) ? 1.0 : 0.0;
  }
  if (a < nb) {
    for (int i = 1; i < len; ++i) {  // H(0) is the identity (length-1 reflector)
      const double t = taus[i];
      if (t == 0.0) continue;
      const double* vi = Vr + i * (i - 1) / 2;
      double w = 0.0;
#pragma unroll
      for (int j = 0; j < 32; ++j) {
        const int c = q * 32 + j;
        if (c < i)
          w += Y[j] * vi[c];
        else if (c == i)
          w += Y[j];
      }
      w = grp4_sum(w) * t;
#pragma unroll
      for (int j = 0; j < 32; ++j) {
        const int c = q * 32 + j;
        if (c < i)
          Y[j] -= w * vi[c];
        else if (c == i)
          Y[j] -= w;
      }
    }
  }
  // ---- DGELQ2 on Y (nb x len)
  for (int i = 0; i < nb; ++i) {
    double* buf = vbuf + (i & 1) * 128;
    if (a == i) {
      double ss = 0.0, al = 0.0;
#pragma unroll
      for (int j = 0; j < 32; ++j) {
        const int c = q * 32 + j;
        if (c > i && c < len) ss += Y[j] * Y[j];
        if (c == i) al = Y[j];
      }
      ss = grp4_sum(ss);
      al = grp4_sum(al);
      const double xnorm = sqrt(ss);  // rows of an orthogonal factor: |entries| <= 1
      const Refl R = make_refl(al, xnorm);
#pragma unroll
      for (int j = 0; j < 32; ++j) {
        const int c = q * 32 + j;
        if (c > i) {
          if (R.tau != 0.0) Y[j] *= R.scal;
          buf[c] = (c < len) ? Y[j] : 0.0;
        } else if (c == i) {
          Y[j] = R.beta;
          buf[c] = 1.0;
        } else {
          buf[c] = 0.0;
        }
      }
      if (q == 0) tau2[i] = R.tau;
    }
    __syncthreads();
    const double t = tau2[i];
    if (a > i && a < nb && t != 0.0) {
      double w = 0.0;
#pragma unroll
      for (int j = 0; j < 32; ++j) {
        const int c = q * 32 + j;
        if (c >= i && c < len) w += Y[j] * buf[c];
      }
      w = grp4_sum(w) * t;
#pragma unroll
      for (int j = 0; j < 32; ++j) {
        const int c = q * 32 + j;
        if (c >= i && c < len) Y[j] -= w * buf[c];
      }
    }
  }
  // explicit V (len x nb): v_a(c) = 0 (c<a), 1 (c=a), Y(a,c) (c>a)
  if (a < nb) {
#pragma unroll
    for (int j = 0; j < 32; ++j) {
      const int c = q * 32 + j;
      if (c < len) {
        const double v = c < a ? 0.0 : (c == a ? 1.0 : Y[j]);
        Ys[a * 129 + c] = v;
        Vout[PHT_IDX(c, a, ldv)] = v;
      }
    }
  }
  __syncthreads();
  // Gram G(a,b) = v_a . v_b (a < b), then T ('F','R': H = H(0)...H(nb-1))
  for (int e = tid; e < nb * nb; e += kT) {
    const int x = e % nb, y = e / nb;
    if (x < y) {
      double s = 0.0;
      for (int c = y; c < len; ++c) s += Ys[x * 129 + c] * Ys[y * 129 + c];
      G[x * 65 + y] = s;
    }
  }
  __syncthreads();
  t_from_gram(G, tau2, nb, Tl, Tout, ldt, nb);
}

hipError_t s1_panel_qr_fast(hipStream_t st, double* A, int lda, int n, int jb,
                            int nb, int p, int nblocks, double* Vt, int ldv,
                            long long vstride, double* Tt, int ldt,
                            long long tstride) {
  if (nb > 64 || p * nb > 128) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_s1_panel_qr_fast, dim3(1), dim3(kT), 0, st, A, lda, n, jb,
                     nb, p, nblocks, Vt, ldv, vstride, Tt, ldt, tstride);
  return hipGetLastError();
}

hipError_t s1_rq_fast(hipStream_t st, const double* B, int ldb, int i0, int len,
                      int nb, double* V, int ldv, double* T, int ldt) {
  if (nb > 64 || len > 128 || len <= nb) return hipErrorInvalidValue;
  const size_t lds = (size_t)(128 * 129 / 2 + 64 * 129 + 256 + 128 + 64) * sizeof(double);
  static bool attr = false;
  if (!attr) {
    hipError_t e = hipFuncSetAttribute((const void*)k_s1_rq_fast,
                                       hipFuncAttributeMaxDynamicSharedMemorySize,
                                       160 * 1024);
    if (e != hipSuccess) return e;
    attr = true;
  }
  hipLaunchKernelGGL(k_s1_rq_fast, dim3(1), dim3(kT), lds, st, B, ldb, i0, len, nb,
                     V, ldv, T, ldt);
  return hipGetLastError();
}

}  // namespace pht
