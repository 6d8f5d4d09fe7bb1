// Device-side helpers shared by the gfx950 kernels of libparaht_hip.so.
// All matrices are FP64 column-major; indices inside kernels are 0-based.
#pragma once
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cstdint>

namespace pht {

typedef double dbl4 __attribute__((ext_vector_type(4)));

#define PHT_IDX(i, j, ld) ((size_t)(j) * (size_t)(ld) + (size_t)(i))

// One v_mfma_f64_16x16x4_f64: D(16x16) += A(16x4) * B(4x16).
// Lane l supplies A[l%16][l/16] and B[l/16][l%16]; D[4*i + l/16][l%16] is
// element i of the lane's dbl4 (checked on gfx950).
__device__ __forceinline__ dbl4 mfma16(double a, double b, dbl4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  return v;
}

// Sum over the whole workgroup (blockDim.x a multiple of 64, <= 1024).
// red must hold >= blockDim.x/64 doubles of LDS.  Contains barriers.
__device__ __forceinline__ double block_sum(double v, double* red) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int nw = blockDim.x >> 6;
  v = wave_sum(v);
  __syncthreads();
  if (lane == 0) red[wave] = v;
  __syncthreads();
  double s = 0.0;
  for (int w = 0; w < nw; ++w) s += red[w];
  return s;
}

__device__ __forceinline__ double block_max(double v, double* red) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int nw = blockDim.x >> 6;
  v = wave_max(v);
  __syncthreads();
  if (lane == 0) red[wave] = v;
  __syncthreads();
  double s = 0.0;
  for (int w = 0; w < nw; ++w) s = fmax(s, red[w]);
  return s;
}

// Householder reflector from alpha and ||x|| (DLARFG semantics:
// beta = -sign(alpha)*||[alpha;x]||, tau = (beta-alpha)/beta,
// v = x * scal with scal = 1/(alpha-beta); tau = 0 and scal = 1 when x = 0).
// Tiny |beta| is handled by rescaling with 1/safmin as DLARFG specifies.
struct Refl {
  double tau, scal, beta;
};

__device__ __forceinline__ double dlapy2_d(double x, double y) {
  double xa = fabs(x), ya = fabs(y);
  double w = fmax(xa, ya), z = fmin(xa, ya);
  if (z == 0.0 || w > DBL_MAX) return w;
  double q = z / w;
  return w * sqrt(1.0 + q * q);
}

__device__ __forceinline__ Refl make_refl(double alpha, double xnorm) {
  Refl R;
  if (xnorm == 0.0) {
    R.tau = 0.0;
    R.scal = 1.0;
    R.beta = alpha;
    return R;
  }
  double beta = -copysign(dlapy2_d(alpha, xnorm), alpha);
  const double safmin = DBL_MIN / (DBL_EPSILON * 0.5);
  double a = alpha, xn = xnorm, mult = 1.0;
  int knt = 0;
  if (fabs(beta) < safmin) {
    const double rsafmn = 1.0 / safmin;
    do {
      ++knt;
      xn *= rsafmn;
      beta *= rsafmn;
      a *= rsafmn;
      mult *= rsafmn;
    } while (fabs(beta) < safmin && knt < 20);
    beta = -copysign(dlapy2_d(a, xn), a);
  }
  R.tau = (beta - a) / beta;
  R.scal = mult / (a - beta);  // applied to the unscaled x
  for (int j = 0; j < knt; ++j) beta *= safmin;
  R.beta = beta;
  return R;
}

// Plane rotation [c s; -s c][f; g] = [r; 0] with c >= 0 and r carrying the
// sign of f (Anderson, "Algorithm 978", ACM TOMS 2017).
__device__ __forceinline__ void dlartg_d(double f, double g, double& c,
                                         double& s, double& r) {
  const double safmin = DBL_MIN;
  const double safmax = 1.0 / DBL_MIN;
  const double rtmin = 1.4916681462400413e-154;  // sqrt(DBL_MIN)
  const double rtmax = 9.4807519081091774e+153;  // sqrt(safmax/2)
  double f1 = fabs(f), g1 = fabs(g);
  if (g == 0.0) {
    c = 1.0;
    s = 0.0;
    r = f;
  } else if (f == 0.0) {
    c = 0.0;
    s = copysign(1.0, g);
    r = g1;
  } else if (f1 > rtmin && f1 < rtmax && g1 > rtmin && g1 < rtmax) {
    double d = sqrt(f * f + g * g);
    c = f1 / d;
    r = copysign(d, f);
    s = g / r;
  } else {
    double u = fmin(safmax, fmax(safmin, fmax(f1, g1)));
    double fs = f / u, gs = g / u;
    double d = sqrt(fs * fs + gs * gs);
    c = fabs(fs) / d;
    double rr = copysign(d, f);
    s = gs / rr;
    r = rr * u;
  }
}

// value of x in lane `src` of the wave (src wave-uniform, ideally a constant)
__device__ __forceinline__ double readlane_d(double x, int src) {
  const unsigned long long u = __builtin_bit_cast(unsigned long long, x);
  const unsigned lo = __builtin_amdgcn_readlane((unsigned)u, src);
  const unsigned hi = __builtin_amdgcn_readlane((unsigned)(u >> 32), src);
  return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

// Givens rotation with dlartg's conventions (c >= 0, r with the sign of f,
// g == 0 gives the identity) without division or square root: f and g are
// scaled by the power of two that brings the larger magnitude into
// [0.5, 1) (exact), 1/d comes from v_rsq_f64 (relative error below 2^-24)
// refined by two Newton steps (below 2^-52), and c = |f|/d, s = g/r, r as
// products.  c, s, r agree with dlartg to a few ulp for all finite inputs.
__device__ __forceinline__ void rot_rsq(double f, double g, double& c, double& s, double& r) {
  const double m = fmax(fabs(f), fabs(g));
  const int k = (m > 0.0) ? __builtin_amdgcn_frexp_exp(m) : 0;
  const double fs = ldexp(f, -k), gs = ldexp(g, -k);
  const double d2 = fma(fs, fs, gs * gs);  // in [0.25, 2] unless f = g = 0
  double y = __builtin_amdgcn_rsq(d2);
  double e = fma(-d2 * y, y, 1.0);
  y = fma(0.5 * y, e, y);
  e = fma(-d2 * y, y, 1.0);
  y = fma(0.5 * y, e, y);
  const double sg = (f < 0.0) ? -y : y;
  c = (g == 0.0) ? 1.0 : fabs(fs) * y;
  s = (g == 0.0) ? 0.0 : gs * sg;
  r = (g == 0.0) ? f : ldexp(d2 * sg, k);
}

// Cross-workgroup hand-off.  An agent-scope acquire or release on gfx950
// also invalidates (buffer_inv sc1) or writes back (buffer_wbl2 sc1) the
// L2 of the XCD it runs on, which every other kernel on that XCD pays for.
// So the spin reads the counter with relaxed agent-scope loads (coherent,
// no invalidation per poll), thread 0 alone issues one acquire fence once
// every wait is over (acquire_waits), and a step is published by one
// release store from thread 0 after every wave's stores have completed
// (publish).  The workgroup barrier between thread 0 and the other waves
// carries the ordering to them.

// Thread 0 spins (bounded by `limit` polls) until *p >= target; a wait that
// gives up sets *errflag and lets the workgroup go on, so every wave exits.
__device__ __forceinline__ void wait_at_least(int* p, int target, int* errflag, int limit) {
  if (threadIdx.x == 0) {
    int spins = 0;
    while (__hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
      __builtin_amdgcn_s_sleep(2);
      if (++spins > limit) {
        __hip_atomic_store(errflag, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        break;
      }
    }
  }
}

// After thread 0's waits: one agent-scope acquire, then the workgroup barrier.
__device__ __forceinline__ void acquire_waits() {
  if (threadIdx.x == 0) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  __syncthreads();
}

// Every wave waits for its own global stores to complete (reach L2), the
// barrier orders that before thread 0's release store, whose L2 write-back
// then carries all of them.
__device__ __forceinline__ void publish(int* p, int value) {
  asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
  __syncthreads();
  if (threadIdx.x == 0) __hip_atomic_store(p, value, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

}  // namespace pht
