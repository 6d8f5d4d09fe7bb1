// Reverse-mode adjoint of the Hessenberg-triangular reduction.
//
// For H = Q^T A Z, T = Q^T B Z (Q e1 = e1) the differentials are
//   dH = E_A - Om_Q H + H Om_Z,   dT = E_B - Om_Q T + T Om_Z,
// with E = Q^T dA Z, Om_Q = Q^T dQ and Om_Z = Z^T dZ skew, Om_Q e1 = 0.  The
// structure of dH (zero below the subdiagonal) and dT (zero below the
// diagonal) determines the strictly lower parts w_Q (column 0 zero), w_Z of
// Om_Q, Om_Z column by column, j = 0..n-2:
//   T(j+1:,j+1:) w_Z(j+1:,j) = -E_B(j+1:,j) + w_Q(j+1:,0:j) T(0:j,j)
//   w_Q(j+2:,j+1) = [E_A(j+2:,j) - w_Q(j+2:,0:j) H(0:j,j)
//                    + H(j+2:,j+1:) w_Z(j+1:,j)] / H(j+1,j).
// Given the cotangents g_Q, g_Z of w_Q, w_Z (the strictly lower parts of
// G - G^T, G_Q = Q^T Qbar - Hbar H^T - Tbar T^T, G_Z = Z^T Zbar + H^T Hbar +
// T^T Tbar, computed with GEMMs by the caller), this file runs the transposed
// recurrence, j = n-2..0, and writes X_A, X_B with
//   <g_Q, w_Q> + <g_Z, w_Z> = <X_A, E_A> + <X_B, E_B>:
//   a    = gbar_Q(j+2:, j+1) / H(j+1,j)            -> X_A(j+2:, j)
//   c_Z  = g_Z(j+1:, j) + H(j+2:, j+1:)^T a
//   b    = T(j+1:,j+1:)^{-T} c_Z                   -> X_B(j+1:, j) = -b
//   gbar_Q(j+1:, 0:j) += b T(0:j, j)^T - a H(0:j, j)^T   (rank-2 update)
// The caller then forms Abar = Q (hess(Hbar) + X_A) Z^T and
// Bbar = Q (triu(Tbar) + X_B) Z^T.
//
// Each step is three data-parallel phases (two GEMV-shaped, one rank-2
// update) separated by barriers.  k_adjoint_wg gives one workgroup to a
// pencil (batches of small pencils); k_adjoint_grid spreads one pencil over a
// cooperative grid with a bounded grid-wide barrier (large pencils), since the
// recurrence moves O(n^3) bytes and one compute unit's bandwidth is too small.
#include "device_util.hpp"
#include "kernels.hpp"

namespace pht {

namespace {
constexpr int kAT = 1024;           // threads of a workgroup-per-pencil group
constexpr int kGridT = 256;         // threads per workgroup of the grid variant
constexpr int kAdjSpinLimit = 1 << 26;
constexpr size_t kAdjWgMaxLds = 156 * 1024;  // dynamic LDS of k_adjoint_wg

// The execution group of one pencil: `size` threads with ranks 0..size-1 in
// waves of 64, and a barrier over all of them.
struct WgGroup {
  __device__ int rank() const { return threadIdx.x; }
  __device__ int size() const { return kAT; }
  __device__ void sync() const {
    __threadfence_block();
    __syncthreads();
  }
};

// Grid-wide barrier for a cooperative launch: an arrival counter and a
// generation number; the last arrival resets the counter and advances the
// generation.  Waits are bounded (errflag on overrun).
struct GridGroup {
  int* count;
  int* gen;
  int* errflag;
  __device__ int rank() const { return blockIdx.x * kGridT + threadIdx.x; }
  __device__ int size() const { return gridDim.x * kGridT; }
  __device__ void sync() const {
    __threadfence();
    __syncthreads();
    if (threadIdx.x == 0) {
      const int g = __hip_atomic_load(gen, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
      if (__hip_atomic_fetch_add(count, 1, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT) ==
          (int)gridDim.x - 1) {
        __hip_atomic_store(count, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(gen, g + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
      } else {
        int spins = 0;
        while (__hip_atomic_load(gen, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == g) {
          __builtin_amdgcn_s_sleep(1);
          if (++spins > kAdjSpinLimit) {
            __hip_atomic_store(errflag, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            break;
          }
        }
      }
    }
    __syncthreads();
  }
};

// The transposed recurrence for one pencil of order n (all matrices n x n,
// column-major, ld n).  GQ holds g_Q on entry and is updated in place; vec
// holds four n-vectors of scratch (LDS or global): a of this step and of the
// next (double-buffered: the rank-2 phase still reads this step's a while it
// produces the next), c_Z and b.  XA, XB are written on their strictly lower
// parts (rows >= j+2 / >= j+1 of column j).
template <typename Group>
__device__ void adjoint_pencil(const Group& grp, int n, const double* __restrict__ H,
                               const double* __restrict__ T, const double* __restrict__ Tinv,
                               double* __restrict__ GQ, const double* __restrict__ GZ,
                               double* __restrict__ XA, double* __restrict__ XB, double* vec) {
  if (n < 2) return;
  const int rk = grp.rank(), G = grp.size();
  const int lane = rk & 63, wave = rk >> 6, nwaves = G >> 6;
  double* av = vec;            // a of step j
  double* an = vec + n;        // a of step j-1
  double* cz = vec + 2 * n;
  double* bv = vec + 3 * n;
  // step j = n-2 has no rows >= j+2: a = 0
  for (int i = rk; i < n; i += G) av[i] = 0.0;
  grp.sync();
  for (int j = n - 2; j >= 0; --j) {
    // ---- c_Z(m) = g_Z(m, j) + sum_{i=j+2}^{min(n-1,m+1)} H(i, m) a(i), m >= j+1
    for (int m = j + 1 + wave; m < n; m += nwaves) {
      const double* hc = H + (size_t)m * n;
      double s = 0.0;
      const int iend = min(n - 1, m + 1);
      for (int i = j + 2 + lane; i <= iend; i += 64) s += hc[i] * av[i];
      s = wave_sum(s);
      if (lane == 0) cz[m] = GZ[(size_t)j * n + m] + s;
    }
    grp.sync();
    // ---- b(m') = sum_{m=j+1}^{m'} Tinv(m, m') c_Z(m), m' >= j+1
    for (int mp = j + 1 + wave; mp < n; mp += nwaves) {
      const double* tc = Tinv + (size_t)mp * n;
      double s = 0.0;
      for (int m = j + 1 + lane; m <= mp; m += 64) s += tc[m] * cz[m];
      s = wave_sum(s);
      if (lane == 0) bv[mp] = s;
    }
    grp.sync();
    // ---- outputs of column j; rank-2 update of gbar_Q(j+1:, 0:j); the next
    //      step's a from the updated column j
    const int rows = n - 1 - j;  // rows j+1..n-1
    const double hnext = (j >= 1) ? H[(size_t)(j - 1) * n + j] : 1.0;  // H(j, j-1)
    for (long long e = rk; e < (long long)rows * (j + 1); e += G) {
      const int i = j + 1 + (int)(e % rows), k = (int)(e / rows);
      const double ai = (i >= j + 2) ? av[i] : 0.0;
      const double g = GQ[(size_t)k * n + i] + bv[i] * T[(size_t)j * n + k] -
                       ai * H[(size_t)j * n + k];
      GQ[(size_t)k * n + i] = g;
      if (k == j && j >= 1) an[i] = g / hnext;  // a of step j-1, rows >= j+1
    }
    for (int i = j + 1 + rk; i < n; i += G) {
      XB[(size_t)j * n + i] = -bv[i];
      if (i >= j + 2) XA[(size_t)j * n + i] = av[i];
    }
    grp.sync();
    double* t = av;
    av = an;
    an = t;
  }
}

__global__ __launch_bounds__(kAT) void k_adjoint_wg(const BatchPencil* __restrict__ pencils,
                                                    int count, int* next, const double* H,
                                                    const double* T, const double* Tinv,
                                                    double* GQ, const double* GZ, double* XA,
                                                    double* XB) {
  extern __shared__ double vecs[];  // 4 vectors of the largest pencil
  __shared__ int cur;
  for (;;) {
    if (threadIdx.x == 0) cur = atomicAdd(next, 1);
    __syncthreads();
    const int idx = cur;
    __syncthreads();
    if (idx >= count) break;
    const BatchPencil pc = pencils[idx];
    const size_t o = (size_t)pc.offset;
    adjoint_pencil(WgGroup{}, pc.n, H + o, T + o, Tinv + o, GQ + o, GZ + o, XA + o, XB + o,
                   vecs);
    __syncthreads();
  }
}

__global__ __launch_bounds__(kGridT) void k_adjoint_grid(int n, const double* H, const double* T,
                                                         const double* Tinv, double* GQ,
                                                         const double* GZ, double* XA,
                                                         double* XB, double* vecs, int* bar,
                                                         int* errflag) {
  const GridGroup grp{bar, bar + 1, errflag};
  adjoint_pencil(grp, n, H, T, Tinv, GQ, GZ, XA, XB, vecs);
}

__global__ void k_adjoint_reset(int* ints, int count) {
  for (int i = threadIdx.x; i < count; i += blockDim.x) ints[i] = 0;
}
}  // namespace

size_t adjoint_wg_lds(int nmax) { return (size_t)4 * nmax * sizeof(double); }
size_t adjoint_wg_max_lds() { return kAdjWgMaxLds; }

hipError_t adjoint_batched(hipStream_t st, const BatchPencil* dpencils, int count, int nmax,
                           int* dnext, const double* H, const double* T, const double* Tinv,
                           double* GQ, const double* GZ, double* XA, double* XB, int grid) {
  const size_t lds = adjoint_wg_lds(nmax);
  if (lds > kAdjWgMaxLds) return hipErrorInvalidValue;
  // the dynamic-LDS limit is raised to what this launch needs (the kernel's
  // static LDS comes on top of it)
  static size_t allowed = 64 * 1024;
  if (lds > allowed) {
    hipError_t e = hipFuncSetAttribute((const void*)k_adjoint_wg,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    allowed = lds;
  }
  hipLaunchKernelGGL(k_adjoint_reset, dim3(1), dim3(64), 0, st, dnext, 1);
  hipLaunchKernelGGL(k_adjoint_wg, dim3(grid), dim3(kAT), lds, st, dpencils, count, dnext, H, T,
                     Tinv, GQ, GZ, XA, XB);
  return hipGetLastError();
}

hipError_t adjoint_grid(hipStream_t st, int n, const double* H, const double* T,
                        const double* Tinv, double* GQ, const double* GZ, double* XA,
                        double* XB, double* vecs, int* ints, int grid) {
  // ints: barrier counter, generation, error flag
  hipLaunchKernelGGL(k_adjoint_reset, dim3(1), dim3(64), 0, st, ints, 3);
  void* args[] = {&n, &H, &T, &Tinv, &GQ, &GZ, &XA, &XB, &vecs, &ints, (void*)nullptr};
  int* err = ints + 2;
  args[10] = &err;
  return hipLaunchCooperativeKernel((const void*)k_adjoint_grid, dim3(grid), dim3(kGridT), args,
                                    0, st);
}

int adjoint_grid_max_blocks(int device) {
  int per_cu = 0, ncu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)k_adjoint_grid, kGridT,
                                                   0) != hipSuccess)
    return 0;
  if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess)
    return 0;
  return per_cu * ncu;
}

}  // namespace pht
