// pht_profile: a small driver for hardware-counter runs (rocprofv3 --pmc)
// of the library, without Python or torch in the process.
//
//   pht_profile copy [doubles]          plain copy kernel, three launches of
//                                       `doubles` FP64 words (default 2^27:
//                                       1 GiB read + 1 GiB written each)
//   pht_profile reduce n r p nq         one reduction (paraht_reduce_dev) of a
//                                       pencil filled on the device: A with
//                                       uniform entries in [-1, 1), B upper
//                                       triangular with uniform entries and a
//                                       diagonal in [1, 2), Q = Z = I
//
// The pencil's values do not change which entries the kernels read and
// write, so the counters of one reduction are those of the benchmark's
// pencil of the same (n, r, p, nq).  Prints one line per step and exits with
// the library's status.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "../../include/paraht.h"

namespace {

__global__ void k_copy(const double* __restrict__ src, double* __restrict__ dst, size_t n) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += stride)
    dst[i] = src[i];
}

// counter-based uniform numbers in [0, 1): splitmix64 of (seed, index)
__device__ __forceinline__ double uniform01(unsigned long long seed, unsigned long long i) {
  unsigned long long z = seed + (i + 1) * 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

// which = 0: A, 1: B (upper triangular), 2: identity; column-major, ld = n
__global__ void k_fill(double* M, int n, int which, unsigned long long seed) {
  const size_t total = (size_t)n * n;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < total; e += stride) {
    const int i = (int)(e % n), j = (int)(e / n);
    double v;
    if (which == 2)
      v = (i == j) ? 1.0 : 0.0;
    else if (which == 1)
      v = (i > j) ? 0.0 : (i == j ? 1.0 + uniform01(seed, e) : 2.0 * uniform01(seed, e) - 1.0);
    else
      v = 2.0 * uniform01(seed, e) - 1.0;
    M[e] = v;
  }
}

int fail(const char* what, hipError_t e) {
  fprintf(stderr, "pht_profile: %s: %s\n", what, hipGetErrorString(e));
  return 1;
}

int run_copy(size_t n) {
  double *src = nullptr, *dst = nullptr;
  hipError_t e;
  if ((e = hipMalloc(&src, n * sizeof(double))) != hipSuccess) return fail("hipMalloc", e);
  if ((e = hipMalloc(&dst, n * sizeof(double))) != hipSuccess) return fail("hipMalloc", e);
  if ((e = hipMemset(src, 0, n * sizeof(double))) != hipSuccess) return fail("hipMemset", e);
  for (int rep = 0; rep < 3; ++rep) {
    hipLaunchKernelGGL(k_copy, dim3(4096), dim3(256), 0, 0, src, dst, n);
    if ((e = hipDeviceSynchronize()) != hipSuccess) return fail("copy", e);
    printf("copy %d: %zu bytes read + %zu bytes written\n", rep, n * sizeof(double),
           n * sizeof(double));
    fflush(stdout);
  }
  (void)hipFree(src);
  (void)hipFree(dst);
  return 0;
}

int run_reduce(int n, int r, int p, int nq) {
  double* M[4] = {nullptr, nullptr, nullptr, nullptr};
  hipError_t e;
  for (int k = 0; k < 4; ++k)
    if ((e = hipMalloc(&M[k], (size_t)n * n * sizeof(double))) != hipSuccess)
      return fail("hipMalloc", e);
  const int which[4] = {0, 1, 2, 2};
  for (int k = 0; k < 4; ++k)
    hipLaunchKernelGGL(k_fill, dim3(2048), dim3(256), 0, 0, M[k], n, which[k], 1302ull + k);
  if ((e = hipDeviceSynchronize()) != hipSuccess) return fail("fill", e);
  printf("pencil n=%d filled\n", n);
  fflush(stdout);
  paraht_ctx* ctx = nullptr;
  int st = paraht_ctx_create(&ctx, 0);
  if (st) {
    fprintf(stderr, "pht_profile: paraht_ctx_create: %s\n", paraht_status_string(st));
    return 2;
  }
  st = paraht_reduce_dev(ctx, n, M[0], n, M[1], n, M[2], n, M[3], n, r, p, nq, 0);
  printf("reduction n=%d r=%d p=%d nq=%d: %s\n", n, r, p, nq, paraht_status_string(st));
  fflush(stdout);
  paraht_ctx_destroy(ctx);
  for (int k = 0; k < 4; ++k) (void)hipFree(M[k]);
  return st ? 2 : 0;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc >= 2 && strcmp(argv[1], "copy") == 0)
    return run_copy(argc >= 3 ? strtoull(argv[2], nullptr, 10) : (size_t)1 << 27);
  if (argc == 6 && strcmp(argv[1], "reduce") == 0)
    return run_reduce(atoi(argv[2]), atoi(argv[3]), atoi(argv[4]), atoi(argv[5]));
  fprintf(stderr, "usage: pht_profile copy [doubles] | pht_profile reduce n r p nq\n");
  return 64;
}
