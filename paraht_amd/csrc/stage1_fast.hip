// Register-resident stage-1 generators (kernels); the device code is in
// s1_dev.hpp, shared with the batched persistent kernel.
#include "s1_dev.hpp"

namespace pht {
using namespace s1;

__global__ __launch_bounds__(kT) void k_s1_panel_qr_fast(
    double* __restrict__ A, int lda, int n, int jb, int nb, int p, int nblocks,
    double* __restrict__ Vt, int ldv, long long vstride, double* __restrict__ Tt,
    int ldt, long long tstride) {
  extern __shared__ double sm[];
  panel_qr_body(sm, A, lda, n, jb, nb, p, nblocks, Vt, ldv, vstride, Tt, ldt, tstride);
}

__global__ __launch_bounds__(kT) void k_s1_panel_qr_tau(
    double* __restrict__ A, int lda, int n, int jb, int nb, int p, int nblocks,
    double* __restrict__ Vt, int ldv, long long vstride, double* __restrict__ tau_out) {
  extern __shared__ double sm[];
  panel_qr_body(sm, A, lda, n, jb, nb, p, nblocks, Vt, ldv, vstride, nullptr, 0, 0, tau_out);
}

__global__ __launch_bounds__(kT) void k_s1_panel_t(
    int n, int jb, int nb, int p, int nblocks, const double* __restrict__ Vt, int ldv,
    long long vstride, const double* __restrict__ tau_in, double* __restrict__ Tt, int ldt,
    long long tstride) {
  extern __shared__ double sm[];
  panel_t_body(sm, n, jb, nb, p, nblocks, Vt, ldv, vstride, tau_in, Tt, ldt, tstride);
}

__global__ __launch_bounds__(kT) void k_s1_rq_fast(
    const double* __restrict__ B, int ldb, int i0, int len, int nb,
    double* __restrict__ Vout, int ldv, double* __restrict__ Tout, int ldt) {
  extern __shared__ double sm[];
  rq_gen_body(sm, B, ldb, i0, len, nb, Vout, ldv, Tout, ldt);
}

namespace {
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
                            long long tstride, double* tau_ws) {
  if (nb > 64 || p * nb > 128) return hipErrorInvalidValue;
  if (tau_ws == nullptr) {
    static bool attr = false;
    if (hipError_t e = allow_lds((const void*)k_s1_panel_qr_fast, kPanelLds, attr)) return e;
    hipLaunchKernelGGL(k_s1_panel_qr_fast, dim3(1), dim3(kT), kPanelLds, st, A, lda, n, jb,
                       nb, p, nblocks, Vt, ldv, vstride, Tt, ldt, tstride);
    return hipGetLastError();
  }
  // the chain of block QRs without the T matrices, then every block's T at once
  static bool attr_q = false, attr_t = false;
  if (hipError_t e = allow_lds((const void*)k_s1_panel_qr_tau, kPanelLds, attr_q)) return e;
  if (hipError_t e = allow_lds((const void*)k_s1_panel_t, kPanelLds, attr_t)) return e;
  hipLaunchKernelGGL(k_s1_panel_qr_tau, dim3(1), dim3(kT), kPanelLds, st, A, lda, n, jb, nb,
                     p, nblocks, Vt, ldv, vstride, tau_ws);
  if (hipError_t e = hipGetLastError()) return e;
  hipLaunchKernelGGL(k_s1_panel_t, dim3(nblocks), dim3(kT), kPanelLds, st, n, jb, nb, p,
                     nblocks, Vt, ldv, vstride, tau_ws, Tt, ldt, tstride);
  return hipGetLastError();
}

hipError_t s1_rq_fast(hipStream_t st, const double* B, int ldb, int i0, int len,
                      int nb, double* V, int ldv, double* T, int ldt) {
  if (nb > 64 || len > 128 || len < 1) return hipErrorInvalidValue;
  static bool attr = false;
  if (hipError_t e = allow_lds((const void*)k_s1_rq_fast, kRqLds, attr)) return e;
  hipLaunchKernelGGL(k_s1_rq_fast, dim3(1), dim3(kT), kRqLds, st, B, ldb, i0, len, nb,
                     V, ldv, T, ldt);
  return hipGetLastError();
}

}  // namespace pht
