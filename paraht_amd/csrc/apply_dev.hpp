// Device code of the compact-WY block-reflector applier, shared by the
// multi-workgroup kernel (applier.hip) and the batched persistent kernel.
#pragma once
#include "device_util.hpp"
#include "kernels.hpp"

namespace pht {

template <int SIDE>
__device__ __forceinline__ size_t cidx(int span, int other, int ldc) {
  return SIDE == kLeftT ? PHT_IDX(span, other, ldc) : PHT_IDX(other, span, ldc);
}

// One wave applies block bk to the 16 indices o0..o0+15 of the other
// dimension (columns for kLeftT, rows for kRightN) of C, on FP64 MFMA:
//   W = V^T C_slice,  W2 = T^T W,  C_slice -= V W2.
// Element i of a lane's accumulator holds row 4*i + lane/16 of its tile, so
// the i-th MFMA of a 16-deep chunk takes its reduction indices 4*i + slot
// from register i directly.  NKT = ceil(k/16).
template <int SIDE, int NKT>
__device__ __forceinline__ void apply_block_slice(double* __restrict__ C, int ldc,
                                                  const double* __restrict__ V, int ldv,
                                                  const double* __restrict__ T, int ldt,
                                                  const ApplyBlock& bk, int o0, int lane) {
  const int l16 = lane & 15;
  const int q = lane >> 4;
  const int oi = o0 + l16;
  const bool omask = (oi >= bk.lo) && (oi < bk.hi);
  const double* Vb = V + bk.voff;
  const double* Tb = T + bk.toff;
  const int len = bk.len, k = bk.k, st = bk.start;
  const int nct = (len + 15) >> 4;

  dbl4 W[NKT];
#pragma unroll
  for (int t = 0; t < NKT; ++t) W[t] = dbl4{0.0, 0.0, 0.0, 0.0};

  // phase 1: W = V^T C_slice
  for (int ct = 0; ct < nct; ++ct) {
    double c[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int s = 16 * ct + 4 * i + q;
      c[i] = (omask && s < len) ? C[cidx<SIDE>(st + s, oi, ldc)] : 0.0;
    }
#pragma unroll
    for (int kt = 0; kt < NKT; ++kt) {
      const int m = 16 * kt + l16;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int s = 16 * ct + 4 * i + q;
        const double a = (s < len && m < k) ? Vb[PHT_IDX(s, m, ldv)] : 0.0;
        W[kt] = mfma16(a, c[i], W[kt]);
      }
    }
  }
  // phase 2: W2 = T^T W
  dbl4 W2[NKT];
#pragma unroll
  for (int mt = 0; mt < NKT; ++mt) {
    W2[mt] = dbl4{0.0, 0.0, 0.0, 0.0};
    const int m = 16 * mt + l16;
#pragma unroll
    for (int kt = 0; kt < NKT; ++kt) {
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int kk = 16 * kt + 4 * i + q;
        const double a = (kk < k && m < k) ? Tb[PHT_IDX(kk, m, ldt)] : 0.0;
        W2[mt] = mfma16(a, W[kt][i], W2[mt]);
      }
    }
  }
  // phase 3: C_slice -= V W2
  for (int ct = 0; ct < nct; ++ct) {
    dbl4 acc;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int s = 16 * ct + 4 * i + q;
      acc[i] = (omask && s < len) ? C[cidx<SIDE>(st + s, oi, ldc)] : 0.0;
    }
    const int srow = 16 * ct + l16;
#pragma unroll
    for (int kt = 0; kt < NKT; ++kt) {
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int m = 16 * kt + 4 * i + q;
        const double a =
            (srow < len && m < k) ? -Vb[PHT_IDX(srow, m, ldv)] : 0.0;
        acc = mfma16(a, W2[kt][i], acc);
      }
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int s = 16 * ct + 4 * i + q;
      if (omask && s < len) C[cidx<SIDE>(st + s, oi, ldc)] = acc[i];
    }
  }
}

}  // namespace pht
