// Compact-WY block-reflector applier on FP64 MFMA (v_mfma_f64_16x16x4_f64).
//
// Replaces the reference's DLARFB call sites of both stages:
//   left  C <- H^T C : ab2rht.f90:161,202; rht2ht.f90:1680,1748,1759,1862,1940,1945
//   right C <- C H   : ab2rht.f90:241,327,335,359,362,412,415,455,458;
//                      rht2ht.f90:1841,1917,1982,2015
// with H = I - V T V^T given explicitly (V: len x k, T: k x k).
//
// Work split: a workgroup of 4 waves owns 64 consecutive indices of the
// "other" dimension (columns for a left update, rows for a right update);
// each wave owns 16 of them and is independent of the other waves.  For every
// block (applied in list order, which is the order the reference applies
// them) a wave computes
//   W  = V^T C_slice          (k x 16, reduction over the block span)
//   W2 = T^T W
//   C_slice -= V W2
// with 16x16 accumulator tiles.  Element i of a lane's accumulator holds row
// 4*i + lane/16 of its tile, so the i-th MFMA of a 16-deep chunk takes its
// reduction indices 4*i + slot from register i directly: the reduction index
// is permuted within each chunk and no data moves across lanes.
#include "apply_dev.hpp"

namespace pht {

template <int SIDE, int NKT>
__global__ __launch_bounds__(256) void k_apply_blocks(
    double* __restrict__ C, int ldc, const double* __restrict__ V, int ldv,
    const double* __restrict__ T, int ldt, const ApplyBlock* __restrict__ blocks,
    int nblocks, int other_begin, int other_end) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int o0 = other_begin + (blockIdx.x * 4 + wave) * 16;

  for (int b = 0; b < nblocks; ++b) {
    ApplyBlock bk = blocks[b];
    // the launch's range [other_begin, other_end) bounds every block's (the
    // last workgroup's rows past other_end belong to another launch)
    bk.lo = bk.lo > other_begin ? bk.lo : other_begin;
    bk.hi = bk.hi < other_end ? bk.hi : other_end;
    if ((o0 < bk.hi) && (o0 + 16 > bk.lo))
      apply_block_slice<SIDE, NKT>(C, ldc, V, ldv, T, ldt, bk, o0, lane);
    // consecutive blocks overlap in the span dimension: make this wave's
    // stores visible before the next block's loads
    __syncthreads();
  }
}

// LDS-staged variant for blocks of len <= 128 rows and k <= 64 reflectors
// (every stage-1 shape of the register-resident generators).  Per block the
// workgroup stages V and T in LDS with coalesced loads (one batch of loads
// per thread, issued together with the wave's C slice), each wave keeps its
// 16 x len slice of C in registers between the two MFMA phases, and all
// operands of the MFMAs come from LDS.  Row strides of 130 and 66 doubles
// (both 2 mod 32) keep the phase-1 and phase-2 operand reads free of bank
// conflicts.
constexpr int kStageLen = 128, kStageK = 64;
constexpr int kLdsV = 130, kLdsT = 66;

template <int SIDE, int NKT>
__global__ __launch_bounds__(256) void k_apply_blocks_lds(
    double* __restrict__ C, int ldc, const double* __restrict__ V, int ldv,
    const double* __restrict__ T, int ldt, const ApplyBlock* __restrict__ blocks,
    int nblocks, int other_begin, int other_end, int zcols) {
  __shared__ double Vs[kStageK * kLdsV];  // V(s, m) at m * kLdsV + s
  __shared__ double Ts[kStageK * kLdsT];  // T(kk, m) at m * kLdsT + kk
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int l16 = lane & 15;
  const int q = lane >> 4;
  const int o0 = other_begin + (blockIdx.x * 4 + wave) * 16;
  const int oi = o0 + l16;

  for (int b = 0; b < nblocks; ++b) {
    ApplyBlock bk = blocks[b];
    // the launch's range [other_begin, other_end) bounds every block's (the
    // last workgroup's rows past other_end belong to another launch)
    bk.lo = bk.lo > other_begin ? bk.lo : other_begin;
    bk.hi = bk.hi < other_end ? bk.hi : other_end;
    const int len = bk.len, k = bk.k, st = bk.start;
    const bool active = (o0 < bk.hi) && (o0 + 16 > bk.lo);  // uniform per wave
    const bool omask = (oi >= bk.lo) && (oi < bk.hi);
    const double* Vb = V + bk.voff;
    const double* Tb = T + bk.toff;
    const int nct = (len + 15) >> 4;

    // one batch of global loads: V, T (staging) and this wave's C slice
    double vr[kStageLen * kStageK / 256];
#pragma unroll
    for (int j = 0; j < kStageLen * kStageK / 256; ++j) {
      const int e = tid + 256 * j;
      const int m = e >> 7, sr = e & 127;
      vr[j] = (sr < len && m < k) ? Vb[PHT_IDX(sr, m, ldv)] : 0.0;
    }
    double tr[kStageK * kStageK / 256];
#pragma unroll
    for (int j = 0; j < kStageK * kStageK / 256; ++j) {
      const int e = tid + 256 * j;
      const int m = e >> 6, kk = e & 63;
      tr[j] = (kk < k && m < k) ? Tb[PHT_IDX(kk, m, ldt)] : 0.0;
    }
    double c[8][4];
#pragma unroll
    for (int ct = 0; ct < 8; ++ct)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int sidx = 16 * ct + 4 * i + q;
        c[ct][i] = (active && omask && sidx < len) ? C[cidx<SIDE>(st + sidx, oi, ldc)] : 0.0;
      }
#pragma unroll
    for (int j = 0; j < kStageLen * kStageK / 256; ++j) {
      const int e = tid + 256 * j;
      Vs[(e >> 7) * kLdsV + (e & 127)] = vr[j];
    }
#pragma unroll
    for (int j = 0; j < kStageK * kStageK / 256; ++j) {
      const int e = tid + 256 * j;
      Ts[(e >> 6) * kLdsT + (e & 63)] = tr[j];
    }
    __syncthreads();

    if (active) {
      // phase 1: W = V^T C_slice
      dbl4 W[NKT];
#pragma unroll
      for (int t = 0; t < NKT; ++t) W[t] = dbl4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int ct = 0; ct < 8; ++ct) {
        if (ct < nct) {
#pragma unroll
          for (int kt = 0; kt < NKT; ++kt)
#pragma unroll
            for (int i = 0; i < 4; ++i)
              W[kt] = mfma16(Vs[(16 * kt + l16) * kLdsV + 16 * ct + 4 * i + q], c[ct][i], W[kt]);
        }
      }
      // phase 2: W2 = T^T W
      dbl4 W2[NKT];
#pragma unroll
      for (int mt = 0; mt < NKT; ++mt) {
        W2[mt] = dbl4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int kt = 0; kt < NKT; ++kt)
#pragma unroll
          for (int i = 0; i < 4; ++i)
            W2[mt] = mfma16(Ts[(16 * mt + l16) * kLdsT + 16 * kt + 4 * i + q], W[kt][i], W2[mt]);
      }
      // phase 3: C_slice -= V W2, from the registers of phase 1
#pragma unroll
      for (int ct = 0; ct < 8; ++ct) {
        if (ct < nct) {
          dbl4 acc = dbl4{c[ct][0], c[ct][1], c[ct][2], c[ct][3]};
#pragma unroll
          for (int kt = 0; kt < NKT; ++kt)
#pragma unroll
            for (int i = 0; i < 4; ++i)
              acc = mfma16(-Vs[(16 * kt + 4 * i + q) * kLdsV + 16 * ct + l16], W2[kt][i], acc);
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            const int sidx = 16 * ct + 4 * i + q;
            // zcols (right side): the entries below the diagonal of the
            // block's first zcols columns are stored as exact zeros
            const bool z = SIDE == kRightN && sidx < zcols && oi > st + sidx && oi < st + len;
            if (omask && sidx < len) C[cidx<SIDE>(st + sidx, oi, ldc)] = z ? 0.0 : acc[i];
          }
        }
      }
    }
    // LDS is restaged for the next block; consecutive blocks overlap in the
    // span dimension, so this wave's stores must land before its next loads
    __syncthreads();
  }
}

template <int SIDE>
static hipError_t launch_side(hipStream_t st, double* C, int ldc,
                              const double* V, int ldv, const double* T,
                              int ldt, const ApplyBlock* dblocks, int nblocks,
                              int kmax, int max_len, int other_begin, int other_end,
                              int zcols) {
  const int extent = other_end - other_begin;
  if (extent <= 0 || nblocks <= 0) return hipSuccess;
  dim3 grid((extent + 63) / 64), block(256);
  const int nkt = (kmax + 15) / 16;
  if (max_len <= kStageLen && kmax <= kStageK) {
    switch (nkt) {
#define PHT_CASE(N)                                                              \
  case N:                                                                        \
    hipLaunchKernelGGL((k_apply_blocks_lds<SIDE, N>), grid, block, 0, st, C, ldc, \
                       V, ldv, T, ldt, dblocks, nblocks, other_begin, other_end, zcols); \
    return hipGetLastError();
      PHT_CASE(1)
      PHT_CASE(2)
      PHT_CASE(3)
      PHT_CASE(4)
#undef PHT_CASE
      default:
        break;
    }
  }
  switch (nkt) {
#define PHT_CASE(N)                                                         \
  case N:                                                                   \
    hipLaunchKernelGGL((k_apply_blocks<SIDE, N>), grid, block, 0, st, C, ldc, \
                       V, ldv, T, ldt, dblocks, nblocks, other_begin, other_end); \
    break;
    PHT_CASE(1)
    PHT_CASE(2)
    PHT_CASE(3)
    PHT_CASE(4)
    PHT_CASE(5)
    PHT_CASE(6)
    PHT_CASE(7)
    PHT_CASE(8)
#undef PHT_CASE
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

bool apply_fuses_zero(int kmax, int max_len) {
  return max_len <= kStageLen && kmax <= kStageK && kmax >= 1;
}

hipError_t apply_blocks(hipStream_t st, Side side, double* C, int ldc,
                        const double* V, int ldv, const double* T, int ldt,
                        const ApplyBlock* dblocks, int nblocks, int kmax, int max_len,
                        int other_begin, int other_end, int zero_cols) {
  if (zero_cols > 0 && (side != kRightN || !apply_fuses_zero(kmax, max_len)))
    return hipErrorInvalidValue;
  if (side == kLeftT)
    return launch_side<kLeftT>(st, C, ldc, V, ldv, T, ldt, dblocks, nblocks,
                               kmax, max_len, other_begin, other_end, 0);
  return launch_side<kRightN>(st, C, ldc, V, ldv, T, ldt, dblocks, nblocks,
                              kmax, max_len, other_begin, other_end, zero_cols);
}

}  // namespace pht
