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
    int nblocks, int other_begin) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int o0 = other_begin + (blockIdx.x * 4 + wave) * 16;

  for (int b = 0; b < nblocks; ++b) {
    const ApplyBlock bk = blocks[b];
    if ((o0 < bk.hi) && (o0 + 16 > bk.lo))
      apply_block_slice<SIDE, NKT>(C, ldc, V, ldv, T, ldt, bk, o0, lane);
    // consecutive blocks overlap in the span dimension: make this wave's
    // stores visible before the next block's loads
    __syncthreads();
  }
}

template <int SIDE>
static hipError_t launch_side(hipStream_t st, double* C, int ldc,
                              const double* V, int ldv, const double* T,
                              int ldt, const ApplyBlock* dblocks, int nblocks,
                              int kmax, int other_begin, int other_end) {
  const int extent = other_end - other_begin;
  if (extent <= 0 || nblocks <= 0) return hipSuccess;
  dim3 grid((extent + 63) / 64), block(256);
  const int nkt = (kmax + 15) / 16;
  switch (nkt) {
#define PHT_CASE(N)                                                         \
  case N:                                                                   \
    hipLaunchKernelGGL((k_apply_blocks<SIDE, N>), grid, block, 0, st, C, ldc, \
                       V, ldv, T, ldt, dblocks, nblocks, other_begin);       \
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

hipError_t apply_blocks(hipStream_t st, Side side, double* C, int ldc,
                        const double* V, int ldv, const double* T, int ldt,
                        const ApplyBlock* dblocks, int nblocks, int kmax,
                        int other_begin, int other_end) {
  if (side == kLeftT)
    return launch_side<kLeftT>(st, C, ldc, V, ldv, T, ldt, dblocks, nblocks,
                               kmax, other_begin, other_end);
  return launch_side<kRightN>(st, C, ldc, V, ldv, T, ldt, dblocks, nblocks,
                              kmax, other_begin, other_end);
}

}  // namespace pht
