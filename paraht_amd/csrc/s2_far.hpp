// Far updates of the stage-2 Givens chase (rht2ht_givens semantics,
// src/rht2ht.f90:1019-1325): the rotations of a sweep applied outside its
// generation window, to the rows of Q and Z, to the rows of A and B above
// the window and to the columns of A and B right of it, block by block in
// the reference's bottom-up order.  Kernels for compile-time (r, nq); the
// pairs are instantiated in s2_far_r<r>.hip (one translation unit per r).
//
// A launch can be limited to a range of rows (columns) per block and
// sequence (S2Range), which splits a sweep's A/B far update into the part the
// next sweep's generation reads (applied first, before that generation) and
// the rest (applied beside it).
#pragma once
#include <cstdlib>

#include "s2_dev.hpp"

namespace pht {
using namespace s2;

// far updates in the split wavefront form (PARAHT_S2_WAVEFRONT=0: the
// sequential form)
inline bool s2_wavefront() {
  static int v = -1;
  if (v < 0) {
    const char* e = getenv("PARAHT_S2_WAVEFRONT");
    v = (e && *e == '0') ? 0 : 1;
  }
  return v == 1;
}

// lanes per row (column) in the split wavefront kernels
constexpr int kLanes = 2;

// windows up to 80 entries in the wavefront kernels
constexpr bool wavefront_fits(int R, int NQ) { return R + NQ - 1 <= 80; }

// ---- ranges (S2Range, kernels.hpp) ----
// thr(x, icm): first row (1-based) of the generation window of sequence icm
// (0-based) of bulge block x, istart of rht2ht.f90:853,937
__device__ __forceinline__ int s2_thr(int s, int R, int NQ, int x, int icm) {
  return imax(s + 1, s + 1 + (x - NQ + icm) * R);
}
// row i (0-based) takes sequence icm of block ib
__device__ __forceinline__ bool row_in_range(int i, int s, int R, int NQ, int ib, int icm,
                                             const S2Range& rg) {
  return i < rg.row_end && i + 1 < s2_thr(s, R, NQ, ib - rg.hi_shift, icm) &&
         (rg.lo_shift < 0 || s2_thr(s, R, NQ, ib - rg.lo_shift, icm) <= i + 1);
}
// some row of [rf, rf+cnt) may take a sequence of block ib
__device__ __forceinline__ bool rows_meet_range(int rf, int cnt, int s, int R, int NQ, int ib,
                                                const S2Range& rg) {
  if (rf >= rg.row_end) return false;
  if (rf + 1 >= s2_thr(s, R, NQ, ib - rg.hi_shift, NQ - 1)) return false;
  if (rg.lo_shift >= 0 && rf + cnt < s2_thr(s, R, NQ, ib - rg.lo_shift, 0)) return false;
  return true;
}
// last column (1-based) of A (which = 0) or B (1) inside block ib's window:
// its left rotations go to the columns after it (rht2ht.f90:1046-1130)
__device__ __forceinline__ int cols_istop(int which, int s, int R, int NQ, int ib) {
  return (which == 0) ? ((ib == 1) ? s + NQ - 1 : s + (ib - 2) * R + NQ)
                      : s + (ib - 1) * R + R + NQ - 1;
}
// column j (0-based) takes block ib
__device__ __forceinline__ bool col_in_range(int j, int istop, int R, const S2Range& rg) {
  return j + 1 > istop + rg.c_lo * R && (rg.c_hi < 0 || j + 1 <= istop + rg.c_hi * R);
}
__device__ __forceinline__ bool cols_meet_range(int j0, int cnt, int istop, int R,
                                                const S2Range& rg) {
  if (j0 + cnt <= istop + rg.c_lo * R) return false;
  if (rg.c_hi >= 0 && j0 + 1 > istop + rg.c_hi * R) return false;
  return true;
}

// Register-window rotation sweeps for compile-time (r, nq).  A thread owns one
// row (k_s2_rows_v) or column (k_s2_cols_v) of a matrix.  For every bulge
// block it loads the block's whole window of R+NQ-1 entries at once (one
// memory latency per block), and the block's NQ x (R-1) rotation
// coefficients are staged in LDS by the workgroup.  The rotations of column
// ic act on window positions (ic+ig-2, ic+ig-1); the window of R registers
// slides by one position per ic (the NQ-1 entries beyond it wait in `ahead`),
// so every register index is static.  Out-of-range rotations are identities
// in the tape and out-of-range entries are zeros, so no bounds tests are
// needed inside the rotation loop.
namespace {
// The NQ x (R-1) rotations of one bulge block, staged in LDS by the
// workgroup: (c, s) of rotation (ic, ig) at cs[ic-1][ig] (ig >= 1).
template <int R, int NQ>
struct TapeBlock {
  double2 cs[NQ][R];
};

// sgn = -1 stages the transposed rotations (left rotations seen from the
// right, or the row form of a left rotation)
template <int R, int NQ, int NT>
__device__ __forceinline__ void stage_tape(TapeBlock<R, NQ>& tb, const double* __restrict__ Ct,
                                           const double* __restrict__ St, int ib, double sgn) {
  constexpr int LDR = R + NQ;
  constexpr int PER = (NQ * R + NT - 1) / NT;
  const size_t boff = (size_t)(ib - 1) * LDR * NQ;
  // all loads first, then the LDS writes: one memory latency per block
  double c[PER], sn[PER];
#pragma unroll
  for (int q = 0; q < PER; ++q) {
    const int e = threadIdx.x + q * NT;
    const int icm = e / R, ig = e % R;
    const size_t ti = boff + (size_t)icm * LDR + (icm + ig - 1);
    const bool ok = e < NQ * R && ig >= 1;
    c[q] = ok ? Ct[ti] : 1.0;
    sn[q] = ok ? St[ti] : 0.0;
  }
#pragma unroll
  for (int q = 0; q < PER; ++q) {
    const int e = threadIdx.x + q * NT;
    if (e < NQ * R) tb.cs[e / R][e % R] = make_double2(c[q], sgn * sn[q]);
  }
}

// copies entries [0, W) of a window of NW entries between a row/column of
// global memory (stride `stride` between entries) and an LDS column, in
// groups of 16 entries whose loads are all in flight together
template <int NW>
__device__ __forceinline__ void window_to_lds(double (*tile)[65], int t, const double* src,
                                              size_t stride, int W, bool ok) {
#pragma unroll
  for (int k0 = 0; k0 < NW; k0 += 16) {
    double tmp[16];
#pragma unroll
    for (int k = 0; k < 16; ++k)
      tmp[k] = (ok && k0 + k < W && k0 + k < NW) ? src[(k0 + k) * stride] : 0.0;
#pragma unroll
    for (int k = 0; k < 16; ++k)
      if (k0 + k < NW) tile[k0 + k][t] = tmp[k];
  }
}

// applies column ic of the block to the sliding window v: rotation ig maps
// (v[ig], v[ig-1]) to (c v[ig] + s v[ig-1], c v[ig-1] - s v[ig]).  The
// coefficient reads are fenced in groups of 8 rotations so that at most a
// group's worth of them is live.
template <int R, int NQ>
__device__ __forceinline__ void rotate_window(double (&v)[R], const TapeBlock<R, NQ>& tb,
                                              int icm) {
#pragma unroll
  for (int ig = R - 1; ig >= 1; --ig) {
    if (ig % 8 == 0) __builtin_amdgcn_sched_barrier(0);
    const double2 cs = tb.cs[icm][ig];
    const double x = v[ig], y = v[ig - 1];
    v[ig] = cs.x * x + cs.y * y;
    v[ig - 1] = cs.x * y - cs.y * x;
  }
}

// One block's window of one row, staged in the workgroup's LDS tile
// (tile[k][t] = entry k of the window of the workgroup's row t).
template <int R, int NQ, bool PARTIAL>
__device__ __forceinline__ void rows_one_block(double (*tile)[65], int i, bool valid, int s,
                                               int ib, int W, const TapeBlock<R, NQ>& tb,
                                               const S2Range& rg) {
  const int t = threadIdx.x;
  double v[R];
#pragma unroll
  for (int k = 0; k < R; ++k) v[k] = (k < W) ? tile[k][t] : 0.0;
#pragma unroll 1
  for (int icm = 0; icm < NQ; ++icm) {
    bool act = valid;
    if (PARTIAL) act = act && row_in_range(i, s, R, NQ, ib, icm, rg);
    if (act) rotate_window<R, NQ>(v, tb, icm);
    if (icm < W) tile[icm][t] = v[0];
#pragma unroll
    for (int k = 0; k < R - 1; ++k) v[k] = v[k + 1];
    v[R - 1] = (R + icm < W) ? tile[R + icm][t] : 0.0;
  }
#pragma unroll
  for (int k = 0; k < R - 1; ++k)
    if (NQ + k < W) tile[NQ + k][t] = v[k];
}
}  // namespace

constexpr int kRowsThreads = 64;

// AB: rows of A (blockIdx.y = 0) and B (1) above the generation window, right
// rotations; otherwise all rows of Q (0, left rotations seen from the right)
// and Z (1, right rotations).
template <int R, int NQ, bool AB>
__global__ __launch_bounds__(kRowsThreads) void k_s2_rows_v(
    double* __restrict__ A, int lda, double* __restrict__ B, int ldb,
    double* __restrict__ Q, int ldq, double* __restrict__ Z, int ldz, int n,
    int s, int nblocks, const double* __restrict__ CU,
    const double* __restrict__ SU, const double* __restrict__ CV,
    const double* __restrict__ SV, S2Range rg) {
  constexpr int NW = R + NQ - 1;
  __shared__ double tile[NW][65];
  __shared__ TapeBlock<R, NQ> tb;
  const int which = AB ? blockIdx.y : 2 + blockIdx.y;  // 0 A, 1 B, 2 Q, 3 Z
  const int row_first = blockIdx.x * kRowsThreads;
  const int t = threadIdx.x;
  const int i = row_first + t;  // 0-based row
  const bool valid = i < n;
  double* M = which == 0 ? A : which == 1 ? B : which == 2 ? Q : Z;
  const int ld = which == 0 ? lda : which == 1 ? ldb : which == 2 ? ldq : ldz;
  const double* Ct = which == 2 ? CU : CV;
  const double* St = which == 2 ? SU : SV;
  const double sgn = which == 2 ? -1.0 : 1.0;
  for (int ib = nblocks; ib >= 1; --ib) {
    const int cw0 = s + (ib - 1) * R;  // 0-based first window column
    const int W = imin(NW, n - cw0);
    if (W <= 0) continue;
    // A, B: only blocks whose row range meets the workgroup's rows
    if (AB && !rows_meet_range(row_first, kRowsThreads, s, R, NQ, ib, rg)) continue;
    __syncthreads();  // the previous block's tile and coefficients are consumed
    stage_tape<R, NQ, kRowsThreads>(tb, Ct, St, ib, sgn);
    window_to_lds<NW>(tile, t, M + PHT_IDX(i, cw0, ld), (size_t)ld, W, valid);
    __syncthreads();
    rows_one_block<R, NQ, AB>(tile, i, valid, s, ib, W, tb, rg);
    // the entries of the row's active sequences [icm0, icm1) go back
    int icm0 = valid ? 0 : NQ, icm1 = NQ;
    if (AB && valid) {
      icm0 = NQ;
      icm1 = 0;
      for (int icm = NQ - 1; icm >= 0; --icm)
        if (row_in_range(i, s, R, NQ, ib, icm, rg)) {
          icm0 = icm;
          if (icm1 == 0) icm1 = icm + 1;
        }
    }
    if (valid && icm0 < icm1)
      for (int k = icm0; k < imin(W, icm1 + R - 1); ++k) M[PHT_IDX(i, cw0 + k, ld)] = tile[k][t];
  }
}

// Columns of A (which = 0) or B (1) right of the generation window: left
// rotations on the rows of each block.  The 64 columns of a workgroup are
// staged through LDS so that global accesses run along the columns.
template <int R, int NQ>
__global__ __launch_bounds__(64) void k_s2_cols_v(
    double* __restrict__ A, int lda, double* __restrict__ B, int ldb, int n,
    int s, int nblocks, const double* __restrict__ CU,
    const double* __restrict__ SU, S2Range rg) {
  constexpr int NW = R + NQ - 1;
  static_assert(NW <= 128, "two 64-row chunks per column");
  __shared__ double tile[NW][65];
  __shared__ TapeBlock<R, NQ> tb;
  const int which = blockIdx.y;
  const int t = threadIdx.x;
  const int j0 = rg.col_begin + blockIdx.x * 64;
  const int j = j0 + t;
  double* M = which == 0 ? A : B;
  const int ld = which == 0 ? lda : ldb;
  for (int ib = nblocks; ib >= 1; --ib) {
    const int istop = cols_istop(which, s, R, NQ, ib);
    if (!cols_meet_range(j0, 64, istop, R, rg)) continue;
    const int rw0 = s + (ib - 1) * R;  // 0-based first row
    const int W = imin(NW, n - rw0);
    if (W <= 0) continue;
    const int cmax = imin(64, n - j0);
    __syncthreads();  // the previous block's tile and coefficients are consumed
    stage_tape<R, NQ, 64>(tb, CU, SU, ib, -1.0);
    // 8 columns x 2 row chunks per group, all 16 loads in flight together
#pragma unroll
    for (int c0 = 0; c0 < 64; c0 += 8) {
      double tmp[16];
#pragma unroll
      for (int q = 0; q < 16; ++q) {
        const int c = c0 + (q >> 1), k = t + 64 * (q & 1);
        tmp[q] = (c < cmax && k < W) ? M[PHT_IDX(rw0 + k, j0 + c, ld)] : 0.0;
      }
#pragma unroll
      for (int q = 0; q < 16; ++q) {
        const int c = c0 + (q >> 1), k = t + 64 * (q & 1);
        if (k < NW) tile[k][c] = tmp[q];
      }
    }
    __syncthreads();
    const bool act = (j < n) && col_in_range(j, istop, R, rg);
    double v[R];
#pragma unroll
    for (int k = 0; k < R; ++k) v[k] = (k < W) ? tile[k][t] : 0.0;
#pragma unroll 1
    for (int icm = 0; icm < NQ; ++icm) {
      // the row form of a left rotation: the transposed coefficients
      if (act) rotate_window<R, NQ>(v, tb, icm);
      if (icm < W) tile[icm][t] = v[0];
#pragma unroll
      for (int k = 0; k < R - 1; ++k) v[k] = v[k + 1];
      v[R - 1] = (R + icm < W) ? tile[R + icm][t] : 0.0;
    }
#pragma unroll
    for (int k = 0; k < R - 1; ++k)
      if (NQ + k < W) tile[NQ + k][t] = v[k];
    __syncthreads();
    for (int c = 0; c < cmax; ++c) {
      if (!col_in_range(j0 + c, istop, R, rg)) continue;
      for (int k = t; k < W; k += 64) M[PHT_IDX(rw0 + k, j0 + c, ld)] = tile[k][c];
    }
  }
}

// ---------------------------------------------------------------------------
// Wavefront form of the far updates.  A block's NQ rotation sequences act on
// a window of R+NQ-1 entries; sequence icm applies ig = R-1..1 to the pairs
// (icm+ig-1, icm+ig).  Rotation (icm+1, ig) only has to follow (icm, ig)
// (every other rotation of sequence icm that shares one of its entries comes
// earlier in that sequence), so rotation (icm, ig) can run at step
// tau = (R-1-ig) + icm, and the rotations of one step act on disjoint pairs:
// R+NQ-2 dependent steps instead of NQ (R-1).  Sequences below icm0 are
// skipped (identity rotations, exact).
// First and last step at which the wavefront touches window entry e:
// rotation (icm, ig) touches entries icm+ig-1 and icm+ig at step
// R-1-ig+icm, so as the upper entry of a pair e is touched for icm in
// [max(0, e-R+1), min(NQ-1, e-1)] at steps R-1-e+2 icm, as the lower entry
// for icm in [max(0, e+2-R), min(NQ-1, e)] at steps R-2-e+2 icm.
constexpr int wf_max(int a, int b) { return a > b ? a : b; }
constexpr int wf_min(int a, int b) { return a < b ? a : b; }
constexpr int wf_first(int R, int NQ, int e) {
  int f = 1 << 30;
  if (wf_max(0, e - R + 1) <= wf_min(NQ - 1, e - 1)) f = R - 1 - e + 2 * wf_max(0, e - R + 1);
  if (wf_max(0, e + 2 - R) <= wf_min(NQ - 1, e))
    f = wf_min(f, R - 2 - e + 2 * wf_max(0, e + 2 - R));
  return f;
}
constexpr int wf_last(int R, int NQ, int e) {
  int l = -1;
  if (wf_max(0, e - R + 1) <= wf_min(NQ - 1, e - 1)) l = R - 1 - e + 2 * wf_min(NQ - 1, e - 1);
  if (wf_max(0, e + 2 - R) <= wf_min(NQ - 1, e))
    l = wf_max(l, R - 2 - e + 2 * wf_min(NQ - 1, e));
  return l;
}

// Window entry e of a thread lives at col[e * LDT] of the LDS tile.  After
// step TAU, the entries whose last touch was step TAU are written back, then
// the entries first touched at step TAU+1 are read (one step ahead).  The
// writes come first: in the split form (several lanes per row, k_s2_rows_w2)
// one lane's last write of an entry is the other lane's first read of it.
template <int R, int NQ, int LDT, int TAU, int E>
__device__ __forceinline__ void wf_store(double (&w)[R + NQ - 1], double* col) {
  if constexpr (wf_last(R, NQ, E) == TAU) col[E * LDT] = w[E];
  if constexpr (E + 1 < R + NQ - 1) wf_store<R, NQ, LDT, TAU, E + 1>(w, col);
}
template <int R, int NQ, int LDT, int TAU, int E>
__device__ __forceinline__ void wf_load(double (&w)[R + NQ - 1], const double* col) {
  if constexpr (wf_first(R, NQ, E) == TAU + 1) w[E] = col[E * LDT];
  if constexpr (E + 1 < R + NQ - 1) wf_load<R, NQ, LDT, TAU, E + 1>(w, col);
}

// Coefficients of step TAU (sequence ICM's rotation ig = R-1-TAU+ICM) into
// cs[ICM]; entries of sequences without a rotation at that step are unused.
template <int R, int NQ, int TAU, int ICM>
__device__ __forceinline__ void wf_coeffs(double2 (&cs)[NQ], const TapeBlock<R, NQ>& tb) {
  constexpr int ig = R - 1 - TAU + ICM;
  if constexpr (ig >= 1 && ig <= R - 1) cs[ICM] = tb.cs[ICM][ig];
  if constexpr (ICM + 1 < NQ) wf_coeffs<R, NQ, TAU, ICM + 1>(cs, tb);
}

// The rotations of step TAU: (icm, ig = R-1-TAU+icm) for each sequence;
// sequences outside [icm0, icm1) are identities.
template <int R, int NQ, int TAU, int ICM>
__device__ __forceinline__ void wf_rotations(double (&w)[R + NQ - 1], const double2 (&cs)[NQ],
                                             int icm0, int icm1) {
  constexpr int ig = R - 1 - TAU + ICM;
  if constexpr (ig >= 1 && ig <= R - 1) {
    const bool on = ICM >= icm0 && ICM < icm1;
    const double c = on ? cs[ICM].x : 1.0, sn = on ? cs[ICM].y : 0.0;
    const double x = w[ICM + ig], y = w[ICM + ig - 1];
    w[ICM + ig] = c * x + sn * y;
    w[ICM + ig - 1] = c * y - sn * x;
  }
  if constexpr (ICM + 1 < NQ) wf_rotations<R, NQ, TAU, ICM + 1>(w, cs, icm0, icm1);
}

// Step TAU with the next step's coefficients and window entries read ahead.
template <int R, int NQ, int LDT, int TAU>
__device__ __forceinline__ void wf_steps(double (&w)[R + NQ - 1], double* col,
                                         const TapeBlock<R, NQ>& tb, int icm0, int icm1,
                                         const double2 (&cs)[NQ]) {
  double2 nx[NQ];
  if constexpr (TAU + 1 < R + NQ - 2) wf_coeffs<R, NQ, TAU + 1, 0>(nx, tb);
  wf_rotations<R, NQ, TAU, 0>(w, cs, icm0, icm1);
  wf_store<R, NQ, LDT, TAU, 0>(w, col);
  // In the split form another lane of the wave reads what these stores
  // write, at addresses this thread never reads: seen from one thread the
  // loads below are independent of the stores, and the compiler may issue
  // them first (it did at four lanes and nq = 4, merging a step's loads into
  // the previous step's).  The barrier keeps program order, which the LDS
  // then keeps within the wave.
  asm volatile("" ::: "memory");
  wf_load<R, NQ, LDT, TAU, 0>(w, col);
  // one step's coefficients and streamed entries in flight at a time
  asm volatile("" ::: "memory");
  __builtin_amdgcn_sched_barrier(0);
  if constexpr (TAU + 1 < R + NQ - 2) wf_steps<R, NQ, LDT, TAU + 1>(w, col, tb, icm0, icm1, nx);
}

// Applies a block's NQ rotation sequences to one window held in an LDS
// column (entry e at col[e * LDT]), in wavefront order; only the band of
// entries the current steps touch is held in registers.
template <int R, int NQ, int LDT>
__device__ __forceinline__ void wavefront_window(double* col, const TapeBlock<R, NQ>& tb,
                                                 int icm0, int icm1) {
  double w[R + NQ - 1];
  double2 cs[NQ];
  wf_coeffs<R, NQ, 0, 0>(cs, tb);
  wf_load<R, NQ, LDT, -1, 0>(w, col);  // entries first touched at step 0
  wf_steps<R, NQ, LDT, 0>(w, col, tb, icm0, icm1, cs);
}

// ---------------------------------------------------------------------------
// Split form: L lanes per row (column).  Lane h applies the sequences
// [h NQL, (h+1) NQL), NQL = NQ/L, at the same global steps as the one-lane
// wavefront: at step tau its slot i holds rotation (h NQL + i,
// ig = R-1-tau+h NQL+i), which touches entries 2i + 2h NQL + R-1-tau (-1).
// With lane h's register e holding window entry e + (2h-(L-1)) NQL, all
// lanes touch the same registers at every step, so they run one instruction
// stream: a wavefront with R' = R + (L-1) NQL, NQ' = NQL, whose slot i
// carries rotation ig' - (L-1-h) NQL of the lane's sequence, the identity
// where that falls outside [1, R-1].  Each lane executes NQL rotations per
// step instead of NQ.  Lane h's last write of an entry and lane h+1's first
// read of it fall after the same step, the write first.  The tile has
// (L-1) NQL zero rows on either side, which only identity rotations touch.
// (One lane per row gives one wave per 64 rows and leaves three of a CU's
// four SIMDs idle; L lanes cut the rotations per lane and step by L and run
// L times as many waves.)
template <int R, int NQ, int L, int NT>
__device__ __forceinline__ void stage_tape_split(
    TapeBlock<R + (L - 1) * (NQ / L), NQ / L> (&tb)[L], const double* __restrict__ Ct,
    const double* __restrict__ St, int ib, double sgn) {
  constexpr int NQH = NQ / L, RP = R + (L - 1) * NQH, LDR = R + NQ;
  constexpr int TOT = L * NQH * RP;
  constexpr int PER = (TOT + NT - 1) / NT;
  const size_t boff = (size_t)(ib - 1) * LDR * NQ;
  double c[PER], sn[PER];
#pragma unroll
  for (int q = 0; q < PER; ++q) {
    const int e = threadIdx.x + q * NT;
    const int h = e / (NQH * RP), rem = e % (NQH * RP);
    const int i = rem / RP, igp = rem % RP;
    const int ig = igp - (L - 1 - h) * NQH, icm = h * NQH + i;
    const bool ok = e < TOT && ig >= 1 && ig <= R - 1;
    const size_t ti = ok ? boff + (size_t)icm * LDR + (icm + ig - 1) : 0;
    c[q] = ok ? Ct[ti] : 1.0;
    sn[q] = ok ? St[ti] : 0.0;
  }
#pragma unroll
  for (int q = 0; q < PER; ++q) {
    const int e = threadIdx.x + q * NT;
    if (e < TOT) {
      const int h = e / (NQH * RP), rem = e % (NQH * RP);
      tb[h].cs[rem / RP][rem % RP] = make_double2(c[q], sgn * sn[q]);
    }
  }
}

// Rows in the split wavefront form, L lanes per row and 64/L rows per
// workgroup.  Each row's window lives in its own column of the LDS tile;
// consecutive blocks share NQ-1 entries (block ib-1's entries [R, R+NQ-1)
// are block ib's [0, NQ-1)), which are carried in the tile, so each later
// block loads R fresh entries.  AB: rows of A (blockIdx.y = 0) and B (1)
// above the generation window; otherwise all rows of Q (0) and Z (1).
// threads of a split-form workgroup: two waves share each block's staged
// coefficients (the staging is per workgroup, the same for every block)
constexpr int kSplitThreads = 128;

template <int R, int NQ, int L, bool AB>
__global__ __launch_bounds__(kSplitThreads) void k_s2_rows_w2(
    double* __restrict__ A, int lda, double* __restrict__ B, int ldb,
    double* __restrict__ Q, int ldq, double* __restrict__ Z, int ldz, int n,
    int s, int nblocks, const double* __restrict__ CU,
    const double* __restrict__ SU, const double* __restrict__ CV,
    const double* __restrict__ SV, S2Range rg) {
  constexpr int NQH = NQ / L, RP = R + (L - 1) * NQH, NW = R + NQ - 1;
  constexpr int PAD = (L - 1) * NQH, NP = NW + 2 * PAD, CPW = kSplitThreads / L, LDT = CPW + 1;
  static_assert(NQ % L == 0, "the lanes of a row take equal shares of the sequences");
  __shared__ double tile[NP][LDT];  // row PAD + e: window entry e
  __shared__ TapeBlock<RP, NQH> tb[L];
  const int which = AB ? blockIdx.y : 2 + blockIdx.y;  // 0 A, 1 B, 2 Q, 3 Z
  const int row_first = blockIdx.x * CPW;
  const int t = threadIdx.x, c = t / L, h = t % L;
  const int i = row_first + c;  // 0-based row
  const bool valid = i < n;
  double* M = which == 0 ? A : which == 1 ? B : which == 2 ? Q : Z;
  const int ld = which == 0 ? lda : which == 1 ? ldb : which == 2 ? ldq : ldz;
  const double* Ct = which == 2 ? CU : CV;
  const double* St = which == 2 ? SU : SV;
  const double sgn = which == 2 ? -1.0 : 1.0;
  for (int e = t; e < PAD * LDT; e += kSplitThreads) {
    tile[e / LDT][e % LDT] = 0.0;
    tile[PAD + NW + e / LDT][e % LDT] = 0.0;
  }
  bool carry = false;
  for (int ib = nblocks; ib >= 1; --ib) {
    const int cw0 = s + (ib - 1) * R;  // 0-based first window column
    const int W = imin(NW, n - cw0);
    if (W <= 0) {
      carry = false;
      continue;
    }
    if (AB && !rows_meet_range(row_first, CPW, s, R, NQ, ib, rg)) {
      carry = false;
      continue;
    }
    __syncthreads();  // the previous block's coefficients are consumed
    stage_tape_split<R, NQ, L, kSplitThreads>(tb, Ct, St, ib, sgn);
    double* src = M + PHT_IDX(valid ? i : 0, cw0, ld);
    if (carry) {
#pragma unroll
      for (int k = 0; k < NQ - 1; k += L)
        if (k + h < NQ - 1) tile[PAD + R + k + h][c] = tile[PAD + k + h][c];
    }
    const int kf = carry ? R : NW;  // fresh entries [0, kf)
#pragma unroll
    for (int k0 = 0; k0 < NW; k0 += 16 * L) {
      double tmp[16];
#pragma unroll
      for (int q = 0; q < 16; ++q) {
        const int k = k0 + L * q + h;
        tmp[q] = (valid && k < W && k < kf) ? src[(size_t)k * ld] : 0.0;
      }
#pragma unroll
      for (int q = 0; q < 16; ++q) {
        const int k = k0 + L * q + h;
        if (k < NW && k < kf) tile[PAD + k][c] = tmp[q];
      }
    }
    __syncthreads();
    // the row's active sequences form one run [icm0, icm1): both bounds of
    // the range grow with icm
    int icm0 = valid ? 0 : NQ, icm1 = NQ;
    if (AB && valid) {
      icm0 = NQ;
      icm1 = 0;
#pragma unroll
      for (int icm = NQ - 1; icm >= 0; --icm)
        if (row_in_range(i, s, R, NQ, ib, icm, rg)) {
          icm0 = icm;
          if (icm1 == 0) icm1 = icm + 1;
        }
    }
    const int icm0h = icm0 - h * NQH, icm1h = icm1 - h * NQH;
    double* col = &tile[2 * h * NQH][c];
    if (__all(icm0h <= 0 && icm1h >= NQH))
      wavefront_window<RP, NQH, LDT>(col, tb[h], 0, NQH);
    else if (icm0h < NQH && icm1h > 0 && icm0h < icm1h)
      wavefront_window<RP, NQH, LDT>(col, tb[h], icm0h, icm1h);
    // only the entries the row's active sequences touched go back (sequence
    // icm touches entries icm..icm+R-1): the rest of the window may belong
    // to work running beside this launch
    if (icm0 < icm1) {
      const int k1 = imin(W, icm1 + R - 1);
      for (int k = h; k < k1; k += L)
        if (k >= icm0) src[(size_t)k * ld] = tile[PAD + k][c];
    }
    carry = true;
  }
}

// Columns of A (which = 0) or B (1) right of the generation window in the
// split wavefront form, L lanes per column and 64/L columns per workgroup
// (left rotations in row form: the transposed coefficients).  Global
// accesses run along the columns through the tile; the tile carries the
// shared entries between blocks as in k_s2_rows_w2.
template <int R, int NQ, int L>
__global__ __launch_bounds__(kSplitThreads) void k_s2_cols_w2(
    double* __restrict__ A, int lda, double* __restrict__ B, int ldb, int n,
    int s, int nblocks, const double* __restrict__ CU,
    const double* __restrict__ SU, S2Range rg) {
  constexpr int NQH = NQ / L, RP = R + (L - 1) * NQH, NW = R + NQ - 1;
  constexpr int PAD = (L - 1) * NQH, NP = NW + 2 * PAD, CPW = kSplitThreads / L, LDT = CPW + 1;
  static_assert(NQ % L == 0, "the lanes of a column take equal shares of the sequences");
  static_assert(NW <= kSplitThreads, "one thread per window row");
  __shared__ double tile[NP][LDT];  // row PAD + k: window row k
  __shared__ TapeBlock<RP, NQH> tb[L];
  const int which = blockIdx.y;
  const int t = threadIdx.x, c = t / L, h = t % L;
  const int j0 = rg.col_begin + blockIdx.x * CPW;
  const int j = j0 + c;
  double* M = which == 0 ? A : B;
  const int ld = which == 0 ? lda : ldb;
  const int cmax = imin(CPW, n - j0);
  for (int e = t; e < PAD * LDT; e += kSplitThreads) {
    tile[e / LDT][e % LDT] = 0.0;
    tile[PAD + NW + e / LDT][e % LDT] = 0.0;
  }
  bool carry = false;
  for (int ib = nblocks; ib >= 1; --ib) {
    const int istop = cols_istop(which, s, R, NQ, ib);
    if (!cols_meet_range(j0, CPW, istop, R, rg)) {  // no column of this workgroup takes part
      carry = false;
      continue;
    }
    const int rw0 = s + (ib - 1) * R;  // 0-based first row
    const int W = imin(NW, n - rw0);
    if (W <= 0) {
      carry = false;
      continue;
    }
    __syncthreads();  // the previous block's coefficients and tile are consumed
    stage_tape_split<R, NQ, L, kSplitThreads>(tb, CU, SU, ib, -1.0);
    if (carry) {
#pragma unroll
      for (int k = 0; k < NQ - 1; k += L)
        if (k + h < NQ - 1) tile[PAD + R + k + h][c] = tile[PAD + k + h][c];
    }
    const int kf = carry ? R : NW;  // fresh rows [0, kf)
    // row t of 16 columns per group, all 16 loads in flight together
#pragma unroll
    for (int c0 = 0; c0 < CPW; c0 += 16) {
      double tmp[16];
#pragma unroll
      for (int q = 0; q < 16; ++q) {
        const int cc = c0 + q, k = t;
        tmp[q] = (cc < cmax && k < W && k < kf) ? M[PHT_IDX(rw0 + k, j0 + cc, ld)] : 0.0;
      }
      // every column's carry copy is complete before any fresh row lands
      if (c0 == 0) __syncthreads();
#pragma unroll
      for (int q = 0; q < 16; ++q) {
        const int cc = c0 + q, k = t;
        if (k < NW && k < kf) tile[PAD + k][cc] = tmp[q];
      }
    }
    __syncthreads();
    const bool act = (j < n) && col_in_range(j, istop, R, rg);
    if (act) wavefront_window<RP, NQH, LDT>(&tile[2 * h * NQH][c], tb[h], 0, NQH);
    __syncthreads();
    for (int cc = 0; cc < cmax; ++cc) {
      if (!col_in_range(j0 + cc, istop, R, rg)) continue;
      for (int k = t; k < W; k += kSplitThreads) M[PHT_IDX(rw0 + k, j0 + cc, ld)] = tile[PAD + k][cc];
    }
    carry = true;
  }
}


// ---- launchers ----
template <int R, int NQ>
hipError_t s2_far_rows(hipStream_t st, const S2FarArgs& a, bool ab, const S2Range& rg,
                       int row_end) {
  if (row_end <= 0) return hipSuccess;
  if (wavefront_fits(R, NQ) && s2_wavefront()) {
    constexpr int kRows = kSplitThreads / kLanes;
    const dim3 g((row_end + kRows - 1) / kRows, 2);
    if (ab)
      hipLaunchKernelGGL((k_s2_rows_w2<R, NQ, kLanes, true>), g, dim3(kSplitThreads), 0, st, a.A,
                         a.lda, a.B, a.ldb, a.Q, a.ldq, a.Z, a.ldz, a.n, a.s, a.nblocks, a.CU,
                         a.SU, a.CV, a.SV, rg);
    else
      hipLaunchKernelGGL((k_s2_rows_w2<R, NQ, kLanes, false>), g, dim3(kSplitThreads), 0, st, a.A,
                         a.lda, a.B, a.ldb, a.Q, a.ldq, a.Z, a.ldz, a.n, a.s, a.nblocks, a.CU,
                         a.SU, a.CV, a.SV, rg);
  } else {
    const dim3 g((row_end + kRowsThreads - 1) / kRowsThreads, 2);
    if (ab)
      hipLaunchKernelGGL((k_s2_rows_v<R, NQ, true>), g, dim3(kRowsThreads), 0, st, a.A, a.lda,
                         a.B, a.ldb, a.Q, a.ldq, a.Z, a.ldz, a.n, a.s, a.nblocks, a.CU, a.SU,
                         a.CV, a.SV, rg);
    else
      hipLaunchKernelGGL((k_s2_rows_v<R, NQ, false>), g, dim3(kRowsThreads), 0, st, a.A, a.lda,
                         a.B, a.ldb, a.Q, a.ldq, a.Z, a.ldz, a.n, a.s, a.nblocks, a.CU, a.SU,
                         a.CV, a.SV, rg);
  }
  return hipGetLastError();
}

template <int R, int NQ>
hipError_t s2_far_cols(hipStream_t st, const S2FarArgs& a, const S2Range& rg) {
  const int ncols = a.n - rg.col_begin;
  if (ncols <= 0) return hipSuccess;
  if (wavefront_fits(R, NQ) && s2_wavefront()) {
    constexpr int kCols = kSplitThreads / kLanes;
    hipLaunchKernelGGL((k_s2_cols_w2<R, NQ, kLanes>), dim3((ncols + kCols - 1) / kCols, 2),
                       dim3(kSplitThreads), 0, st, a.A, a.lda, a.B, a.ldb, a.n, a.s, a.nblocks,
                       a.CU, a.SU, rg);
  } else {
    hipLaunchKernelGGL((k_s2_cols_v<R, NQ>), dim3((ncols + 63) / 64, 2), dim3(64), 0, st, a.A,
                       a.lda, a.B, a.ldb, a.n, a.s, a.nblocks, a.CU, a.SU, rg);
  }
  return hipGetLastError();
}

// explicit instantiation of the launchers of one (r, nq) pair
#define PHT_S2_FAR_INST(R, NQ)                                                           \
  template hipError_t s2_far_rows<R, NQ>(hipStream_t, const S2FarArgs&, bool, const S2Range&, \
                                         int);                                             \
  template hipError_t s2_far_cols<R, NQ>(hipStream_t, const S2FarArgs&, const S2Range&);

}  // namespace pht
