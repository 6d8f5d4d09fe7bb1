// Stage-2 kernels: r-Hessenberg-triangular -> Hessenberg-triangular with
// Givens rotations, following rht2ht_givens (src/rht2ht.f90:696-1344).
//
// A sweep reduces nq columns starting at column s (1-based, `jblock` in the
// reference).  Step (ic, ib) of the sweep works on the bulge block ib of
// column ic; in the reference's notation (rht2ht.f90:828-953)
//   row0 = s + ic + (ib-1) r      first row/column touched by the step
//   col  = row0 - 1 (ib = 1) or row0 - r (ib > 1)   column being reduced
//   nb   = min(r, n - row0 + 1)
// Left rotations ig = nb-1..1 act on rows (row0+ig-1, row0+ig), right
// rotations on columns (row0+ig-1, row0+ig); both are stored at local index
// ic+ig-1 of block ib of the tape CU,SU,CV,SV(r+nq, nq, nblocks)
// (rht2ht.f90:735-738).
//
// k_s2_gen performs the steps in the reference's order.  The reference defers
// a step's left rotations and applies them to each later column of the sweep
// just before that column is reduced (rht2ht.f90:803-825, 887-909), which
// keeps every pair of columns that a right rotation mixes in the same
// left-rotation state; here the deferred rotations of a block are kept
// multiplied out (M, (r+nq) x (r+nq)) and applied as one mat-vec.  For the
// last column (nb < 2) the reference skips the deferred rotations owed by
// later bulge blocks (rht2ht.f90:827), which loses an update when nq > r;
// this kernel applies them.  The rest of each rotation's action --
// rows above the reference's window and columns right of it, plus Q and Z --
// is done by k_s2_rows / k_s2_cols after the sweep, in the reference's
// bottom-up block order (rht2ht.f90:1019-1325).
#include "s2_dev.hpp"

namespace pht {

namespace {
constexpr int kGenThreads = 256;
constexpr int kTile = 128;
}  // namespace
using namespace s2;

// ---------------------------------------------------------------------------
// Generation of one sweep (rht2ht.f90:787-957).
//
// k_s2_prep resets the sweep's tape (identity rotations), the deferred-rotation
// matrices M (identity) and the progress counters.  k_s2_sweep runs one
// workgroup per column ic of the sweep; the workgroup chases its column's
// bulge down block by block.  Step (ic, ib) reads what step (ic-1, ib+1)
// writes and writes what step (ic-1, ib+2) reads, so a workgroup starts a step
// only when the previous column's workgroup has finished step ib+2 (or all of
// its steps); the wait is a bounded spin on a release/acquire counter.
//
// M holds, per bulge block ib, the product of the left rotations of the
// columns already reduced in this sweep (local rows of the block); the
// reference's lazily-applied left rotations of a column (rht2ht.f90:803-825,
// 887-909) are applied as one mat-vec with it.
constexpr int kMaxHelpers = kS2MaxHelpers;
constexpr int kDone = 1 << 29;
constexpr int kParts = kGenThreads / 128;    // reduction parts of the mat-vec

__global__ void k_s2_prep(double* CU, double* SU, double* CV, double* SV,
                          size_t tsz, double* Mw, int ldr, int nblocks,
                          int* progress, int nq, int* errflag, int clear_err) {
  const size_t gid = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  const size_t gsz = (size_t)gridDim.x * blockDim.x;
  for (size_t e = gid; e < tsz; e += gsz) {
    CU[e] = 1.0;
    CV[e] = 1.0;
    SU[e] = 0.0;
    SV[e] = 0.0;
  }
  const size_t msz = (size_t)ldr * ldr;
  for (size_t e = gid; e < msz * nblocks; e += gsz) {
    const size_t k = e % msz;
    Mw[e] = ((k % ldr) == (k / ldr)) ? 1.0 : 0.0;
  }
  for (size_t e = gid; e < (size_t)nq * (1 + kMaxHelpers); e += gsz) progress[e] = 0;
  // the flag collects every sweep of a call: only the call's first sweep
  // clears it
  if (gid == 0 && clear_err) *errflag = 0;
}

// Helper workgroup h of column ic: after its column's step ib is published,
// applies the step's right rotations to the window rows above the bulge,
// rows istart..row0-1 of A and of B (rht2ht.f90:853-867,937-951), and
// publishes hprog.  Its column's next steps never read these rows (nq <= r+1),
// the next column's steps wait for them.
template <int RMAX>
__device__ void sweep_helper(double* __restrict__ A, int lda, double* __restrict__ B, int ldb,
                             int n, int r, int nq, int s, const double* __restrict__ CV,
                             const double* __restrict__ SV, int ldr, int* progress, int* hprog,
                             int* errflag, int spin, int ic, int h, int nhelp, double* hc,
                             double* hs) {
  const int tid = threadIdx.x;
  const int jc = s + ic - 1;
  int* mine = &hprog[(ic - 1) * nhelp + h];
  for (int ib = 1;; ++ib) {
    int row0;
    if (ib == 1) {
      if (jc > n) break;
      row0 = jc + 1;
    } else {
      const int col = jc + 1 + (ib - 2) * r;
      if (col > n) break;
      row0 = col + r;
    }
    wait_at_least(&progress[ic - 1], ib, errflag, spin);
    acquire_waits();
    const int nb = imin(r, n - row0 + 1);
    const int istart = imax(s + 1, s + 1 + (ib - 1 - nq + ic) * r);
    const int nrows = row0 - istart;  // rows istart..row0-1 (1-based)
    if (nb >= 2 && nrows > 0) {
      if (tid >= 1 && tid < nb) {
        const size_t ti = tape_idx(ic + tid - 1, ic, ib, ldr, nq);
        hc[tid] = CV[ti];
        hs[tid] = SV[ti];
      }
      __syncthreads();
      for (int t = h * kGenThreads + tid; t < 2 * nrows; t += nhelp * kGenThreads) {
        const bool inA = t < nrows;
        rotate_row<RMAX>(inA ? A : B, inA ? lda : ldb, istart - 1 + (inA ? t : t - nrows),
                         row0 - 1, nb, hc, hs);
      }
    }
    publish(mine, ib);
  }
  publish(mine, kDone);
}

template <int RMAX>
__global__ __launch_bounds__(kGenThreads) void k_s2_sweep(
    double* __restrict__ A, int lda, double* __restrict__ B, int ldb, int n,
    int r, int nq, int s, double* __restrict__ CU, double* __restrict__ SU,
    double* __restrict__ CV, double* __restrict__ SV, int ldr,
    double* __restrict__ Mw, int* progress, int* errflag, int spin, int nhelp, int* hprog,
    int mbufs) {
  __shared__ double Bb[RMAX][RMAX + 1];
  __shared__ double lc[RMAX], ls[RMAX], rc[RMAX], rs[RMAX];
  __shared__ double xa[kMaxLdr], xb[kMaxLdr];
  __shared__ double red[2][kParts][128];
  // M of the current block: M(i, j) at Ml[j*(ldr+1) + i]; with mbufs = 2 a
  // second buffer receives the next block's M during the current step
  extern __shared__ double Ml[];
  if ((int)blockIdx.x >= nq) {  // helper workgroups follow the nq leaders
    const int hidx = blockIdx.x - nq;
    sweep_helper<RMAX>(A, lda, B, ldb, n, r, nq, s, CV, SV, ldr, progress, hprog, errflag, spin,
                       hidx / nhelp + 1, hidx % nhelp, nhelp, rc, rs);
    return;
  }
  const int ic = blockIdx.x + 1;
  const int jc = s + ic - 1;
  const size_t mstride = (size_t)ldr * (ldr + 1);
  bool have_m = false;
  WgThreads wg;

  for (int ib = 1;; ++ib) {
    int col, row0;
    if (ib == 1) {
      if (jc > n) break;
      col = jc;
      row0 = jc + 1;
    } else {
      col = jc + 1 + (ib - 2) * r;  // jb of the reference
      if (col > n) break;
      row0 = col + r;
    }
    if (ic > 1) {  // the previous column's leader and helpers are 2 steps ahead
      wait_at_least(&progress[ic - 2], ib + 2, errflag, spin);
      for (int h = 0; h < nhelp; ++h)
        wait_at_least(&hprog[(ic - 2) * nhelp + h], ib + 2, errflag, spin);
      acquire_waits();
    }
    // the next block exists iff its column jc + 1 + (ib-1) r is inside the matrix
    const bool next = jc + 1 + (ib - 1) * r <= n;
    double* Mcur = Ml + (mbufs == 2 ? (size_t)(ib & 1) * mstride : 0);
    double* Mnxt = (mbufs == 2 && next) ? Ml + (size_t)((ib + 1) & 1) * mstride : nullptr;
    have_m = chase_step<RMAX, kGenThreads>(wg, A, lda, B, ldb, n, r, nq, s, ic, ib, col, row0, CU,
                                           SU, CV, SV, ldr, Mw, nhelp == 0, Bb, lc, ls, rc, rs,
                                           xa, xb, &red[0][0][0], Mcur, have_m, Mnxt);
    publish(&progress[ic - 1], ib);
  }
  publish(&progress[ic - 1], kDone);
}

// ---------------------------------------------------------------------------
// Rows: which = 0 A, 1 B (rows above the generation window, right rotations),
// 2 Q (all rows, left rotations seen from the right), 3 Z (all rows, right).
__global__ __launch_bounds__(kTile) void k_s2_rows(
    double* __restrict__ A, int lda, double* __restrict__ B, int ldb,
    double* __restrict__ Q, int ldq, double* __restrict__ Z, int ldz, int n,
    int r, int nq, int s, int nblocks, const double* __restrict__ CU,
    const double* __restrict__ SU, const double* __restrict__ CV,
    const double* __restrict__ SV, int ldr, int which0) {
  extern __shared__ double win[];  // [ldr][kTile]
  const int which = which0 + blockIdx.y;
  const int t = threadIdx.x;
  const int i = blockIdx.x * kTile + t;  // 0-based row
  double* M;
  int ld;
  const double *Ct, *St;
  switch (which) {
    case 0: M = A; ld = lda; Ct = CV; St = SV; break;
    case 1: M = B; ld = ldb; Ct = CV; St = SV; break;
    case 2: M = Q; ld = ldq; Ct = CU; St = SU; break;
    default: M = Z; ld = ldz; Ct = CV; St = SV; break;
  }
  const bool partial = which < 2;
  const bool valid = i < n;
  for (int ib = nblocks; ib >= 1; --ib) {
    const int cw0 = s + (ib - 1) * r + 1;  // 1-based first window column
    const int W = imin(r + nq - 1, n - cw0 + 1);
    if (W <= 0) continue;
    // rows of A,B updated here: 1-based row i+1 < istart(ic, ib)
    const int istart_max = imax(s + 1, s + 1 + (ib - 1) * r);  // ic = nq
    if (partial && blockIdx.x * kTile + 1 >= istart_max) continue;
    if (valid)
      for (int w = 0; w < W; ++w) win[w * kTile + t] = M[PHT_IDX(i, cw0 - 1 + w, ld)];
    for (int ic = 1; ic <= nq; ++ic) {
      if (partial) {
        const int istart = imax(s + 1, s + 1 + (ib - 1 - nq + ic) * r);
        if (!(i + 1 < istart)) continue;
      }
      for (int ig = r - 1; ig >= 1; --ig) {
        const int hi = ic + ig - 1;  // 0-based window index of local column ic+ig
        if (hi >= W) continue;
        const size_t ti = tape_idx(ic + ig - 1, ic, ib, ldr, nq);
        const double c = Ct[ti], sn = St[ti];
        if (c == 1.0 && sn == 0.0) continue;
        double* px = &win[hi * kTile + t];
        double* py = &win[(hi - 1) * kTile + t];
        const double x = *px, y = *py;
        if (which == 2) {
          // columns (ic-1+ig, ic+ig) of Q <- [c*a + s*b, c*b - s*a]
          *py = c * y + sn * x;
          *px = c * x - sn * y;
        } else {
          *px = c * x + sn * y;
          *py = c * y - sn * x;
        }
      }
    }
    if (valid)
      for (int w = 0; w < W; ++w) M[PHT_IDX(i, cw0 - 1 + w, ld)] = win[w * kTile + t];
  }
}

// Columns: which = 0 A, 1 B; left rotations on the rows of each block for the
// columns right of the generation window.
__global__ __launch_bounds__(kTile) void k_s2_cols(
    double* __restrict__ A, int lda, double* __restrict__ B, int ldb, int n,
    int r, int nq, int s, int nblocks, const double* __restrict__ CU,
    const double* __restrict__ SU, int ldr) {
  extern __shared__ double win[];  // [ldr][kTile]
  const int which = blockIdx.y;
  const int t = threadIdx.x;
  const int j = blockIdx.x * kTile + t;  // 0-based column
  double* M = which == 0 ? A : B;
  const int ld = which == 0 ? lda : ldb;
  const bool valid = j < n;
  for (int ib = nblocks; ib >= 1; --ib) {
    const int istop = (which == 0) ? ((ib == 1) ? s + nq - 1 : s + (ib - 2) * r + nq)
                                   : s + (ib - 1) * r + r + nq - 1;
    if ((blockIdx.x + 1) * kTile <= istop) continue;  // whole tile left of it
    const bool act = valid && (j + 1 > istop);
    const int rw0 = s + (ib - 1) * r + 1;
    const int W = imin(r + nq - 1, n - rw0 + 1);
    if (W <= 0) continue;
    if (act)
      for (int w = 0; w < W; ++w) win[w * kTile + t] = M[PHT_IDX(rw0 - 1 + w, j, ld)];
    for (int ic = 1; ic <= nq; ++ic) {
      for (int ig = r - 1; ig >= 1; --ig) {
        const int hi = ic + ig - 1;
        if (hi >= W) continue;
        const size_t ti = tape_idx(ic + ig - 1, ic, ib, ldr, nq);
        const double c = CU[ti], sn = SU[ti];
        if (c == 1.0 && sn == 0.0) continue;
        double* pb = &win[hi * kTile + t];
        double* pa = &win[(hi - 1) * kTile + t];
        const double a = *pa, b = *pb;
        *pa = c * a + sn * b;
        *pb = c * b - sn * a;
      }
    }
    if (act)
      for (int w = 0; w < W; ++w) M[PHT_IDX(rw0 - 1 + w, j, ld)] = win[w * kTile + t];
  }
}

// ---------------------------------------------------------------------------
// ---------------------------------------------------------------------------
namespace {
constexpr size_t kLdsBudget = 160 * 1024;
// static LDS of k_s2_sweep<R> (Bb, rotation vectors, x vectors, partial sums)
constexpr size_t sweep_static_lds(int R) {
  return sizeof(double) * ((size_t)R * (R + 1) + 4 * R + 2 * kMaxLdr + 2 * kParts * 128);
}
int sweep_rmax(int r) { return r <= 8 ? 8 : r <= 16 ? 16 : r <= 32 ? 32 : 64; }
}  // namespace

size_t s2_sweep_lds(int r, int nq) {
  const size_t ldr = (size_t)r + nq;
  if (r > 64 || ldr > (size_t)kMaxLdr) return 0;
  const size_t dyn = ldr * (ldr + 1) * sizeof(double);
  return (sweep_static_lds(sweep_rmax(r)) + dyn <= kLdsBudget) ? dyn : 0;
}

hipError_t s2_prepare(hipStream_t st, const S2Params& P, double* CU, double* SU,
                      double* CV, double* SV) {
  const size_t tsz = (size_t)P.ldr * P.nq * P.nblocks;
  hipLaunchKernelGGL(k_s2_prep, dim3(128), dim3(256), 0, st, CU, SU, CV, SV, tsz,
                     P.M, P.ldr, P.nblocks, P.progress, P.nq, P.errflag, P.clear_err);
  return hipGetLastError();
}

hipError_t s2_generate(hipStream_t st, double* A, int lda, double* B, int ldb,
                       const S2Params& P, double* CU, double* SU, double* CV,
                       double* SV) {
  const dim3 g(P.nq * (1 + P.nhelp)), b(kGenThreads);
  const size_t one = s2_sweep_lds(P.r, P.nq);
  if (one == 0) return hipErrorInvalidValue;
  // two M buffers (the next block's M staged during the current step) when
  // they fit beside the static LDS
  const int mbufs = sweep_static_lds(sweep_rmax(P.r)) + 2 * one <= kLdsBudget ? 2 : 1;
  const size_t lds = mbufs * one;
#define PHT_GEN(R)                                                                \
  do {                                                                            \
    static bool attr = false;                                                     \
    if (!attr) {                                                                  \
      hipError_t e = hipFuncSetAttribute((const void*)k_s2_sweep<R>,              \
                                         hipFuncAttributeMaxDynamicSharedMemorySize, \
                                         (int)(kLdsBudget - sweep_static_lds(R)));   \
      if (e != hipSuccess) return e;                                              \
      attr = true;                                                                \
    }                                                                             \
    hipLaunchKernelGGL(k_s2_sweep<R>, g, b, lds, st, A, lda, B, ldb, P.n, P.r, P.nq, \
                       P.jblock, CU, SU, CV, SV, P.ldr, P.M, P.progress, P.errflag, P.spin, \
                       P.nhelp, P.hprog, mbufs);                                   \
  } while (0)
  if (P.r <= 8)
    PHT_GEN(8);
  else if (P.r <= 16)
    PHT_GEN(16);
  else if (P.r <= 32)
    PHT_GEN(32);
  else if (P.r <= 64)
    PHT_GEN(64);
  else
    return hipErrorInvalidValue;
#undef PHT_GEN
  return hipGetLastError();
}

bool s2_far_split(int r, int nq) {
#define PHT_HAS(RR, NN) \
  if (r == RR && nq == NN) return true;
  PHT_S2_PAIRS(PHT_HAS)
#undef PHT_HAS
  return false;
}

hipError_t s2_apply_rows(hipStream_t st, double* A, int lda, double* B,
                         int ldb, double* Q, int ldq, double* Z, int ldz,
                         const S2Params& P, const double* CU, const double* SU,
                         const double* CV, const double* SV, RowsPart part,
                         const S2Range& rg, int row_end) {
  const size_t lds = (size_t)P.ldr * kTile * sizeof(double);
  if (lds > 160 * 1024) return hipErrorInvalidValue;
  static bool attr = false;
  if (!attr) {
    hipError_t e = hipFuncSetAttribute((const void*)k_s2_rows,
                                       hipFuncAttributeMaxDynamicSharedMemorySize,
                                       160 * 1024);
    if (e != hipSuccess) return e;
    attr = true;
  }
  const S2FarArgs fa{A, B, Q, Z, lda, ldb, ldq, ldz, P.n, P.jblock, P.nblocks, CU, SU, CV, SV};
  if (row_end < 0) row_end = P.n;
#define PHT_ROWS(RR, NN)                                                        \
  if (P.r == RR && P.nq == NN) {                                                \
    if (part != kRowsQZ) {                                                      \
      const int re = row_end < P.n ? row_end : P.n;                            \
      const hipError_t e = s2_far_rows<RR, NN>(st, fa, true, rg, re < rg.row_end ? re : rg.row_end); \
      if (e != hipSuccess) return e;                                            \
    }                                                                           \
    if (part != kRowsAB) return s2_far_rows<RR, NN>(st, fa, false, S2Range{}, P.n); \
    return hipSuccess;                                                          \
  }
  PHT_S2_PAIRS(PHT_ROWS)
#undef PHT_ROWS
  // the generic kernels apply a whole far update (s2_far_split is false)
  if (rg.lo_shift >= 0 || rg.hi_shift != 0 || rg.row_end < P.n) return hipErrorInvalidValue;
  dim3 grid((P.n + kTile - 1) / kTile, part == kRowsAll ? 4 : 2), block(kTile);
  hipLaunchKernelGGL(k_s2_rows, grid, block, lds, st, A, lda, B, ldb, Q, ldq, Z,
                     ldz, P.n, P.r, P.nq, P.jblock, P.nblocks, CU, SU, CV, SV,
                     P.ldr, part == kRowsQZ ? 2 : 0);
  return hipGetLastError();
}

hipError_t s2_apply_cols(hipStream_t st, double* A, int lda, double* B,
                         int ldb, const S2Params& P, const double* CU,
                         const double* SU, const S2Range& rg) {
  const size_t lds = (size_t)P.ldr * kTile * sizeof(double);
  if (lds > 160 * 1024) return hipErrorInvalidValue;
  static bool attr = false;
  if (!attr) {
    hipError_t e = hipFuncSetAttribute((const void*)k_s2_cols,
                                       hipFuncAttributeMaxDynamicSharedMemorySize,
                                       160 * 1024);
    if (e != hipSuccess) return e;
    attr = true;
  }
  const S2FarArgs fa{A, B, nullptr, nullptr, lda, ldb, 0, 0, P.n, P.jblock, P.nblocks, CU, SU,
                     nullptr, nullptr};
#define PHT_COLS(RR, NN) \
  if (P.r == RR && P.nq == NN) return s2_far_cols<RR, NN>(st, fa, rg);
  PHT_S2_PAIRS(PHT_COLS)
#undef PHT_COLS
  if (rg.c_lo != 0 || rg.c_hi >= 0 || rg.col_begin != 0) return hipErrorInvalidValue;
  dim3 grid((P.n + kTile - 1) / kTile, 2), block(kTile);
  hipLaunchKernelGGL(k_s2_cols, grid, block, lds, st, A, lda, B, ldb, P.n, P.r,
                     P.nq, P.jblock, P.nblocks, CU, SU, P.ldr);
  return hipGetLastError();
}

}  // namespace pht
