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
// far updates in the split wavefront form (PARAHT_S2_WAVEFRONT=0: the
// sequential form)
bool s2_wavefront() {
  static int v = -1;
  if (v < 0) {
    const char* e = getenv("PARAHT_S2_WAVEFRONT");
    v = (e && *e == '0') ? 0 : 1;
  }
  return v == 1;
}
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
constexpr int kSpinLimit = 1 << 26;
constexpr int kMaxHelpers = kS2MaxHelpers;
constexpr int kDone = 1 << 29;
constexpr int kParts = kGenThreads / 128;    // reduction parts of the mat-vec

__global__ void k_s2_prep(double* CU, double* SU, double* CV, double* SV,
                          size_t tsz, double* Mw, int ldr, int nblocks,
                          int* progress, int nq, int* errflag) {
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
  if (gid == 0) *errflag = 0;
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

// Thread 0 spins (bounded) until *p >= target.
__device__ __forceinline__ void wait_at_least(int* p, int target, int* errflag) {
  if (threadIdx.x == 0) {
    int spins = 0;
    while (__hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
      __builtin_amdgcn_s_sleep(2);
      if (++spins > kSpinLimit) {
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

// Helper workgroup h of column ic: after its column's step ib is published,
// applies the step's right rotations to the window rows above the bulge,
// rows istart..row0-1 of A and of B (rht2ht.f90:853-867,937-951), and
// publishes hprog.  Its column's next steps never read these rows (nq <= r+1),
// the next column's steps wait for them.
template <int RMAX>
__device__ void sweep_helper(double* __restrict__ A, int lda, double* __restrict__ B, int ldb,
                             int n, int r, int nq, int s, const double* __restrict__ CV,
                             const double* __restrict__ SV, int ldr, int* progress, int* hprog,
                             int* errflag, int ic, int h, int nhelp, double* hc, double* hs) {
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
    wait_at_least(&progress[ic - 1], ib, errflag);
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
    double* __restrict__ Mw, int* progress, int* errflag, int nhelp, int* hprog,
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
    sweep_helper<RMAX>(A, lda, B, ldb, n, r, nq, s, CV, SV, ldr, progress, hprog, errflag,
                       hidx / nhelp + 1, hidx % nhelp, nhelp, rc, rs);
    return;
  }
  const int ic = blockIdx.x + 1;
  const int jc = s + ic - 1;
  const size_t mstride = (size_t)ldr * (ldr + 1);
  bool have_m = false;

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
      wait_at_least(&progress[ic - 2], ib + 2, errflag);
      for (int h = 0; h < nhelp; ++h) wait_at_least(&hprog[(ic - 2) * nhelp + h], ib + 2, errflag);
      acquire_waits();
    }
    // the next block exists iff its column jc + 1 + (ib-1) r is inside the matrix
    const bool next = jc + 1 + (ib - 1) * r <= n;
    double* Mcur = Ml + (mbufs == 2 ? (size_t)(ib & 1) * mstride : 0);
    double* Mnxt = (mbufs == 2 && next) ? Ml + (size_t)((ib + 1) & 1) * mstride : nullptr;
    have_m = chase_step<RMAX, kGenThreads>(A, lda, B, ldb, n, r, nq, s, ic, ib, col, row0, CU,
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
                                               int ib, int W, const TapeBlock<R, NQ>& tb) {
  const int t = threadIdx.x;
  double v[R];
#pragma unroll
  for (int k = 0; k < R; ++k) v[k] = (k < W) ? tile[k][t] : 0.0;
#pragma unroll 1
  for (int icm = 0; icm < NQ; ++icm) {
    bool act = valid;
    if (PARTIAL) act = act && (i + 1 < imax(s + 1, s + 1 + (ib - NQ + icm) * R));
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
    const double* __restrict__ SV) {
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
    // A, B: rows at or below istart(nq, ib) are the generation's
    if (AB && row_first + 1 >= imax(s + 1, s + 1 + (ib - 1) * R)) continue;
    __syncthreads();  // the previous block's tile and coefficients are consumed
    stage_tape<R, NQ, kRowsThreads>(tb, Ct, St, ib, sgn);
    window_to_lds<NW>(tile, t, M + PHT_IDX(i, cw0, ld), (size_t)ld, W, valid);
    __syncthreads();
    rows_one_block<R, NQ, AB>(tile, i, valid, s, ib, W, tb);
    if (valid)
      for (int k = 0; k < W; ++k) M[PHT_IDX(i, cw0 + k, ld)] = tile[k][t];
  }
}

// Columns of A (which = 0) or B (1) right of the generation window: left
// rotations on the rows of each block.  The 64 columns of a workgroup are
// staged through LDS so that global accesses run along the columns.
template <int R, int NQ>
__global__ __launch_bounds__(64) void k_s2_cols_v(
    double* __restrict__ A, int lda, double* __restrict__ B, int ldb, int n,
    int s, int nblocks, const double* __restrict__ CU,
    const double* __restrict__ SU) {
  constexpr int NW = R + NQ - 1;
  static_assert(NW <= 128, "two 64-row chunks per column");
  __shared__ double tile[NW][65];
  __shared__ TapeBlock<R, NQ> tb;
  const int which = blockIdx.y;
  const int t = threadIdx.x;
  const int j0 = blockIdx.x * 64;
  const int j = j0 + t;
  double* M = which == 0 ? A : B;
  const int ld = which == 0 ? lda : ldb;
  for (int ib = nblocks; ib >= 1; --ib) {
    const int istop = (which == 0) ? ((ib == 1) ? s + NQ - 1 : s + (ib - 2) * R + NQ)
                                   : s + (ib - 1) * R + R + NQ - 1;  // 1-based
    if (j0 + 64 <= istop) continue;
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
    const bool act = (j < n) && (j + 1 > istop);
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
      if (j0 + c + 1 <= istop) continue;
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

// The rotations of step TAU: (icm, ig = R-1-TAU+icm) for each sequence.
template <int R, int NQ, int TAU, int ICM>
__device__ __forceinline__ void wf_rotations(double (&w)[R + NQ - 1], const double2 (&cs)[NQ],
                                             int icm0) {
  constexpr int ig = R - 1 - TAU + ICM;
  if constexpr (ig >= 1 && ig <= R - 1) {
    const bool on = ICM >= icm0;
    const double c = on ? cs[ICM].x : 1.0, sn = on ? cs[ICM].y : 0.0;
    const double x = w[ICM + ig], y = w[ICM + ig - 1];
    w[ICM + ig] = c * x + sn * y;
    w[ICM + ig - 1] = c * y - sn * x;
  }
  if constexpr (ICM + 1 < NQ) wf_rotations<R, NQ, TAU, ICM + 1>(w, cs, icm0);
}

// Step TAU with the next step's coefficients and window entries read ahead.
template <int R, int NQ, int LDT, int TAU>
__device__ __forceinline__ void wf_steps(double (&w)[R + NQ - 1], double* col,
                                         const TapeBlock<R, NQ>& tb, int icm0,
                                         const double2 (&cs)[NQ]) {
  double2 nx[NQ];
  if constexpr (TAU + 1 < R + NQ - 2) wf_coeffs<R, NQ, TAU + 1, 0>(nx, tb);
  wf_rotations<R, NQ, TAU, 0>(w, cs, icm0);
  wf_store<R, NQ, LDT, TAU, 0>(w, col);
  wf_load<R, NQ, LDT, TAU, 0>(w, col);
  // one step's coefficients and streamed entries in flight at a time
  asm volatile("" ::: "memory");
  __builtin_amdgcn_sched_barrier(0);
  if constexpr (TAU + 1 < R + NQ - 2) wf_steps<R, NQ, LDT, TAU + 1>(w, col, tb, icm0, nx);
}

// Applies a block's NQ rotation sequences to one window held in an LDS
// column (entry e at col[e * LDT]), in wavefront order; only the band of
// entries the current steps touch is held in registers.
template <int R, int NQ, int LDT>
__device__ __forceinline__ void wavefront_window(double* col, const TapeBlock<R, NQ>& tb,
                                                 int icm0) {
  double w[R + NQ - 1];
  double2 cs[NQ];
  wf_coeffs<R, NQ, 0, 0>(cs, tb);
  wf_load<R, NQ, LDT, -1, 0>(w, col);  // entries first touched at step 0
  wf_steps<R, NQ, LDT, 0>(w, col, tb, icm0, cs);
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
    const double* __restrict__ SV) {
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
    if (AB && row_first + 1 >= imax(s + 1, s + 1 + (ib - 1) * R)) break;
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
    int icm0 = valid ? 0 : NQ;
    if (AB && valid) {
      icm0 = NQ;
#pragma unroll
      for (int icm = NQ - 1; icm >= 0; --icm)
        if (i + 1 < imax(s + 1, s + 1 + (ib - NQ + icm) * R)) icm0 = icm;
    }
    const int icm0h = icm0 - h * NQH;
    double* col = &tile[2 * h * NQH][c];
    if (__all(icm0h <= 0))
      wavefront_window<RP, NQH, LDT>(col, tb[h], 0);
    else if (icm0h < NQH)
      wavefront_window<RP, NQH, LDT>(col, tb[h], icm0h);
    if (valid)
      for (int k = h; k < W; k += L) src[(size_t)k * ld] = tile[PAD + k][c];
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
    const double* __restrict__ SU) {
  constexpr int NQH = NQ / L, RP = R + (L - 1) * NQH, NW = R + NQ - 1;
  constexpr int PAD = (L - 1) * NQH, NP = NW + 2 * PAD, CPW = kSplitThreads / L, LDT = CPW + 1;
  static_assert(NQ % L == 0, "the lanes of a column take equal shares of the sequences");
  static_assert(NW <= kSplitThreads, "one thread per window row");
  __shared__ double tile[NP][LDT];  // row PAD + k: window row k
  __shared__ TapeBlock<RP, NQH> tb[L];
  const int which = blockIdx.y;
  const int t = threadIdx.x, c = t / L, h = t % L;
  const int j0 = blockIdx.x * CPW;
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
    const int istop = (which == 0) ? ((ib == 1) ? s + NQ - 1 : s + (ib - 2) * R + NQ)
                                   : s + (ib - 1) * R + R + NQ - 1;  // 1-based
    if (j0 + CPW <= istop) {  // no column of this workgroup takes part
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
    const bool act = (j < n) && (j + 1 > istop);
    if (act) wavefront_window<RP, NQH, LDT>(&tile[2 * h * NQH][c], tb[h], 0);
    __syncthreads();
    for (int cc = 0; cc < cmax; ++cc) {
      if (j0 + cc + 1 <= istop) continue;
      for (int k = t; k < W; k += kSplitThreads) M[PHT_IDX(rw0 + k, j0 + cc, ld)] = tile[PAD + k][cc];
    }
    carry = true;
  }
}

// lanes per row (column) in the split wavefront kernels
constexpr int kLanes = 2;

// windows up to 80 entries in the wavefront kernels
constexpr bool wavefront_fits(int R, int NQ) { return R + NQ - 1 <= 80; }

// (r, nq) pairs with register-window kernels; others use the LDS kernels
#define PHT_S2_PAIRS(X) \
  X(8, 4) X(8, 8) X(16, 4) X(16, 8) X(16, 16) X(32, 8) X(32, 16) X(32, 32) \
  X(64, 8) X(64, 16) X(64, 32)

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
                     P.M, P.ldr, P.nblocks, P.progress, P.nq, P.errflag);
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
                       P.jblock, CU, SU, CV, SV, P.ldr, P.M, P.progress, P.errflag, \
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

hipError_t s2_apply_rows(hipStream_t st, double* A, int lda, double* B,
                         int ldb, double* Q, int ldq, double* Z, int ldz,
                         const S2Params& P, const double* CU, const double* SU,
                         const double* CV, const double* SV, RowsPart part) {
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
#define PHT_ROWS(RR, NN)                                                        \
  if (P.r == RR && P.nq == NN) {                                                \
    const dim3 g((P.n + kRowsThreads - 1) / kRowsThreads, 2);                   \
    if (wavefront_fits(RR, NN) && s2_wavefront()) {                             \
      constexpr int kRows = kSplitThreads / kLanes;                             \
      const dim3 g2((P.n + kRows - 1) / kRows, 2);                              \
      if (part != kRowsQZ)                                                      \
        hipLaunchKernelGGL((k_s2_rows_w2<RR, NN, kLanes, true>), g2, dim3(kSplitThreads), 0, st, A, \
                           lda, B, ldb, Q, ldq, Z, ldz, P.n, P.jblock,          \
                           P.nblocks, CU, SU, CV, SV);                          \
      if (part != kRowsAB)                                                      \
        hipLaunchKernelGGL((k_s2_rows_w2<RR, NN, kLanes, false>), g2, dim3(kSplitThreads), 0, st, A, \
                           lda, B, ldb, Q, ldq, Z, ldz, P.n, P.jblock,          \
                           P.nblocks, CU, SU, CV, SV);                          \
      return hipGetLastError();                                                 \
    }                                                                           \
    if (part != kRowsQZ)                                                        \
      hipLaunchKernelGGL((k_s2_rows_v<RR, NN, true>), g, dim3(kRowsThreads), 0, \
                         st, A, lda, B, ldb, Q, ldq, Z, ldz, P.n, P.jblock,     \
                         P.nblocks, CU, SU, CV, SV);                            \
    if (part != kRowsAB)                                                        \
      hipLaunchKernelGGL((k_s2_rows_v<RR, NN, false>), g, dim3(kRowsThreads), 0, \
                         st, A, lda, B, ldb, Q, ldq, Z, ldz, P.n, P.jblock,     \
                         P.nblocks, CU, SU, CV, SV);                            \
    return hipGetLastError();                                                   \
  }
  PHT_S2_PAIRS(PHT_ROWS)
#undef PHT_ROWS
  dim3 grid((P.n + kTile - 1) / kTile, part == kRowsAll ? 4 : 2), block(kTile);
  hipLaunchKernelGGL(k_s2_rows, grid, block, lds, st, A, lda, B, ldb, Q, ldq, Z,
                     ldz, P.n, P.r, P.nq, P.jblock, P.nblocks, CU, SU, CV, SV,
                     P.ldr, part == kRowsQZ ? 2 : 0);
  return hipGetLastError();
}

hipError_t s2_apply_cols(hipStream_t st, double* A, int lda, double* B,
                         int ldb, const S2Params& P, const double* CU,
                         const double* SU) {
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
#define PHT_COLS(RR, NN)                                                     \
  if (P.r == RR && P.nq == NN) {                                             \
    if (wavefront_fits(RR, NN) && s2_wavefront())                            \
      hipLaunchKernelGGL((k_s2_cols_w2<RR, NN, kLanes>),                     \
                         dim3((P.n + kSplitThreads / kLanes - 1) / (kSplitThreads / kLanes), 2), \
                         dim3(kSplitThreads), 0, st, A, lda, B, ldb, P.n, P.jblock, \
                         P.nblocks, CU, SU);                                 \
    else                                                                     \
      hipLaunchKernelGGL((k_s2_cols_v<RR, NN>), dim3((P.n + 63) / 64, 2),    \
                         dim3(64), 0, st, A, lda, B, ldb, P.n, P.jblock,     \
                         P.nblocks, CU, SU);                                 \
    return hipGetLastError();                                                \
  }
  PHT_S2_PAIRS(PHT_COLS)
#undef PHT_COLS
  dim3 grid((P.n + kTile - 1) / kTile, 2), block(kTile);
  hipLaunchKernelGGL(k_s2_cols, grid, block, lds, st, A, lda, B, ldb, P.n, P.r,
                     P.nq, P.jblock, P.nblocks, CU, SU, P.ldr);
  return hipGetLastError();
}

}  // namespace pht
