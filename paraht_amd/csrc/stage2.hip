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
#include "device_util.hpp"
#include "kernels.hpp"

namespace pht {

namespace {
constexpr int kGenThreads = 256;
constexpr int kTile = 128;

__device__ __forceinline__ size_t tape_idx(int loc1, int ic1, int ib1, int ldr,
                                           int nq) {
  // 1-based (local row, ic, ib) -> offset
  return (size_t)(loc1 - 1) + (size_t)(ic1 - 1) * ldr +
         (size_t)(ib1 - 1) * ldr * nq;
}

__device__ __forceinline__ int imin(int a, int b) { return a < b ? a : b; }
__device__ __forceinline__ int imax(int a, int b) { return a > b ? a : b; }

}  // namespace

// ---------------------------------------------------------------------------
// Generation of one sweep (rht2ht.f90:787-957), one workgroup.
// M holds, per bulge block ib, the product of the left rotations of the
// columns already reduced in this sweep (local rows of the block); the
// reference's lazily-applied left rotations of a column (rht2ht.f90:803-825,
// 887-909) are applied as one mat-vec with it.
template <int RMAX>
__global__ __launch_bounds__(kGenThreads) void k_s2_gen(
    double* __restrict__ A, int lda, double* __restrict__ B, int ldb, int n,
    int r, int nq, int s, int nblocks, double* __restrict__ CU,
    double* __restrict__ SU, double* __restrict__ CV, double* __restrict__ SV,
    int ldr, double* __restrict__ Mw) {
  __shared__ double lc[RMAX], ls[RMAX], rc[RMAX], rs[RMAX], colbuf[RMAX];
  __shared__ double xa[256], xb[256];
  const int tid = threadIdx.x;
  const int lane = tid & 63;

  const size_t tsz = (size_t)ldr * nq * nblocks;
  for (size_t e = tid; e < tsz; e += kGenThreads) {
    CU[e] = 1.0;
    CV[e] = 1.0;
    SU[e] = 0.0;
    SV[e] = 0.0;
  }
  const size_t msz = (size_t)ldr * ldr;
  for (size_t e = tid; e < msz * nblocks; e += kGenThreads) {
    const size_t k = e % msz;
    Mw[e] = ((k % ldr) == (k / ldr)) ? 1.0 : 0.0;
  }
  __syncthreads();

  for (int ic = 1; ic <= nq; ++ic) {
    const int jc = s + ic - 1;
    for (int ib = 1;; ++ib) {
      // step (ic, ib): column `col` is reduced; rows/columns from row0 on
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
      const int nb = imin(r, n - row0 + 1);
      const int base = s + (ib - 1) * r;  // local row l of block ib = row base+l
      const int L = imin(ldr - 1, n - base);
      double* M = Mw + (size_t)(ib - 1) * msz;

      // ---- delayed left rotations of the sweep's earlier columns:
      //      A(:, col) and B(:, row0+r-1), local rows 1..L of block ib
      if (ic > 1 && L > 0) {
        const bool doB = row0 + r - 1 <= n;
        for (int t = tid; t < L; t += kGenThreads) {
          xa[t] = A[PHT_IDX(base + t, col - 1, lda)];
          if (doB) xb[t] = B[PHT_IDX(base + t, row0 + r - 2, ldb)];
        }
        __syncthreads();
        for (int t = tid; t < 2 * L; t += kGenThreads) {
          const bool isB = t >= L;
          if (isB && !doB) continue;
          const int i = isB ? t - L : t;
          const double* x = isB ? xb : xa;
          double acc = 0.0;
          for (int j = 0; j < L; ++j) acc += M[i + (size_t)j * ldr] * x[j];
          if (isB)
            B[PHT_IDX(base + i, row0 + r - 2, ldb)] = acc;
          else
            A[PHT_IDX(base + i, col - 1, lda)] = acc;
        }
        __syncthreads();
      }
      // the reference leaves the column here when nb < 2 (rht2ht.f90:827,911);
      // later blocks still owe this column their delayed rotations
      if (nb < 2) continue;
      const int istart = imax(s + 1, s + 1 + (ib - 1 - nq + ic) * r);
      const int R0 = row0 - 1, C0 = col - 1;  // 0-based anchors

      // ---- left chain: zero A(row0+1:row0+nb-1, col) (rht2ht.f90:830-835,914-919)
      if (tid < nb) colbuf[tid] = A[PHT_IDX(R0 + tid, C0, lda)];
      __syncthreads();
      if (tid == 0) {
        double g = colbuf[nb - 1];
        for (int ig = nb - 1; ig >= 1; --ig) {
          double c, sn, rr;
          dlartg_d(colbuf[ig - 1], g, c, sn, rr);
          lc[ig] = c;
          ls[ig] = sn;
          g = rr;
        }
        colbuf[0] = g;
      }
      __syncthreads();
      if (tid < nb) A[PHT_IDX(R0 + tid, C0, lda)] = (tid == 0) ? colbuf[0] : 0.0;
      if (tid >= 1 && tid < nb) {
        const size_t ti = tape_idx(ic + tid - 1, ic, ib, ldr, nq);
        CU[ti] = lc[tid];
        SU[ti] = ls[tid];
      }

      // ---- left rotations on the B block (rht2ht.f90:837-840,921-924) and
      //      accumulation into M (local rows ic-1 .. ic+nb-2, 0-based)
      for (int t = tid; t < nb + ldr; t += kGenThreads) {
        double* p;
        size_t stride;
        if (t < nb) {
          p = B + PHT_IDX(R0, R0 + t, ldb);
          stride = 1;
        } else {
          p = M + (ic - 1) + (size_t)(t - nb) * ldr;
          stride = 1;
        }
        double seg[RMAX];
#pragma unroll
        for (int i = 0; i < RMAX; ++i) seg[i] = (i < nb) ? p[i * stride] : 0.0;
#pragma unroll
        for (int ig = RMAX - 1; ig >= 1; --ig) {
          if (ig < nb) {
            const double c = lc[ig], sn = ls[ig];
            const double a = seg[ig - 1], b = seg[ig];
            seg[ig - 1] = c * a + sn * b;
            seg[ig] = c * b - sn * a;
          }
        }
#pragma unroll
        for (int i = 0; i < RMAX; ++i)
          if (i < nb) p[i * stride] = seg[i];
      }
      __syncthreads();

      // ---- right chain on the B block: zero the subdiagonal it now has
      //      (rht2ht.f90:843-851,927-935); wave 0, lane = block row
      if (tid < 64) {
        double brow[RMAX];
#pragma unroll
        for (int j = 0; j < RMAX; ++j)
          brow[j] = (lane < nb && j < nb) ? B[PHT_IDX(R0 + lane, R0 + j, ldb)] : 0.0;
#pragma unroll
        for (int ig = RMAX - 1; ig >= 1; --ig) {
          if (ig < nb) {
            const double f = __shfl(brow[ig], ig, 64);
            const double g = __shfl(brow[ig - 1], ig, 64);
            double c, sn, rr;
            dlartg_d(f, g, c, sn, rr);
            if (lane == ig) {
              brow[ig] = rr;
              brow[ig - 1] = 0.0;
            } else if (lane < ig) {
              const double x = brow[ig], y = brow[ig - 1];
              brow[ig] = c * x + sn * y;
              brow[ig - 1] = c * y - sn * x;
            }
            if (lane == 0) {
              rc[ig] = c;
              rs[ig] = sn;
            }
          }
        }
        if (lane < nb) {
#pragma unroll
          for (int j = 0; j < RMAX; ++j)
            if (j < nb) B[PHT_IDX(R0 + lane, R0 + j, ldb)] = brow[j];
        }
      }
      __syncthreads();
      if (tid >= 1 && tid < nb) {
        const size_t ti = tape_idx(ic + tid - 1, ic, ib, ldr, nq);
        CV[ti] = rc[tid];
        SV[ti] = rs[tid];
      }

      // ---- right rotations on the window rows: A rows istart..row0+r+nb-1,
      //      B rows istart..row0-1 (rht2ht.f90:853-867,937-951)
      {
        const int aend = imin(n, row0 + r + nb - 1);  // 1-based inclusive
        const int nAr = aend - istart + 1;
        const int nBr = imax(0, row0 - istart);
        for (int t = tid; t < nAr + nBr; t += kGenThreads) {
          double* Mx;
          int ld, rr0;
          if (t < nAr) {
            Mx = A;
            ld = lda;
            rr0 = istart - 1 + t;
          } else {
            Mx = B;
            ld = ldb;
            rr0 = istart - 1 + (t - nAr);
          }
          double seg[RMAX];
#pragma unroll
          for (int j = 0; j < RMAX; ++j)
            seg[j] = (j < nb) ? Mx[PHT_IDX(rr0, R0 + j, ld)] : 0.0;
#pragma unroll
          for (int ig = RMAX - 1; ig >= 1; --ig) {
            if (ig < nb) {
              const double c = rc[ig], sn = rs[ig];
              const double x = seg[ig], y = seg[ig - 1];
              seg[ig] = c * x + sn * y;
              seg[ig - 1] = c * y - sn * x;
            }
          }
#pragma unroll
          for (int j = 0; j < RMAX; ++j)
            if (j < nb) Mx[PHT_IDX(rr0, R0 + j, ld)] = seg[j];
        }
      }
      __syncthreads();
    }
  }
}

// ---------------------------------------------------------------------------
// Rows: which = 0 A, 1 B (rows above the generation window, right rotations),
// 2 Q (all rows, left rotations seen from the right), 3 Z (all rows, right).
__global__ __launch_bounds__(kTile) void k_s2_rows(
    double* __restrict__ A, int lda, double* __restrict__ B, int ldb,
    double* __restrict__ Q, int ldq, double* __restrict__ Z, int ldz, int n,
    int r, int nq, int s, int nblocks, const double* __restrict__ CU,
    const double* __restrict__ SU, const double* __restrict__ CV,
    const double* __restrict__ SV, int ldr) {
  extern __shared__ double win[];  // [ldr][kTile]
  const int which = blockIdx.y;
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
hipError_t s2_generate(hipStream_t st, double* A, int lda, double* B, int ldb,
                       const S2Params& P, double* CU, double* SU, double* CV,
                       double* SV) {
  const dim3 g(1), b(kGenThreads);
#define PHT_GEN(R)                                                             \
  hipLaunchKernelGGL(k_s2_gen<R>, g, b, 0, st, A, lda, B, ldb, P.n, P.r, P.nq, \
                     P.jblock, P.nblocks, CU, SU, CV, SV, P.ldr, P.M)
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
                         const double* CV, const double* SV) {
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
  dim3 grid((P.n + kTile - 1) / kTile, 4), block(kTile);
  hipLaunchKernelGGL(k_s2_rows, grid, block, lds, st, A, lda, B, ldb, Q, ldq, Z,
                     ldz, P.n, P.r, P.nq, P.jblock, P.nblocks, CU, SU, CV, SV,
                     P.ldr);
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
  dim3 grid((P.n + kTile - 1) / kTile, 2), block(kTile);
  hipLaunchKernelGGL(k_s2_cols, grid, block, lds, st, A, lda, B, ldb, P.n, P.r,
                     P.nq, P.jblock, P.nblocks, CU, SU, P.ldr);
  return hipGetLastError();
}

}  // namespace pht
