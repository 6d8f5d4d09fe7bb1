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
constexpr int kDone = 1 << 29;

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
  for (size_t e = gid; e < (size_t)nq; e += gsz) progress[e] = 0;
  if (gid == 0) *errflag = 0;
}

// Givens rotation from (f, g) when |r| = d is already known (d > 0):
// [c s; -s c][f; g] = [r; 0], c >= 0, r with the sign of f (f == 0: c = 0,
// s = sign g) -- the conventions of dlartg_d.
__device__ __forceinline__ void rot_from_norm(double f, double g, double d,
                                              double& c, double& sn) {
  if (g == 0.0) {
    c = 1.0;
    sn = 0.0;
  } else if (f == 0.0) {
    c = 0.0;
    sn = copysign(1.0, g);
  } else {
    const double rinv = 1.0 / d;
    c = fabs(f) * rinv;
    sn = ((f < 0.0) ? -g : g) * rinv;  // s = g / r with r = sign(f) d
  }
}

template <int RMAX>
__global__ __launch_bounds__(kGenThreads) void k_s2_sweep(
    double* __restrict__ A, int lda, double* __restrict__ B, int ldb, int n,
    int r, int nq, int s, double* __restrict__ CU, double* __restrict__ SU,
    double* __restrict__ CV, double* __restrict__ SV, int ldr,
    double* __restrict__ Mw, int* progress, int* errflag) {
  __shared__ double Bb[RMAX][RMAX + 1];
  __shared__ double lc[RMAX], ls[RMAX], rc[RMAX], rs[RMAX];
  __shared__ double xa[160], xb[160];
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int ic = blockIdx.x + 1;
  const int jc = s + ic - 1;
  const size_t msz = (size_t)ldr * ldr;

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
    if (ic > 1) {
      if (tid == 0) {
        int spins = 0;
        while (__hip_atomic_load(&progress[ic - 2], __ATOMIC_ACQUIRE,
                                 __HIP_MEMORY_SCOPE_AGENT) < ib + 2) {
          __builtin_amdgcn_s_sleep(2);
          if (++spins > kSpinLimit) {
            __hip_atomic_store(errflag, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            break;
          }
        }
      }
      __syncthreads();
      __threadfence();
    }
    const int nb = imin(r, n - row0 + 1);
    const int base = s + (ib - 1) * r;  // local row l of block ib = row base+l
    const int L = imin(ldr - 1, n - base);
    double* M = Mw + (size_t)(ib - 1) * msz;

    // ---- deferred left rotations of the sweep's earlier columns:
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
        double acc0 = 0.0, acc1 = 0.0;
        int j = 0;
        for (; j + 1 < L; j += 2) {
          acc0 += M[i + (size_t)j * ldr] * x[j];
          acc1 += M[i + (size_t)(j + 1) * ldr] * x[j + 1];
        }
        if (j < L) acc0 += M[i + (size_t)j * ldr] * x[j];
        if (isB)
          B[PHT_IDX(base + i, row0 + r - 2, ldb)] = acc0 + acc1;
        else
          A[PHT_IDX(base + i, col - 1, lda)] = acc0 + acc1;
      }
      __syncthreads();
    }

    // the reference leaves the column here when nb < 2 (rht2ht.f90:827,911);
    // later blocks still owe this column their deferred rotations
    if (nb >= 2) {
      const int istart = imax(s + 1, s + 1 + (ib - 1 - nq + ic) * r);
      const int R0 = row0 - 1, C0 = col - 1;  // 0-based anchors

      // ---- left chain: zero A(row0+1:row0+nb-1, col) (rht2ht.f90:830-835,
      //      914-919).  The chain's partial results are signed suffix norms of
      //      the column, so every rotation follows from a suffix sum of
      //      squares (wave 0, lane = row).
      if (wave == 0) {
        const double x = (lane < nb) ? A[PHT_IDX(R0 + lane, C0, lda)] : 0.0;
        const double mx = wave_max(fabs(x));
        const double xs = (mx > 0.0) ? x / mx : 0.0;
        double S = xs * xs;  // inclusive suffix sum over lanes >= lane
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
          const double t = __shfl_down(S, off, 64);
          if (lane + off < 64) S += t;
        }
        const double Snext = __shfl_down(S, 1, 64);  // suffix from lane+1
        // g: the value the chain leaves at row row0+lane
        double g;
        if (lane >= nb - 1 || Snext == 0.0)
          g = x;
        else
          g = (x < 0.0 ? -1.0 : 1.0) * (sqrt(S) * mx);
        const double f_above = __shfl_up(x, 1, 64);        // x_{lane-1}
        const double d_above = __shfl_up(fabs(g), 1, 64);  // |g_{lane-1}|
        if (lane >= 1 && lane < nb) {
          double c, sn;
          rot_from_norm(f_above, g, d_above, c, sn);
          lc[lane] = c;
          ls[lane] = sn;
          const size_t ti = tape_idx(ic + lane - 1, ic, ib, ldr, nq);
          CU[ti] = c;
          SU[ti] = sn;
        }
        if (lane < nb) A[PHT_IDX(R0 + lane, C0, lda)] = (lane == 0) ? g : 0.0;
      }
      // B block into LDS (coalesced by rows)
      for (int e = tid; e < nb * nb; e += kGenThreads) {
        const int k = e % nb, j = e / nb;
        Bb[k][j] = B[PHT_IDX(R0 + k, R0 + j, ldb)];
      }
      __syncthreads();

      // ---- left rotations on the B block (rht2ht.f90:837-840,921-924):
      //      threads 0..nb-1, one block column each; meanwhile threads
      //      64.. fold the same rotations into M (local rows ic-1..ic+nb-2)
      if (tid < nb) {
        const int j = tid;
        double carry = Bb[nb - 1][j];
        for (int ig = nb - 1; ig >= 1; --ig) {
          const double c = lc[ig], sn = ls[ig];
          const double a = Bb[ig - 1][j];
          Bb[ig][j] = c * carry - sn * a;
          carry = c * a + sn * carry;
        }
        Bb[0][j] = carry;
      } else if (tid >= 64 && tid < 64 + ldr) {
        double* p = M + (ic - 1) + (size_t)(tid - 64) * ldr;
        double carry = p[nb - 1];
        for (int ig = nb - 1; ig >= 1; --ig) {
          const double c = lc[ig], sn = ls[ig];
          const double a = p[ig - 1];
          p[ig] = c * carry - sn * a;
          carry = c * a + sn * carry;
        }
        p[0] = carry;
      }
      __syncthreads();

      // ---- right chain on the B block (rht2ht.f90:843-851,927-935):
      //      wave 0, lane = block row, rotations broadcast by shuffles
      if (wave == 0) {
        double brow[RMAX];
#pragma unroll
        for (int j = 0; j < RMAX; ++j) brow[j] = (lane < nb && j < nb) ? Bb[lane][j] : 0.0;
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
            if (j < nb) Bb[lane][j] = brow[j];
        }
      }
      __syncthreads();
      for (int e = tid; e < nb * nb; e += kGenThreads) {
        const int k = e % nb, j = e / nb;
        B[PHT_IDX(R0 + k, R0 + j, ldb)] = Bb[k][j];
      }
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
          double carry = 0.0;
#pragma unroll
          for (int ig = RMAX - 1; ig >= 1; --ig) {
            if (ig < nb) {
              if (ig == nb - 1) carry = seg[ig];
              const double c = rc[ig], sn = rs[ig];
              const double y = seg[ig - 1];
              seg[ig] = c * carry + sn * y;
              carry = c * y - sn * carry;
            }
          }
          seg[0] = carry;
#pragma unroll
          for (int j = 0; j < RMAX; ++j)
            if (j < nb) Mx[PHT_IDX(rr0, R0 + j, ld)] = seg[j];
        }
      }
    }
    // ---- publish the step
    __threadfence();
    __syncthreads();
    if (tid == 0)
      __hip_atomic_store(&progress[ic - 1], ib, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
  }
  __threadfence();
  __syncthreads();
  if (tid == 0)
    __hip_atomic_store(&progress[ic - 1], kDone, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
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
// Register-window rotation sweeps for compile-time (r, nq).  A thread keeps a
// window of R consecutive entries of its row (or column) in registers; the
// rotations of column ic of a block act on window positions (ig-1, ig)
// (local indices ic+ig-1, ic+ig), after which local index ic is final, is
// stored, and the window slides by one.  Rotation coefficients are uniform
// across the workgroup and come from the tape through scalar loads.
template <int R, int NQ>
__global__ __launch_bounds__(256) void k_s2_rows_v(
    double* __restrict__ A, int lda, double* __restrict__ B, int ldb,
    double* __restrict__ Q, int ldq, double* __restrict__ Z, int ldz, int n,
    int s, int nblocks, const double* __restrict__ CU,
    const double* __restrict__ SU, const double* __restrict__ CV,
    const double* __restrict__ SV) {
  constexpr int LDR = R + NQ;
  const int which = blockIdx.y;
  const int row_first = blockIdx.x * 256;
  const int i = row_first + threadIdx.x;  // 0-based row
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
    const int cw0 = s + (ib - 1) * R;  // 0-based first window column
    const int W = imin(LDR - 1, n - cw0);
    if (W <= 0) continue;
    if (partial && row_first + 1 >= imax(s + 1, s + 1 + (ib - 1) * R)) continue;
    double v[R];
#pragma unroll
    for (int k = 0; k < R; ++k)
      v[k] = (valid && k < W) ? M[PHT_IDX(i, cw0 + k, ld)] : 0.0;
    const size_t boff = (size_t)(ib - 1) * LDR * NQ;
    for (int ic = 1; ic <= NQ; ++ic) {
      bool act = valid;
      if (partial) act = act && (i + 1 < imax(s + 1, s + 1 + (ib - 1 - NQ + ic) * R));
      const double* cc = Ct + boff + (size_t)(ic - 1) * LDR + (ic - 2);
      const double* ss = St + boff + (size_t)(ic - 1) * LDR + (ic - 2);
      if (__any(act)) {
#pragma unroll
        for (int ig = R - 1; ig >= 1; --ig) {
          if (ic + ig - 1 < W) {
            const double c = cc[ig], sn = ss[ig];
            const double x = v[ig], y = v[ig - 1];
            double nx, ny;
            if (which == 2) {
              ny = c * y + sn * x;
              nx = c * x - sn * y;
            } else {
              nx = c * x + sn * y;
              ny = c * y - sn * x;
            }
            v[ig] = act ? nx : x;
            v[ig - 1] = act ? ny : y;
          }
        }
      }
      if (valid && ic - 1 < W) M[PHT_IDX(i, cw0 + ic - 1, ld)] = v[0];
#pragma unroll
      for (int k = 0; k < R - 1; ++k) v[k] = v[k + 1];
      v[R - 1] = (valid && ic - 1 + R < W) ? M[PHT_IDX(i, cw0 + ic - 1 + R, ld)] : 0.0;
    }
#pragma unroll
    for (int k = 0; k < R - 1; ++k)
      if (valid && NQ + k < W) M[PHT_IDX(i, cw0 + NQ + k, ld)] = v[k];
  }
}

template <int R, int NQ>
__global__ __launch_bounds__(64) void k_s2_cols_v(
    double* __restrict__ A, int lda, double* __restrict__ B, int ldb, int n,
    int s, int nblocks, const double* __restrict__ CU,
    const double* __restrict__ SU) {
  constexpr int LDR = R + NQ;
  __shared__ double tile[LDR][65];
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
    const int W = imin(LDR - 1, n - rw0);
    if (W <= 0) continue;
    const int cmax = imin(64, n - j0);
    for (int c = 0; c < cmax; ++c)
      for (int k = t; k < W; k += 64) tile[k][c] = M[PHT_IDX(rw0 + k, j0 + c, ld)];
    __syncthreads();
    const bool act = (j < n) && (j + 1 > istop);
    double v[R];
#pragma unroll
    for (int k = 0; k < R; ++k) v[k] = (k < W) ? tile[k][t] : 0.0;
    const size_t boff = (size_t)(ib - 1) * LDR * NQ;
    for (int ic = 1; ic <= NQ; ++ic) {
      const double* cc = CU + boff + (size_t)(ic - 1) * LDR + (ic - 2);
      const double* ss = SU + boff + (size_t)(ic - 1) * LDR + (ic - 2);
      if (act) {
#pragma unroll
        for (int ig = R - 1; ig >= 1; --ig) {
          if (ic + ig - 1 < W) {
            const double c = cc[ig], sn = ss[ig];
            const double a = v[ig - 1], b = v[ig];
            v[ig - 1] = c * a + sn * b;
            v[ig] = c * b - sn * a;
          }
        }
      }
      if (ic - 1 < W) tile[ic - 1][t] = v[0];
#pragma unroll
      for (int k = 0; k < R - 1; ++k) v[k] = v[k + 1];
      v[R - 1] = (ic - 1 + R < W) ? tile[ic - 1 + R][t] : 0.0;
    }
#pragma unroll
    for (int k = 0; k < R - 1; ++k)
      if (NQ + k < W) tile[NQ + k][t] = v[k];
    __syncthreads();
    for (int c = 0; c < cmax; ++c) {
      if (j0 + c + 1 <= istop) continue;
      for (int k = t; k < W; k += 64) M[PHT_IDX(rw0 + k, j0 + c, ld)] = tile[k][c];
    }
    __syncthreads();
  }
}

// (r, nq) pairs with register-window kernels; others use the LDS kernels
#define PHT_S2_PAIRS(X) \
  X(8, 4) X(8, 8) X(16, 4) X(16, 8) X(16, 16) X(32, 8) X(32, 16) X(32, 32) \
  X(64, 8) X(64, 16) X(64, 32)

// ---------------------------------------------------------------------------
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
  const dim3 g(P.nq), b(kGenThreads);
#define PHT_GEN(R)                                                              \
  hipLaunchKernelGGL(k_s2_sweep<R>, g, b, 0, st, A, lda, B, ldb, P.n, P.r, P.nq, \
                     P.jblock, CU, SU, CV, SV, P.ldr, P.M, P.progress, P.errflag)
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
#define PHT_ROWS(RR, NN)                                                      \
  if (P.r == RR && P.nq == NN) {                                              \
    hipLaunchKernelGGL((k_s2_rows_v<RR, NN>), dim3((P.n + 255) / 256, 4),     \
                       dim3(256), 0, st, A, lda, B, ldb, Q, ldq, Z, ldz, P.n, \
                       P.jblock, P.nblocks, CU, SU, CV, SV);                  \
    return hipGetLastError();                                                 \
  }
  PHT_S2_PAIRS(PHT_ROWS)
#undef PHT_ROWS
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
#define PHT_COLS(RR, NN)                                                     \
  if (P.r == RR && P.nq == NN) {                                             \
    hipLaunchKernelGGL((k_s2_cols_v<RR, NN>), dim3((P.n + 63) / 64, 2),      \
                       dim3(64), 0, st, A, lda, B, ldb, P.n, P.jblock,       \
                       P.nblocks, CU, SU);                                   \
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
