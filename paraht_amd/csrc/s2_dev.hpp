// Device code of the stage-2 Givens chase (rht2ht_givens semantics,
// src/rht2ht.f90:696-1344), shared by the pipelined sweep kernel
// (stage2.hip) and the batched persistent kernel.
#pragma once
#include "device_util.hpp"
#include "kernels.hpp"

namespace pht {
namespace s2 {

__device__ __forceinline__ size_t tape_idx(int loc1, int ic1, int ib1, int ldr, int nq) {
  // 1-based (local row, ic, ib) -> offset in CU, SU, CV, SV(ldr, nq, nblocks)
  return (size_t)(loc1 - 1) + (size_t)(ic1 - 1) * ldr + (size_t)(ib1 - 1) * ldr * nq;
}
__device__ __forceinline__ int imin(int a, int b) { return a < b ? a : b; }
__device__ __forceinline__ int imax(int a, int b) { return a > b ? a : b; }

constexpr int kMaxLdr = 160;  // r + nq (host checks)

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

// Right rotations (rc, rs)[ig], ig = nb-1..1, of one step applied to row rr0
// (0-based) of Mx on the step's columns R0..R0+nb-1.
// A window row with fewer than RMAX columns (the last block of a sweep): kept
// out of line so that its per-column tests and addresses do not take
// registers in the kernels that inline the full-block path.
// Mx is in global memory.  Saying so makes its accesses global_* rather than
// flat_* instructions: a flat access counts against the LDS counter too, so
// every later wait for an LDS read would also wait for it to complete.
typedef __attribute__((address_space(1))) double gdouble;
__device__ __forceinline__ gdouble* as_global(double* p) { return (gdouble*)p; }

template <int RMAX>
__device__ __noinline__ void rotate_row_partial(double* __restrict__ Mg, int ld, int rr0, int R0,
                                                int nb, const double* rc, const double* rs) {
  gdouble* Mx = as_global(Mg);
  double seg[RMAX];
#pragma unroll
  for (int j = 0; j < RMAX; ++j) seg[j] = (j < nb) ? Mx[PHT_IDX(rr0, R0 + j, ld)] : 0.0;
  double carry = 0.0;
#pragma unroll
  for (int ig = RMAX - 1; ig >= 1; --ig) {
    if (ig % 8 == 0) __builtin_amdgcn_sched_barrier(0);  // bounds live coefficients
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

// Row rr0 (0-based) of Mx, columns R0..R0+nb-1, times the right rotations
// (rc, rs)[ig], ig = nb-1..1, each on columns (ig-1, ig).
template <int RMAX>
__device__ __forceinline__ void rotate_row(double* __restrict__ Mx, int ld, int rr0, int R0,
                                           int nb, const double* rc, const double* rs) {
  if (nb != RMAX) {
    rotate_row_partial<RMAX>(Mx, ld, rr0, R0, nb, rc, rs);
    return;
  }
  // the row's RMAX entries by a pointer walk (no per-column tests or address
  // tables), all loads in flight together
  double seg[RMAX];
  gdouble* p = as_global(Mx + PHT_IDX(rr0, R0, ld));
#pragma unroll
  for (int j = 0; j < RMAX; ++j) {
    seg[j] = *p;
    p += ld;
  }
  double carry = seg[RMAX - 1];
#pragma unroll
  for (int ig = RMAX - 1; ig >= 1; --ig) {
    if (ig % 8 == 0) __builtin_amdgcn_sched_barrier(0);  // bounds live coefficients
    const double c = rc[ig], sn = rs[ig];
    const double y = seg[ig - 1];
    seg[ig] = c * carry + sn * y;
    carry = c * y - sn * carry;
  }
  seg[0] = carry;
  // a fresh walk for the stores: the load addresses are not kept alive
  asm volatile("" : "+v"(p));
#pragma unroll
  for (int j = RMAX - 1; j >= 0; --j) {
    p -= ld;
    *p = seg[j];
  }
}

// Right chain on the B block Bb (one wave, lane = block row): rotations
// ig = nb-1..1 on columns (ig-1, ig) zero B(ig, ig-1), the pivot values being
// read from lane ig.  Rotation ig only mixes columns ig-1 and ig, and column
// ig-1 is untouched before it, so a lane carries one running value (`cur`,
// its entry in column ig) and streams the original column ig-1 from LDS;
// column ig is final after rotation ig.  (rc, rs)[ig] receive the rotations.
template <int RMAX>
__device__ __forceinline__ void right_chain(double (*Bb)[RMAX + 1], int nb, int lane,
                                            double* rc, double* rs) {
  const bool row = lane < nb;
  double cur = row ? Bb[lane][nb - 1] : 0.0;
  double myc = 1.0, mys = 0.0;  // the rotation pivoted on this lane
  // chunks of 8 rotations: column ig-1 is only written by rotation ig-1,
  // after rotation ig read it, so a chunk's original columns are read up
  // front and the chain never waits on LDS
  for (int ig0 = nb - 1; ig0 >= 1; ig0 -= 8) {
    double yv[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int ig = ig0 - u;
      yv[u] = (row && ig >= 1) ? Bb[lane][ig >= 1 ? ig - 1 : 0] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int ig = ig0 - u;
      if (ig < 1) break;
      const double y = yv[u];
      const double f = readlane_d(cur, ig);
      const double g = readlane_d(y, ig);
      double c, sn, rr;
      rot_rsq(f, g, c, sn, rr);
      const double nx = c * cur + sn * y, ny = c * y - sn * cur;
      const bool piv = lane == ig, above = lane < ig;
      if (row) Bb[lane][ig] = piv ? rr : (above ? nx : cur);
      cur = piv ? 0.0 : (above ? ny : y);
      myc = piv ? c : myc;
      mys = piv ? sn : mys;
    }
  }
  if (row) Bb[lane][0] = cur;
  if (lane >= 1 && lane < nb) {
    rc[lane] = myc;
    rs[lane] = mys;
  }
}

// Left rotations (lc, ls)[ig], ig = nb-1..1, on the entries p[k*st], k < nb,
// of one LDS column: (a, b) = (p[ig-1], p[ig]) <- (c a + s b, c b - s a).
// The reads do not depend on the carried value and run ahead of it.
__device__ __forceinline__ void left_chain_lds(double* p, int st, int nb, const double* lc,
                                               const double* ls) {
  double carry = p[(nb - 1) * st];
#pragma unroll 8
  for (int ig = nb - 1; ig >= 1; --ig) {
    const double c = lc[ig], sn = ls[ig];
    const double a = p[(ig - 1) * st];
    p[ig * st] = c * carry - sn * a;
    carry = c * a + sn * carry;
  }
  p[0] = carry;
}

// M of one block (ldr x ldr, column-major, identity for the first column
// ic = 1) into LDS (ld ldr+1) by nw waves starting at wave w0 (of the thread
// group whose thread tid calls it): a wave per set of columns, lanes along
// the rows (ldr <= LDRMAX rows: (LDRMAX+63)/64 chunks of 64, three at
// kMaxLdr = 160), 8 columns' loads in flight together.
template <int LDRMAX = kMaxLdr>
__device__ __forceinline__ void stage_m(double* Ml, const double* __restrict__ M, int ldr,
                                        int ic, int w0, int nw, int tid) {
  constexpr int CH = (LDRMAX + 63) / 64;
  const int lane = tid & 63, wave = (tid >> 6) - w0;
  const int ldl = ldr + 1;
  for (int cm0 = wave; cm0 < ldr; cm0 += 8 * nw) {
    double tmp[8][CH];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int cm = cm0 + u * nw;
#pragma unroll
      for (int h = 0; h < CH; ++h) {
        const int rw = lane + 64 * h;
        tmp[u][h] = (cm < ldr && rw < ldr)
                        ? ((ic > 1) ? M[(size_t)cm * ldr + rw] : (rw == cm ? 1.0 : 0.0))
                        : 0.0;
      }
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int cm = cm0 + u * nw;
#pragma unroll
      for (int h = 0; h < CH; ++h) {
        const int rw = lane + 64 * h;
        if (cm < ldr && rw < ldr) Ml[cm * ldl + rw] = tmp[u][h];
      }
    }
  }
}

// The threads that run a chase step: a whole workgroup, or a group of whole
// waves of it (the batched kernel runs several columns of a sweep at once,
// one group each).  tid() is the thread's index in the group, sync() a
// barrier of the group that orders its LDS and global memory accesses.
struct WgThreads {
  __device__ int tid() const { return threadIdx.x; }
  __device__ void sync() { __syncthreads(); }
};

struct WaveGroup {
  int base;     // first thread of the group (a multiple of 64)
  int nwaves;   // waves in the group
  int* count;   // the group's arrival counter in LDS (0 when the group starts)
  int arrived;  // barriers this thread has passed
  __device__ int tid() const { return (int)threadIdx.x - base; }
  __device__ void sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    ++arrived;
    if ((threadIdx.x & 63) == 0)
      __hip_atomic_fetch_add(count, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    const int goal = arrived * nwaves;
    // bounded: the group's waves arrive within microseconds; a missing one
    // ends the wait (and corrupts the result) instead of hanging the kernel
    for (int spin = 0;
         __hip_atomic_load(count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) < goal &&
         spin < (1 << 24);
         ++spin)
      __builtin_amdgcn_s_sleep(1);
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
  }
};

// One step (ic, ib) of a sweep starting at column s (rht2ht.f90:787-957), by
// the NT threads of a group (grp: WgThreads or WaveGroup):
//   * M of block ib (the product of the deferred left rotations of columns
//     1..ic-1, identity for ic = 1) is staged in LDS (Ml, ld ldr+1) and
//     applied to A(:, col) and B(:, row0+r-1) as one mat-vec;
//   * the left chain zeroes A(row0+1:row0+nb-1, col) (tape CU, SU), its
//     rotations are applied to the B block and folded into M (written back);
//   * the right chain restores the B block's triangularity (tape CV, SV) and
//     its rotations are applied to the window rows: A rows row0..row0+r+nb-1
//     and, with `above`, A and B rows istart..row0-1.
// col, row0 are 1-based as in the reference.  LDS: Bb[RMAX][RMAX+1], lc, ls,
// rc, rs [RMAX], xa, xb [kMaxLdr], red [2][NT/128][128], Ml [ldr][ldr+1].
// With have_m, Ml already holds M of block ib.  With Mnext (a second LDS
// buffer), waves 1.. stage M of block ib+1 into it while wave 0 runs the right
// chain; the caller must have acquired that block's M (the previous column's
// step ib+1 is complete).  Returns whether Mnext was filled.
// LDRMAX bounds ldr = r + nq (the M staging holds (LDRMAX+63)/64 chunks).
template <int RMAX, int NT, class Grp, int LDRMAX = kMaxLdr>
__device__ inline bool chase_step(Grp& grp, double* __restrict__ A, int lda, double* __restrict__ B,
                                  int ldb, int n, int r, int nq, int s, int ic, int ib, int col,
                                  int row0, double* __restrict__ CU, double* __restrict__ SU,
                                  double* __restrict__ CV, double* __restrict__ SV, int ldr,
                                  double* __restrict__ Mw, bool above, double (*Bb)[RMAX + 1],
                                  double* lc, double* ls, double* rc, double* rs, double* xa,
                                  double* xb, double* red, double* Ml, bool have_m = false,
                                  double* Mnext = nullptr) {
  constexpr int kParts = NT / 128;
  const int tid = grp.tid();
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int ldl = ldr + 1;
  const size_t msz = (size_t)ldr * ldr;
  const int nb = imin(r, n - row0 + 1);
  const int base = s + (ib - 1) * r;  // local row l of block ib = row base+l
  const int L = imin(ldr - 1, n - base);
  double* M = Mw + (size_t)(ib - 1) * msz;

  // ---- M of this block into LDS (unless the previous step staged it)
  if (!have_m) stage_m<LDRMAX>(Ml, M, ldr, ic, 0, NT / 64, tid);
  grp.sync();
  bool staged_next = false;

  // ---- deferred left rotations of the sweep's earlier columns:
  //      A(:, col) and B(:, row0+r-1), local rows 1..L of block ib.
  //      Thread t: output rows i = t%128 (+128), half h = t/128 of the
  //      reduction; its loads of M are all in flight together.
  if (ic > 1 && L > 0) {
    const bool doB = row0 + r - 1 <= n;
    for (int t = tid; t < L; t += NT) {
      xa[t] = A[PHT_IDX(base + t, col - 1, lda)];
      xb[t] = doB ? B[PHT_IDX(base + t, row0 + r - 2, ldb)] : 0.0;
    }
    grp.sync();
    const int h = tid >> 7, il = tid & 127;
    const int jp = (L + kParts - 1) / kParts, j0 = h * jp, j1 = imin(L, j0 + jp);
    for (int i0 = 0; i0 < L; i0 += 128) {
      const int i = i0 + il;
      double ya = 0.0, yb = 0.0;
      if (i < L) {
        for (int j = j0; j < j1; ++j) {
          const double m = Ml[j * ldl + i];
          ya += m * xa[j];
          yb += m * xb[j];
        }
      }
      red[(0 * kParts + h) * 128 + il] = ya;
      red[(1 * kParts + h) * 128 + il] = yb;
      grp.sync();
      if (h == 0 && i < L) {
        double sa = 0.0, sb = 0.0;
#pragma unroll
        for (int q = 0; q < kParts; ++q) {
          sa += red[(0 * kParts + q) * 128 + il];
          sb += red[(1 * kParts + q) * 128 + il];
        }
        A[PHT_IDX(base + i, col - 1, lda)] = sa;
        if (doB) B[PHT_IDX(base + i, row0 + r - 2, ldb)] = sb;
      }
      grp.sync();
    }
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
    // B block into LDS (coalesced by rows; all loads in flight together)
    {
      constexpr int PER = (RMAX * RMAX + NT - 1) / NT;
      double tmp[PER];
#pragma unroll
      for (int q = 0; q < PER; ++q) {
        const int e = tid + q * NT;
        const int k = e % RMAX, j = e / RMAX;
        tmp[q] = (k < nb && j < nb) ? B[PHT_IDX(R0 + k, R0 + j, ldb)] : 0.0;
      }
#pragma unroll
      for (int q = 0; q < PER; ++q) {
        const int e = tid + q * NT;
        if (e < RMAX * RMAX) Bb[e % RMAX][e / RMAX] = tmp[q];
      }
    }
    grp.sync();

    // ---- left rotations on the B block (rht2ht.f90:837-840,921-924):
    //      threads 0..nb-1, one block column each; meanwhile threads
    //      64.. fold the same rotations into M (local rows ic-1..ic+nb-2)
    if (tid < nb)
      left_chain_lds(&Bb[0][tid], RMAX + 1, nb, lc, ls);
    else if (tid >= 64 && tid < 64 + ldr)
      left_chain_lds(&Ml[(tid - 64) * ldl + (ic - 1)], 1, nb, lc, ls);
    grp.sync();
    // the updated M rows go back while wave 0 runs the right chain
    if (wave > 0) {
      constexpr int PM = (RMAX * kMaxLdr + NT - 64 - 1) / (NT - 64);
#pragma unroll
      for (int q = 0; q < PM; ++q) {
        const int e = (tid - 64) + q * (NT - 64);
        const int k = e % RMAX, cm = e / RMAX;
        if (k < nb && cm < ldr) M[(ic - 1) + k + (size_t)cm * ldr] = Ml[cm * ldl + (ic - 1) + k];
      }
      if (NT > 64 && Mnext != nullptr) stage_m<LDRMAX>(Mnext, M + msz, ldr, ic, 1, NT / 64 - 1, tid);
    }
    staged_next = NT > 64 && Mnext != nullptr;

    // ---- right chain on the B block (rht2ht.f90:843-851,927-935):
    //      wave 0, lane = block row; the pivot values are read from lane ig
    if (wave == 0) right_chain<RMAX>(Bb, nb, lane, rc, rs);
    grp.sync();

    // ---- the B block and the tape go back, and the right rotations are
    //      applied to the window rows (rht2ht.f90:853-867,937-951): A rows
    //      istart..row0+r+nb-1 and B rows istart..row0-1; without `above` only
    //      A rows row0..row0+r+nb-1 (helpers take the rest).  The window rows
    //      never include the B block's rows, so when they fit in half the
    //      workgroup the other half stores the block and the tape meanwhile.
    const int aend = imin(n, row0 + r + nb - 1);  // 1-based inclusive
    const int first = above ? istart : row0;
    const int nAr = aend - first + 1;
    const int nBr = above ? row0 - istart : 0;
    const bool halves = nAr + nBr <= NT / 2;
    const int wt = halves ? NT / 2 : NT;      // threads on the window rows
    const int st0 = halves ? NT / 2 : 0;      // first thread of the stores
    if (tid >= st0) {
      for (int e = tid - st0; e < nb * nb; e += NT - st0) {
        const int k = e % nb, j = e / nb;
        B[PHT_IDX(R0 + k, R0 + j, ldb)] = Bb[k][j];
      }
      const int tt = tid - st0;
      if (tt >= 1 && tt < nb) {
        const size_t ti = tape_idx(ic + tt - 1, ic, ib, ldr, nq);
        CV[ti] = rc[tt];
        SV[ti] = rs[tt];
      }
    }
    if (tid < wt) {
      for (int t = tid; t < nAr + nBr; t += wt) {
        const bool inA = t < nAr;
        rotate_row<RMAX>(inA ? A : B, inA ? lda : ldb, first - 1 + (inA ? t : t - nAr), R0,
                         nb, rc, rs);
      }
    }
  }
  return staged_next;
}

}  // namespace s2
}  // namespace pht
