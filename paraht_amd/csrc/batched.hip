// Batched persistent reduction of many independent small pencils.
//
// One launch of k_batched reduces a whole batch.  There is one 1024-thread
// workgroup per compute unit; each takes pencils from an atomic work queue
// (the host orders the queue by decreasing n^3) and runs the complete
// two-stage reduction of that pencil by itself, with the same algorithms as
// the single-pencil path:
//  * stage 1 (ab2rht_simple, src/ab2rht.f90:8-495): the panel QR and RQ
//    generators of s1_dev.hpp and the compact-WY MFMA applier of
//    apply_dev.hpp, its 16 waves splitting the other dimension of each update;
//  * stage 2 (rht2ht_givens semantics, src/rht2ht.f90:696-1344): the chase
//    steps of s2_dev.hpp in the reference's order (column by column, block by
//    block, each with its whole window), then the far updates of the sweep
//    with register windows.
// Every pencil has its own (r, p, nq), normally from the size-bucket table
// (batched_defaults) or the autotuner; r <= 16, p*r <= 128 and (r, nq) one of
// the pairs of PHT_BATCH_PAIRS, or r = 1: the one-stage reduction of a pencil
// of order <= 64 in LDS (onestage_wg).
#include "apply_dev.hpp"
#include "s1_dev.hpp"
#include "s2_dev.hpp"

namespace pht {

namespace {
constexpr int kBT = 1024;      // threads of a batched workgroup (= s1::kT)
constexpr int kBRmax = 16;     // largest r of the batched path
constexpr int kBWaves = kBT / 64;
static_assert(kBT == s1::kT, "the stage-1 bodies assume 1024 threads");

#define PHT_BATCH_PAIRS(X) X(8, 4) X(8, 8) X(16, 4) X(16, 8) X(16, 16)

// ---- stage 1 -------------------------------------------------------------

// Block bl of the panel at 0-based column jb (bottom-up order, ab2rht.f90:94-99)
struct PanelGeom {
  int n, nb, p, jb, nblocks, P;
  __device__ ApplyBlock block(int bl, int lo, int hi, bool right) const {
    const int i0 = jb + nb + (nblocks - 1 - bl) * (p - 1) * nb;
    const int len = min(n - i0, P);
    const int k = right ? ((len > nb) ? nb : len) : min(nb, len);
    ApplyBlock b;
    b.start = i0;
    b.len = len;
    b.k = k;
    b.lo = lo < 0 ? i0 : lo;
    b.hi = hi < 0 ? i0 + len : hi;
    b.pad_ = 0;
    b.voff = (long long)bl * P * nb;
    b.toff = (long long)bl * nb * nb;
    return b;
  }
};

// Blocks [b0, b1) of a panel applied to C (other dimension [olo, ohi)) by
// the 16 waves of the workgroup, each owning 16-index slices; lo/hi of -1
// take the block's own start / end.
template <int SIDE>
__device__ void apply_wg(double* C, int ldc, const double* V, int ldv, const double* T, int ldt,
                         const PanelGeom& g, int b0, int b1, int lo, int hi, bool right,
                         int olo, int ohi) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int o0 = olo + 16 * wave; o0 < ohi; o0 += 16 * kBWaves) {
    for (int bl = b0; bl < b1; ++bl) {
      const ApplyBlock bk = g.block(bl, lo, hi, right);
      if (o0 < bk.hi && o0 + 16 > bk.lo)
        apply_block_slice<SIDE, 1>(C, ldc, V, ldv, T, ldt, bk, o0, lane);
      __threadfence_block();  // the next block reads what this one wrote
    }
  }
  __syncthreads();
}

// The panel QR and the RQ generator out of line: inlined into stage1_wg they
// shared its register allocation and spilled (1.5k scratch reloads in the
// block loop at 128 registers); on their own, like the single-pencil
// generator kernels, they do not.
__device__ __noinline__ void panel_qr_wg(double* sm, double* A, int n, int jb, int nb, int p,
                                         int nblocks, double* VL, int P, double* TL) {
  s1::panel_qr_body(sm, A, n, n, jb, nb, p, nblocks, VL, P, (long long)P * nb, TL, nb,
                    (long long)nb * nb);
}
__device__ __noinline__ void rq_gen_wg(double* sm, const double* B, int n, int i0, int len, int nb,
                                       double* V, int P, double* T) {
  s1::rq_gen_body(sm, B, n, i0, len, nb, V, P, T, nb);
}

__device__ __noinline__ void stage1_wg(double* sm, double* A, double* B, double* Q, double* Z, int n, int nb,
                          int p, double* ws) {
  const int P = p * nb;
  const int nbmax = (n - nb + (p - 1) * nb - 1) / ((p - 1) * nb) + 1;
  double* VL = ws;
  double* VR = VL + (size_t)nbmax * P * nb;
  double* TL = VR + (size_t)nbmax * P * nb;
  double* TR = TL + (size_t)nbmax * nb * nb;
  for (int jb = 0; jb <= n - nb - 2; jb += nb) {
    int nblocks = (n - nb - jb) / ((p - 1) * nb);
    if ((n - nb - jb) % ((p - 1) * nb)) ++nblocks;
    const PanelGeom g{n, nb, p, jb, nblocks, P};
    // left generation and application (ab2rht.f90:85-250)
    panel_qr_wg(sm, A, n, jb, nb, p, nblocks, VL, P, TL);
    __syncthreads();
    apply_wg<kLeftT>(B, n, VL, P, TL, nb, g, 0, nblocks, -1, n, false, jb + nb, n);
    apply_wg<kLeftT>(A, n, VL, P, TL, nb, g, 0, nblocks, jb + nb, n, false, jb + nb, n);
    apply_wg<kRightN>(Q, n, VL, P, TL, nb, g, 0, nblocks, 0, n, false, 0, n);
    // right generation, block by block, with B's update (ab2rht.f90:270-336)
    for (int bl = 0; bl < nblocks; ++bl) {
      const ApplyBlock bk = g.block(bl, 0, -1, true);
      rq_gen_wg(sm, B, n, bk.start, bk.len, nb, VR + bk.voff, P, TR + bk.toff);
      __threadfence_block();
      __syncthreads();
      apply_wg<kRightN>(B, n, VR, P, TR, nb, g, bl, bl + 1, 0, -1, true, 0, bk.start + bk.len);
      for (int e = threadIdx.x; e < bk.k * bk.len; e += kBT) {
        const int c = e / bk.len, r = e % bk.len;  // B(i0+r, i0+c) = 0 below the diagonal
        if (r > c) B[PHT_IDX(bk.start + r, bk.start + c, n)] = 0.0;
      }
      __threadfence_block();
      __syncthreads();
    }
    // right application to A and Z (ab2rht.f90:371-467)
    apply_wg<kRightN>(A, n, VR, P, TR, nb, g, 0, nblocks, 0, n, true, 0, n);
    apply_wg<kRightN>(Z, n, VR, P, TR, nb, g, 0, nblocks, 0, n, true, 0, n);
  }
}

// ---- stage 2 -------------------------------------------------------------

// Far updates of one sweep (rht2ht.f90:1019-1325) for compile-time (R, NQ):
// rows of A, B above each block's window and all rows of Q, Z (right
// rotations; Q takes the transposed ones), then columns of A, B right of
// the window (left rotations).  A thread keeps a window of R entries in
// registers and slides it by one per column ic; the block's rotations (both
// sides) are staged in LDS.  The rows of the four matrices, and the columns
// of the two, are one task list per block: the rotation chains are latency
// bound, so the pass is as fast as its most loaded wave and wants every wave.
// Window entries in flight per row or column, and rotation coefficients per
// compiler fence (16: one fence per sequence).  Measured on 2048 pencils of
// order 32..512, far updates per workgroup: (4, 16) 93 ms, (8, 4) and (8, 16)
// 100 ms; 262 ms with each entry loaded as it is needed.
constexpr int kFarAhead = 4;
constexpr int kCsAhead = 16;

// The NQ rotation sequences of a block applied to one row or column of a
// matrix: entry e of the window is p[e * stride], entries e < W exist, and
// sequence icm (skipped where bit icm of `on` is clear) applies rotations
// ig = R-1..1 to the pairs (icm+ig-1, icm+ig).  A window of R entries slides
// by one per sequence; the entries it takes in are loaded D sequences ahead,
// so the chain of rotations does not wait on memory.
template <int R, int NQ, int D, bool FULL>
__device__ __forceinline__ void apply_window_w(double* pg, size_t stride, int W, const double2* cs,
                                               unsigned on) {
  if (FULL) W = R + NQ - 1;  // every entry exists: no guard on any load or store
  // The matrices are in global memory: global (not flat) loads and stores
  // count only against vmcnt, so the waits for the LDS coefficients do not
  // also drain the loads in flight.
  __attribute__((address_space(1))) double* p = (__attribute__((address_space(1))) double*)pg;
  typedef const __attribute__((address_space(3))) double lds_double;
  lds_double* csl = (lds_double*)cs;  // (c, s) pairs
  double v[R], pf[D];
#pragma unroll
  for (int k = 0; k < R; ++k) v[k] = k < W ? p[k * stride] : 0.0;
#pragma unroll
  for (int d = 0; d < D; ++d) pf[d] = R + d < W ? p[(R + d) * stride] : 0.0;
#pragma unroll
  for (int icm = 0; icm < NQ; ++icm) {
    if ((on >> icm) & 1u) {
#pragma unroll
      for (int ig = R - 1; ig >= 1; --ig) {
        // at most kCsAhead coefficients loaded ahead: the compiler would
        // otherwise load far more (all NQ * R of them, ahead of the loop
        // over tasks, where they are the same for every task) and spill
        if ((R - 1 - ig) % kCsAhead == 0) asm volatile("" : "+v"(csl));
        const double cc = csl[2 * (icm * R + ig)], sn = csl[2 * (icm * R + ig) + 1];
        const double x = v[ig], y = v[ig - 1];
        v[ig] = cc * x + sn * y;
        v[ig - 1] = cc * y - sn * x;
      }
    }
    if (icm < W) p[icm * stride] = v[0];
#pragma unroll
    for (int k = 0; k < R - 1; ++k) v[k] = v[k + 1];
    v[R - 1] = pf[icm % D];  // entry R + icm
    const int e = R + icm + D;
    pf[icm % D] = e < W ? p[e * stride] : 0.0;
  }
#pragma unroll
  for (int k = 0; k < R - 1; ++k)
    if (NQ + k < W) p[(NQ + k) * stride] = v[k];
}

template <int R, int NQ, int D>
__device__ __forceinline__ void apply_window(double* p, size_t stride, int W, const double2* cs,
                                             unsigned on) {
  if (W == R + NQ - 1)
    apply_window_w<R, NQ, D, true>(p, stride, W, cs, on);
  else
    apply_window_w<R, NQ, D, false>(p, stride, W, cs, on);
}

template <int R, int NQ>
__device__ __noinline__ void far_updates_wg(double* sm, double* A, double* B, double* Q, double* Z, int n,
                               int s, int nblocks, const double* CU, const double* SU,
                               const double* CV, const double* SV) {
  constexpr int LDR = R + NQ;
  double2* csv = reinterpret_cast<double2*>(sm);  // [NQ][R] right rotations (rows of A, B, Z)
  double2* csu = csv + NQ * R;                     // [NQ][R] left rotations (rows of Q, columns)
  const int tid = threadIdx.x;
  // One task per row, over the rows of all four matrices at once, so that a
  // block's pass keeps the whole workgroup busy: tasks [0, n) are Z's rows,
  // [n, 2n) Q's, then the rows of A and of B above the block window (only
  // those meet a rotation of the block).  Each row sees the blocks in the
  // order ib = nblocks..1 as in the sequential update.
  for (int ib = nblocks; ib >= 1; --ib) {
    const int cw0 = s + (ib - 1) * R;  // 0-based first window column
    const int W = min(R + NQ - 1, n - cw0);
    if (W <= 0) continue;
    const int na = min(n, s + (ib - 1) * R);  // rows of A (and of B) above the window
    __syncthreads();  // the previous block's coefficients are consumed
    for (int e = tid; e < NQ * R; e += kBT) {
      const int icm = e / R, ig = e % R;
      const size_t ti = (size_t)(ib - 1) * LDR * NQ + (size_t)icm * LDR + (icm + ig - 1);
      csv[e] = (ig >= 1) ? make_double2(CV[ti], SV[ti]) : make_double2(1.0, 0.0);
      csu[e] = (ig >= 1) ? make_double2(CU[ti], -SU[ti]) : make_double2(1.0, 0.0);
    }
    __syncthreads();
    const int ntask = 2 * n + 2 * na;
    for (int t = tid; t < ntask; t += kBT) {
      const int which = t < n ? 3 : t < 2 * n ? 2 : t < 2 * n + na ? 0 : 1;  // 0 A, 1 B, 2 Q, 3 Z
      const int i = t < 2 * n ? t - (which == 3 ? 0 : n) : t - 2 * n - (which == 0 ? 0 : na);
      double* M = which == 0 ? A : which == 1 ? B : which == 2 ? Q : Z;
      const double2* cs = which == 2 ? csu : csv;
      unsigned on = 0;
#pragma unroll
      for (int icm = 0; icm < NQ; ++icm)
        on |= (which >= 2 || i + 1 < max(s + 1, s + 1 + (ib - NQ + icm) * R)) ? (1u << icm) : 0u;
      apply_window<R, NQ, kFarAhead>(&M[PHT_IDX(i, cw0, n)], (size_t)n, W, cs, on);
    }
  }
  __syncthreads();
  // Columns of A and of B right of each block's stop column, one task per
  // column of both matrices, blocks in the order ib = nblocks..1.
  for (int ib = nblocks; ib >= 1; --ib) {
    const int rw0 = s + (ib - 1) * R;  // 0-based first row
    const int W = min(R + NQ - 1, n - rw0);
    if (W <= 0) continue;
    const int istopa = (ib == 1) ? s + NQ - 1 : s + (ib - 2) * R + NQ;  // 1-based
    const int istopb = s + (ib - 1) * R + R + NQ - 1;
    const int nca = max(0, n - istopa), ncb = max(0, n - istopb);  // columns j >= istop (0-based)
    if (nca + ncb == 0) continue;
    __syncthreads();
    for (int e = tid; e < NQ * R; e += kBT) {
      const int icm = e / R, ig = e % R;
      const size_t ti = (size_t)(ib - 1) * LDR * NQ + (size_t)icm * LDR + (icm + ig - 1);
      csu[e] = (ig >= 1) ? make_double2(CU[ti], -SU[ti]) : make_double2(1.0, 0.0);
    }
    __syncthreads();
    for (int t = tid; t < nca + ncb; t += kBT) {
      double* M = t < nca ? A : B;
      const int j = t < nca ? istopa + t : istopb + (t - nca);
      apply_window<R, NQ, kFarAhead>(&M[PHT_IDX(rw0, j, n)], 1, W, csu, ~0u);
    }
  }
  __threadfence_block();
  __syncthreads();
}

// Stage 2 of one pencil.  The generation runs kGroups columns of a sweep at
// once, one group of kBT / kGroups threads (whole waves) each: group g takes
// the columns ic = g+1, g+1+kGroups, ..., and step (ic, ib) waits until column
// ic-1 has finished step ib+2 (LDS progress counters), the order of the
// single-pencil sweep whose leaders also take the rows above their window
// (rht2ht.f90:787-957 in the reference's column order otherwise).  Each group
// has its own LDS arena for the steps.
constexpr int kGroups = 4;
constexpr int kGT = kBT / kGroups;
static_assert(kGT % 128 == 0, "a group holds whole pairs of waves");
constexpr int kArena = kBRmax * (kBRmax + 1) + 4 * kBRmax + 2 * s2::kMaxLdr + 2 * (kGT / 128) * 128 +
                       (2 * kBRmax) * (2 * kBRmax + 1);
static_assert((kGroups * kArena + 64) * sizeof(double) <= s1::kRqLds, "stage-2 arenas fit");

__device__ __noinline__ void stage2_wg(double* sm, double* A, double* B, double* Q, double* Z, int n, int r,
                          int nq, double* ws) {
  using namespace s2;
  const int ldr = r + nq;
  const int nbl = (n - 1) / r + ((n - 1) % r != 0);
  const size_t tsz = (size_t)ldr * nq * nbl;
  double* CU = ws;
  double* SU = CU + tsz;
  double* CV = SU + tsz;
  double* SV = CV + tsz;
  double* Mw = SV + tsz;
  // this thread's group and its LDS arena
  const int g = threadIdx.x / kGT;
  double* arena = sm + (size_t)g * kArena;
  double(*Bb)[kBRmax + 1] = reinterpret_cast<double(*)[kBRmax + 1]>(arena);
  double* lc = arena + kBRmax * (kBRmax + 1);
  double* ls = lc + kBRmax;
  double* rc = ls + kBRmax;
  double* rs = rc + kBRmax;
  double* xa = rs + kBRmax;
  double* xb = xa + kMaxLdr;
  double* red = xb + kMaxLdr;
  double* Ml = red + 2 * (kGT / 128) * 128;
  // column progress (index ic, 0 unused) and the groups' barrier counters
  int* prog = reinterpret_cast<int*>(sm + (size_t)kGroups * kArena);
  int* gcount = prog + 40;
  constexpr int kDoneStep = 1 << 30;
  for (int s = 1; s <= n - 2; s += nq) {
    const int nblocks = (n - s) / r + ((n - s) % r != 0);
    const size_t tz = (size_t)ldr * nq * nblocks, msz = (size_t)ldr * ldr;
    for (size_t e = threadIdx.x; e < tz; e += kBT) {
      CU[e] = 1.0;
      CV[e] = 1.0;
      SU[e] = 0.0;
      SV[e] = 0.0;
    }
    for (size_t e = threadIdx.x; e < msz * nblocks; e += kBT) {
      const size_t k = e % msz;
      Mw[e] = ((k % ldr) == (k / ldr)) ? 1.0 : 0.0;
    }
    for (int e = threadIdx.x; e < 40 + kGroups; e += kBT) prog[e] = 0;
    __threadfence_block();
    __syncthreads();
    WaveGroup grp{g * kGT, kGT / 64, &gcount[g], 0};
    for (int ic = g + 1; ic <= nq; ic += kGroups) {
      const int jc = s + ic - 1;
      for (int ib = 1;; ++ib) {
        int col, row0;
        if (ib == 1) {
          if (jc > n) break;
          col = jc;
          row0 = jc + 1;
        } else {
          col = jc + 1 + (ib - 2) * r;
          if (col > n) break;
          row0 = col + r;
        }
        if (ic > 1) {  // column ic-1 is two steps ahead (or done); bounded wait
          for (int spin = 0; __hip_atomic_load(&prog[ic - 1], __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_WORKGROUP) < ib + 2 &&
                             spin < (1 << 26);
               ++spin)
            __builtin_amdgcn_s_sleep(1);
          __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        }
        chase_step<kBRmax, kGT, WaveGroup, 2 * kBRmax>(grp, A, n, B, n, n, r, nq, s, ic, ib, col,
                                                       row0, CU, SU, CV, SV, ldr, Mw, true, Bb, lc,
                                                       ls, rc, rs, xa, xb, red, Ml);
        grp.sync();  // the step's stores are done
        if (grp.tid() == 0)
          __hip_atomic_store(&prog[ic], ib, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
      }
      grp.sync();
      if (grp.tid() == 0)
        __hip_atomic_store(&prog[ic], kDoneStep, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
    __threadfence_block();
    __syncthreads();
    // far updates of the sweep
    bool done = false;
#define PHT_FAR(RR, NN)                                                                      \
    if (!done && r == RR && nq == NN) {                                                      \
      far_updates_wg<RR, NN>(sm, A, B, Q, Z, n, s, nblocks, CU, SU, CV, SV);                 \
      done = true;                                                                           \
    }
    PHT_BATCH_PAIRS(PHT_FAR)
#undef PHT_FAR
  }
}

// ---- one stage for tiny pencils (r = 1) ---------------------------------
// A pencil of order n <= kOneStageMax is reduced straight to HT form with
// LAPACK DGGHRD's Givens sequence (Moler & Stewart 1973): for each column j,
// from the bottom, a left rotation of rows i-1, i zeroes A(i, j) and the fill
// B(i, i-1) it makes is removed by a right rotation of columns i-1, i.  The
// whole pencil sits in LDS (ld = n + 1).  Wave 0 runs the rotation chain on
// A and B (one lane per row or column, no barrier inside a column); waves 1
// and 2 apply the previous column's rotations to Q and Z from a
// double-buffered LDS tape.  Rotations with dlartg's conventions (rot_rsq),
// so the result equals DGGHRD's up to rounding.
constexpr int kOneStageMax = 64;

__device__ __forceinline__ void lds_order() { __asm__ volatile("" ::: "memory"); }

__device__ __noinline__ void onestage_wg(double* sm, double* A, double* B, double* Q, double* Z,
                                         int n) {
  const int ld = n + 1;
  double* a = sm;
  double* b = a + (size_t)n * ld;
  double* q = b + (size_t)n * ld;
  double* z = q + (size_t)n * ld;
  double* tape = z + (size_t)n * ld;  // [2 buffers][4][n]: cl, sl, cr, sr
  const int tid = threadIdx.x, wave = tid >> 6, l = tid & 63;
  for (int x = tid; x < n * n; x += kBT) {
    const int i = x % n, j = x / n;
    a[i + j * ld] = A[x];
    b[i + j * ld] = B[x];
    q[i + j * ld] = Q[x];
    z[i + j * ld] = Z[x];
  }
  __syncthreads();
  for (int j = 0; j <= n - 2; ++j) {
    double* tp = tape + (size_t)(j & 1) * 4 * n;  // column j's rotations
    double* tq = tape + (size_t)((j + 1) & 1) * 4 * n;  // column j-1's
    if (wave == 0 && j <= n - 3) {
      for (int i = n - 1; i >= j + 2; --i) {
        double c, s, r;
        rot_rsq(a[i - 1 + j * ld], a[i + j * ld], c, s, r);
        lds_order();
        if (l > j && l < n) {  // columns j+1.. of A, rows i-1, i
          const double x = a[i - 1 + l * ld], y = a[i + l * ld];
          a[i - 1 + l * ld] = c * x + s * y;
          a[i + l * ld] = c * y - s * x;
        }
        if (l >= i - 1 && l < n) {  // columns i-1.. of B
          const double x = b[i - 1 + l * ld], y = b[i + l * ld];
          b[i - 1 + l * ld] = c * x + s * y;
          b[i + l * ld] = c * y - s * x;
        }
        if (l == 0) {
          a[i - 1 + j * ld] = r;
          a[i + j * ld] = 0.0;
          tp[i] = c;
          tp[n + i] = s;
        }
        lds_order();
        double c2, s2, r2;
        rot_rsq(b[i + i * ld], b[i + (i - 1) * ld], c2, s2, r2);
        lds_order();
        if (l < n) {  // rows of A, columns i (x) and i-1 (y)
          const double x = a[l + i * ld], y = a[l + (i - 1) * ld];
          a[l + i * ld] = c2 * x + s2 * y;
          a[l + (i - 1) * ld] = c2 * y - s2 * x;
        }
        if (l < i) {  // rows 0..i-1 of B
          const double x = b[l + i * ld], y = b[l + (i - 1) * ld];
          b[l + i * ld] = c2 * x + s2 * y;
          b[l + (i - 1) * ld] = c2 * y - s2 * x;
        }
        if (l == 0) {
          b[i + i * ld] = r2;
          b[i + (i - 1) * ld] = 0.0;
          tp[2 * n + i] = c2;
          tp[3 * n + i] = s2;
        }
        lds_order();
      }
    } else if ((wave == 1 || wave == 2) && j >= 1 && j - 1 <= n - 3 && l < n) {
      // column j-1's rotations on the rows of Q (wave 1) or Z (wave 2)
      const int jp = j - 1;
      if (wave == 1) {
        for (int i = n - 1; i >= jp + 2; --i) {
          const double c = tq[i], s = tq[n + i];
          const double x = q[l + (i - 1) * ld], y = q[l + i * ld];
          q[l + (i - 1) * ld] = c * x + s * y;
          q[l + i * ld] = c * y - s * x;
        }
      } else {
        for (int i = n - 1; i >= jp + 2; --i) {
          const double c = tq[2 * n + i], s = tq[3 * n + i];
          const double x = z[l + i * ld], y = z[l + (i - 1) * ld];
          z[l + i * ld] = c * x + s * y;
          z[l + (i - 1) * ld] = c * y - s * x;
        }
      }
    }
    __syncthreads();
  }
  for (int x = tid; x < n * n; x += kBT) {
    const int i = x % n, j = x / n;
    A[x] = a[i + j * ld];
    B[x] = b[i + j * ld];
    Q[x] = q[i + j * ld];
    Z[x] = z[i + j * ld];
  }
}

static_assert((4 * kOneStageMax * (kOneStageMax + 1) + 8 * kOneStageMax) * sizeof(double) <=
                  s1::kRqLds,
              "a tiny pencil fits the batched kernel's LDS");

__global__ __launch_bounds__(kBT) void k_batched(const BatchPencil* __restrict__ pencils,
                                                 int count, int* next, double* Ap, double* Bp,
                                                 double* Qp, double* Zp, double* ws_all,
                                                 size_t ws_per_wg) {
  extern __shared__ double sm[];
  __shared__ int cur;
  double* ws = ws_all + (size_t)blockIdx.x * ws_per_wg;
  for (;;) {
    if (threadIdx.x == 0) cur = atomicAdd(next, 1);
    __syncthreads();
    const int idx = cur;
    __syncthreads();  // every thread has read cur before it is replaced
    if (idx >= count) break;
    const BatchPencil pc = pencils[idx];
    double* A = Ap + pc.offset;
    double* B = Bp + pc.offset;
    double* Q = Qp + pc.offset;
    double* Z = Zp + pc.offset;
    if (pc.r == 1) {
      onestage_wg(sm, A, B, Q, Z, pc.n);
    } else {
      if (pc.n > pc.r + 1) stage1_wg(sm, A, B, Q, Z, pc.n, pc.r, pc.p, ws);
      if (pc.n > 2) stage2_wg(sm, A, B, Q, Z, pc.n, pc.r, pc.nq, ws);
    }
    __threadfence();
    __syncthreads();
  }
}

__global__ void k_batch_reset(int* next) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *next = 0;
}
}  // namespace

bool batched_supported(int n, int r, int p, int nq) {
  if (r == 1) return n >= 1 && n <= kOneStageMax;  // one stage; p, nq unused
  if (n < 1 || r < 1 || r > kBRmax || p < 2 || p * r > 128) return false;
  bool ok = false;
#define PHT_OK(RR, NN) ok = ok || (r == RR && nq == NN);
  PHT_BATCH_PAIRS(PHT_OK)
#undef PHT_OK
  return ok && s2_sweep_lds(r, nq) != 0;
}

size_t batched_ws_doubles(int n, int r, int p, int nq) {
  if (r == 1) return 64;
  const size_t P = (size_t)p * r;
  const size_t nbmax = (n > r) ? (size_t)((n - r + (p - 1) * r - 1) / ((p - 1) * r) + 1) : 1;
  const size_t s1 = 2 * nbmax * P * r + 2 * nbmax * r * r;
  const size_t ldr = (size_t)r + nq;
  const size_t nbl = (size_t)((n - 1) / r + 1);
  const size_t s2 = 4 * ldr * nq * nbl + ldr * ldr * nbl;
  return (s1 > s2 ? s1 : s2) + 64;
}

hipError_t batched_launch(hipStream_t st, const BatchPencil* dpencils, int count, int* dnext,
                          double* Ap, double* Bp, double* Qp, double* Zp, double* ws,
                          size_t ws_per_wg, int grid) {
  static bool attr = false;
  const size_t lds = s1::kRqLds;
  if (!attr) {
    hipError_t e = hipFuncSetAttribute((const void*)k_batched,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    attr = true;
  }
  hipLaunchKernelGGL(k_batch_reset, dim3(1), dim3(64), 0, st, dnext);
  hipLaunchKernelGGL(k_batched, dim3(grid), dim3(kBT), lds, st, dpencils, count, dnext, Ap, Bp,
                     Qp, Zp, ws, ws_per_wg);
  return hipGetLastError();
}

}  // namespace pht
