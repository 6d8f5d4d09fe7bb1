// Host-side declarations of the kernel launchers of libparaht_hip.so.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace pht {

// One compact-WY block reflector H = I - V*T*V^T to apply to a matrix C.
// The block acts on the index span [start, start+len) of the transformed
// dimension (rows for a left update, columns for a right update) and is
// applied on the range [lo, hi) of the other dimension.  V is explicit
// (len x k, unit entries and zeros stored), T is k x k with zeros stored
// outside its triangle.  All indices are 0-based.
struct ApplyBlock {
  int start, len, k, lo, hi, pad_;
  long long voff, toff;
};

enum Side { kLeftT = 0, kRightN = 1 };

// Applies blocks[0..nblocks) in order: C <- H^T C (kLeftT) or C <- C H
// (kRightN), on FP64 MFMA tiles.  [other_begin, other_end) bounds the other
// dimension over all blocks (it sets the grid).  kmax = max k and max_len =
// max len over the blocks (len <= 128 and k <= 64 take the LDS-staged kernel).
// zero_cols > 0 (right side, one block, and only where apply_fuses_zero
// holds): the block's entries below the diagonal in its first zero_cols
// columns are stored as exact zeros, as zero_below would set them.
hipError_t apply_blocks(hipStream_t st, Side side, double* C, int ldc,
                        const double* V, int ldv, const double* T, int ldt,
                        const ApplyBlock* dblocks, int nblocks, int kmax, int max_len,
                        int other_begin, int other_end, int zero_cols = 0);
bool apply_fuses_zero(int kmax, int max_len);

// ---- stage 1 (ab2rht_simple) ----
// The panel at column jb0 (0-based; jb = jb0+1 of ab2rht.f90:71) has
// nblocks blocks of p*nb rows overlapping by nb rows; block bl (0-based) is
// at row i0 = jb0 + nb + (nblocks-1-bl) (p-1) nb, bottom-up
// (ab2rht.f90:74-75,94-99).
__host__ __device__ inline int s1_nblocks(int n, int nb, int p, int jb0) {
  const int m = n - nb - jb0, d = (p - 1) * nb;
  return m / d + (m % d != 0);
}
__host__ __device__ inline int s1_npanels(int n, int nb) {
  return n - nb - 2 >= 0 ? (n - nb - 2) / nb + 1 : 0;
}
struct S1Block {
  int i0, len, k, kr;  // first row, rows, left reflectors, right reflectors
};
__host__ __device__ inline S1Block s1_block(int n, int nb, int p, int jb0, int nblocks, int bl) {
  S1Block b;
  b.i0 = jb0 + nb + (nblocks - 1 - bl) * (p - 1) * nb;
  b.len = n - b.i0 < p * nb ? n - b.i0 : p * nb;
  b.k = nb < b.len ? nb : b.len;
  b.kr = b.len > nb ? nb : b.len;
  return b;
}
// Writes the compact-WY block lists of every panel into five consecutive
// arrays of nd blocks (nd = the sum of the panels' nblocks): left updates of
// B (columns i0..n), of A (columns jb0+nb..n), of Q (rows 0..n), right
// updates of A and Z (rows 0..n) and of B (rows 0..i0+len).  V and T of
// block bl sit at bl*p*nb*nb and bl*nb*nb of the panel's tapes.
hipError_t s1_block_lists(hipStream_t st, ApplyBlock* lists, int nd, int n, int nb, int p);
// QR of a general B by panels of nb columns (the reference driver's dgeqrf,
// src/profilereduction.f90:117): panel q covers columns
// [qr_panel_col(q), +nb) and rows from its first column down, factored by the
// stage-1 panel kernels on B shifted up by nb rows (order n + nb).  Three
// block lists of nd blocks each: the left update of B's trailing columns,
// the left update of A (all columns), the right update of Q (all rows).
__host__ __device__ inline int qr_npanels(int n, int nb) { return (n + nb - 1) / nb; }
__host__ __device__ inline int qr_panel_col(int n, int nb, int q) {
  return q * nb < n - nb ? q * nb : n - nb;
}
hipError_t qr_block_lists(hipStream_t st, ApplyBlock* lists, int nd, int n, int nb, int p);

// Bottom-up block QRs of the panel A(jb+nb:n, jb:jb+nb-1) (0-based jb);
// writes explicit V (ldv = Pmax rows per block, vstride per block) and T.
hipError_t s1_panel_qr(hipStream_t st, double* A, int lda, int n, int jb,
                       int nb, int p, int nblocks, double* Vt, int ldv,
                       long long vstride, double* Tt, int ldt,
                       long long tstride);
// RQ-based right generator for one B block B(i0:i0+len, i0:i0+len); writes
// explicit column-form V (len x k) and T (k x k); k = nb if len > nb else len.
// scratch: device buffer of >= len*len + nb*len doubles (used when the block
// does not fit in LDS).
hipError_t s1_rq_gen(hipStream_t st, const double* B, int ldb, int i0,
                     int len, int nb, double* V, int ldv, double* T, int ldt,
                     double* scratch);
// Register-resident versions for P = p*nb <= 128 and nb <= 64 (rq: any len; len <= nb is the reference's 'B','R' tail block).
// With tau_ws (nblocks * 64 doubles), the panel's chain of block QRs leaves
// the T matrices to a second launch that builds them all at once.
hipError_t s1_panel_qr_fast(hipStream_t st, double* A, int lda, int n, int jb,
                            int nb, int p, int nblocks, double* Vt, int ldv,
                            long long vstride, double* Tt, int ldt,
                            long long tstride, double* tau_ws = nullptr);
hipError_t s1_rq_fast(hipStream_t st, const double* B, int ldb, int i0, int len,
                      int nb, double* V, int ldv, double* T, int ldt);
// Whether the stage-1 generators take blocks of p*nb rows and nb reflectors.
bool s1_supported(int nb, int p);
// Sets B(j+1:i0+len-1, j) = 0 for j in [i0, i0+ncols).
hipError_t zero_below(hipStream_t st, double* B, int ldb, int i0, int len,
                      int ncols);

// ---- stage 2 (rht2ht_givens semantics) ----
struct S2Params {
  int n, r, nq, jblock;  // jblock 1-based like the reference
  int nblocks;           // bulge blocks of this sweep
  int ldr;               // r + nq
  double* M;             // nblocks x ldr x ldr workspace (deferred rotations)
  int* progress;         // nq step counters of the sweep's leader workgroups
  int* hprog;            // nq x kS2MaxHelpers step counters of the helper workgroups
  int nhelp;             // helper workgroups per column, 0..kS2MaxHelpers
  int* errflag;          // set when a workgroup's wait hit the spin limit
  int spin;              // polls before a wait gives up (kS2SpinLimit)
  int clear_err;         // the call's first sweep: s2_prepare clears *errflag
};
constexpr int kS2SpinLimit = 1 << 26;
constexpr int kS2MaxHelpers = 4;
// ints of the sweep's counters: progress, hprog (contiguous) and errflag
inline size_t s2_counter_ints(int nq) { return (size_t)nq * (1 + kS2MaxHelpers) + 64; }
// Dynamic LDS bytes of the sweep kernel for (r, nq), 0 if (r, nq) does not
// fit (r <= 64 and M of (r+nq)^2 doubles in LDS beside the B block).
size_t s2_sweep_lds(int r, int nq);
// Resets the sweep's tape to identity rotations, M to identity and the
// progress counters; must precede s2_generate.
hipError_t s2_prepare(hipStream_t st, const S2Params& P, double* CU, double* SU,
                      double* CV, double* SV);
hipError_t s2_generate(hipStream_t st, double* A, int lda, double* B, int ldb,
                       const S2Params& P, double* CU, double* SU, double* CV,
                       double* SV);
// Ranges of a sweep's A/B far update.  Rows above the window: block ib's
// sequence icm (0-based) applies to the rows i (0-based) with i < row_end and
//   thr(ib - lo_shift, icm) <= i+1 < thr(ib - hi_shift, icm),
//   thr(x, icm) = max(s+1, s+1+(x-nq+icm) r)   (the window's first row),
// lo_shift < 0 meaning no lower bound.  Columns right of the window: block ib
// applies to the columns j with istop(ib) + c_lo r < j+1 <= istop(ib) + c_hi r
// (c_hi < 0: no upper bound); the grid starts at column col_begin (0-based).
// The default is the whole far update.
struct S2Range {
  int lo_shift = -1, hi_shift = 0, row_end = 1 << 30;
  int c_lo = 0, c_hi = -1;
  int col_begin = 0;
};
// Remaining right rotations (rows of A,B above the generation window, all
// rows of Z) and left rotations seen from the right (all rows of Q); part
// selects the A/B rows, the Q/Z rows or both.  The A/B rows can be limited
// to a range (s2_far_split) and to the grid's rows [0, row_end).
enum RowsPart { kRowsAll = 0, kRowsAB = 1, kRowsQZ = 2 };
hipError_t s2_apply_rows(hipStream_t st, double* A, int lda, double* B,
                         int ldb, double* Q, int ldq, double* Z, int ldz,
                         const S2Params& P, const double* CU, const double* SU,
                         const double* CV, const double* SV, RowsPart part,
                         const S2Range& rg = S2Range{}, int row_end = -1);
// Remaining left rotations (columns of A,B right of the generation window),
// limited to a range (s2_far_split).
hipError_t s2_apply_cols(hipStream_t st, double* A, int lda, double* B,
                         int ldb, const S2Params& P, const double* CU,
                         const double* SU, const S2Range& rg = S2Range{});
// Whether (r, nq) has far-update kernels that take ranges (the pairs of
// PHT_S2_PAIRS); the others run the whole far update in one launch.
bool s2_far_split(int r, int nq);

// (r, nq) pairs with register-window far-update kernels
#define PHT_S2_PAIRS(X) \
  X(8, 4) X(8, 8) X(16, 4) X(16, 8) X(16, 16) X(32, 8) X(32, 16) X(32, 32) \
  X(64, 8) X(64, 16) X(64, 32)
// Arguments of the far-update launchers for one sweep.
struct S2FarArgs {
  double *A, *B, *Q, *Z;
  int lda, ldb, ldq, ldz;
  int n, s, nblocks;
  const double *CU, *SU, *CV, *SV;
};
template <int R, int NQ>
hipError_t s2_far_rows(hipStream_t st, const S2FarArgs& a, bool ab, const S2Range& rg,
                       int row_end);
template <int R, int NQ>
hipError_t s2_far_cols(hipStream_t st, const S2FarArgs& a, const S2Range& rg);

// ---- stage 2, Householder variant (rht2ht_householder semantics) ----
// Tape of one sweep, doubles from `base`: U, V ((r+nq) x nq x nblocks each),
// tauU, tauV (nq x nblocks each) -- zeroed by hh_prepare -- then TU, TV
// (nq x nq x nblocks each).
struct HhTape {
  double *U, *V, *tauU, *tauV, *TU, *TV;
};
inline size_t hh_tape_doubles(int r, int nq, int nbl) {
  return (size_t)2 * (r + nq) * nq * nbl + (size_t)2 * nq * nbl + (size_t)2 * nq * nq * nbl;
}
inline HhTape hh_tape(double* base, int r, int nq, int nbl) {
  HhTape t;
  const size_t u = (size_t)(r + nq) * nq * nbl, q = (size_t)nq * nbl;
  t.U = base;
  t.V = t.U + u;
  t.tauU = t.V + u;
  t.tauV = t.tauU + q;
  t.TU = t.tauV + q;
  t.TV = t.TU + (size_t)nq * nq * nbl;
  return t;
}
// (r, nq) the Householder kernels take (r <= 64, nq <= 32, r + nq <= 96).
bool hh_supported(int r, int nq);
// (r, nq) with MFMA far-update kernels; the others use a generic kernel.
bool hh_far_supported(int r, int nq);
// Dynamic LDS of the sweep kernel.
size_t hh_sweep_lds(int r);
// Zeroes nd doubles of the tape (U, V, tauU, tauV) and nc counters; the
// call's first sweep clears *errflag.
hipError_t hh_prepare(hipStream_t st, double* tape, size_t nd, int* counters, int nc,
                      int* errflag, int clear_err);
// Generation of the sweep at column s (1-based): nq leader workgroups, each
// with nhelp helper workgroups (0: the leaders do the rows above the window).
hipError_t hh_generate(hipStream_t st, double* A, int lda, double* B, int ldb, int n, int r,
                       int nq, int s, double* U, double* V, double* tauU, double* tauV,
                       int* progress, int* hprog, int* errflag, int spin, int nhelp);
// T factors of every block of the sweep (DLARFT 'F','C').
hipError_t hh_factors(hipStream_t st, const double* U, const double* V, const double* tauU,
                      const double* tauV, double* TU, double* TV, int ldu, int nq, int nblocks);
// Phases of the far updates: Q and Z (all rows, all reflectors); A/B rows or
// columns in one pass; the lookahead part; the rest.
enum HhPhase { kHhFull = 0, kHhAll = 1, kHhLook = 2, kHhFar = 3 };
// Far updates of one or two matrices (M1 = nullptr: one) of the sweep at s:
// cols = false: rows (right reflectors Y with factor T), cols = true: columns
// right of the window (left reflectors; isA selects A's stop columns).
// generic forces the non-MFMA kernel.
hipError_t hh_far(hipStream_t st, bool cols, double* M0, int ld0, double* M1, int ld1,
                  const double* Y0, const double* T0, const double* tau0, const double* Y1,
                  const double* T1, const double* tau1, int isA0, int isA1, int n, int r, int nq,
                  int s, int nblocks, int phase, bool generic = false);

// ---- batched reduction of many small pencils ----
// One pencil of a batch: its n x n column-major matrices (ld = n) start at
// element `offset` of each of the A, B, Q, Z pools.
struct BatchPencil {
  long long offset;
  int n, r, p, nq;
};
// (n, r, p, nq) that the batched kernel handles.
bool batched_supported(int n, int r, int p, int nq);
// Workspace doubles one workgroup needs for a pencil.
size_t batched_ws_doubles(int n, int r, int p, int nq);
// Resets the work queue counter and launches the persistent kernel (grid
// workgroups, ws_per_wg doubles of workspace each).
hipError_t batched_launch(hipStream_t st, const BatchPencil* dpencils, int count, int* dnext,
                          double* Ap, double* Bp, double* Qp, double* Zp, double* ws,
                          size_t ws_per_wg, int grid);

// ---- adjoint (reverse mode) of the reduction ----
// Dynamic LDS of the workgroup-per-pencil adjoint kernel for pencils of
// order <= nmax, and the most it accepts (4 vectors of n doubles: n <= 4992).
size_t adjoint_wg_lds(int nmax);
size_t adjoint_wg_max_lds();
// Transposed recurrence for a queue of pencils (largest first), one
// workgroup each; matrices in pools at the pencils' offsets (ld = n).
hipError_t adjoint_batched(hipStream_t st, const BatchPencil* dpencils, int count, int nmax,
                           int* dnext, const double* H, const double* T, const double* Tinv,
                           double* GQ, const double* GZ, double* XA, double* XB, int grid);
// The same for one pencil over a cooperative grid (vecs: 4n doubles; ints:
// barrier counter, generation, error flag).
hipError_t adjoint_grid(hipStream_t st, int n, const double* H, const double* T,
                        const double* Tinv, double* GQ, const double* GZ, double* XA,
                        double* XB, double* vecs, int* ints, int grid);
// Largest co-resident grid of the cooperative adjoint kernel (0 on failure).
int adjoint_grid_max_blocks(int device);

// ---- QZ iteration on Hessenberg-triangular pencils ----
// One task per pencil: H, T, Q, Z at element offset `off` of their pools
// (leading dimensions ldh..ldz), eigenvalues at `eoff` of alphar/alphai/beta,
// status at dinfo[idx].
struct QzTask {
  long long off, eoff;
  int n, ldh, ldt, ldq, ldz, idx;
};
constexpr int kQzSchur = 0x1;  // Schur form (whole rows/columns, Q and Z)
// Workgroup size for a pencil of order n.
int qz_threads(int n);
// One workgroup per task; dinfo[idx] = 0 or the 1-based index of the last
// eigenvalue that did not converge in 30 n sweeps.
hipError_t qz_launch(hipStream_t st, const QzTask* dtasks, int count, int threads, double* H,
                     double* T, double* Q, double* Z, double* ar, double* ai, double* be,
                     int* dinfo, int flags);

}  // namespace pht
