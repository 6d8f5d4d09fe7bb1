/*
 * paraht.h -- C-ABI of libparaht_hip.so, the MI355X (gfx950) implementation of
 * ParaHT's two-stage Hessenberg-triangular (HT) reduction of a pencil (A,B).
 *
 * All matrices are FP64, column-major, with leading dimensions, and are
 * updated in place exactly as the reference's Fortran routines update theirs:
 * on return Q_out^T * A_in * Z_out == A_out (with Q_in, Z_in accumulated), and
 * the same for B.  Arguments are passed by value (the Fortran drop-in modules
 * pass them with iso_c_binding `value`).
 *
 * Return value: 0 on success; -i if the i-th argument had an illegal value
 * (LAPACK convention, counting from 1 in the order of the reference's Fortran
 * argument list); a positive PARAHT_ERR_* code on a device or allocation
 * failure.  Calls block the host until the device work has finished.
 */
#ifndef PARAHT_H
#define PARAHT_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes (positive = failure outside the arguments) ---- */
#define PARAHT_OK 0
#define PARAHT_ERR_HIP 1      /* a HIP runtime call failed */
#define PARAHT_ERR_NOMEM 2    /* device or host allocation failed */
#define PARAHT_ERR_NODEV 3    /* no usable gfx950 device */
#define PARAHT_ERR_UNSUPPORTED 4 /* parameters outside what the kernels support */

/* ---- flags ---- */
#define PARAHT_TIME_COMPONENTS 0x1 /* print per-phase times like the reference's
                                       time_components argument */
#define PARAHT_STAGE2_GIVENS 0x2   /* stage 2 with Givens rotations
                                       (rht2ht_givens); without it stage 2 is
                                       the Householder variant
                                       (rht2ht_householder, r <= 64, nq <= 32,
                                       r + nq <= 96) */

/* Opaque per-device context: stream, workspaces and phase timers.
 * A context is not re-entrant; use one per host thread. */
typedef struct paraht_ctx paraht_ctx;

int paraht_ctx_create(paraht_ctx** ctx, int device);
int paraht_ctx_destroy(paraht_ctx* ctx);
/* Use an existing hipStream_t (as void*) for all work of this context;
 * NULL restores the context's own stream.  The library also runs the
 * updates of Q and Z (and, in stage 1, the left update of A) on a second,
 * non-blocking stream of its own, ordered against this one by events; every
 * call makes this stream wait for that work before it returns. */
int paraht_ctx_set_stream(paraht_ctx* ctx, void* stream);
/* Phase times (milliseconds) of the last call on this context:
 * [0] stage 1 total, [1] stage-1 left generation, [2] stage-1 right
 * generation chain, [3] stage 2 total, [4] stage-2 generation.
 * Returns the number of entries written. */
int paraht_ctx_last_times(paraht_ctx* ctx, double* ms, int nmax);

/* ---- Per-kernel timing ----
 * The device counterpart of the reference's per-task timers (time_tasks,
 * src/ab2rht.f90:26, src/rht2ht.f90:27,1366): with profiling enabled every
 * kernel launch of the context is bracketed by HIP events on the context's
 * stream and summed per kernel class, together with the launch's algorithmic
 * flops (compact-WY appliers: 4*len*k*width per block) and bytes (stage-2
 * row sweeps: every touched element read and written once). */
#define PARAHT_K_WY_APPLY 0    /* FP64-MFMA compact-WY block applier */
#define PARAHT_K_S1_PANEL_QR 1 /* stage-1 left generation (panel QRs) */
#define PARAHT_K_S1_RQ_GEN 2   /* stage-1 right generation (block RQ + LQ) */
#define PARAHT_K_S2_GEN 3      /* stage-2 rotation generation (band window) */
#define PARAHT_K_S2_ROWS 4     /* stage-2 row sweeps (A,B above window; Q, Z) */
#define PARAHT_K_S2_COLS 5     /* stage-2 column sweeps (A,B right of window) */
#define PARAHT_K_OTHER 6       /* small helpers (explicit zeroing) */
#define PARAHT_NKERNELS 7
/* Enable (1) or disable (0) profiling; both reset the accumulated sums. */
int paraht_ctx_profile(paraht_ctx* ctx, int enable);
/* Accumulated time, launch count and algorithmic flops/bytes of a kernel
 * class since profiling was enabled (waits for the context's stream). */
int paraht_ctx_kernel_stats(paraht_ctx* ctx, int kernel, double* total_ms,
                            long long* launches, double* flops, double* bytes);

/* ---- Stage 1: (A,B) -> r-Hessenberg-triangular ----
 * Replaces ab2rht_simple(n, A, lda, B, ldb, Q, ldq, Z, ldz, nb, p,
 * time_components), src/ab2rht.f90:8-13.  nb is the bandwidth r, p the number
 * of nb-row blocks per reflector block (p >= 2).
 * Host pointers: the arrays are copied to the device and back. */
int paraht_ab2rht(int n, double* A, int lda, double* B, int ldb, double* Q,
                  int ldq, double* Z, int ldz, int nb, int p, int flags);

/* ---- Stage 2: r-Hessenberg-triangular -> Hessenberg-triangular ----
 * Replaces rht2ht_householder(n, A, lda, B, ldb, Q, ldq, Z, ldz, r, nq,
 * time_components), src/rht2ht.f90:1346-1351, and with PARAHT_STAGE2_GIVENS
 * rht2ht_givens(n, A, lda, B, ldb, Q, ldq, Z, ldz, r, nq),
 * src/rht2ht.f90:696-700.  A must have lower bandwidth <= r, B upper
 * triangular.  Host pointers. */
int paraht_rht2ht(int n, double* A, int lda, double* B, int ldb, double* Q,
                  int ldq, double* Z, int ldz, int r, int nq, int flags);

/* Replaces rht2ht_householder_unblocked(n, A, lda, B, ldb, Q, ldq, Z, ldz, r),
 * src/rht2ht.f90:2053-2057 (the unblocked specification of stage 2). */
int paraht_rht2ht_unblocked(int n, double* A, int lda, double* B, int ldb,
                            double* Q, int ldq, double* Z, int ldz, int r);

/* Both stages with one host<->device round trip (the reference driver's
 * ab2rht_simple + rht2ht_* sequence, src/profilereduction.f90:136,144). */
int paraht_reduce(int n, double* A, int lda, double* B, int ldb, double* Q,
                  int ldq, double* Z, int ldz, int r, int p, int nq, int flags);

/* ---- Device-pointer variants (arrays already resident in HBM) ----
 * Same semantics as the host versions; pointers are device addresses.  The
 * work is enqueued on the context's stream and the call waits for it. */
int paraht_ab2rht_dev(paraht_ctx* ctx, int n, double* dA, int lda, double* dB,
                      int ldb, double* dQ, int ldq, double* dZ, int ldz,
                      int nb, int p, int flags);
int paraht_rht2ht_dev(paraht_ctx* ctx, int n, double* dA, int lda, double* dB,
                      int ldb, double* dQ, int ldq, double* dZ, int ldz, int r,
                      int nq, int flags);
int paraht_reduce_dev(paraht_ctx* ctx, int n, double* dA, int lda, double* dB,
                      int ldb, double* dQ, int ldq, double* dZ, int ldz, int r,
                      int p, int nq, int flags);

/* ---- QR of a general B (before stage 1) ----
 * The reference's driver makes B upper triangular with LAPACK dgeqrf on the
 * host before the reduction and keeps only R (src/profilereduction.f90:117-123;
 * its A is random).  These do it on the GPU and carry the orthogonal factor:
 * B <- R = Q_B^T B, A <- Q_B^T A, Q <- Q Q_B, so that a later reduction of
 * (A, B) with this Q gives Q^T A0 Z = H, Q^T B0 Z = T for the original pencil
 * (A0, B0).  nb: the columns per panel, p: panel blocks of p*nb rows (as for
 * stage 1; nb <= 64 and p*nb <= 128 take the register-resident generators).
 * Z is not involved.  Returns -k for an illegal k-th argument,
 * PARAHT_ERR_UNSUPPORTED where the generators cannot take (nb, p). */
int paraht_qr_b(int n, double* A, int lda, double* B, int ldb, double* Q,
                int ldq, int nb, int p);
int paraht_qr_b_dev(paraht_ctx* ctx, int n, double* dA, int lda, double* dB,
                    int ldb, double* dQ, int ldq, int nb, int p);

/* ---- Allocation-free reduction ----
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the two stages of paraht_reduce_dev, on the caller's
 * stream (a hipStream_t as void*, NULL for the null stream) and with no
 * hipMalloc, hipFree, stream or event of their own.  The workspace is a
 * 64 MiB array in the library's device data, which each GPU the library runs
 * on holds until the process ends (r = 64, p = 2 fits n <= 19392; r = 16,
 * p = 8 fits n <= 35120);
 * larger pencils return PARAHT_ERR_UNSUPPORTED before any work.  Every call on
 * one GPU shares that array, so calls that overlap on one GPU race.  flags
 * may hold PARAHT_STAGE2_GIVENS only (per-phase timing needs events).
 * Arguments as paraht_reduce_dev, the context replaced by the stream; the
 * call waits for the stream. */
int paraht_reduce_dev_noalloc(void* stream, int n, double* dA, int lda, double* dB, int ldb,
                              double* dQ, int ldq, double* dZ, int ldz, int r, int p, int nq,
                              int flags);

/* ---- Batched reduction of many small pencils ----
 * No counterpart in the reference.  Reduces `count` independent pencils with
 * both stages in one persistent launch: one workgroup per compute unit takes
 * pencils from a work queue ordered by decreasing n^3 and reduces each by
 * itself.  pencils[i].offset is the element offset of pencil i's n x n
 * column-major matrices (leading dimension n) in each of the device pools
 * dA, dB, dQ, dZ; the matrices are reduced in place (Q, Z accumulated).
 * r, p, nq of 0 take the library's size-bucket defaults
 * (paraht_batched_defaults); supported: r <= 16, p*r <= 128,
 * (r, nq) in {(8,4), (8,8), (16,4), (16,8), (16,16)}.
 * PARAHT_BATCH_GRAPH captures the launch in a HIP graph and replays it while
 * the arguments repeat. */
typedef struct {
  long long offset;
  int n, r, p, nq;
} paraht_pencil;
#define PARAHT_BATCH_GRAPH 0x4
int paraht_batched_dev(paraht_ctx* ctx, int count, const paraht_pencil* pencils, double* dA,
                       double* dB, double* dQ, double* dZ, int flags);
/* The (r, p, nq) the batched path uses for a pencil of order n when the
 * caller passes zeros (the size-bucket table). */
int paraht_batched_defaults(int n, int* r, int* p, int* nq);

/* ---- Reverse-mode adjoint of the reduction ----
 * No counterpart in the reference.  For reduced pencils (H, T) with
 * cotangents folded into g_Q, g_Z (strictly lower parts of G - G^T, with
 * G_Q = Q^T Qbar - Hbar H^T - Tbar T^T and G_Z = Z^T Zbar + H^T Hbar +
 * T^T Tbar, column 0 of g_Q zero), runs the transposed recurrence of the
 * structure constraints on dH, dT and writes X_A (rows >= j+2 of each column
 * j) and X_B (rows >= j+1); then Abar = Q (hess(Hbar) + X_A) Z^T and
 * Bbar = Q (triu(Tbar) + X_B) Z^T.  Tinv is T^{-1}.  All arrays are device
 * pools holding each pencil's n x n column-major matrix (ld = n) at its
 * offset; dGQ is overwritten (workspace); dXA, dXB must be zero on entry.
 * One pencil of order >= 512 runs over a cooperative grid, otherwise one
 * workgroup per pencil.  Requires H(j+1, j) != 0 and T nonsingular. */
int paraht_adjoint_dev(paraht_ctx* ctx, int count, const paraht_pencil* pencils,
                       const double* dH, const double* dT, const double* dTinv, double* dGQ,
                       const double* dGZ, double* dXA, double* dXB);

/* ---- QZ iteration on a Hessenberg-triangular pencil ----
 * No counterpart in the reference: its driver stops at the HT form and checks
 * it against LAPACK DGGHD3 (src/profilereduction.f90:152-162).  These take the
 * place of the LAPACK routine that follows it, DHGEQZ(JOB, COMPQ, COMPZ, N,
 * ILO=1, IHI=N, H, LDH, T, LDT, ALPHAR, ALPHAI, BETA, Q, LDQ, Z, LDZ, ...):
 * the implicit double-shift QZ iteration on (H, T) with H upper Hessenberg
 * and T upper triangular (only those parts are read).  The generalized
 * eigenvalues are (alphar[j] + i alphai[j]) / beta[j]; beta[j] = 0 marks an
 * infinite eigenvalue; beta >= 0; complex pairs are adjacent with
 * alphai[j] > 0.
 * Without PARAHT_QZ_SCHUR (DHGEQZ's JOB='E') only the eigenvalues are
 * computed and H, T are left in an unspecified state.  With it (JOB='S')
 * H becomes quasi-upper-triangular S (1x1 and 2x2 blocks, the latter for
 * complex pairs, whose T block is diagonal and positive), T upper triangular
 * P, and Q <- Q Qz, Z <- Z Zz when Q, Z are not NULL (COMPQ/COMPZ='V'), so
 * that Q^T A Z = S and Q^T B Z = P for the pencil (A, B) whose reduction gave
 * (H, T, Q, Z).  *info: 0, or the 1-based index of the last eigenvalue that
 * did not converge within 30 n sweeps (DHGEQZ's INFO = 1..N); eigenvalues
 * info+1..n are valid then (a pencil with NaN or Inf entries returns n at
 * once).  One workgroup per pencil on the GPU. */
#define PARAHT_QZ_SCHUR 0x8
int paraht_hgeqz(int n, double* H, int ldh, double* T, int ldt, double* Q, int ldq, double* Z,
                 int ldz, double* alphar, double* alphai, double* beta, int flags, int* info);
/* The same on device arrays (alphar, alphai, beta: n doubles each in HBM;
 * info on the host). */
int paraht_hgeqz_dev(paraht_ctx* ctx, int n, double* dH, int ldh, double* dT, int ldt,
                     double* dQ, int ldq, double* dZ, int ldz, double* dalphar, double* dalphai,
                     double* dbeta, int flags, int* info);
/* Many pencils in one launch, packed as for paraht_batched_dev (pencil i's
 * n x n matrices at pencils[i].offset of each pool, ld = n; r, p, nq are
 * ignored; dQ, dZ may be NULL).  Pencil i's eigenvalues are at element
 * offset n_0 + ... + n_{i-1} of dalphar, dalphai, dbeta; info[i] (host array
 * of count ints) is its status. */
int paraht_hgeqz_batched_dev(paraht_ctx* ctx, int count, const paraht_pencil* pencils,
                             double* dH, double* dT, double* dQ, double* dZ, double* dalphar,
                             double* dalphai, double* dbeta, int flags, int* info);

/* Library version (major*10000 + minor*100 + patch). */
int paraht_version(void);
/* Human-readable text of a status code. */
const char* paraht_status_string(int status);

#ifdef __cplusplus
}
#endif
#endif /* PARAHT_H */
