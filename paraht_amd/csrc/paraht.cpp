// Host orchestration and C-ABI of libparaht_hip.so (see include/paraht.h).
//
// Stage 1 follows ab2rht_simple (src/ab2rht.f90:8-495) panel by panel:
// left generation -> left updates of B, A and Q -> the bottom-up chain of
// right generators, each immediately applied to rows 1..ib2 of B (the
// reference's lookahead + rest of B) -> right updates of A and Z.
// Stage 2 follows rht2ht_givens (src/rht2ht.f90:696-1344) sweep by sweep:
// generation in the band window -> remaining row updates (A, B above the
// window; Q, Z) -> remaining column updates (A, B right of the window).
// All work of a call is enqueued on the context's stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <vector>

#include "../../include/paraht.h"
#include "kernels.hpp"

using pht::ApplyBlock;

// Per-kernel-class event timing (enabled by paraht_ctx_profile).
struct KernelRec {
  int kid;
  int e0, e1;  // indices into the event pool
  double flops, bytes;
};

struct paraht_ctx {
  int device = 0;
  int cu_count = 256;  // compute units of the device
  hipStream_t own = nullptr;
  hipStream_t stream = nullptr;
  // second stream for work off the critical path (stage 2's Q/Z updates),
  // joined back into `stream` before a call returns
  hipStream_t aux = nullptr;
  // third stream: stage 2's Q/Z updates, beside the A/B far updates on aux
  hipStream_t aux2 = nullptr;
  hipEvent_t ev_tape = nullptr;   // a sweep's tape is complete (stream -> aux2)
  hipEvent_t ev_free[2] = {nullptr, nullptr};  // aux2 (aux in stage 1) is done with tape buffer b
  hipEvent_t ev_near = nullptr;   // a sweep's lookahead updates are done (stream -> aux)
  hipEvent_t ev_far = nullptr;    // a sweep's remaining A/B updates are done (aux -> stream)
  hipEvent_t ev_left = nullptr;   // a panel's left tapes are complete (stream -> aux)
  hipEvent_t ev_aleft = nullptr;  // aux has applied a panel's left update to A
  void* ws = nullptr;
  size_t ws_bytes = 0;
  // a context over a caller's fixed workspace, running everything on the
  // call's stream (the allocation-free entry point): no streams or events
  // are created, ws is never reallocated, and larger needs are refused
  bool single = false;
  double times[8] = {0};
  bool profiling = false;
  std::vector<hipEvent_t> pool;
  size_t pool_used = 0;
  std::vector<KernelRec> recs;
  // accumulated per kernel id
  double k_ms[PARAHT_NKERNELS] = {0}, k_flops[PARAHT_NKERNELS] = {0},
         k_bytes[PARAHT_NKERNELS] = {0};
  long long k_n[PARAHT_NKERNELS] = {0};
  // the batched path's captured launch and the arguments it was captured for
  hipGraphExec_t batch_exec = nullptr;
  std::vector<long long> batch_key;
};

namespace {

bool verbose() {
  static int v = -1;
  if (v < 0) {
    const char* e = getenv("PARAHT_VERBOSE");
    v = (e && *e && *e != '0') ? 1 : 0;
  }
  return v == 1;
}

int hip_fail(hipError_t e, const char* what, int line) {
  if (verbose())
    fprintf(stderr, "paraht: HIP error %d (%s) in %s (paraht.cpp:%d)\n", (int)e,
            hipGetErrorString(e), what, line);
  return e == hipErrorOutOfMemory ? PARAHT_ERR_NOMEM : PARAHT_ERR_HIP;
}

#define CK(expr)                                           \
  do {                                                     \
    hipError_t _e = (expr);                                \
    if (_e != hipSuccess) return hip_fail(_e, #expr, __LINE__); \
  } while (0)

// grow-only device workspace of the context
int ensure_ws(paraht_ctx* c, size_t bytes) {
  if (c->ws_bytes >= bytes) return 0;
  if (c->single) return PARAHT_ERR_UNSUPPORTED;  // a fixed workspace: never reallocated
  if (c->ws) {
    CK(hipStreamSynchronize(c->stream));
    CK(hipFree(c->ws));
    c->ws = nullptr;
    c->ws_bytes = 0;
  }
  CK(hipMalloc(&c->ws, bytes));
  c->ws_bytes = bytes;
  return 0;
}

int check_common(int n, int lda, int ldb, int ldq, int ldz) {
  if (n < 0) return -1;
  if (lda < std::max(1, n)) return -3;
  if (ldb < std::max(1, n)) return -5;
  if (ldq < std::max(1, n)) return -7;
  if (ldz < std::max(1, n)) return -9;
  return 0;
}

// Event pair around one launch when profiling; folds finished records into
// the per-kernel sums whenever the pool runs low.
int prof_flush(paraht_ctx* c) {
  if (c->recs.empty()) return 0;
  CK(hipStreamSynchronize(c->stream));
  if (c->aux) CK(hipStreamSynchronize(c->aux));
  if (c->aux2) CK(hipStreamSynchronize(c->aux2));
  for (const KernelRec& r : c->recs) {
    float t = 0.f;
    CK(hipEventElapsedTime(&t, c->pool[r.e0], c->pool[r.e1]));
    c->k_ms[r.kid] += t;
    c->k_n[r.kid] += 1;
    c->k_flops[r.kid] += r.flops;
    c->k_bytes[r.kid] += r.bytes;
  }
  c->recs.clear();
  c->pool_used = 0;
  return 0;
}

// Makes room for `need` more events (flushing finished records if the pool
// is at its cap), so both events of one launch come from the same pool round.
int prof_reserve(paraht_ctx* c, size_t need) {
  if (c->pool_used + need <= c->pool.size()) return 0;
  if (c->pool.size() >= 8192) return prof_flush(c);
  size_t grow = c->pool.size() ? c->pool.size() : 512;
  for (size_t i = 0; i < grow; ++i) {
    hipEvent_t ev;
    CK(hipEventCreate(&ev));
    c->pool.push_back(ev);
  }
  return 0;
}

int prof_event(paraht_ctx* c, int* idx, hipStream_t st) {
  *idx = (int)c->pool_used++;
  CK(hipEventRecord(c->pool[*idx], st));
  return 0;
}

struct Prof {
  paraht_ctx* c;
  hipStream_t st;
  int kid, e0 = -1;
  double flops, bytes;
  Prof(paraht_ctx* cx, hipStream_t s, int k, double f, double b)
      : c(cx), st(s), kid(k), flops(f), bytes(b) {}
  int begin() {
    if (!c->profiling) return 0;
    if (int e = prof_reserve(c, 2)) return e;
    return prof_event(c, &e0, st);
  }
  int end() {
    if (!c->profiling) return 0;
    int e1 = -1;
    if (int e = prof_event(c, &e1, st)) return e;
    c->recs.push_back({kid, e0, e1, flops, bytes});
    return 0;
  }
};

// Runs `launch` (enqueued on stream `strm`) between the profiling events of
// kernel class `kid`.
#define PROF_ON(strm, kid, flops, bytes, launch)          \
  do {                                                    \
    Prof _p(c, (strm), (kid), (flops), (bytes));          \
    if (int _e = _p.begin()) return _e;                   \
    CK(launch);                                           \
    if (int _e = _p.end()) return _e;                     \
  } while (0)
#define PROF(kid, flops, bytes, launch) PROF_ON(c->stream, kid, flops, bytes, launch)

// Creates the context's second and third streams (default priority) and
// their events; a single-stream context has none.
int ensure_aux(paraht_ctx* c) {
  if (c->aux || c->single) return 0;
  CK(hipEventCreateWithFlags(&c->ev_tape, hipEventDisableTiming));
  CK(hipEventCreateWithFlags(&c->ev_free[0], hipEventDisableTiming));
  CK(hipEventCreateWithFlags(&c->ev_free[1], hipEventDisableTiming));
  CK(hipEventCreateWithFlags(&c->ev_left, hipEventDisableTiming));
  CK(hipEventCreateWithFlags(&c->ev_aleft, hipEventDisableTiming));
  CK(hipEventCreateWithFlags(&c->ev_near, hipEventDisableTiming));
  CK(hipEventCreateWithFlags(&c->ev_far, hipEventDisableTiming));
  CK(hipStreamCreateWithFlags(&c->aux, hipStreamNonBlocking));
  CK(hipStreamCreateWithFlags(&c->aux2, hipStreamNonBlocking));
  return 0;
}

// The streams of a call: off-critical-path work goes to the context's second
// (aux) and third (aux2) streams, ordered against the call's stream by
// events; a single-stream context runs it on the call's stream, in program
// order, and records no events.
hipStream_t side1(paraht_ctx* c) { return c->single ? c->stream : c->aux; }
hipStream_t side2(paraht_ctx* c) { return c->single ? c->stream : c->aux2; }
int rec(paraht_ctx* c, hipEvent_t ev, hipStream_t s) {
  if (c->single) return 0;
  CK(hipEventRecord(ev, s));
  return 0;
}
// s waits for the work before the latest record of ev
int wait_on(paraht_ctx* c, hipStream_t s, hipEvent_t ev) {
  if (c->single) return 0;
  CK(hipStreamWaitEvent(s, ev, 0));
  return 0;
}

// algorithmic flops of a panel's compact-WY updates: 4*len*k*width per block,
// width = the part of [lo, hi) inside the launch's range
double panel_wy_flops(int n, int nb, int p, int jb0, int nblocks, int which, int lo, int hi) {
  double f = 0;
  for (int bl = 0; bl < nblocks; ++bl) {
    const pht::S1Block b = pht::s1_block(n, nb, p, jb0, nblocks, bl);
    int blo = 0, bhi = n, k = b.kr;
    if (which == 0) { blo = b.i0; k = b.k; }          // B left
    else if (which == 1) { blo = jb0 + nb; k = b.k; } // A left
    else if (which == 2) { k = b.k; }                 // Q
    else if (which == 4) { bhi = b.i0 + b.len; }      // B right
    f += 4.0 * b.len * k * std::max(0, std::min(bhi, hi) - std::max(blo, lo));
  }
  return f;
}

// algorithmic flops of the stage-1 generators (leading terms of the LAPACK
// counts of the routines they restate)
double qr_flops(int m, int k) {  // DGEQR2 (m x k) + DLARFT
  return 2.0 * m * k * k - (2.0 / 3.0) * k * k * k + (double)m * k * k;
}
double rq_flops(int len, int nb) {  // DGERQ2 + first nb rows of Q + DGELQ2 + DLARFT
  const double l = len;
  return (4.0 / 3.0) * l * l * l + 2.0 * nb * l * l + 2.0 * l * nb * nb -
         (2.0 / 3.0) * nb * nb * nb + l * nb * nb;
}

// algorithmic flops of one stage-2 sweep's generation (6 per rotation and
// element pair, 2 per mat-vec term); mirrors the step loop of k_s2_sweep
double s2_gen_flops(int n, int r, int nq, int s) {
  const int ldr = r + nq;
  double f = 0.0;
  for (int ic = 1; ic <= nq; ++ic) {
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
      const int nb = std::min(r, n - row0 + 1);
      const int L = std::min(ldr - 1, n - (s + (ib - 1) * r));
      if (ic > 1 && L > 0) f += 2.0 * L * L * (row0 + r - 1 <= n ? 2 : 1);
      if (nb >= 2) {
        const int istart = std::max(s + 1, s + 1 + (ib - 1 - nq + ic) * r);
        const int nAr = std::min(n, row0 + r + nb - 1) - istart + 1;
        const int nBr = std::max(0, row0 - istart);
        f += 6.0 * (nb - 1) * (2.0 * nb + ldr + nAr + nBr);
      }
    }
  }
  return f;
}

// algorithmic bytes of one sweep's column updates: each block's rows of A
// and B right of the window, read and written once
double s2_cols_bytes(int n, int r, int nq, int s, int nblocks, const pht::S2Range& rg) {
  double b = 0.0;
  auto span = [&](int istop) {
    const int lo = istop + rg.c_lo * r;
    const int hi = rg.c_hi < 0 ? n : std::min(n, istop + rg.c_hi * r);
    return std::max(0, hi - std::max(0, lo));
  };
  for (int ib = 1; ib <= nblocks; ++ib) {
    const int W = std::min(r + nq - 1, n - (s + (ib - 1) * r));
    if (W <= 0) continue;
    const int istopA = (ib == 1) ? s + nq - 1 : s + (ib - 2) * r + nq;
    const int istopB = s + (ib - 1) * r + r + nq - 1;
    b += 16.0 * W * (span(istopA) + span(istopB));
  }
  return b;
}

// algorithmic bytes of one sweep's A/B row updates in a range: each row's
// entries of a block's window, read and written once per block that touches
// it (rows thr(ib - lo_shift, 0) - 1 <= i < thr(ib - hi_shift, nq-1) - 1)
double s2_rows_ab_bytes(int n, int r, int nq, int s, int nblocks, const pht::S2Range& rg) {
  double b = 0.0;
  auto thr = [&](int x, int icm) { return std::max(s + 1, s + 1 + (x - nq + icm) * r); };
  for (int ib = 1; ib <= nblocks; ++ib) {
    const int W = std::min(r + nq - 1, n - (s + (ib - 1) * r));
    if (W <= 0) continue;
    const int hi = std::min({n, rg.row_end, thr(ib - rg.hi_shift, nq - 1) - 1});
    const int lo = rg.lo_shift < 0 ? 0 : thr(ib - rg.lo_shift, 0) - 1;
    b += 32.0 * W * std::max(0, hi - lo);
  }
  return b;
}

struct Timer {
  paraht_ctx* c;
  bool on;
  std::vector<hipEvent_t> ev;
  explicit Timer(paraht_ctx* cx, bool enable) : c(cx), on(enable) {}
  ~Timer() {
    for (auto e : ev) (void)hipEventDestroy(e);
  }
  int mark() {
    if (!on) return -1;
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return -1;
    if (hipEventRecord(e, c->stream) != hipSuccess) return -1;
    ev.push_back(e);
    return (int)ev.size() - 1;
  }
  double ms(int a, int b) {
    if (!on || a < 0 || b < 0) return 0.0;
    float t = 0.f;
    if (hipEventElapsedTime(&t, ev[a], ev[b]) != hipSuccess) return 0.0;
    return t;
  }
};

// ------------------------------------------------------------------------
// Workspace of the two stages (nbuf = 2: tapes double-buffered between
// streams; 1: a single stream)
size_t s1_ws_bytes(int n, int nb, int p, int nbuf) {
  if (n <= nb + 1) return 0;
  const int P = p * nb;
  const int nbmax = pht::s1_nblocks(n, nb, p, 0);
  const int np = pht::s1_npanels(n, nb);
  size_t nd = 0;
  for (int q = 0; q < np; ++q) nd += pht::s1_nblocks(n, nb, p, q * nb);
  const size_t vsz = (size_t)nbmax * P * nb, tsz = (size_t)nbmax * nb * nb;
  const size_t scr = (size_t)P * P + (size_t)nb * P + (size_t)nbmax * 64;
  return (2 * nbuf * vsz + 2 * nbuf * tsz + scr) * sizeof(double) + 5 * nd * sizeof(ApplyBlock) +
         256;
}
size_t s2_ws_bytes(int n, int r, int nq, int nbuf) {
  if (n <= 2) return 0;
  const int ldr = r + nq;
  const int nbl = (n - 1) / r + ((n - 1) % r != 0);
  const size_t tsz = (size_t)ldr * nq * nbl;
  const size_t msz = (size_t)ldr * ldr * nbl;
  return (4 * nbuf * tsz + msz) * sizeof(double) + pht::s2_counter_ints(nq) * sizeof(int);
}

// ------------------------------------------------------------------------
// Stage 1
// ------------------------------------------------------------------------
// Per panel (ab2rht.f90:71-467): the chain of block QRs, the left updates of
// B (call stream), A (aux) and Q (aux), then the chain of right generators,
// each block's reflector applied at once to rows 0..i0+len-1 of B (the
// reference's lookahead and rest of B, ab2rht.f90:317-367, in one launch: the
// update is latency-bound, so the rows above the lookahead cost nothing on
// the chain).  Then the right updates of A (the rows the next panel reads on
// the chain, the rows above on aux) and Z (aux).  The panel tapes are
// double-buffered so that panel k's Q and Z updates overlap panel k+1.
int stage1(paraht_ctx* c, int n, double* A, int lda, double* B, int ldb,
           double* Q, int ldq, double* Z, int ldz, int nb, int p, int flags) {
  if (n <= nb + 1) return 0;  // no panel: ab2rht.f90:71 loop is empty
  hipStream_t st = c->stream;
  if (int e = ensure_aux(c)) return e;
  hipStream_t sa = side1(c);
  const int P = p * nb;
  const int nbmax = pht::s1_nblocks(n, nb, p, 0);
  const int np = pht::s1_npanels(n, nb);
  int nd = 0;
  for (int q = 0; q < np; ++q) nd += pht::s1_nblocks(n, nb, p, q * nb);
  const int nbuf = c->single ? 1 : 2;  // one stream needs no second tape
  const size_t vsz = (size_t)nbmax * P * nb, tsz = (size_t)nbmax * nb * nb;
  const size_t scr = (size_t)P * P + (size_t)nb * P + (size_t)nbmax * 64;  // + panel taus
  if (int e = ensure_ws(c, s1_ws_bytes(n, nb, p, nbuf))) return e;
  double* w = (double*)c->ws;
  double *VLb[2], *VRb[2], *TLb[2], *TRb[2];
  for (int b = 0; b < nbuf; ++b) {
    VLb[b] = w; w += vsz;
    VRb[b] = w; w += vsz;
    TLb[b] = w; w += tsz;
    TRb[b] = w; w += tsz;
  }
  if (nbuf == 1) {
    VLb[1] = VLb[0]; VRb[1] = VRb[0]; TLb[1] = TLb[0]; TRb[1] = TRb[0];
  }
  double* gen_ws = w;  // the generators' scratch and the panel's taus
  ApplyBlock* dD = (ApplyBlock*)(((uintptr_t)(gen_ws + scr) + 255) & ~(uintptr_t)255);
  ApplyBlock* dBL = dD;
  ApplyBlock* dAL = dD + nd;
  ApplyBlock* dQL = dD + 2 * (size_t)nd;
  ApplyBlock* dRR = dD + 3 * (size_t)nd;
  ApplyBlock* dBR = dD + 4 * (size_t)nd;
  CK(pht::s1_block_lists(st, dD, nd, n, nb, p));

  const bool tc = (flags & PARAHT_TIME_COMPONENTS) != 0;
  // register-resident generators cover blocks of <= 128 rows, <= 64 reflectors
  const bool fast_gen = nb <= 64 && P <= 128 && getenv("PARAHT_GENERIC_S1") == nullptr;
  Timer tm(c, tc);
  std::vector<int> marks;
  for (int kp = 0, jb0 = 0, first = 0; kp < np; ++kp, jb0 += nb) {
    const int nblocks = pht::s1_nblocks(n, nb, p, jb0);
    const int buf = kp % nbuf;
    double* VL = VLb[buf];
    double* VR = VRb[buf];
    double* TL = TLb[buf];
    double* TR = TRb[buf];
    // panel kp-2's Q and Z updates (aux) have read these tapes
    if (kp >= 2 && nbuf == 2)
      if (int e = wait_on(c, st, c->ev_free[buf])) return e;
    const int m0 = tm.mark();
    double qf = 0.0;
    for (int bl = 0; bl < nblocks; ++bl) {
      const pht::S1Block b = pht::s1_block(n, nb, p, jb0, nblocks, bl);
      qf += qr_flops(b.len, b.k);
    }
    if (fast_gen)
      PROF(PARAHT_K_S1_PANEL_QR, qf, 0.0,
           pht::s1_panel_qr_fast(st, A, lda, n, jb0, nb, p, nblocks, VL, P,
                                 (long long)P * nb, TL, nb, (long long)nb * nb,
                                 gen_ws + (size_t)P * P + (size_t)nb * P));
    else
      PROF(PARAHT_K_S1_PANEL_QR, qf, 0.0,
           pht::s1_panel_qr(st, A, lda, n, jb0, nb, p, nblocks, VL, P,
                            (long long)P * nb, TL, nb, (long long)nb * nb));
    const int m1 = tm.mark();
    // the RQ chain reads only B: the left update of A runs beside it on aux,
    // and the right update of A waits for it
    if (int e = rec(c, c->ev_left, st)) return e;
    if (int e = wait_on(c, sa, c->ev_left)) return e;
    PROF_ON(sa, PARAHT_K_WY_APPLY, panel_wy_flops(n, nb, p, jb0, nblocks, 1, 0, n), 0.0,
            pht::apply_blocks(sa, pht::kLeftT, A, lda, VL, P, TL, nb, dAL + first, nblocks, nb,
                              P, jb0 + nb, n));
    if (int e = rec(c, c->ev_aleft, sa)) return e;
    PROF(PARAHT_K_WY_APPLY, panel_wy_flops(n, nb, p, jb0, nblocks, 0, 0, n), 0.0,
         pht::apply_blocks(st, pht::kLeftT, B, ldb, VL, P, TL, nb, dBL + first, nblocks, nb, P,
                           jb0 + nb, n));
    const int m2 = tm.mark();
    for (int bl = 0; bl < nblocks; ++bl) {
      const pht::S1Block b = pht::s1_block(n, nb, p, jb0, nblocks, bl);
      const double rf = b.len > nb ? rq_flops(b.len, nb) : (4.0 / 3.0) * b.len * b.len * b.len;
      if (fast_gen)
        PROF(PARAHT_K_S1_RQ_GEN, rf, 0.0,
             pht::s1_rq_fast(st, B, ldb, b.i0, b.len, nb, VR + (size_t)bl * P * nb, P,
                             TR + (size_t)bl * nb * nb, nb));
      else
        PROF(PARAHT_K_S1_RQ_GEN, rf, 0.0,
             pht::s1_rq_gen(st, B, ldb, b.i0, b.len, nb, VR + (size_t)bl * P * nb, P,
                            TR + (size_t)bl * nb * nb, nb, gen_ws));
      // the subdiagonal part of the block's first columns becomes exact
      // zeros inside the update where the applier can (ab2rht.f90:329-333)
      const int zc = b.len > nb ? nb : b.len;
      const bool fuse = pht::apply_fuses_zero(b.kr, P);
      PROF(PARAHT_K_WY_APPLY, 4.0 * b.len * b.kr * (b.i0 + b.len), 0.0,
           pht::apply_blocks(st, pht::kRightN, B, ldb, VR, P, TR, nb, dBR + first + bl, 1,
                             b.kr, P, 0, b.i0 + b.len, fuse ? zc : 0));
      if (!fuse)
        PROF(PARAHT_K_OTHER, 0.0, 0.0, pht::zero_below(st, B, ldb, b.i0, b.len, zc));
    }
    const int m3 = tm.mark();
    // this panel's right tapes are complete: aux takes the rows of A above
    // the next panel's first row (jb0 + nb), which no later panel generation
    // reads and no later call-stream update touches before the next panel's
    // tapes exist, then Q and Z (written by these updates only)
    if (int e = rec(c, c->ev_tape, st)) return e;
    if (int e = wait_on(c, sa, c->ev_tape)) return e;
    const int split = std::min(n, jb0 + nb);
    if (split > 0)
      PROF_ON(sa, PARAHT_K_WY_APPLY, panel_wy_flops(n, nb, p, jb0, nblocks, 3, 0, split), 0.0,
              pht::apply_blocks(sa, pht::kRightN, A, lda, VR, P, TR, nb, dRR + first, nblocks,
                                nb, P, 0, split));
    // the rows the next panel's QR reads stay on the chain, after this
    // panel's left update of A (aux)
    if (int e = wait_on(c, st, c->ev_aleft)) return e;
    PROF(PARAHT_K_WY_APPLY, panel_wy_flops(n, nb, p, jb0, nblocks, 3, split, n), 0.0,
         pht::apply_blocks(st, pht::kRightN, A, lda, VR, P, TR, nb, dRR + first, nblocks, nb, P,
                           split, n));
    const int m4 = tm.mark();
    PROF_ON(sa, PARAHT_K_WY_APPLY, panel_wy_flops(n, nb, p, jb0, nblocks, 2, 0, n), 0.0,
            pht::apply_blocks(sa, pht::kRightN, Q, ldq, VL, P, TL, nb, dQL + first, nblocks, nb,
                              P, 0, n));
    PROF_ON(sa, PARAHT_K_WY_APPLY, panel_wy_flops(n, nb, p, jb0, nblocks, 3, 0, n), 0.0,
            pht::apply_blocks(sa, pht::kRightN, Z, ldz, VR, P, TR, nb, dRR + first, nblocks, nb,
                              P, 0, n));
    if (int e = rec(c, c->ev_free[buf], sa)) return e;
    if (tc) marks.insert(marks.end(), {m0, m1, m2, m3, m4});
    first += nblocks;
  }
  // join: the last panel's Q and Z updates (aux runs its updates in order)
  if (np > 0)
    if (int e = wait_on(c, st, c->ev_free[(np - 1) % nbuf])) return e;
  if (tc) {
    CK(hipStreamSynchronize(st));
    double gl = 0, gr = 0, mul = 0;
    for (size_t i = 0; i < marks.size(); i += 5) {
      gl += tm.ms(marks[i], marks[i + 1]);
      mul += tm.ms(marks[i + 1], marks[i + 2]) + tm.ms(marks[i + 3], marks[i + 4]);
      gr += tm.ms(marks[i + 2], marks[i + 3]);
    }
    c->times[1] = gl;
    c->times[2] = gr;
    // same line layout as ab2rht.f90:488-489 (seconds)
    printf("      genlefttime: %16.6f ;genrighttime: %16.6f ;multime: %16.6f\n",
           gl * 1e-3, gr * 1e-3, mul * 1e-3);
  }
  return 0;
}

// ------------------------------------------------------------------------
// QR of a general B (the step before stage 1)
// ------------------------------------------------------------------------
// The reference's driver makes B upper triangular with LAPACK dgeqrf on the
// host and keeps R (src/profilereduction.f90:117-123).  Here B = Q_B R on the GPU by
// panels of nb columns: each panel's rows from its first column down are
// factored bottom-up in blocks of p*nb rows that overlap by nb rows (the
// stage-1 panel kernels, on B shifted up by nb rows), and the panel's block
// reflectors are applied to B's trailing columns, to A from the left and to
// Q from the right: B <- R, A <- Q_B^T A, Q <- Q Q_B, so a pencil (A, B)
// with any B reduces with Q^T A0 Z = H, Q^T B0 Z = T.  The last panel starts
// at column n - nb; its columns already triangular get identity reflectors.
size_t qr_ws_bytes(int n, int nb, int p) {
  const int P = p * nb;
  const int nbmax = pht::s1_nblocks(n + nb, nb, p, 0);
  const int np = pht::qr_npanels(n, nb);
  size_t nd = 0;
  for (int q = 0; q < np; ++q) nd += pht::s1_nblocks(n + nb, nb, p, pht::qr_panel_col(n, nb, q));
  const size_t vsz = (size_t)nbmax * P * nb, tsz = (size_t)nbmax * nb * nb;
  return (vsz + tsz + (size_t)nbmax * 64) * sizeof(double) + 3 * nd * sizeof(ApplyBlock) + 256;
}

int qr_of_b(paraht_ctx* c, int n, double* A, int lda, double* B, int ldb, double* Q, int ldq,
            int nb, int p) {
  if (n < 2) return 0;
  nb = std::min(nb, n);
  hipStream_t st = c->stream;
  const int P = p * nb;
  const int nbmax = pht::s1_nblocks(n + nb, nb, p, 0);
  const int np = pht::qr_npanels(n, nb);
  int nd = 0;
  for (int q = 0; q < np; ++q) nd += pht::s1_nblocks(n + nb, nb, p, pht::qr_panel_col(n, nb, q));
  if (int e = ensure_ws(c, qr_ws_bytes(n, nb, p))) return e;
  double* V = (double*)c->ws;
  double* T = V + (size_t)nbmax * P * nb;
  double* taus = T + (size_t)nbmax * nb * nb;
  ApplyBlock* dD =
      (ApplyBlock*)(((uintptr_t)(taus + (size_t)nbmax * 64) + 255) & ~(uintptr_t)255);
  CK(pht::qr_block_lists(st, dD, nd, n, nb, p));
  // the panel kernels factor rows [jb + nb, n') of columns [jb, jb + nb): on
  // B shifted up by nb rows (every row they touch is >= nb there) with
  // n' = n + nb, those are rows [jb, n) of B
  double* Bup = B - nb;
  const bool fast_gen = nb <= 64 && P <= 128 && getenv("PARAHT_GENERIC_S1") == nullptr;
  for (int q = 0, first = 0; q < np; ++q) {
    const int jb = pht::qr_panel_col(n, nb, q);
    const int nblocks = pht::s1_nblocks(n + nb, nb, p, jb);
    double qf = 0.0;
    for (int bl = 0; bl < nblocks; ++bl) {
      const pht::S1Block b = pht::s1_block(n + nb, nb, p, jb, nblocks, bl);
      qf += qr_flops(b.len, b.k);
    }
    if (fast_gen)
      PROF(PARAHT_K_S1_PANEL_QR, qf, 0.0,
           pht::s1_panel_qr_fast(st, Bup, ldb, n + nb, jb, nb, p, nblocks, V, P,
                                 (long long)P * nb, T, nb, (long long)nb * nb, taus));
    else
      PROF(PARAHT_K_S1_PANEL_QR, qf, 0.0,
           pht::s1_panel_qr(st, Bup, ldb, n + nb, jb, nb, p, nblocks, V, P, (long long)P * nb,
                            T, nb, (long long)nb * nb));
    PROF(PARAHT_K_WY_APPLY, 0.0, 0.0,
         pht::apply_blocks(st, pht::kLeftT, B, ldb, V, P, T, nb, dD + first, nblocks, nb, P,
                           jb + nb, n));
    PROF(PARAHT_K_WY_APPLY, 0.0, 0.0,
         pht::apply_blocks(st, pht::kLeftT, A, lda, V, P, T, nb, dD + nd + first, nblocks, nb,
                           P, 0, n));
    PROF(PARAHT_K_WY_APPLY, 0.0, 0.0,
         pht::apply_blocks(st, pht::kRightN, Q, ldq, V, P, T, nb, dD + 2 * (size_t)nd + first,
                           nblocks, nb, P, 0, n));
    first += nblocks;
  }
  return 0;
}

// ------------------------------------------------------------------------
// Stage 2 (Givens)
// ------------------------------------------------------------------------
// helper workgroups per stage-2 sweep column (PARAHT_S2_HELPERS overrides, 0..4);
// they take the window rows above each bulge off the chase's critical path
int s2_helpers() {
  static int h = [] {
    const char* e = getenv("PARAHT_S2_HELPERS");
    int v = e ? atoi(e) : 3;
    return v < 0 ? 0 : (v > pht::kS2MaxHelpers ? pht::kS2MaxHelpers : v);
  }();
  return h;
}

// Stage-2 lookahead (rht2ht.f90:1630-1772): a sweep's A/B far update is split
// into a lookahead part, applied on the main stream before the next sweep's
// generation, and the rest, applied on the second stream beside it (the
// reference's bulk update under the A(2,2)/B(2,2) tokens,
// rht2ht.f90:1803-1958).
//  * Rows above the window: the lookahead takes, for block ib's sequence icm,
//    the rows within s2_look_rows(nq) blocks of the window's first row
//    (thr(ib - Lr, icm) <= i+1 < thr(ib, icm)), the rest the rows above.
//  * Columns right of the window: the lookahead takes the columns within
//    s2_look_cols(nq) blocks of each block's istop, the rest the columns
//    beyond.
// Order: lookahead rows, lookahead columns (main stream), then rest rows,
// rest columns (second stream).  A right rotation on columns (y, y+1) and a
// left rotation on rows (x, x+1) that share entries must each be applied
// whole on one side of the other; with Lr >= Lc + 1 every pair a boundary
// cuts has its partner on the consistent side, and the split reproduces the
// one-pass update bit for bit.  nq and nq+1 blocks keep the next
// generation's entries out of the rest (tests/test_lookahead_split.py checks
// the index sets and the order).
int s2_look_cols(int nq) { return nq + 2; }
int s2_look_rows(int nq) { return s2_look_cols(nq) + 1; }
int s2_lookahead() {
  const char* e = getenv("PARAHT_S2_LOOKAHEAD");
  if (e && *e == '0') return 0;
  if (e && *e == '2') return 2;
  return 1;
}
// polls before a wait of the call's first sweep gives up
// (PARAHT_DEBUG_SPIN0; the error path's test sets it tiny)
int s2_first_spin() {
  const char* e = getenv("PARAHT_DEBUG_SPIN0");
  if (e && *e) {
    const int v = atoi(e);
    if (v > 0) return v;
  }
  return pht::kS2SpinLimit;
}

// the limits of the stage-2 kernels (checked before any work of a call)
int s2_check(int r, int nq) {
  if (r > 64) return PARAHT_ERR_UNSUPPORTED;
  if (pht::s2_sweep_lds(r, nq) == 0) return PARAHT_ERR_UNSUPPORTED;  // M of a block in LDS
  return 0;
}

int stage2(paraht_ctx* c, int n, double* A, int lda, double* B, int ldb,
           double* Q, int ldq, double* Z, int ldz, int r, int nq, int flags) {
  if (n <= 2) return 0;  // rht2ht.f90:1417 loop is empty
  if (int e = s2_check(r, nq)) return e;
  hipStream_t st = c->stream;
  const int ldr = r + nq;
  int nbl = (n - 1) / r + ((n - 1) % r != 0);
  const size_t tsz = (size_t)ldr * nq * nbl;
  const size_t msz = (size_t)ldr * ldr * nbl;
  // two tape buffers: the Q/Z updates (and the A/B rest) of sweep k run on
  // the other streams while sweep k+1 is generated into the other buffer
  // (one stream needs one)
  const int nbuf = c->single ? 1 : 2;
  if (int e = ensure_ws(c, s2_ws_bytes(n, r, nq, nbuf))) return e;
  if (int e = ensure_aux(c)) return e;
  double* tape[2] = {(double*)c->ws, (double*)c->ws + (nbuf - 1) * 4 * tsz};
  double* Mws = (double*)c->ws + nbuf * 4 * tsz;
  hipStream_t sq = side2(c);  // Q and Z
  const bool tc = (flags & PARAHT_TIME_COMPONENTS) != 0;
  const int look = c->single ? 0 : s2_lookahead();
  const bool split = pht::s2_far_split(r, nq) && look != 0;
  hipStream_t rest = look == 2 ? st : c->aux;  // the stream of the A/B rest
  const int Lc = s2_look_cols(nq), Lr = s2_look_rows(nq);
  Timer tm(c, tc);
  std::vector<int> marks;
  int buf = 0;
  bool far_pending = false;
  // aux2 starts after everything enqueued on the call's stream so far
  // (stage 1's updates of Q and Z included)
  if (int e = rec(c, c->ev_tape, st)) return e;
  if (int e = wait_on(c, sq, c->ev_tape)) return e;
  for (int s = 1, k = 0; s <= n - 2; s += nq, ++k) {
    buf = k & 1;
    double* CU = tape[buf];
    double* SU = CU + tsz;
    double* CV = SU + tsz;
    double* SV = CV + tsz;
    pht::S2Params P;
    P.n = n;
    P.r = r;
    P.nq = nq;
    P.jblock = s;
    P.nblocks = (n - s) / r + ((n - s) % r != 0);
    P.ldr = ldr;
    P.M = Mws;
    P.progress = (int*)(P.M + msz);
    P.hprog = P.progress + nq;
    // helpers need every workgroup resident and nq <= r+1 (a column's steps
    // then never read the rows above its earlier bulges)
    P.nhelp = (nq <= r + 1) ? std::min(s2_helpers(), c->cu_count / nq - 1) : 0;
    if (P.nhelp < 0) P.nhelp = 0;
    P.errflag = P.progress + (size_t)nq * (1 + pht::kS2MaxHelpers);
    P.spin = (k == 0) ? s2_first_spin() : pht::kS2SpinLimit;
    P.clear_err = (k == 0);
    const int m0 = tm.mark();
    // the buffer's previous tape (sweep k-2) must be consumed by its Q/Z update
    // (its A/B rest finished before sweep k-1's lookahead update)
    if (k >= 2)
      if (int e = wait_on(c, st, c->ev_free[buf])) return e;
    PROF(PARAHT_K_OTHER, 0.0, 0.0, pht::s2_prepare(st, P, CU, SU, CV, SV));
    PROF(PARAHT_K_S2_GEN, s2_gen_flops(n, r, nq, s), 0.0,
         pht::s2_generate(st, A, lda, B, ldb, P, CU, SU, CV, SV));
    const int m1 = tm.mark();
    const double w = (double)(n - s);
    if (int e = rec(c, c->ev_tape, st)) return e;
    if (int e = wait_on(c, sq, c->ev_tape)) return e;
    // algorithmic traffic: every element of the sweep's columns s+1..n of Q
    // and Z is read and written once
    PROF_ON(sq, PARAHT_K_S2_ROWS, 0.0, 16.0 * w * 2.0 * n,
            pht::s2_apply_rows(sq, A, lda, B, ldb, Q, ldq, Z, ldz, P, CU, SU, CV, SV,
                               pht::kRowsQZ));
    if (int e = rec(c, c->ev_free[buf], sq)) return e;
    const int rows_all = std::min(n, s + (P.nblocks - 1) * r);  // rows above some window
    const int col0 = std::min(n, s + nq - 1);                  // first column right of one
    if (split) {
      // the previous sweep's rest must be done before this lookahead
      if (far_pending) CK(hipStreamWaitEvent(st, c->ev_far, 0));
      pht::S2Range nr, nc;
      nr.lo_shift = Lr;
      nr.hi_shift = 0;
      nr.row_end = rows_all;
      nc.c_lo = 0;
      nc.c_hi = Lc;
      nc.col_begin = col0;
      PROF(PARAHT_K_S2_ROWS, 0.0, s2_rows_ab_bytes(n, r, nq, s, P.nblocks, nr),
           pht::s2_apply_rows(st, A, lda, B, ldb, Q, ldq, Z, ldz, P, CU, SU, CV, SV,
                              pht::kRowsAB, nr, rows_all));
      PROF(PARAHT_K_S2_COLS, 0.0, s2_cols_bytes(n, r, nq, s, P.nblocks, nc),
           pht::s2_apply_cols(st, A, lda, B, ldb, P, CU, SU, nc));
      CK(hipEventRecord(c->ev_near, st));
      CK(hipStreamWaitEvent(rest, c->ev_near, 0));
      pht::S2Range fr, fc;
      fc.c_lo = Lc;
      fc.c_hi = -1;
      fc.col_begin = std::min(n, col0 + Lc * r);
      // rows i+1 < thr(nblocks - Lr, nq - 1)
      const int far_rows = std::min(n, std::max(s, s + (P.nblocks - 1 - Lr) * r));
      fr.hi_shift = Lr;
      fr.row_end = far_rows;
      PROF_ON(rest, PARAHT_K_S2_ROWS, 0.0, s2_rows_ab_bytes(n, r, nq, s, P.nblocks, fr),
              pht::s2_apply_rows(rest, A, lda, B, ldb, Q, ldq, Z, ldz, P, CU, SU, CV, SV,
                                 pht::kRowsAB, fr, far_rows));
      PROF_ON(rest, PARAHT_K_S2_COLS, 0.0, s2_cols_bytes(n, r, nq, s, P.nblocks, fc),
              pht::s2_apply_cols(rest, A, lda, B, ldb, P, CU, SU, fc));
      CK(hipEventRecord(c->ev_far, rest));
      far_pending = true;
    } else {
      pht::S2Range all, allc;
      allc.col_begin = pht::s2_far_split(r, nq) ? col0 : 0;
      PROF(PARAHT_K_S2_ROWS, 0.0, s2_rows_ab_bytes(n, r, nq, s, P.nblocks, all),
           pht::s2_apply_rows(st, A, lda, B, ldb, Q, ldq, Z, ldz, P, CU, SU, CV, SV,
                              pht::kRowsAB, all, rows_all));
      PROF(PARAHT_K_S2_COLS, 0.0, s2_cols_bytes(n, r, nq, s, P.nblocks, allc),
           pht::s2_apply_cols(st, A, lda, B, ldb, P, CU, SU, allc));
    }
    const int m2 = tm.mark();
    if (tc) marks.insert(marks.end(), {m0, m1, m2});
  }
  // join: the last A/B rest and the last Q/Z update (each stream runs its
  // updates in order)
  if (far_pending) CK(hipStreamWaitEvent(st, c->ev_far, 0));
  if (int e = wait_on(c, st, c->ev_free[buf])) return e;
  {
    int err = 0;
    CK(hipMemcpyAsync(&err, (int*)(Mws + msz) + (size_t)nq * (1 + pht::kS2MaxHelpers),
                       sizeof(int), hipMemcpyDeviceToHost, st));
    CK(hipStreamSynchronize(st));
    if (err) {
      if (verbose()) fprintf(stderr, "paraht: stage-2 sweep wait exceeded its spin limit\n");
      return PARAHT_ERR_HIP;
    }
  }
  if (tc) {
    CK(hipStreamSynchronize(st));
    double gen = 0, app = 0;
    for (size_t i = 0; i < marks.size(); i += 3) {
      gen += tm.ms(marks[i], marks[i + 1]);
      app += tm.ms(marks[i + 1], marks[i + 2]);
    }
    c->times[4] = gen;
    // same line layout as rht2ht.f90:2045-2046 (seconds): the lookahead is
    // the main stream's share of the A/B far updates, the application the
    // rest (measured on the main stream only when the split is off)
    printf("      gen: %16.6f ;lookahead: %16.6f ;application: %16.6f\n",
           gen * 1e-3, split ? app * 1e-3 : 0.0, split ? 0.0 : app * 1e-3);
  }
  return 0;
}

// ------------------------------------------------------------------------
// Stage 2 (Householder, rht2ht_householder, src/rht2ht.f90:1346-2051)
// ------------------------------------------------------------------------
size_t s2h_ws_bytes(int n, int r, int nq, int nbuf) {
  if (n <= 2) return 0;
  const int nbl = (n - 1) / r + ((n - 1) % r != 0);
  return (size_t)nbuf * pht::hh_tape_doubles(r, nq, nbl) * sizeof(double) +
         pht::s2_counter_ints(nq) * sizeof(int);
}

// algorithmic work of a sweep's Householder generation: per step the left
// reflector on the B block (4 nb^2), the block's RQ (4/3 nb^3) and the first
// row of its orthogonal factor (2 nb^2), the right reflector on the A window,
// the B block and the rows above (4 nb (rows)); the deferred reflectors
// (4 r per earlier column, on two vectors)
double s2h_gen_flops(int n, int r, int nq, int s) {
  double f = 0.0;
  for (int ic = 1; ic <= nq; ++ic) {
    const int jc = s + ic - 1;
    for (int ib = 1;; ++ib) {
      if (ib == 1 ? jc > n : jc + 1 + (ib - 2) * r > n) break;
      const int row0 = s + (ib - 1) * r + ic;
      const int nb = std::min(r, n - row0 + 1);
      f += 8.0 * r * (ic - 1);
      if (nb < 2) continue;
      const int istart = ib == 1 ? s + 1 : std::max(s + 1, s + 1 + (ib - 1 - nq + ic) * r);
      const int rows = std::min(n, row0 + r + nb - 1) - row0 + 1 + nb + 2 * (row0 - istart);
      f += 6.0 * nb * nb + (4.0 / 3.0) * nb * nb * nb + 4.0 * nb * rows;
    }
  }
  return f;
}

// algorithmic bytes of a far-update launch: every entry of the windows it
// updates read and written once per block (the rows or columns in the
// phase's range; 16 bytes per entry and block)
double s2h_far_bytes(int n, int r, int nq, int s, int nblocks, bool cols, int phase, int nmat) {
  double b = 0.0;
  const int W = r + nq;
  for (int ib = 1; ib <= nblocks; ++ib) {
    const int w0 = s + (ib - 1) * r;
    const int w = std::min(W, n - w0);
    if (w <= 0) continue;
    long long cnt = 0;
    if (phase == pht::kHhFull) {
      cnt = n;
    } else if (cols) {
      const int istop = std::min(n, s + (ib - 1) * r + r + nq - 1);  // B's
      const int istop2 = std::min(n, s + (ib - 1) * r + r + nq + r * nq - 1);
      cnt = phase == pht::kHhLook ? istop2 - istop : (phase == pht::kHhFar ? n - istop2 : n - istop);
    } else {
      const int base = s + 1 + (ib - 1) * r;
      const int top = std::min(n, base - 1);
      const int look_lo = std::max(s, base - 1 - 2 * nq * r);
      cnt = phase == pht::kHhLook ? top - look_lo : (phase == pht::kHhFar ? look_lo : top);
    }
    b += 16.0 * w * std::max(0LL, cnt);
  }
  return b * nmat;
}

// far-update flops of a launch: 4 W nq (+ 2 nq^2) per row or column and block
double s2h_far_flops(double bytes, int r, int nq) {
  const double W = r + nq;
  return bytes / (16.0 * W) * (4.0 * W * nq + 2.0 * nq * nq);
}

// The generic (non-MFMA) far-update kernels for every (r, nq): PARAHT_HH_GENERIC=1
bool s2h_generic() {
  const char* e = getenv("PARAHT_HH_GENERIC");
  return e && *e == '1';
}

// Per sweep: the generation (leaders + helpers), the blocks' T factors, then
//  * Q and Z on the third stream beside everything (rht2ht.f90:1976-2018);
//  * the A/B lookahead on the call's stream: rows within 2 nq blocks above
//    each window, then columns within r nq of it (rht2ht.f90:1644-1765);
//  * the rest of A and B on the second stream beside the next sweep's
//    generation (rht2ht.f90:1818-1951), which touches none of its entries;
//    the next lookahead waits for it.
// A single-stream context runs the row and column updates of A and B in one
// pass each.
int stage2_hh(paraht_ctx* c, int n, double* A, int lda, double* B, int ldb, double* Q, int ldq,
              double* Z, int ldz, int r, int nq, int flags) {
  if (n <= 2) return 0;  // rht2ht.f90:1417 loop is empty
  if (!pht::hh_supported(r, nq)) return PARAHT_ERR_UNSUPPORTED;
  hipStream_t st = c->stream;
  const int ldu = r + nq;
  const int nbl = (n - 1) / r + ((n - 1) % r != 0);
  const int nbuf = c->single ? 1 : 2;
  if (int e = ensure_ws(c, s2h_ws_bytes(n, r, nq, nbuf))) return e;
  if (int e = ensure_aux(c)) return e;
  const size_t td = pht::hh_tape_doubles(r, nq, nbl);
  const pht::HhTape tape[2] = {pht::hh_tape((double*)c->ws, r, nq, nbl),
                               pht::hh_tape((double*)c->ws + (nbuf - 1) * td, r, nq, nbl)};
  int* counters = (int*)((double*)c->ws + nbuf * td);
  int* progress = counters;
  int* hprog = counters + nq;
  int* errflag = counters + (size_t)nq * (1 + pht::kS2MaxHelpers);
  const size_t zero_d = (size_t)2 * ldu * nq * nbl + (size_t)2 * nq * nbl;  // U, V, tauU, tauV
  hipStream_t sq = side2(c);
  const bool tc = (flags & PARAHT_TIME_COMPONENTS) != 0;
  // a single-stream context keeps the split order on its one stream, so its
  // results are bit-identical to the two-stream schedule's
  const int look = c->single ? 2 : s2_lookahead();
  const bool split = look != 0;
  hipStream_t rest = look == 2 ? st : c->aux;
  const bool gen_far = s2h_generic();
  Timer tm(c, tc);
  std::vector<int> marks;
  int buf = 0;
  bool far_pending = false;
  // the Q/Z stream starts after everything on the call's stream so far
  if (int e = rec(c, c->ev_tape, st)) return e;
  if (int e = wait_on(c, sq, c->ev_tape)) return e;
  for (int s = 1, k = 0; s <= n - 2; s += nq, ++k) {
    buf = k & 1;
    const pht::HhTape& T = tape[buf];
    const int nblocks = (n - s) / r + ((n - s) % r != 0);
    // helpers need every workgroup resident and nq <= r+1 (a column's steps
    // then never read the rows above its earlier bulges)
    int nhelp = (nq <= r + 1) ? std::min(s2_helpers(), c->cu_count / nq - 1) : 0;
    if (nhelp < 0) nhelp = 0;
    const int spin = (k == 0) ? s2_first_spin() : pht::kS2SpinLimit;
    const int m0 = tm.mark();
    // the buffer's previous tape (sweep k-2) must be consumed by its Q/Z update
    // (its A/B rest finished before sweep k-1's lookahead)
    if (k >= 2)
      if (int e = wait_on(c, st, c->ev_free[buf])) return e;
    PROF(PARAHT_K_OTHER, 0.0, 0.0,
         pht::hh_prepare(st, T.U, zero_d, counters, (int)((size_t)nq * (1 + pht::kS2MaxHelpers)),
                         errflag, k == 0));
    PROF(PARAHT_K_S2_GEN, s2h_gen_flops(n, r, nq, s), 0.0,
         pht::hh_generate(st, A, lda, B, ldb, n, r, nq, s, T.U, T.V, T.tauU, T.tauV, progress,
                          hprog, errflag, spin, nhelp));
    PROF(PARAHT_K_OTHER, 0.0, 0.0,
         pht::hh_factors(st, T.U, T.V, T.tauU, T.tauV, T.TU, T.TV, ldu, nq, nblocks));
    const int m1 = tm.mark();
    if (int e = rec(c, c->ev_tape, st)) return e;
    if (int e = wait_on(c, sq, c->ev_tape)) return e;
    {
      const double by = s2h_far_bytes(n, r, nq, s, nblocks, false, pht::kHhFull, 2);
      PROF_ON(sq, PARAHT_K_S2_ROWS, s2h_far_flops(by, r, nq), by,
              pht::hh_far(sq, false, Q, ldq, Z, ldz, T.U, T.TU, T.tauU, T.V, T.TV, T.tauV, 0, 0,
                          n, r, nq, s, nblocks, pht::kHhFull, gen_far));
    }
    if (int e = rec(c, c->ev_free[buf], sq)) return e;
    auto ab = [&](hipStream_t strm, bool cols, int phase) -> int {
      const double by = s2h_far_bytes(n, r, nq, s, nblocks, cols, phase, 2);
      PROF_ON(strm, cols ? PARAHT_K_S2_COLS : PARAHT_K_S2_ROWS, s2h_far_flops(by, r, nq), by,
              cols ? pht::hh_far(strm, true, A, lda, B, ldb, T.U, T.TU, T.tauU, T.U, T.TU,
                                 T.tauU, 1, 0, n, r, nq, s, nblocks, phase, gen_far)
                   : pht::hh_far(strm, false, A, lda, B, ldb, T.V, T.TV, T.tauV, T.V, T.TV,
                                 T.tauV, 0, 0, n, r, nq, s, nblocks, phase, gen_far));
      return 0;
    };
    if (split) {
      // the previous sweep's rest must be done before this lookahead
      if (far_pending)
        if (int e = wait_on(c, st, c->ev_far)) return e;
      if (int e = ab(st, false, pht::kHhLook)) return e;
      if (int e = ab(st, true, pht::kHhLook)) return e;
      if (int e = rec(c, c->ev_near, st)) return e;
      if (int e = wait_on(c, rest, c->ev_near)) return e;
      if (int e = ab(rest, false, pht::kHhFar)) return e;
      if (int e = ab(rest, true, pht::kHhFar)) return e;
      if (int e = rec(c, c->ev_far, rest)) return e;
      far_pending = true;
    } else {
      if (int e = ab(st, false, pht::kHhAll)) return e;
      if (int e = ab(st, true, pht::kHhAll)) return e;
    }
    const int m2 = tm.mark();
    if (tc) marks.insert(marks.end(), {m0, m1, m2});
  }
  // join: the last A/B rest and the last Q/Z update
  if (far_pending)
    if (int e = wait_on(c, st, c->ev_far)) return e;
  if (int e = wait_on(c, st, c->ev_free[buf])) return e;
  {
    int err = 0;
    CK(hipMemcpyAsync(&err, errflag, sizeof(int), hipMemcpyDeviceToHost, st));
    CK(hipStreamSynchronize(st));
    if (err) {
      if (verbose()) fprintf(stderr, "paraht: stage-2 sweep wait exceeded its spin limit\n");
      return PARAHT_ERR_HIP;
    }
  }
  if (tc) {
    CK(hipStreamSynchronize(st));
    double gen = 0, app = 0;
    for (size_t i = 0; i < marks.size(); i += 3) {
      gen += tm.ms(marks[i], marks[i + 1]);
      app += tm.ms(marks[i + 1], marks[i + 2]);
    }
    c->times[4] = gen;
    // the reference's line (rht2ht.f90:2045-2046, seconds): the lookahead is
    // the call stream's share of the A/B updates
    printf("      gen: %16.6f ;lookahead: %16.6f ;application: %16.6f\n", gen * 1e-3,
           split ? app * 1e-3 : 0.0, split ? 0.0 : app * 1e-3);
  }
  return 0;
}

// The stage-2 variant of a call: PARAHT_STAGE2_GIVENS selects rht2ht_givens,
// otherwise rht2ht_householder.
int s2_check_variant(int r, int nq, int flags) {
  if (flags & PARAHT_STAGE2_GIVENS) return s2_check(r, nq);
  return pht::hh_supported(r, nq) ? 0 : PARAHT_ERR_UNSUPPORTED;
}
int stage2_any(paraht_ctx* c, int n, double* A, int lda, double* B, int ldb, double* Q, int ldq,
               double* Z, int ldz, int r, int nq, int flags) {
  if (flags & PARAHT_STAGE2_GIVENS) return stage2(c, n, A, lda, B, ldb, Q, ldq, Z, ldz, r, nq, flags);
  return stage2_hh(c, n, A, lda, B, ldb, Q, ldq, Z, ldz, r, nq, flags);
}

// ------------------------------------------------------------------------
std::mutex g_mu;
paraht_ctx* g_default[64] = {nullptr};

int default_ctx(paraht_ctx** out) {
  int dev = 0, ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return PARAHT_ERR_NODEV;
  CK(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64) return PARAHT_ERR_NODEV;
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_default[dev]) {
    int e = paraht_ctx_create(&g_default[dev], dev);
    if (e) return e;
  }
  *out = g_default[dev];
  return 0;
}

struct DevMat {
  double* d = nullptr;
  ~DevMat() {
    if (d) (void)hipFree(d);
  }
};

// copy an n x n host matrix (leading dimension ld) to a dense device matrix
int upload(hipStream_t st, DevMat& m, const double* h, int n, int ld) {
  CK(hipMalloc(&m.d, std::max<size_t>(1, (size_t)n * n) * sizeof(double)));
  if (n > 0)
    CK(hipMemcpy2DAsync(m.d, (size_t)n * sizeof(double), h, (size_t)ld * sizeof(double),
                        (size_t)n * sizeof(double), n, hipMemcpyHostToDevice, st));
  return 0;
}

int download(hipStream_t st, double* h, int ld, const DevMat& m, int n) {
  if (n > 0)
    CK(hipMemcpy2DAsync(h, (size_t)ld * sizeof(double), m.d, (size_t)n * sizeof(double),
                        (size_t)n * sizeof(double), n, hipMemcpyDeviceToHost, st));
  return 0;
}

enum Which { kStage1 = 1, kStage2 = 2, kBoth = 3, kQrB = 4 };

int run_dev(paraht_ctx* c, Which w, int n, double* A, int lda, double* B,
            int ldb, double* Q, int ldq, double* Z, int ldz, int r, int p,
            int nq, int flags) {
  hipEvent_t e0 = nullptr, e1 = nullptr, e2 = nullptr;
  CK(hipSetDevice(c->device));
  CK(hipEventCreate(&e0));
  CK(hipEventCreate(&e1));
  CK(hipEventCreate(&e2));
  CK(hipEventRecord(e0, c->stream));
  int st = 0;
  if (w & kQrB) st = qr_of_b(c, n, A, lda, B, ldb, Q, ldq, r, p);
  if (!st && (w & kStage1)) st = stage1(c, n, A, lda, B, ldb, Q, ldq, Z, ldz, r, p, flags);
  if (!st) {
    hipError_t he = hipEventRecord(e1, c->stream);
    if (he != hipSuccess) st = hip_fail(he, "hipEventRecord", __LINE__);
  }
  if (!st && (w & kStage2))
    st = stage2_any(c, n, A, lda, B, ldb, Q, ldq, Z, ldz, r, nq, flags);
  if (!st) {
    hipError_t he = hipEventRecord(e2, c->stream);
    if (he == hipSuccess) he = hipStreamSynchronize(c->stream);
    if (he != hipSuccess) st = hip_fail(he, "synchronize", __LINE__);
  }
  if (!st) {
    float t1 = 0.f, t2 = 0.f;
    if (hipEventElapsedTime(&t1, e0, e1) == hipSuccess) c->times[0] = t1;
    if (hipEventElapsedTime(&t2, e1, e2) == hipSuccess) c->times[3] = t2;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipEventDestroy(e2);
  return st;
}

int run_host(Which w, int n, double* A, int lda, double* B, int ldb, double* Q,
             int ldq, double* Z, int ldz, int r, int p, int nq, int flags) {
  paraht_ctx* c = nullptr;
  if (int e = default_ctx(&c)) return e;
  if (n == 0) return 0;
  DevMat dA, dB, dQ, dZ;
  if (int e = upload(c->stream, dA, A, n, lda)) return e;
  if (int e = upload(c->stream, dB, B, n, ldb)) return e;
  if (int e = upload(c->stream, dQ, Q, n, ldq)) return e;
  // the QR of B has no Z: its callers pass Q in Z's place
  const bool has_z = Z != Q;
  if (has_z)
    if (int e = upload(c->stream, dZ, Z, n, ldz)) return e;
  double* zd = has_z ? dZ.d : dQ.d;
  int st = run_dev(c, w, n, dA.d, n, dB.d, n, dQ.d, n, zd, n, r, p, nq, flags);
  if (st) return st;
  if (int e = download(c->stream, A, lda, dA, n)) return e;
  if (int e = download(c->stream, B, ldb, dB, n)) return e;
  if (int e = download(c->stream, Q, ldq, dQ, n)) return e;
  if (has_z)
    if (int e = download(c->stream, Z, ldz, dZ, n)) return e;
  CK(hipStreamSynchronize(c->stream));
  return 0;
}

}  // namespace

// ------------------------------------------------------------------------
// Batched path
// ------------------------------------------------------------------------
namespace {

// size-bucket table of (r, p, nq) for the batched path, from
// paraht_amd.batched.autotune on MI355X over 2048 pencils with n uniform in
// [32, 512] (buckets n <= 96, 192, 320, 512: every bucket chose the widest
// band, the largest blocks and the most columns per sweep) and timings of
// 512 pencils of one order each: for n <= 64 the one-stage reduction (r = 1)
// takes 2.9 ms at n = 64 and 0.5 ms at n = 8, the best two-stage parameters
// 11.4 ms and 1.1 ms
void batch_bucket(int n, int& r, int& p, int& nq) {
  if (n <= 64) {
    r = 1;
    p = 2;
    nq = 1;
    return;
  }
  r = 16;
  p = 8;
  nq = 16;
}

int batched(paraht_ctx* c, int count, const paraht_pencil* pencils, double* dA, double* dB,
            double* dQ, double* dZ, int flags) {
  std::vector<pht::BatchPencil> q;
  q.reserve(count);
  size_t wsmax = 64;
  for (int i = 0; i < count; ++i) {
    const paraht_pencil& pc = pencils[i];
    if (pc.n < 0 || pc.offset < 0 || pc.r < 0 || pc.p < 0 || pc.nq < 0) return -3;
    if (pc.n == 0) continue;
    int r = pc.r, p = pc.p, nq = pc.nq;
    if (r == 0 || p == 0 || nq == 0) {
      int dr, dp, dn;
      batch_bucket(pc.n, dr, dp, dn);
      if (r == 0) r = dr;
      if (p == 0) p = dp;
      if (nq == 0) nq = dn;
    }
    if (!pht::batched_supported(pc.n, r, p, nq)) {
      if (verbose())
        fprintf(stderr, "paraht: batched pencil %d (n=%d r=%d p=%d nq=%d) unsupported\n", i,
                pc.n, r, p, nq);
      return PARAHT_ERR_UNSUPPORTED;
    }
    q.push_back({pc.offset, pc.n, r, p, nq});
    wsmax = std::max(wsmax, pht::batched_ws_doubles(pc.n, r, p, nq));
  }
  if (q.empty()) return 0;
  // the work queue: largest first (longest-processing-time order)
  std::stable_sort(q.begin(), q.end(), [](const pht::BatchPencil& a, const pht::BatchPencil& b) {
    return a.n > b.n;
  });
  const int nq_ = (int)q.size();
  const int grid = std::min(nq_, c->cu_count);
  const size_t pbytes = ((nq_ * sizeof(pht::BatchPencil) + 255) / 256) * 256;
  if (int e = ensure_ws(c, pbytes + 256 + (size_t)grid * wsmax * sizeof(double))) return e;
  auto* dP = (pht::BatchPencil*)c->ws;
  int* dnext = (int*)((char*)c->ws + pbytes);
  double* wsb = (double*)((char*)c->ws + pbytes + 256);
  hipStream_t st = c->stream;
  CK(hipMemcpyAsync(dP, q.data(), nq_ * sizeof(pht::BatchPencil), hipMemcpyHostToDevice, st));
  CK(hipStreamSynchronize(st));  // q is a host temporary
  if (flags & PARAHT_BATCH_GRAPH) {
    std::vector<long long> key = {nq_, (long long)(uintptr_t)dA, (long long)(uintptr_t)dB,
                                  (long long)(uintptr_t)dQ, (long long)(uintptr_t)dZ,
                                  (long long)(uintptr_t)c->ws, (long long)wsmax, grid};
    for (const auto& b : q) key.insert(key.end(), {b.offset, b.n, b.r, b.p, b.nq});
    if (!c->batch_exec || c->batch_key != key) {
      if (c->batch_exec) {
        CK(hipGraphExecDestroy(c->batch_exec));
        c->batch_exec = nullptr;
      }
      hipGraph_t g = nullptr;
      CK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
      const hipError_t le = pht::batched_launch(st, dP, nq_, dnext, dA, dB, dQ, dZ, wsb, wsmax, grid);
      const hipError_t ce = hipStreamEndCapture(st, &g);
      if (le != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return hip_fail(le, "batched capture", __LINE__);
      }
      CK(ce);
      const hipError_t ie = hipGraphInstantiate(&c->batch_exec, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      CK(ie);
      c->batch_key = key;
    }
    CK(hipGraphLaunch(c->batch_exec, st));
  } else {
    CK(pht::batched_launch(st, dP, nq_, dnext, dA, dB, dQ, dZ, wsb, wsmax, grid));
  }
  CK(hipStreamSynchronize(st));
  return 0;
}

}  // namespace

// ------------------------------------------------------------------------
// Adjoint
// ------------------------------------------------------------------------
namespace {

constexpr int kAdjGridMinN = 512;  // one pencil this large uses the whole grid

int adjoint(paraht_ctx* c, int count, const paraht_pencil* pencils, const double* H,
            const double* T, const double* Tinv, double* GQ, const double* GZ, double* XA,
            double* XB) {
  std::vector<pht::BatchPencil> q;
  q.reserve(count);
  for (int i = 0; i < count; ++i) {
    const paraht_pencil& pc = pencils[i];
    if (pc.n < 0 || pc.offset < 0) return -3;
    if (pc.n >= 2) q.push_back({pc.offset, pc.n, 0, 0, 0});
  }
  if (q.empty()) return 0;
  std::stable_sort(q.begin(), q.end(), [](const pht::BatchPencil& a, const pht::BatchPencil& b) {
    return a.n > b.n;
  });
  hipStream_t st = c->stream;
  const int nmax = q[0].n;
  if (q.size() == 1 && nmax >= kAdjGridMinN) {
    const int maxb = pht::adjoint_grid_max_blocks(c->device);
    const int grid = std::min(c->cu_count, maxb);
    if (grid <= 0) return PARAHT_ERR_UNSUPPORTED;
    if (int e = ensure_ws(c, (size_t)4 * nmax * sizeof(double) + 256)) return e;
    double* vecs = (double*)c->ws;
    int* ints = (int*)(vecs + 4 * (size_t)nmax);  // in the workspace's 256 spare bytes
    const size_t o = (size_t)q[0].offset;
    CK(pht::adjoint_grid(st, nmax, H + o, T + o, Tinv + o, GQ + o, GZ + o, XA + o, XB + o, vecs,
                         ints, grid));
    int err = 0;
    CK(hipMemcpyAsync(&err, ints + 2, sizeof(int), hipMemcpyDeviceToHost, st));
    CK(hipStreamSynchronize(st));
    if (err) {
      if (verbose()) fprintf(stderr, "paraht: adjoint grid barrier exceeded its spin limit\n");
      return PARAHT_ERR_HIP;
    }
    return 0;
  }
  if (pht::adjoint_wg_lds(nmax) > pht::adjoint_wg_max_lds()) return PARAHT_ERR_UNSUPPORTED;
  const int nq_ = (int)q.size();
  const int grid = std::min(nq_, c->cu_count);
  const size_t pbytes = ((nq_ * sizeof(pht::BatchPencil) + 255) / 256) * 256;
  if (int e = ensure_ws(c, pbytes + 256)) return e;
  auto* dP = (pht::BatchPencil*)c->ws;
  int* dnext = (int*)((char*)c->ws + pbytes);
  CK(hipMemcpyAsync(dP, q.data(), nq_ * sizeof(pht::BatchPencil), hipMemcpyHostToDevice, st));
  CK(pht::adjoint_batched(st, dP, nq_, nmax, dnext, H, T, Tinv, GQ, GZ, XA, XB, grid));
  CK(hipStreamSynchronize(st));
  return 0;
}

}  // namespace

// ------------------------------------------------------------------------
// QZ iteration
// ------------------------------------------------------------------------
namespace {

int qz_threads_for(int nmax, int count) {
  if (const char* e = getenv("PARAHT_QZ_THREADS")) {
    const int t = atoi(e);
    if (t >= 64 && t <= 1024 && t % 64 == 0) return t;
  }
  // a batch takes the single-pencil rule for its largest pencil, capped at
  // 512 threads so that two workgroups can share a CU: measured on 2048
  // pencils of order 32..512, 2.36 s at 128 threads, 2.22 s at 256, 1.92 s
  // at 384, 1.90 s at 512 and at 1024
  return count == 1 ? pht::qz_threads(nmax) : std::min(pht::qz_threads(nmax), 512);
}

// tasks in the caller's order (idx = position); runs them largest first and
// writes each pencil's status to info[idx] (host)
int qz_run(paraht_ctx* c, std::vector<pht::QzTask>& q, double* H, double* T, double* Q,
           double* Z, double* ar, double* ai, double* be, int flags, int* info) {
  const int count = (int)q.size();
  for (int i = 0; i < count; ++i) info[i] = 0;
  std::vector<pht::QzTask> run;
  run.reserve(count);
  for (const auto& t : q)
    if (t.n > 0) run.push_back(t);
  if (run.empty()) return 0;
  std::stable_sort(run.begin(), run.end(),
                   [](const pht::QzTask& a, const pht::QzTask& b) { return a.n > b.n; });
  const int nrun = (int)run.size();
  const size_t tbytes = ((nrun * sizeof(pht::QzTask) + 255) / 256) * 256;
  if (int e = ensure_ws(c, tbytes + (size_t)count * sizeof(int))) return e;
  auto* dT = (pht::QzTask*)c->ws;
  int* dinfo = (int*)((char*)c->ws + tbytes);
  hipStream_t st = c->stream;
  CK(hipMemsetAsync(dinfo, 0, (size_t)count * sizeof(int), st));
  CK(hipMemcpyAsync(dT, run.data(), nrun * sizeof(pht::QzTask), hipMemcpyHostToDevice, st));
  const int kflags = (flags & PARAHT_QZ_SCHUR) ? pht::kQzSchur : 0;
  CK(pht::qz_launch(st, dT, nrun, qz_threads_for(run[0].n, nrun), H, T,
                    (flags & PARAHT_QZ_SCHUR) ? Q : nullptr, (flags & PARAHT_QZ_SCHUR) ? Z : nullptr,
                    ar, ai, be, dinfo, kflags));
  CK(hipMemcpyAsync(info, dinfo, (size_t)count * sizeof(int), hipMemcpyDeviceToHost, st));
  CK(hipStreamSynchronize(st));  // run (a host temporary) and info
  return 0;
}

}  // namespace

// ------------------------------------------------------------------------
// Allocation-free reduction
// ------------------------------------------------------------------------
// The workspace of paraht_reduce_dev_noalloc: kScratchDoubles FP64 words
// (64 MiB) per device in the library's device data, which each GPU the
// library runs on holds until the process ends.  Host code reaches it only
// through the address hipGetSymbolAddress returns for the current device
// (the array's name is a host-side stand-in there).
constexpr size_t kScratchDoubles = (size_t)8 << 20;
static __device__ __align__(256) double scratch[kScratchDoubles];

namespace {
struct ScratchDev {
  bool ready = false;
  double* ptr = nullptr;  // the device's copy of `scratch`
  int cu_count = 0;
};
ScratchDev g_scratch[64];
std::mutex g_scratch_mu;

// Once per device, with the device current: checks the architecture and
// keeps the address of the device's copy of the scratch array.
int scratch_for_device(int dev, ScratchDev* out) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  ScratchDev& d = g_scratch[dev];
  if (!d.ready) {
    hipDeviceProp_t prop;
    CK(hipGetDeviceProperties(&prop, dev));
    if (strncmp(prop.gcnArchName, "gfx950", 6) != 0) return PARAHT_ERR_NODEV;
    void* p = nullptr;
    CK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    d.ptr = (double*)p;
    d.cu_count = prop.multiProcessorCount;
    d.ready = true;
  }
  *out = d;
  return 0;
}
}  // namespace

// ------------------------------------------------------------------------
// C-ABI
// ------------------------------------------------------------------------
extern "C" {

int paraht_ctx_create(paraht_ctx** ctx, int device) {
  if (!ctx) return -1;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return PARAHT_ERR_NODEV;
  if (device < 0 || device >= ndev) return -2;
  hipDeviceProp_t prop;
  CK(hipGetDeviceProperties(&prop, device));
  if (strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
    if (verbose()) fprintf(stderr, "paraht: device %d is %s, not gfx950\n", device, prop.gcnArchName);
    return PARAHT_ERR_NODEV;
  }
  CK(hipSetDevice(device));
  paraht_ctx* c = new paraht_ctx();
  c->device = device;
  c->cu_count = prop.multiProcessorCount;
  hipError_t e = hipStreamCreateWithFlags(&c->own, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete c;
    return hip_fail(e, "hipStreamCreate", __LINE__);
  }
  c->stream = c->own;
  *ctx = c;
  return 0;
}

int paraht_ctx_destroy(paraht_ctx* ctx) {
  if (!ctx) return -1;
  (void)hipSetDevice(ctx->device);
  if (ctx->own) (void)hipStreamSynchronize(ctx->own);
  if (ctx->aux) (void)hipStreamSynchronize(ctx->aux);
  if (ctx->aux2) (void)hipStreamSynchronize(ctx->aux2);
  if (ctx->batch_exec) (void)hipGraphExecDestroy(ctx->batch_exec);
  if (ctx->ws) (void)hipFree(ctx->ws);
  for (hipEvent_t e : ctx->pool) (void)hipEventDestroy(e);
  if (ctx->own) (void)hipStreamDestroy(ctx->own);
  if (ctx->aux) (void)hipStreamDestroy(ctx->aux);
  if (ctx->aux2) (void)hipStreamDestroy(ctx->aux2);
  if (ctx->ev_near) (void)hipEventDestroy(ctx->ev_near);
  if (ctx->ev_far) (void)hipEventDestroy(ctx->ev_far);
  if (ctx->ev_tape) (void)hipEventDestroy(ctx->ev_tape);
  if (ctx->ev_left) (void)hipEventDestroy(ctx->ev_left);
  if (ctx->ev_aleft) (void)hipEventDestroy(ctx->ev_aleft);
  for (hipEvent_t e : ctx->ev_free)
    if (e) (void)hipEventDestroy(e);
  delete ctx;
  return 0;
}

int paraht_ctx_set_stream(paraht_ctx* ctx, void* stream) {
  if (!ctx) return -1;
  ctx->stream = stream ? (hipStream_t)stream : ctx->own;
  return 0;
}

int paraht_ctx_last_times(paraht_ctx* ctx, double* ms, int nmax) {
  if (!ctx) return -1;
  if (!ms || nmax < 0) return -2;
  int k = std::min(nmax, 5);
  for (int i = 0; i < k; ++i) ms[i] = ctx->times[i];
  return k;
}

int paraht_ctx_profile(paraht_ctx* ctx, int enable) {
  if (!ctx) return -1;
  if (int e = prof_flush(ctx)) return e;
  for (int k = 0; k < PARAHT_NKERNELS; ++k) {
    ctx->k_ms[k] = ctx->k_flops[k] = ctx->k_bytes[k] = 0.0;
    ctx->k_n[k] = 0;
  }
  ctx->profiling = enable != 0;
  return 0;
}

int paraht_ctx_kernel_stats(paraht_ctx* ctx, int kernel, double* total_ms,
                            long long* launches, double* flops, double* bytes) {
  if (!ctx) return -1;
  if (kernel < 0 || kernel >= PARAHT_NKERNELS) return -2;
  if (int e = prof_flush(ctx)) return e;
  if (total_ms) *total_ms = ctx->k_ms[kernel];
  if (launches) *launches = ctx->k_n[kernel];
  if (flops) *flops = ctx->k_flops[kernel];
  if (bytes) *bytes = ctx->k_bytes[kernel];
  return 0;
}

int paraht_ab2rht(int n, double* A, int lda, double* B, int ldb, double* Q,
                  int ldq, double* Z, int ldz, int nb, int p, int flags) {
  if (int e = check_common(n, lda, ldb, ldq, ldz)) return e;
  if (nb < 1) return -10;
  if (p < 2) return -11;
  if (!pht::s1_supported(nb, p)) return PARAHT_ERR_UNSUPPORTED;
  return run_host(kStage1, n, A, lda, B, ldb, Q, ldq, Z, ldz, nb, p, 1, flags);
}

int paraht_rht2ht(int n, double* A, int lda, double* B, int ldb, double* Q,
                  int ldq, double* Z, int ldz, int r, int nq, int flags) {
  if (int e = check_common(n, lda, ldb, ldq, ldz)) return e;
  if (r < 1) return -10;
  if (nq < 1) return -11;
  if (int e = s2_check_variant(r, nq, flags)) return e;
  return run_host(kStage2, n, A, lda, B, ldb, Q, ldq, Z, ldz, r, 2, nq, flags);
}

int paraht_rht2ht_unblocked(int n, double* A, int lda, double* B, int ldb,
                            double* Q, int ldq, double* Z, int ldz, int r) {
  if (int e = check_common(n, lda, ldb, ldq, ldz)) return e;
  if (r < 1) return -10;
  if (int e = s2_check_variant(r, 1, 0)) return e;
  // one column per sweep: every reflector is applied before the next column
  return run_host(kStage2, n, A, lda, B, ldb, Q, ldq, Z, ldz, r, 2, 1, 0);
}

int paraht_reduce(int n, double* A, int lda, double* B, int ldb, double* Q,
                  int ldq, double* Z, int ldz, int r, int p, int nq, int flags) {
  if (int e = check_common(n, lda, ldb, ldq, ldz)) return e;
  if (r < 1) return -10;
  if (p < 2) return -11;
  if (nq < 1) return -12;
  if (!pht::s1_supported(r, p)) return PARAHT_ERR_UNSUPPORTED;
  if (int e = s2_check_variant(r, nq, flags)) return e;  // before stage 1 changes anything
  return run_host(kBoth, n, A, lda, B, ldb, Q, ldq, Z, ldz, r, p, nq, flags);
}

int paraht_ab2rht_dev(paraht_ctx* ctx, int n, double* dA, int lda, double* dB,
                      int ldb, double* dQ, int ldq, double* dZ, int ldz,
                      int nb, int p, int flags) {
  if (!ctx) return -1;
  if (int e = check_common(n, lda, ldb, ldq, ldz)) return e - 1;
  if (nb < 1) return -11;
  if (p < 2) return -12;
  if (!pht::s1_supported(nb, p)) return PARAHT_ERR_UNSUPPORTED;
  return run_dev(ctx, kStage1, n, dA, lda, dB, ldb, dQ, ldq, dZ, ldz, nb, p, 1, flags);
}

int paraht_rht2ht_dev(paraht_ctx* ctx, int n, double* dA, int lda, double* dB,
                      int ldb, double* dQ, int ldq, double* dZ, int ldz, int r,
                      int nq, int flags) {
  if (!ctx) return -1;
  if (int e = check_common(n, lda, ldb, ldq, ldz)) return e - 1;
  if (r < 1) return -11;
  if (nq < 1) return -12;
  if (int e = s2_check_variant(r, nq, flags)) return e;
  return run_dev(ctx, kStage2, n, dA, lda, dB, ldb, dQ, ldq, dZ, ldz, r, 2, nq, flags);
}

int paraht_reduce_dev(paraht_ctx* ctx, int n, double* dA, int lda, double* dB,
                      int ldb, double* dQ, int ldq, double* dZ, int ldz, int r,
                      int p, int nq, int flags) {
  if (!ctx) return -1;
  if (int e = check_common(n, lda, ldb, ldq, ldz)) return e - 1;
  if (r < 1) return -11;
  if (p < 2) return -12;
  if (nq < 1) return -13;
  if (!pht::s1_supported(r, p)) return PARAHT_ERR_UNSUPPORTED;
  if (int e = s2_check_variant(r, nq, flags)) return e;  // before stage 1 changes anything
  return run_dev(ctx, kBoth, n, dA, lda, dB, ldb, dQ, ldq, dZ, ldz, r, p, nq, flags);
}

int paraht_qr_b(int n, double* A, int lda, double* B, int ldb, double* Q, int ldq, int nb,
                int p) {
  if (int e = check_common(n, lda, ldb, ldq, ldq)) return e;
  if (nb < 1) return -8;
  if (p < 2) return -9;
  if (!pht::s1_supported(std::min(nb, std::max(n, 1)), p)) return PARAHT_ERR_UNSUPPORTED;
  // Z is not touched: Q stands in for it in the transfer
  return run_host(kQrB, n, A, lda, B, ldb, Q, ldq, Q, ldq, nb, p, 1, 0);
}

int paraht_qr_b_dev(paraht_ctx* ctx, int n, double* dA, int lda, double* dB, int ldb,
                    double* dQ, int ldq, int nb, int p) {
  if (!ctx) return -1;
  if (int e = check_common(n, lda, ldb, ldq, ldq)) return e - 1;
  if (nb < 1) return -9;
  if (p < 2) return -10;
  if (!pht::s1_supported(std::min(nb, std::max(n, 1)), p)) return PARAHT_ERR_UNSUPPORTED;
  return run_dev(ctx, kQrB, n, dA, lda, dB, ldb, dQ, ldq, dQ, ldq, nb, p, 1, 0);
}

int paraht_reduce_dev_noalloc(void* stream, int n, double* dA, int lda, double* dB, int ldb,
                              double* dQ, int ldq, double* dZ, int ldz, int r, int p, int nq,
                              int flags) {
  if (int e = check_common(n, lda, ldb, ldq, ldz)) return e - 1;
  if (r < 1) return -11;
  if (p < 2) return -12;
  if (nq < 1) return -13;
  if (flags & ~PARAHT_STAGE2_GIVENS) return -14;  // per-phase timing needs events
  if (!pht::s1_supported(r, p)) return PARAHT_ERR_UNSUPPORTED;
  if (int e = s2_check_variant(r, nq, flags)) return e;
  const size_t need = std::max(s1_ws_bytes(n, r, p, 1), (flags & PARAHT_STAGE2_GIVENS)
                                                            ? s2_ws_bytes(n, r, nq, 1)
                                                            : s2h_ws_bytes(n, r, nq, 1));
  if (need > kScratchDoubles * sizeof(double)) {
    if (verbose())
      fprintf(stderr, "paraht: n=%d needs %zu bytes of workspace, the allocation-free path has %zu\n",
              n, need, kScratchDoubles * sizeof(double));
    return PARAHT_ERR_UNSUPPORTED;
  }
  int dev = 0, ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return PARAHT_ERR_NODEV;
  CK(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64) return PARAHT_ERR_NODEV;
  ScratchDev sd;
  if (int e = scratch_for_device(dev, &sd)) return e;
  if (n == 0) return 0;
  // a context on the stack over the scratch array: the call's stream only,
  // no streams, events or allocations of its own
  paraht_ctx c;
  c.device = dev;
  c.cu_count = sd.cu_count;
  c.stream = (hipStream_t)stream;
  c.single = true;
  c.ws = sd.ptr;
  c.ws_bytes = kScratchDoubles * sizeof(double);
  int st = stage1(&c, n, dA, lda, dB, ldb, dQ, ldq, dZ, ldz, r, p, flags);
  if (!st) st = stage2_any(&c, n, dA, lda, dB, ldb, dQ, ldq, dZ, ldz, r, nq, flags);
  if (!st) CK(hipStreamSynchronize(c.stream));
  return st;
}

int paraht_batched_dev(paraht_ctx* ctx, int count, const paraht_pencil* pencils, double* dA,
                       double* dB, double* dQ, double* dZ, int flags) {
  if (!ctx) return -1;
  if (count < 0) return -2;
  if (count == 0) return 0;
  if (!pencils) return -3;
  if (!dA) return -4;
  if (!dB) return -5;
  if (!dQ) return -6;
  if (!dZ) return -7;
  CK(hipSetDevice(ctx->device));
  return batched(ctx, count, pencils, dA, dB, dQ, dZ, flags);
}

int paraht_adjoint_dev(paraht_ctx* ctx, int count, const paraht_pencil* pencils,
                       const double* dH, const double* dT, const double* dTinv, double* dGQ,
                       const double* dGZ, double* dXA, double* dXB) {
  if (!ctx) return -1;
  if (count < 0) return -2;
  if (count == 0) return 0;
  if (!pencils) return -3;
  if (!dH) return -4;
  if (!dT) return -5;
  if (!dTinv) return -6;
  if (!dGQ) return -7;
  if (!dGZ) return -8;
  if (!dXA) return -9;
  if (!dXB) return -10;
  CK(hipSetDevice(ctx->device));
  return adjoint(ctx, count, pencils, dH, dT, dTinv, dGQ, dGZ, dXA, dXB);
}

int paraht_hgeqz_dev(paraht_ctx* ctx, int n, double* dH, int ldh, double* dT, int ldt,
                     double* dQ, int ldq, double* dZ, int ldz, double* dalphar, double* dalphai,
                     double* dbeta, int flags, int* info) {
  if (!ctx) return -1;
  if (n < 0) return -2;
  if (!dH && n > 0) return -3;
  if (ldh < std::max(1, n)) return -4;
  if (!dT && n > 0) return -5;
  if (ldt < std::max(1, n)) return -6;
  const bool schur = (flags & PARAHT_QZ_SCHUR) != 0;
  if (schur && dQ && ldq < std::max(1, n)) return -8;
  if (schur && dZ && ldz < std::max(1, n)) return -10;
  if ((!dalphar || !dalphai || !dbeta) && n > 0) return -11;
  if (flags & ~PARAHT_QZ_SCHUR) return -14;
  if (!info) return -15;
  *info = 0;
  if (n == 0) return 0;
  CK(hipSetDevice(ctx->device));
  std::vector<pht::QzTask> q(1);
  q[0] = {0, 0, n, ldh, ldt, dQ ? ldq : n, dZ ? ldz : n, 0};
  return qz_run(ctx, q, dH, dT, dQ, dZ, dalphar, dalphai, dbeta, flags, info);
}

int paraht_hgeqz_batched_dev(paraht_ctx* ctx, int count, const paraht_pencil* pencils,
                             double* dH, double* dT, double* dQ, double* dZ, double* dalphar,
                             double* dalphai, double* dbeta, int flags, int* info) {
  if (!ctx) return -1;
  if (count < 0) return -2;
  if (count == 0) return 0;
  if (!pencils) return -3;
  if (!dH) return -4;
  if (!dT) return -5;
  if (!dalphar || !dalphai || !dbeta) return -8;
  if (flags & ~PARAHT_QZ_SCHUR) return -11;
  if (!info) return -12;
  std::vector<pht::QzTask> q(count);
  long long eoff = 0;
  for (int i = 0; i < count; ++i) {
    const paraht_pencil& pc = pencils[i];
    if (pc.n < 0 || pc.offset < 0) return -3;
    q[i] = {pc.offset, eoff, pc.n, std::max(pc.n, 1), std::max(pc.n, 1), std::max(pc.n, 1),
            std::max(pc.n, 1), i};
    eoff += pc.n;
  }
  CK(hipSetDevice(ctx->device));
  return qz_run(ctx, q, dH, dT, dQ, dZ, dalphar, dalphai, dbeta, flags, info);
}

int paraht_hgeqz(int n, double* H, int ldh, double* T, int ldt, double* Q, int ldq, double* Z,
                 int ldz, double* alphar, double* alphai, double* beta, int flags, int* info) {
  if (n < 0) return -1;
  if (!H && n > 0) return -2;
  if (ldh < std::max(1, n)) return -3;
  if (!T && n > 0) return -4;
  if (ldt < std::max(1, n)) return -5;
  const bool schur = (flags & PARAHT_QZ_SCHUR) != 0;
  if (schur && Q && ldq < std::max(1, n)) return -7;
  if (schur && Z && ldz < std::max(1, n)) return -9;
  if ((!alphar || !alphai || !beta) && n > 0) return -10;
  if (flags & ~PARAHT_QZ_SCHUR) return -13;
  if (!info) return -14;
  *info = 0;
  if (n == 0) return 0;
  paraht_ctx* c = nullptr;
  if (int e = default_ctx(&c)) return e;
  hipStream_t st = c->stream;
  const bool vq = schur && Q, vz = schur && Z;
  DevMat dH, dT, dQ, dZ, dE;
  if (int e = upload(st, dH, H, n, ldh)) return e;
  if (int e = upload(st, dT, T, n, ldt)) return e;
  if (vq)
    if (int e = upload(st, dQ, Q, n, ldq)) return e;
  if (vz)
    if (int e = upload(st, dZ, Z, n, ldz)) return e;
  CK(hipMalloc(&dE.d, 3 * (size_t)n * sizeof(double)));
  std::vector<pht::QzTask> q(1);
  q[0] = {0, 0, n, n, n, n, n, 0};
  if (int e = qz_run(c, q, dH.d, dT.d, dQ.d, dZ.d, dE.d, dE.d + n, dE.d + 2 * n, flags, info))
    return e;
  if (int e = download(st, H, ldh, dH, n)) return e;
  if (int e = download(st, T, ldt, dT, n)) return e;
  if (vq)
    if (int e = download(st, Q, ldq, dQ, n)) return e;
  if (vz)
    if (int e = download(st, Z, ldz, dZ, n)) return e;
  CK(hipMemcpyAsync(alphar, dE.d, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, st));
  CK(hipMemcpyAsync(alphai, dE.d + n, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, st));
  CK(hipMemcpyAsync(beta, dE.d + 2 * n, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, st));
  CK(hipStreamSynchronize(st));
  return 0;
}

int paraht_batched_defaults(int n, int* r, int* p, int* nq) {
  if (n < 0) return -1;
  if (!r) return -2;
  if (!p) return -3;
  if (!nq) return -4;
  batch_bucket(n, *r, *p, *nq);
  return 0;
}

int paraht_version(void) { return 100; }

const char* paraht_status_string(int status) {
  if (status == 0) return "success";
  if (status < 0) return "illegal argument";
  switch (status) {
    case PARAHT_ERR_HIP: return "HIP runtime error";
    case PARAHT_ERR_NOMEM: return "out of memory";
    case PARAHT_ERR_NODEV: return "no usable gfx950 device";
    case PARAHT_ERR_UNSUPPORTED: return "unsupported parameters";
    default: return "unknown status";
  }
}

}  // extern "C"
