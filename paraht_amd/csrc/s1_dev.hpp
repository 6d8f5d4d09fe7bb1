// Device code of the register-resident stage-1 generators for blocks of at most 128 rows and
// 64 reflectors (P = p*nb <= 128, nb <= 64), the shapes of the benchmark
// configurations (nb = 64, p = 2; nb = 16, p = 8).  Same mathematics as the
// generic kernels in stage1.hip (DGEQR2 / DGERQ2 + DORGRQ + DGELQ2 / DLARFT
// semantics, src/ab2rht.f90:105-116,290-311); the difference is data
// placement and control flow, and the right generator takes the DGELQ2
// reflectors of its orthonormal Y from an LU factorisation instead of a
// second Householder loop (the same reflectors and factors, see rq_gen_body).
//  * A workgroup of 1024 threads holds the block in registers with G = 8 or
//    16 threads per row (column): every thread keeps 16 (8) entries, element
//    j of a thread in group position q being index G*j + q, so the lanes that
//    read one LDS vector touch consecutive words.  Building a reflector (the
//    serial part of every step) costs a thread 8-16 entries, not a row.
//  * The reflector loops run over the chunk index j of the pivot (static, the
//    outer loop is unrolled) and the group position inside the chunk
//    (runtime), so every register index and every chunk bound is a
//    compile-time constant: no select chains to find the pivot entry, no
//    runtime tests to skip the chunks a reflector does not reach.
//  * A reflector goes to one of two LDS broadcast rows (one barrier per
//    reflector); the lines keep their final entries in registers, and the
//    explicit V rows (zero-padded to 128) are written once after the loop.
//  * The reflector comes from the group's sum of squares with v_rsq_f64 and
//    v_rcp_f64 refined by Newton steps (refl_from_ss); DLARFG's scaled path
//    takes over when the sum of squares leaves the safe range.
//  * G = V^T V runs on FP64 MFMA (one 16x16 tile per wave); T comes from the
//    DLARFT forward recurrence T(0:j-1,j) = -tau_j T(0:j-1,0:j-1) G(0:j-1,j),
//    each lane of one wave owning a row of T (no barriers).
#pragma once
#include "device_util.hpp"
#include "kernels.hpp"

namespace pht {
namespace s1 {
constexpr int kT = 1024;
constexpr int kLd = 129;  // LDS row length of stored reflectors (odd: spreads banks)

// x from another lane of the same DPP row (16 lanes), by DPP control CTRL
template <int CTRL>
__device__ __forceinline__ double dpp_d(double x) {
  const unsigned long long u = __builtin_bit_cast(unsigned long long, x);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)u, CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(u >> 32), CTRL, 0xF, 0xF, false);
  return __builtin_bit_cast(double, ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}

// sum over aligned groups of G = 2, 4, 8 or 16 lanes with DPP swizzles (no
// LDS round trip): quad_perm xor 1, quad_perm xor 2, then the half-row
// mirror pairs the two quads of an 8-lane half and the row mirror the two
// halves of a 16-lane row
template <int G>
__device__ __forceinline__ double grp_sum(double v) {
  static_assert(G == 2 || G == 4 || G == 8 || G == 16, "DPP group");
  v += dpp_d<0xB1>(v);                // quad_perm(1,0,3,2)
  if (G > 2) v += dpp_d<0x4E>(v);     // quad_perm(2,3,0,1)
  if (G > 4) v += dpp_d<0x141>(v);    // row_half_mirror
  if (G > 8) v += dpp_d<0x140>(v);    // row_mirror
  return v;
}
template <int G>
__device__ __forceinline__ double grp_max(double v) {
#pragma unroll
  for (int o = 1; o < G; o <<= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}

// ||x|| over a group of G lanes holding NJ values each, entries selected by
// `sel`; the scaled evaluation runs only when the plain sum of squares
// over/underflowed.
template <int G, int NJ, typename Sel>
__device__ __forceinline__ double grp_norm(const double* X, Sel sel) {
  double ss = 0.0;
#pragma unroll
  for (int j = 0; j < NJ; ++j) ss += sel(j) ? X[j] * X[j] : 0.0;
  ss = grp_sum<G>(ss);
  if (ss == 0.0 || (ss > 1e-290 && ss < 1e290)) return sqrt(ss);
  double mx = 0.0;
#pragma unroll
  for (int j = 0; j < NJ; ++j) mx = fmax(mx, sel(j) ? fabs(X[j]) : 0.0);
  mx = grp_max<G>(mx);
  if (mx == 0.0) return 0.0;
  double s2 = 0.0;
#pragma unroll
  for (int j = 0; j < NJ; ++j) {
    const double t = sel(j) ? X[j] / mx : 0.0;
    s2 += t * t;
  }
  return mx * sqrt(grp_sum<G>(s2));
}

// DLARFG from alpha and ss = ||x||^2 (already summed over the group):
// beta = -sign(alpha) sqrt(alpha^2 + ss), tau = (beta - alpha) / beta =
// 1 + |alpha| / |beta|, scal = 1 / (alpha - beta) = sign(alpha) / (|alpha| + |beta|).
// 1/|beta| is v_rsq_f64 of alpha^2 + ss refined by two Newton steps, and
// 1/(|alpha| + |beta|) v_rcp_f64 refined by two more (about 1 ulp each).
// Outside the safe range (ss tiny or huge, |alpha| huge) the scaled norm and
// make_refl's rescaling loop take over, as DLARFG does.
template <int G, int NJ, typename Sel>
__device__ __forceinline__ Refl refl_from_ss(double al, double ss, const double* X, Sel sel) {
  Refl R;
  if (ss == 0.0) {
    R.tau = 0.0;
    R.scal = 1.0;
    R.beta = al;
    return R;
  }
  if (ss > 1e-290 && ss < 1e290 && fabs(al) < 1e140) {
    const double n2 = fma(al, al, ss);
    double y = __builtin_amdgcn_rsq(n2);
    double e = fma(-n2 * y, y, 1.0);
    y = fma(0.5 * y, e, y);
    e = fma(-n2 * y, y, 1.0);
    y = fma(0.5 * y, e, y);
    const double nrm = n2 * y, aa = fabs(al);
    R.tau = fma(aa, y, 1.0);
    const double d = aa + nrm;
    double z = __builtin_amdgcn_rcp(d);
    double f = fma(-d, z, 1.0);
    z = fma(z, f, z);
    f = fma(-d, z, 1.0);
    z = fma(z, f, z);
    R.scal = copysign(z, al);
    R.beta = -copysign(nrm, al);
    return R;
  }
  return make_refl(al, grp_norm<G, NJ>(X, sel));
}

// G(x,y) = v_x . v_y (x < y < k) for reflectors stored as zero-padded rows of
// Vs (ldvs apart, 128 entries), then T (kfull x kfull, ld ldt, zeros below the
// diagonal and beyond k) of H = H(0) ... H(k-1) = I - V T V^T.  Called by all
// kT threads after a barrier that completed Vs and tau.
//
// T is built from 16 x 16 blocks: the four diagonal blocks by the DLARFT
// forward recurrence T(0:j-1,j) = -tau_j T(0:j-1,0:j-1) G(0:j-1,j) (one wave
// each, a lane owning a row in registers), then the compact-WY product rule
// for H = H_A H_B, T_AB = -T_AA G_AB T_BB, on FP64 MFMA: first within the two
// 32-column halves, then between them.  Entries beyond k are zero (their
// Gram rows and tau are masked).
__device__ __forceinline__ void gram_and_t(const double* Vs, int k, const double* tau,
                                           double* G, double* Tl, double* Tout, int ldt,
                                           int kfull, int ldvs = kLd) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int l16 = lane & 15, q = lane >> 4;
  {
    // one 16x16 tile of G per wave: rows 16*I.., columns 16*J.. (I <= J)
    const int I = wave >> 2, J = wave & 3;
    if (I <= J) {
      const bool oka = 16 * I + l16 < k, okb = 16 * J + l16 < k;
      const double* va = Vs + (16 * I + (oka ? l16 : 0)) * ldvs + q;
      const double* vb = Vs + (16 * J + (okb ? l16 : 0)) * ldvs + q;
      // the operands of eight MFMAs are loaded at once (rows past k are
      // masked after the load), and two accumulators halve the chain
      dbl4 acc = {0.0, 0.0, 0.0, 0.0}, acc1 = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int k0 = 0; k0 < 128; k0 += 32) {
        double x[8], y[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          x[u] = va[k0 + 4 * u];
          y[u] = vb[k0 + 4 * u];
        }
#pragma unroll
        for (int u = 0; u < 8; u += 2) {
          acc = mfma16(oka ? x[u] : 0.0, okb ? y[u] : 0.0, acc);
          acc1 = mfma16(oka ? x[u + 1] : 0.0, okb ? y[u + 1] : 0.0, acc1);
        }
      }
#pragma unroll
      for (int i = 0; i < 4; ++i) acc[i] += acc1[i];
#pragma unroll
      for (int i = 0; i < 4; ++i) G[(16 * I + 4 * i + q) * 65 + 16 * J + l16] = acc[i];
    }
  }
  __syncthreads();
  if (wave < 4 && lane < 16) {
    // diagonal block w: lane a holds row a of T(base.., base..)
    const int base = 16 * wave, a = lane;
    double trow[16];
#pragma unroll
    for (int l = 0; l < 16; ++l) trow[l] = (l == a && base + l < k) ? tau[base + l] : 0.0;
#pragma unroll
    for (int j = 1; j < 16; ++j) {
      double g[16];
#pragma unroll
      for (int l = 0; l < j; ++l) g[l] = G[(base + l) * 65 + base + j];
      double s0 = 0.0, s1 = 0.0;
#pragma unroll
      for (int l = 0; l < j; l += 2) {
        s0 += trow[l] * g[l];
        if (l + 1 < j) s1 += trow[l + 1] * g[l + 1];
      }
      if (j > a && base + j < k) trow[j] = -tau[base + j] * (s0 + s1);
    }
#pragma unroll
    for (int l = 0; l < 16; ++l) Tl[(base + a) * 65 + base + l] = trow[l];
  }
  __syncthreads();
  if (wave < 2) {
    // T(b0, b1) = -T(b0, b0) [G(b0, b1) T(b1, b1)] for the blocks of a half
    const int b0 = 32 * wave, b1 = b0 + 16;
    dbl4 w = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int i = 0; i < 4; ++i)
      w = mfma16(G[(b0 + l16) * 65 + b1 + 4 * i + q], Tl[(b1 + 4 * i + q) * 65 + b1 + l16], w);
    dbl4 t = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int i = 0; i < 4; ++i) t = mfma16(Tl[(b0 + l16) * 65 + b0 + 4 * i + q], w[i], t);
#pragma unroll
    for (int i = 0; i < 4; ++i) Tl[(b0 + 4 * i + q) * 65 + b1 + l16] = -t[i];
  }
  __syncthreads();
  // T(0:32, 32:64) = -T(0:32, 0:32) [G(0:32, 32:64) T(32:64, 32:64)]; the
  // bracket's block (I, J) goes to G's unused block (J, I)
  const int I = wave >> 1, J = 2 + (wave & 1);
  if (wave < 4) {
    dbl4 w = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int K = 2; K < 4; ++K) {
      if (K > J) continue;
#pragma unroll
      for (int i = 0; i < 4; ++i)
        w = mfma16(G[(16 * I + l16) * 65 + 16 * K + 4 * i + q],
                   Tl[(16 * K + 4 * i + q) * 65 + 16 * J + l16], w);
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) G[(16 * J + 4 * i + q) * 65 + 16 * I + l16] = w[i];
  }
  __syncthreads();
  if (wave < 4) {
    dbl4 t = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int K = 0; K < 2; ++K) {
      if (K < I) continue;
#pragma unroll
      for (int i = 0; i < 4; ++i)
        t = mfma16(Tl[(16 * I + l16) * 65 + 16 * K + 4 * i + q],
                   G[(16 * J + 4 * i + q) * 65 + 16 * K + l16], t);
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) Tl[(16 * I + 4 * i + q) * 65 + 16 * J + l16] = -t[i];
  }
  __syncthreads();
  for (int e = threadIdx.x; e < kfull * kfull; e += kT) {
    const int r = e % kfull, c = e / kfull;
    Tout[PHT_IDX(r, c, ldt)] = (r <= c && c < k) ? Tl[r * 65 + c] : 0.0;
  }
}
// ---------------------------------------------------------------------------
// Left generation of one panel (ab2rht.f90:93-118), blocks bottom-up.
// Thread t owns column t/16 of the block, rows 16*j + t%16 (j < 8).
// With tau_out, each block's tau goes there (64 per block) and T is left to
// panel_t_body, one workgroup per block after the chain; otherwise T is
// built here, block by block.
__device__ __forceinline__ void panel_qr_body(
    double* sm, double* __restrict__ A, int lda, int n, int jb, int nb, int p, int nblocks,
    double* __restrict__ Vt, int ldv, long long vstride, double* __restrict__ Tt,
    int ldt, long long tstride, double* __restrict__ tau_out = nullptr) {
  constexpr int NJ = 8;
  double* Vc = sm;               // [64][kLd]: reflector i in row i
  double* G = Vc + 64 * kLd;     // [64][65]
  double* Tl = G + 64 * 65;      // [64][65]
  double* tau = Tl + 64 * 65;    // [64]
  double* bc = tau + 64;         // [2][128]: broadcast rows of the reflectors
  const int tid = threadIdx.x;
  const int col = tid >> 4, q = tid & 15;
  const int P = p * nb;

  for (int bl = 0; bl < nblocks; ++bl) {
    const int i0 = jb + nb + (nblocks - 1 - bl) * (p - 1) * nb;
    const int len = min(n - i0, P);
    const int k = min(nb, len);
    double X[NJ];
#pragma unroll
    for (int j = 0; j < NJ; ++j) {
      const int r = 16 * j + q;
      X[j] = (col < nb && r < len) ? A[PHT_IDX(i0 + r, jb + col, lda)] : 0.0;
    }
    // DGEQR2: reflector i = 16 jj + qp from column i (16 lanes), applied to
    // the columns right of it; it is zero above row i, so only chunks j >= jj
    // take part.
    double mysc = 1.0, mybeta = 0.0;
#pragma unroll
    for (int jj = 0; jj < NJ; ++jj) {
      for (int qp = 0; qp < 16; ++qp) {
        const int i = 16 * jj + qp;
        if (i >= k) break;
        double* vb = bc + (i & 1) * 128;
        if (col == i) {
          const double al = grp_sum<16>(q == qp ? X[jj] : 0.0);
          double s0 = (q > qp) ? X[jj] * X[jj] : 0.0, s1 = 0.0;
#pragma unroll
          for (int j = jj + 1; j < NJ; j += 2) {
            s0 += X[j] * X[j];
            if (j + 1 < NJ) s1 += X[j + 1] * X[j + 1];
          }
          const double ss = grp_sum<16>(s0 + s1);
          const Refl R = refl_from_ss<16, NJ>(al, ss, X, [&](int j) { return 16 * j + q > i; });
          const double sc = (R.tau != 0.0) ? R.scal : 1.0;
          mysc = sc;
          mybeta = R.beta;
          vb[16 * jj + q] = (q > qp) ? X[jj] * sc : (q == qp ? 1.0 : 0.0);
#pragma unroll
          for (int j = jj + 1; j < NJ; ++j) vb[16 * j + q] = X[j] * sc;
          if (q == 0) tau[i] = R.tau;
        }
        __syncthreads();
        const double t = tau[i];
        if (col > i && col < nb && t != 0.0) {
          double v[NJ];
#pragma unroll
          for (int j = jj; j < NJ; ++j) v[j] = vb[16 * j + q];
          double w0 = 0.0, w1 = 0.0;
#pragma unroll
          for (int j = jj; j < NJ; j += 2) {
            w0 += X[j] * v[j];
            if (j + 1 < NJ) w1 += X[j + 1] * v[j + 1];
          }
          const double w = grp_sum<16>(w0 + w1) * t;
#pragma unroll
          for (int j = jj; j < NJ; ++j) X[j] -= w * v[j];
        }
      }
    }
    // R (upper trapezoid) and zeros back into A; the explicit reflector rows
    // (zero above the diagonal, 1 on it) into Vc
    if (col < nb) {
#pragma unroll
      for (int j = 0; j < NJ; ++j) {
        const int r = 16 * j + q;
        if (r < len)
          A[PHT_IDX(i0 + r, jb + col, lda)] =
              (r < col) ? X[j] : (r == col ? (col < k ? mybeta : X[j]) : 0.0);
      }
    }
    if (col < k) {
#pragma unroll
      for (int j = 0; j < NJ; ++j) {
        const int r = 16 * j + q;
        Vc[col * kLd + r] = (r > col) ? X[j] * mysc : (r == col ? 1.0 : 0.0);
      }
    }
    __syncthreads();
    // explicit V to the tape (coalesced along the rows)
    double* Vb = Vt + (long long)bl * vstride;
    for (int e = tid; e < len * nb; e += kT) {
      const int r = e % len, c = e / len;
      Vb[PHT_IDX(r, c, ldv)] = (c < k) ? Vc[c * kLd + r] : 0.0;
    }
    if (tau_out != nullptr) {
      if (tid < k) tau_out[(size_t)bl * 64 + tid] = tau[tid];
      // the next block (above) re-reads the R rows just written to A by
      // this workgroup (the barrier orders them), and overwrites Vc / tau
      __syncthreads();
    } else {
      gram_and_t(Vc, k, tau, G, Tl, Tt + (long long)bl * tstride, ldt, nb);
      __threadfence();
      __syncthreads();
    }
  }
}

// T of block bl = blockIdx.x of a panel from its stored V and tau (the
// deferred part of panel_qr_body): V back into zero-padded LDS rows, then
// gram_and_t.  The blocks are independent, so a panel's T matrices are
// built in one launch of nblocks workgroups.
__device__ inline void panel_t_body(double* sm, int n, int jb, int nb, int p, int nblocks,
                                    const double* __restrict__ Vt, int ldv, long long vstride,
                                    const double* __restrict__ tau_in,
                                    double* __restrict__ Tt, int ldt, long long tstride) {
  double* Vc = sm;               // [64][kLd]
  double* G = Vc + 64 * kLd;     // [64][65]
  double* Tl = G + 64 * 65;      // [64][65]
  double* tau = Tl + 64 * 65;    // [64]
  const int tid = threadIdx.x, bl = blockIdx.x;
  const int P = p * nb;
  const int i0 = jb + nb + (nblocks - 1 - bl) * (p - 1) * nb;
  const int len = min(n - i0, P);
  const int k = min(nb, len);
  const double* Vb = Vt + (long long)bl * vstride;
  for (int e = tid; e < 64 * 128; e += kT) {
    const int c = e >> 7, r = e & 127;
    Vc[c * kLd + r] = (c < k && r < len) ? Vb[PHT_IDX(r, c, ldv)] : 0.0;
  }
  if (tid < 64) tau[tid] = (tid < k) ? tau_in[(size_t)bl * 64 + tid] : 0.0;
  __syncthreads();
  gram_and_t(Vc, k, tau, G, Tl, Tt + (long long)bl * tstride, ldt, nb);
}

// ---------------------------------------------------------------------------
// Right generator for one B block of len <= 128 rows (ab2rht.f90:290-311).
// len > nb: RQ, the first nb rows of its Q, LQ of those (nb reflectors).
// len <= nb: the len RQ reflectors themselves, applied as H(len-1)...H(0)
// (the 'B','R' branch, ab2rht.f90:305-309).
// RQ phase: thread t owns row t/8 of the block, columns 8*j + t%8 (j < 16).
// Y / LQ phase: thread t owns row t/16 of Y = Q(0:nb-1,:), columns 16*j + t%16.
__device__ __forceinline__ void rq_gen_body(
    double* sm, const double* __restrict__ B, int ldb, int i0, int len, int nb,
    double* __restrict__ Vout, int ldv, double* __restrict__ Tout, int ldt) {
  double* Vr = sm;                  // [128][kLd]: RQ reflector i in row i
  double* taus = Vr + 128 * kLd;    // [128]
  double* tau2 = taus + 128;        // [64]
  double* Ys = Vr;                  // [64][kLd]: LQ reflectors, after the Y phase
  double* G = Vr + 64 * kLd;        // [64][65]
  double* Tl = G + 64 * 65;         // [64][65] (ends inside taus, dead by then)
  double* Tb = tau2 + 64;           // [8][16][17]: T of each block of 16 RQ reflectors
  const int tid = threadIdx.x;

  // ---- DGERQ2 on the len x len block: reflector i = 8 jj + hp from row i,
  // zero right of column i, applied to the rows above it; only chunks
  // j <= jj take part.  The broadcast rows live in Tb (free until the Y phase).
  {
    constexpr int NJ = 16;
    const int row = tid >> 3, h = tid & 7;
    double* bc = Tb;
    double X[NJ];
#pragma unroll
    for (int j = 0; j < NJ; ++j) {
      const int c = 8 * j + h;
      X[j] = (row < len && c < len) ? B[PHT_IDX(i0 + row, i0 + c, ldb)] : 0.0;
    }
    double mysc = 1.0;
#pragma unroll
    for (int jj = NJ - 1; jj >= 0; --jj) {
      for (int hp = 7; hp >= 0; --hp) {
        const int i = 8 * jj + hp;
        if (i >= len) continue;
        double* vb = bc + (i & 1) * 128;
        if (row == i) {
          const double al = grp_sum<8>(h == hp ? X[jj] : 0.0);
          double s0 = (h < hp) ? X[jj] * X[jj] : 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
#pragma unroll
          for (int j = 0; j < jj; j += 4) {
            s0 += X[j] * X[j];
            if (j + 1 < jj) s1 += X[j + 1] * X[j + 1];
            if (j + 2 < jj) s2 += X[j + 2] * X[j + 2];
            if (j + 3 < jj) s3 += X[j + 3] * X[j + 3];
          }
          const double ss = grp_sum<8>((s0 + s1) + (s2 + s3));
          const Refl R = refl_from_ss<8, NJ>(al, ss, X, [&](int j) { return 8 * j + h < i; });
          const double sc = (R.tau != 0.0) ? R.scal : 1.0;
          mysc = sc;
#pragma unroll
          for (int j = 0; j < jj; ++j) vb[8 * j + h] = X[j] * sc;
          vb[8 * jj + h] = (h < hp) ? X[jj] * sc : (h == hp ? 1.0 : 0.0);
          if (h == 0) taus[i] = R.tau;
        }
        __syncthreads();
        const double t = taus[i];
        if (row < i && t != 0.0) {
          double v[NJ];
#pragma unroll
          for (int j = 0; j <= jj; ++j) v[j] = vb[8 * j + h];
          double w0 = 0.0, w1 = 0.0, w2 = 0.0, w3 = 0.0;
#pragma unroll
          for (int j = 0; j <= jj; j += 4) {
            w0 += X[j] * v[j];
            if (j + 1 <= jj) w1 += X[j + 1] * v[j + 1];
            if (j + 2 <= jj) w2 += X[j + 2] * v[j + 2];
            if (j + 3 <= jj) w3 += X[j + 3] * v[j + 3];
          }
          const double w = grp_sum<8>((w0 + w1) + (w2 + w3)) * t;
#pragma unroll
          for (int j = 0; j <= jj; ++j) X[j] -= w * v[j];
        }
      }
    }
    // reflector i in row i of Vr: zero right of column i, 1 on it
#pragma unroll
    for (int j = 0; j < NJ; ++j) {
      const int c = 8 * j + h;
      Vr[row * kLd + c] = (row >= len) ? 0.0 : (c < row) ? X[j] * mysc : (c == row ? 1.0 : 0.0);
    }
  }
  __syncthreads();

  if (len <= nb) {
    // B Q^T = B H(len-1) ... H(0): the reflectors in reverse order form a
    // forward product; they go to rows 64.. of the buffer (free here)
    double* Vb = Vr + 64 * kLd;
    for (int e = tid; e < len * 128; e += kT) {
      const int j = e / 128, c = e % 128;
      Vb[j * kLd + c] = Vr[(len - 1 - j) * kLd + c];
      if (c == 0) tau2[j] = taus[len - 1 - j];
    }
    __syncthreads();
    for (int e = tid; e < len * nb; e += kT) {
      const int c = e % len, j = e / len;
      Vout[PHT_IDX(c, j, ldv)] = (j < len) ? Vb[j * kLd + c] : 0.0;
    }
    // G and T go to rows 0..64 of the buffer (dead once copied); T's last
    // row reaches the first reflector row only after the Gram phase read it
    double* G2 = Vr;
    gram_and_t(Vb, len, tau2, G2, G2 + 64 * 65, Tout, ldt, nb);
    return;
  }

  // ---- Y = E^T H(0) H(1) ... H(len-1): the first nb rows of the RQ's Q, in
  // blocks of 16 reflectors.  Block m = H(16m) ... H(16m+15) = I - V^T T V
  // (forward T); Y^T <- (I - V^T T^T V) Y^T for m = 0, 1, ...  First one wave
  // per block builds T from the block's Gram matrix (on MFMA, into the block's
  // LDS slot, which then receives T); then one wave per 16 columns of Y^T,
  // held as 8 MFMA accumulator tiles (element i of tile ct: row 16 ct + 4 i
  // + lane/16, column lane%16), applies the blocks in order.
  {
    const int wave = tid >> 6, lane = tid & 63, l16 = lane & 15, qq = lane >> 4;
    const int nblk = (len + 15) >> 4;
    if (wave < nblk) {
      const int b = 16 * wave;
      double* slot = Tb + wave * 272;
      const bool ok = b + l16 < len;
      const double* va = Vr + (ok ? b + l16 : 0) * kLd + qq;
      dbl4 acc = {0.0, 0.0, 0.0, 0.0}, acc1 = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int k0 = 0; k0 < 128; k0 += 32) {
        double x[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) x[u] = va[k0 + 4 * u];
#pragma unroll
        for (int u = 0; u < 8; u += 2) {
          const double x0 = ok ? x[u] : 0.0, x1 = ok ? x[u + 1] : 0.0;
          acc = mfma16(x0, x0, acc);
          acc1 = mfma16(x1, x1, acc1);
        }
      }
#pragma unroll
      for (int i = 0; i < 4; ++i) acc[i] += acc1[i];
#pragma unroll
      for (int i = 0; i < 4; ++i) slot[(4 * i + qq) * 17 + l16] = acc[i];
      __builtin_amdgcn_wave_barrier();
      if (lane < 16) {
        const int a0 = lane;
        double trow[16];
#pragma unroll
        for (int l = 0; l < 16; ++l) trow[l] = (l == a0 && b + l < len) ? taus[b + l] : 0.0;
#pragma unroll
        for (int j = 1; j < 16; ++j) {
          double g[16];
#pragma unroll
          for (int l = 0; l < j; ++l) g[l] = slot[l * 17 + j];
          double s0 = 0.0, s1 = 0.0;
#pragma unroll
          for (int l = 0; l < j; l += 2) {
            s0 += trow[l] * g[l];
            if (l + 1 < j) s1 += trow[l + 1] * g[l + 1];
          }
          if (j > a0 && b + j < len) trow[j] = -taus[b + j] * (s0 + s1);
        }
        // every lane's Gram reads precede these writes in the wave's order
#pragma unroll
        for (int l = 0; l < 16; ++l) slot[a0 * 17 + l] = trow[l];
      }
    }
    __syncthreads();
    dbl4 Yt[8];
    if (wave < ((nb + 15) >> 4)) {
      const int col = 16 * wave + l16;
#pragma unroll
      for (int ct = 0; ct < 8; ++ct)
#pragma unroll
        for (int i = 0; i < 4; ++i)
          Yt[ct][i] = (16 * ct + 4 * i + qq == col && col < nb) ? 1.0 : 0.0;
      for (int m = 0; m < nblk; ++m) {
        const int b = 16 * m;  // the block's entries lie in columns [0, b+16)
        const double* slot = Tb + m * 272;
        const bool vok = b + l16 < len;
        const double* va = Vr + (vok ? b + l16 : 0) * kLd;
        dbl4 wt = {0.0, 0.0, 0.0, 0.0}, wt1 = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int ct = 0; ct < 8; ++ct) {
          if (ct > m) continue;
          double av[4];
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            const double x = va[16 * ct + 4 * i + qq];
            av[i] = vok ? x : 0.0;
          }
          wt = mfma16(av[0], Yt[ct][0], wt);
          wt1 = mfma16(av[1], Yt[ct][1], wt1);
          wt = mfma16(av[2], Yt[ct][2], wt);
          wt1 = mfma16(av[3], Yt[ct][3], wt1);
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) wt[i] += wt1[i];
        double tv[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) tv[i] = slot[(4 * i + qq) * 17 + l16];
        dbl4 w2 = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int i = 0; i < 4; ++i) w2 = mfma16(tv[i], wt[i], w2);
#pragma unroll
        for (int ct = 0; ct < 8; ++ct) {
          if (ct > m) continue;
          double av[4];
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            const int kk = 4 * i + qq;  // row b + kk < 128: in the buffer
            const double x = Vr[(b + kk) * kLd + 16 * ct + l16];
            av[i] = (b + kk < len) ? -x : 0.0;
          }
#pragma unroll
          for (int i = 0; i < 4; ++i) Yt[ct] = mfma16(av[i], w2[i], Yt[ct]);
        }
      }
    }
    __syncthreads();  // Ys overwrites Vr from here on
    if (wave < ((nb + 15) >> 4)) {
      double* yrow = Ys + (16 * wave + l16) * kLd;
#pragma unroll
      for (int ct = 0; ct < 8; ++ct)
#pragma unroll
        for (int i = 0; i < 4; ++i) yrow[16 * ct + 4 * i + qq] = Yt[ct][i];
    }
    __syncthreads();
  }
  // ---- The reflectors of DGELQ2 on Y (nb x len, orthonormal rows), from
  // the LU factorisation without pivoting of W = [S; 0] - Y^T, S diagonal
  // +-1 (Householder reconstruction, Ballard, Demmel, Grigori, Jacquelin,
  // Nguyen and Solomonik, "Reconstructing Householder vectors from
  // tall-skinny QR", IPDPS 2014): S_jj is the sign of W's current diagonal
  // entry, so every pivot has magnitude 1 + |W_jj| >= 1; then V is L (unit
  // lower trapezoidal) and tau_j = S_jj U_jj, which are DLARFG's reflectors
  // and factors for the same signs.  Step j needs no reduction: the owners of
  // W's column j and row j publish them in LDS (two alternating slots), and
  // after one barrier every lane forms its multiplier and updates its
  // entries.  Lane layout: W row i (a column of Y) on 8 lanes, entry k = 8m+h.
  {
    constexpr int NM = 8;
    const int i = tid >> 3, h = tid & 7;
    const bool live = i < len;
    double Wr[NM];
#pragma unroll
    for (int m = 0; m < NM; ++m) {
      const int k = 8 * m + h;
      Wr[m] = (live && k < nb) ? -Ys[k * kLd + i] : 0.0;
    }
    double* colb = Tb;        // [2][128]: W(:, j)
    double* rowb = Tb + 256;  // [2][64]: W(j, :)
#pragma unroll
    for (int m = 0; m < NM; ++m) {
      if (8 * m >= nb) break;
      for (int hp = 0; hp < 8; ++hp) {
        const int j = 8 * m + hp;
        if (j >= nb) break;
        double* cb = colb + (j & 1) * 128;
        double* rb = rowb + (j & 1) * 64;
        if (h == hp && live && i >= j) cb[i] = Wr[m];
        if (i == j) {
#pragma unroll
          for (int mm = m; mm < NM; ++mm) rb[8 * mm + h] = Wr[mm];
        }
        __syncthreads();
        // every read of the step is issued at once (all indices are in range)
        const double c = cb[j], ci = cb[i];
        double rv[NM];
#pragma unroll
        for (int mm = m; mm < NM; ++mm) rv[mm] = rb[8 * mm + h];
        const double sj = (c >= 0.0) ? 1.0 : -1.0;
        const double pv = c + sj;  // |pv| in [1, 2]: v_rcp_f64 and two Newton steps
        double inv = __builtin_amdgcn_rcp(pv);
        double e = fma(-pv, inv, 1.0);
        inv = fma(inv, e, inv);
        e = fma(-pv, inv, 1.0);
        inv = fma(inv, e, inv);
        if (live && i > j) {
          const double l = ci * inv;
#pragma unroll
          for (int mm = m; mm < NM; ++mm) {
            const int k = 8 * mm + h;
            Wr[mm] = (k > j && k < nb) ? fma(-l, rv[mm], Wr[mm]) : Wr[mm];
          }
          if (h == hp) Wr[m] = l;
        }
        if (i == j && h == hp) tau2[j] = pv * sj;
      }
    }
    __syncthreads();  // the slots are read; Ys is free
    // reflector k in row k of Ys: zero left of column k, 1 on it
#pragma unroll
    for (int m = 0; m < NM; ++m) {
      const int k = 8 * m + h;
      if (k < nb) Ys[k * kLd + i] = (live && i > k) ? Wr[m] : (i == k ? 1.0 : 0.0);
    }
  }
  __syncthreads();
  // explicit V (len x nb), coalesced along the rows
  for (int e = tid; e < len * nb; e += kT) {
    const int c = e % len, r = e / len;
    Vout[PHT_IDX(c, r, ldv)] = Ys[r * kLd + c];
  }
  gram_and_t(Ys, nb, tau2, G, Tl, Tout, ldt, nb);
}


constexpr size_t kPanelLds = (size_t)(64 * kLd + 2 * 64 * 65 + 64 + 256) * sizeof(double);
constexpr size_t kRqLds = (size_t)(128 * kLd + 128 + 64 + 8 * 272) * sizeof(double);
static_assert(kPanelLds <= 160 * 1024 && kRqLds <= 160 * 1024, "LDS budget");
static_assert(kPanelLds <= kRqLds, "the batched kernel sizes its LDS for both bodies by kRqLds");
static_assert(8 * 272 >= 256, "two broadcast rows fit in the Y phase's T slots");
static_assert(8 * 272 >= 2 * 128 + 2 * 64, "the LU's column and row slots fit there too");
static_assert(64 * kLd + 2 * 64 * 65 <= 128 * kLd + 128, "Gram/T alias in the RQ kernel");

}  // namespace s1
}  // namespace pht
