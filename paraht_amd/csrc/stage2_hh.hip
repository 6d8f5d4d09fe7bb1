// Stage-2 kernels of the Householder variant, rht2ht_householder
// (src/rht2ht.f90:1346-2051): r-Hessenberg-triangular -> Hessenberg-triangular
// with one Householder reflector per column and bulge block on each side.
//
// A sweep reduces the nq columns s..s+nq-1 (1-based, `jblock` in the
// reference).  Step (ic, ib) of the sweep (rht2ht.f90:1450-1596) works on the
// bulge block ib of column ic.  In this file's notation (1-based rows)
//   g0   = s + (ib-1) r     the block's local index loc (1-based) is row g0+loc
//   row0 = g0 + ic          first row/column of the step's r x r B block
//   col  = row0-1 (ib = 1) or row0-r (ib > 1)   the column of A being reduced
//   nb   = min(r, n - row0 + 1)
// The step first applies the reflectors of the earlier columns of the block
// (which the reference defers, rht2ht.f90:1455-1464, 1537-1546) to A's column
// col and to B's column row0+r-1, then generates the left reflector from A's
// column, applies it to the B block, finds the first row of the orthogonal
// factor of the block's RQ factorization (DGERQ2 + DORGR2), generates the right
// reflector from that row, and applies it to the A window (rows row0 ..
// row0+r+nb-1), the B block and the rows istart..row0-1 of A and B above them.
// Reflector ic of block ib is column ic of U(:, :, ib) / V(:, :, ib)
// ((r+nq) x nq, local rows ic..ic+nb-1, unit first entry), as in the
// reference's tape (rht2ht.f90:1389-1394).
//
// The rest of each reflector's action -- rows of A and B above the window,
// columns right of it, and all of Q and Z -- is applied after the sweep from
// the blocks' compact-WY factors (rht2ht.f90:1599-2024) by k_hh_far on FP64
// MFMA tiles.
#include "s1_dev.hpp"
#include "s2_dev.hpp"

namespace pht {

using s2::imax;
using s2::imin;

namespace {
constexpr int kHT = 256;                 // threads of a sweep workgroup
constexpr int kHhDone = 1 << 29;
constexpr int kHhMaxLdu = 96;            // r + nq
constexpr int kHhMaxNq = 32;

// sum over the wave: DPP within each 16-lane row, then the four row sums
__device__ __forceinline__ double wsum(double v) {
  v = s1::grp_sum<16>(v);
  return (readlane_d(v, 0) + readlane_d(v, 16)) + (readlane_d(v, 32) + readlane_d(v, 48));
}
__device__ __forceinline__ double wmax(double v) {
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}

// DLARFG of (alpha, x) held one entry per lane (x0 for lanes where in0, x1
// for lanes where in1); returns tau, scal (v = x * scal) and beta.
// beta = -sign(alpha) sqrt(alpha^2 + ss), tau = 1 + |alpha| / |beta|,
// scal = sign(alpha) / (|alpha| + |beta|), with 1/|beta| from v_rsq_f64 and
// 1/(|alpha| + |beta|) from v_rcp_f64, each refined by two Newton steps;
// outside the safe range DLARFG's scaled norm and rescaling loop take over.
__device__ __forceinline__ Refl wave_refl(double alpha, double x0, bool in0, double x1, bool in1) {
  const double a0 = in0 ? x0 : 0.0, a1 = in1 ? x1 : 0.0;
  const double ss = wsum(fma(a0, a0, a1 * a1));
  Refl R;
  if (ss == 0.0) {
    R.tau = 0.0;
    R.scal = 1.0;
    R.beta = alpha;
    return R;
  }
  if (ss > 1e-290 && ss < 1e290 && fabs(alpha) < 1e140) {
    const double n2 = fma(alpha, alpha, ss);
    double y = __builtin_amdgcn_rsq(n2);
    double e = fma(-n2 * y, y, 1.0);
    y = fma(0.5 * y, e, y);
    e = fma(-n2 * y, y, 1.0);
    y = fma(0.5 * y, e, y);
    const double nrm = n2 * y, aa = fabs(alpha);
    R.tau = fma(aa, y, 1.0);
    const double d = aa + nrm;
    double z = __builtin_amdgcn_rcp(d);
    double f = fma(-d, z, 1.0);
    z = fma(z, f, z);
    f = fma(-d, z, 1.0);
    z = fma(z, f, z);
    R.scal = copysign(z, alpha);
    R.beta = -copysign(nrm, alpha);
    return R;
  }
  const double mx = wmax(fmax(fabs(a0), fabs(a1)));
  const double b0 = a0 / mx, b1 = a1 / mx;
  return make_refl(alpha, mx * sqrt(wsum(fma(b0, b0, b1 * b1))));
}

}  // namespace

struct HhSweepArgs {
  double* A;
  double* B;
  int lda, ldb, n, r, nq, s, ldu;
  double *U, *V, *tauU, *tauV;  // the sweep's tape
  int* progress;                // nq leader step counters
  int* hprog;                   // nq x kS2MaxHelpers helper step counters
  int* errflag;
  int spin, nhelp;
  int nbl;  // blocks of the tape; a step past them has no earlier reflectors
};

// LDS of a sweep workgroup (doubles)
template <int RMAX>
struct HhLds {
  static constexpr int LD = RMAX + 1;                  // B block and RQ reflectors
  static constexpr int LDA = 2 * RMAX + 1;             // A window
  static constexpr int kBl = 0;                        // B block, column c at c*LD
  static constexpr int kVr = kBl + RMAX * LD;          // RQ reflector of row i at i*LD;
                                                       // before the RQ: earlier columns' U
  static constexpr int kAl = kVr + RMAX * LD;          // A window, column c at c*LDA
  static constexpr int kVb = kAl + RMAX * LDA;         // [2][RMAX] reflector broadcast
  static constexpr int kTauR = kVb + 2 * RMAX;         // RQ taus
  static constexpr int kTb = kTauR + RMAX;             // [2] broadcast taus (+pad)
  static constexpr int kU = kTb + 4;                   // left reflector u
  static constexpr int kV = kU + RMAX;                 // right reflector v
  static constexpr int kPr = kV + RMAX;                // the RQ pivot row (64 entries)
  static constexpr int kTauU = kPr + 64;               // earlier columns' tau
  static constexpr int kSc = kTauU + kHhMaxNq;         // scalars
  // earlier columns' U ((r+nq) x (nq-1), before the RQ): in the RQ reflector
  // area where it fits, else after the scalars
  static constexpr int kUcLen = (RMAX + kHhMaxNq) * kHhMaxNq;
  static constexpr bool kUcShared = RMAX * LD >= kUcLen;
  static constexpr int kUc = kUcShared ? kVr : kSc + 8;
  static constexpr int kSize = kSc + 8 + (kUcShared ? 0 : kUcLen);
};

// rows istart..row0-1 (1-based) of A and B, columns row0..row0+nb-1, times
// (I - tau v v^T); rows split over `nthr` threads starting at thread `first`
template <int RMAX>
__device__ __noinline__ void hh_rows_above(const HhSweepArgs& a, int istart, int row0, int nb,
                                              const double* v, double tau, int first, int nthr) {
  const int nr = row0 - istart;
  for (int e = first; e < 2 * nr; e += nthr) {
    const bool inA = e < nr;
    s2::gdouble* M = s2::as_global(inA ? a.A : a.B);
    const int ld = inA ? a.lda : a.ldb;
    const size_t base = (size_t)(istart - 1 + (inA ? e : e - nr)) + (size_t)(row0 - 1) * ld;
    double x[RMAX];
#pragma unroll
    for (int c = 0; c < RMAX; ++c) x[c] = (c < nb) ? M[base + (size_t)c * ld] : 0.0;
    double d = 0.0;
#pragma unroll
    for (int c = 0; c < RMAX; ++c) d = fma(x[c], v[c], d);
    d *= tau;
#pragma unroll
    for (int c = 0; c < RMAX; ++c)
      if (c < nb) M[base + (size_t)c * ld] = fma(-d, v[c], x[c]);
  }
}

// One leader step (ic, ib); `rows_above`: also the rows above the window
// (when no helper workgroups run).
template <int RMAX>
__device__ void hh_step(const HhSweepArgs& a, int ic, int ib, double* L, bool rows_above) {
  using Lay = HhLds<RMAX>;
  constexpr int LD = Lay::LD, LDA = Lay::LDA;
  constexpr int E = RMAX / 4;   // RQ: columns 4k+q of row t>>2
  constexpr int AW = RMAX / 2;  // A window: columns 2k+h of row t>>1
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int n = a.n, r = a.r, nq = a.nq, ldu = a.ldu;
  const int g0 = a.s + (ib - 1) * r;
  const int row0 = g0 + ic;
  const int col = (ib == 1) ? row0 - 1 : row0 - r;
  const int nb = imin(r, n - row0 + 1);
  const int bcol = row0 + r - 1;
  const bool gen = nb >= 2;
  // tape and matrices through global-typed pointers: global_* rather than
  // flat_* accesses, whose completion every LDS wait would also wait for
  s2::gdouble* Ub = s2::as_global(a.U) + (size_t)(ib - 1) * ldu * nq;
  s2::gdouble* Vb = s2::as_global(a.V) + (size_t)(ib - 1) * ldu * nq;
  double* Bl = L + Lay::kBl;
  double* Vr = L + Lay::kVr;
  double* vb = L + Lay::kVb;
  double* tauR = L + Lay::kTauR;
  double* tb = L + Lay::kTb;
  double* uvec = L + Lay::kU;
  double* vvec = L + Lay::kV;
  double* Uc = L + Lay::kUc;
  double* Al = L + Lay::kAl;
  double* tauUc = L + Lay::kTauU;
  double* sc = L + Lay::kSc;

  // earlier columns' reflectors of this block: none exist past the tape's
  // blocks (their steps had nb < 2)
  const int nlazy = ib <= a.nbl ? ic - 1 : 0;

  // ---- loads: earlier columns' reflectors, B block, A window, the two columns.
  // Addresses are clamped into the matrix and the entries outside the tiles
  // replaced by zeros afterwards, so every load is unconditional and all of
  // them are in flight together.
  for (int e = t; e < nlazy * ldu; e += kHT) Uc[e] = Ub[e];
  if (t < nlazy) tauUc[t] = s2::as_global(a.tauU)[(size_t)(ib - 1) * nq + t];
  const int nrw = gen ? imin(n, row0 + r + nb - 1) - row0 + 1 : 0;  // rows of the A window
  if (gen) {
    // all loads of both tiles in flight together, then the LDS stores
    constexpr int KB = RMAX * RMAX / kHT, KA = 2 * RMAX * RMAX / kHT;
    double vb_[KB], va_[KA];
    const s2::gdouble* Bb = s2::as_global(a.B) + (size_t)(row0 - 1) + (size_t)(row0 - 1) * a.ldb;
    const s2::gdouble* Ab = s2::as_global(a.A) + (size_t)(row0 - 1) + (size_t)(row0 - 1) * a.lda;
#pragma unroll
    for (int k = 0; k < KB; ++k) {
      const int e = t + k * kHT;
      const int m = e % RMAX, c = e / RMAX;
      vb_[k] = Bb[(size_t)imin(m, nb - 1) + (size_t)imin(c, nb - 1) * a.ldb];
    }
#pragma unroll
    for (int k = 0; k < KA; ++k) {
      const int e = t + k * kHT;
      const int m = e % (2 * RMAX), c = e / (2 * RMAX);
      va_[k] = Ab[(size_t)imin(m, nrw - 1) + (size_t)imin(c, nb - 1) * a.lda];
    }
#pragma unroll
    for (int k = 0; k < KB; ++k) {
      const int e = t + k * kHT;
      const int m = e % RMAX, c = e / RMAX;
      Bl[c * LD + m] = (m < nb && c < nb) ? vb_[k] : 0.0;
    }
#pragma unroll
    for (int k = 0; k < KA; ++k) {
      const int e = t + k * kHT;
      const int m = e % (2 * RMAX), c = e / (2 * RMAX);
      Al[c * LDA + m] = (m < nrw && c < nb) ? va_[k] : 0.0;
    }
  }
  // lazy vectors: A(g0+1 .. g0+la, col) in wave 0, B(g0+1 .. g0+lb, bcol) in wave 1
  const int la = imin(ic + r - 1, n - g0);
  const int lb = (bcol <= n) ? ic + r - 2 : 0;
  double x0 = 0.0, x1 = 0.0;
  if ((wave == 0 && la > 0) || (wave == 1 && lb > 0)) {
    const s2::gdouble* cp = s2::as_global(wave == 0 ? a.A + (size_t)g0 + (size_t)(col - 1) * a.lda
                                                    : a.B + (size_t)g0 + (size_t)(bcol - 1) * a.ldb);
    const int len = wave == 0 ? la : lb;
    const double v0 = cp[imin(lane, len - 1)], v1 = cp[imin(lane + 64, len - 1)];
    x0 = lane < len ? v0 : 0.0;
    x1 = lane + 64 < len ? v1 : 0.0;
  }
  __syncthreads();

  // ---- the earlier columns' left reflectors on the two vectors (rht2ht.f90:1455-1464)
  if (wave == 0 || (wave == 1 && lb > 0)) {
    for (int i = 0; i < nlazy; ++i) {
      const double u0 = lane < ldu ? Uc[i * ldu + lane] : 0.0;
      const double u1 = lane + 64 < ldu ? Uc[i * ldu + lane + 64] : 0.0;
      const double d = wsum(fma(u0, x0, u1 * x1)) * tauUc[i];
      x0 = fma(-d, u0, x0);
      x1 = fma(-d, u1, x1);
    }
  }
  if (wave == 0) {
    // the left reflector from A(row0 .. row0+nb-1, col) (rht2ht.f90:1466-1474)
    const int p = ic - 1;  // index of A(row0, col) in the vector
    if (gen) {
      const double al = readlane_d(p < 64 ? x0 : x1, p & 63);
      const bool in0 = lane > p && lane < p + nb, in1 = lane + 64 > p && lane + 64 < p + nb;
      const Refl F = wave_refl(al, x0, in0, x1, in1);
      s2::gdouble* ucol = Ub + (size_t)(ic - 1) * ldu;
      // u: 1 at index p, x * scal below; A's entries below p become zeros
      if (in0) {
        const double u = x0 * F.scal;
        uvec[lane - p] = u;
        ucol[lane] = u;
        x0 = 0.0;
      }
      if (in1) {
        const double u = x1 * F.scal;
        uvec[lane + 64 - p] = u;
        ucol[lane + 64] = u;
        x1 = 0.0;
      }
      if (lane == (p & 63)) {
        if (p < 64)
          x0 = F.beta;
        else
          x1 = F.beta;
        ucol[p] = 1.0;
        uvec[0] = 1.0;
        s2::as_global(a.tauU)[(size_t)(ib - 1) * nq + ic - 1] = F.tau;
        sc[0] = F.tau;
      }
      if (lane >= nb && lane < RMAX) uvec[lane] = 0.0;
    }
    s2::gdouble* cp = s2::as_global(a.A + (size_t)g0 + (size_t)(col - 1) * a.lda);
    if (lane < la) cp[lane] = x0;
    if (lane + 64 < la) cp[lane + 64] = x1;
  } else if (wave == 1 && lb > 0) {
    // B's column bcol: rows above the block to memory, block rows (the block's
    // last column, nb == r here) into the LDS block
    s2::gdouble* cp = s2::as_global(a.B + (size_t)g0 + (size_t)(bcol - 1) * a.ldb);
    const int above = ic - 1;
    if (lane < above) cp[lane] = x0;
    if (lane + 64 < above) cp[lane + 64] = x1;
    if (gen) {
      if (lane >= above && lane < lb) Bl[(nb - 1) * LD + lane - above] = x0;
      if (lane + 64 >= above && lane + 64 < lb) Bl[(nb - 1) * LD + lane + 64 - above] = x1;
    }
  }
  if (!gen) return;  // nothing else in this step (rht2ht.f90:1465)
  __syncthreads();

  // ---- B block <- (I - tau u u^T) B block (rht2ht.f90:1479-1480)
  {
    const double tau = sc[0];
    const int c = t >> 2, qr = t & 3;
    if (c < RMAX) {
      double col_v[E];
      double d = 0.0;
#pragma unroll
      for (int k = 0; k < E; ++k) {
        col_v[k] = Bl[c * LD + 4 * k + qr];
        d = fma(uvec[4 * k + qr], col_v[k], d);
      }
      d = s1::grp_sum<4>(d) * tau;
#pragma unroll
      for (int k = 0; k < E; ++k) Bl[c * LD + 4 * k + qr] = fma(-d, uvec[4 * k + qr], col_v[k]);
    }
  }
  __syncthreads();

  // ---- RQ factorization of the block (DGERQ2, rht2ht.f90:1484): reflector of
  // row i (0-based, nb-1 down to 1) from its entries 0..i, applied to rows < i
  {
    const int j = t >> 2, q = t & 3;
    double R[E];
#pragma unroll
    for (int k = 0; k < E; ++k) R[k] = (j < RMAX) ? Bl[(4 * k + q) * LD + j] : 0.0;
    double* prow = L + Lay::kPr;
    for (int i = nb - 1; i >= 1; --i) {
      double* vbi = vb + (i & 1) * RMAX;
      // the wave holding row i builds its reflector with one lane per column
      if (wave == (i >> 4)) {
        if (j == i) {
#pragma unroll
          for (int k = 0; k < E; ++k) prow[4 * k + q] = R[k];
        }
        asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
        const double x = prow[lane];
        const double al = readlane_d(x, i);
        const Refl F = wave_refl(al, x, lane < i, 0.0, false);
        if (lane < RMAX) {
          const double v = lane < i ? x * F.scal : (lane == i ? 1.0 : 0.0);
          vbi[lane] = v;
          Vr[i * LD + lane] = v;
        }
        if (lane == 0) {
          tb[i & 1] = F.tau;
          tauR[i] = F.tau;
        }
      }
      __syncthreads();
      if (j < i) {
        const double tau = tb[i & 1];
        double dp[4] = {0.0, 0.0, 0.0, 0.0};  // four chains for the FMA latency
#pragma unroll
        for (int k = 0; k < E; ++k) dp[k & 3] = fma(R[k], vbi[4 * k + q], dp[k & 3]);
        double d = (dp[0] + dp[1]) + (dp[2] + dp[3]);
        d = s1::grp_sum<4>(d) * tau;
#pragma unroll
        for (int k = 0; k < E; ++k) R[k] = fma(-d, vbi[4 * k + q], R[k]);
      }
    }
  }
  __syncthreads();

  // ---- first row of the RQ's orthogonal factor, e1^T H(1) ... H(nb-1)
  // (DORGR2, rht2ht.f90:1485), and the right reflector from it
  // (rht2ht.f90:1486-1488)
  if (wave == 0) {
    double x = (lane == 0) ? 1.0 : 0.0;
    for (int i = 1; i < nb; ++i) {
      const double v = (lane <= i) ? Vr[i * LD + lane] : 0.0;
      const double d = wsum(x * v) * tauR[i];
      x = fma(-d, v, x);
    }
    const double al = readlane_d(x, 0);
    const bool in0 = lane >= 1 && lane < nb;
    const Refl F = wave_refl(al, x, in0, 0.0, false);
    s2::gdouble* vcol = Vb + (size_t)(ic - 1) * ldu + (ic - 1);
    const double v = (lane == 0) ? 1.0 : (in0 ? x * F.scal : 0.0);
    if (lane < RMAX) vvec[lane] = v;
    if (lane < nb) vcol[lane] = v;
    if (lane == 0) {
      s2::as_global(a.tauV)[(size_t)(ib - 1) * nq + ic - 1] = F.tau;
      sc[1] = F.tau;
    }
  }
  __syncthreads();

  // ---- right reflector on the B block rows, the A window rows
  // (rht2ht.f90:1489-1494)
  {
    const double tau = sc[1];
    const int j = t >> 2, q = t & 3;
    if (j < RMAX) {
      double rv[E];
      double d = 0.0;
#pragma unroll
      for (int k = 0; k < E; ++k) {
        rv[k] = Bl[(4 * k + q) * LD + j];
        d = fma(rv[k], vvec[4 * k + q], d);
      }
      d = s1::grp_sum<4>(d) * tau;
#pragma unroll
      for (int k = 0; k < E; ++k) {
        double y = fma(-d, vvec[4 * k + q], rv[k]);
        if (k == 0 && q == 0 && j >= 1) y = 0.0;  // B(row0+1.., row0) = 0 exactly
        Bl[(4 * k + q) * LD + j] = y;
      }
    }
    const int arow = t >> 1, ah = t & 1;
    if (arow < 2 * RMAX) {
      double aw[AW];
      double d = 0.0;
#pragma unroll
      for (int k = 0; k < AW; ++k) {
        aw[k] = Al[(2 * k + ah) * LDA + arow];
        d = fma(aw[k], vvec[2 * k + ah], d);
      }
      d = s1::grp_sum<2>(d) * tau;
#pragma unroll
      for (int k = 0; k < AW; ++k)
        Al[(2 * k + ah) * LDA + arow] = fma(-d, vvec[2 * k + ah], aw[k]);
    }
  }
  __syncthreads();

  // ---- stores
  {
    s2::gdouble* Bb = s2::as_global(a.B + (size_t)(row0 - 1) + (size_t)(row0 - 1) * a.ldb);
#pragma unroll 4
    for (int k = 0; k < RMAX * RMAX / kHT; ++k) {
      const int e = t + k * kHT;
      const int m = e % RMAX, c = e / RMAX;
      if (m < nb && c < nb) Bb[(size_t)m + (size_t)c * a.ldb] = Bl[c * LD + m];
    }
    s2::gdouble* Ab = s2::as_global(a.A + (size_t)(row0 - 1) + (size_t)(row0 - 1) * a.lda);
#pragma unroll 4
    for (int k = 0; k < 2 * RMAX * RMAX / kHT; ++k) {
      const int e = t + k * kHT;
      const int m = e % (2 * RMAX), c = e / (2 * RMAX);
      if (m < nrw && c < nb) Ab[(size_t)m + (size_t)c * a.lda] = Al[c * LDA + m];
    }
  }
  if (rows_above) {
    const int istart = (ib == 1) ? a.s + 1 : imax(a.s + 1, a.s + 1 + (ib - 1 - nq + ic) * r);
    hh_rows_above<RMAX>(a, istart, row0, nb, vvec, sc[1], t, kHT);
  }
}

// Helper workgroup h of column ic: once its column's step ib is published,
// the rows istart..row0-1 of A and B above the window (rht2ht.f90:1495-1496,
// 1592-1596); the next column's steps wait for it too.
template <int RMAX>
__device__ void hh_helper(const HhSweepArgs& a, int ic, int h, double* L) {
  const int t = threadIdx.x;
  const int n = a.n, r = a.r, nq = a.nq;
  int* mine = &a.hprog[(ic - 1) * a.nhelp + h];
  double* hv = L;
  for (int ib = 1;; ++ib) {
    const int jc = a.s + ic - 1;
    if (ib == 1 ? jc > n : jc + 1 + (ib - 2) * r > n) break;
    wait_at_least(&a.progress[ic - 1], ib, a.errflag, a.spin);
    acquire_waits();
    const int row0 = a.s + (ib - 1) * r + ic;
    const int nb = imin(r, n - row0 + 1);
    const int istart = (ib == 1) ? a.s + 1 : imax(a.s + 1, a.s + 1 + (ib - 1 - nq + ic) * r);
    if (nb >= 2 && istart < row0) {
      const s2::gdouble* vcol =
          s2::as_global(a.V + (size_t)(ib - 1) * a.ldu * nq + (size_t)(ic - 1) * a.ldu + (ic - 1));
      if (t < RMAX) hv[t] = t < nb ? vcol[t] : 0.0;
      __syncthreads();
      const double tau = s2::as_global(a.tauV)[(size_t)(ib - 1) * nq + ic - 1];
      hh_rows_above<RMAX>(a, istart, row0, nb, hv, tau, h * kHT + t, a.nhelp * kHT);
    }
    publish(mine, ib);
  }
  publish(mine, kHhDone);
}

template <int RMAX>
__global__ __launch_bounds__(kHT) void k_hh_sweep(HhSweepArgs a) {
  extern __shared__ double L[];
  if ((int)blockIdx.x >= a.nq) {
    const int hidx = blockIdx.x - a.nq;
    hh_helper<RMAX>(a, hidx / a.nhelp + 1, hidx % a.nhelp, L);
    return;
  }
  const int ic = blockIdx.x + 1;
  const int jc = a.s + ic - 1;
  for (int ib = 1;; ++ib) {
    if (ib == 1 ? jc > a.n : jc + 1 + (ib - 2) * a.r > a.n) break;
    if (ic > 1) {  // the previous column, leader and helpers, two steps ahead
      wait_at_least(&a.progress[ic - 2], ib + 2, a.errflag, a.spin);
      for (int h = 0; h < a.nhelp; ++h)
        wait_at_least(&a.hprog[(ic - 2) * a.nhelp + h], ib + 2, a.errflag, a.spin);
      acquire_waits();
    }
    hh_step<RMAX>(a, ic, ib, L, a.nhelp == 0);
    publish(&a.progress[ic - 1], ib);
  }
  publish(&a.progress[ic - 1], kHhDone);
}

// Tape reset: U, V, tauU, tauV zero; counters zero; the call's first sweep
// clears the error flag.
__global__ void k_hh_prep(double* tape, size_t nd, int* counters, int nc, int* errflag,
                          int clear_err) {
  const size_t gid = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  const size_t gsz = (size_t)gridDim.x * blockDim.x;
  for (size_t e = gid; e < nd; e += gsz) tape[e] = 0.0;
  for (size_t e = gid; e < (size_t)nc; e += gsz) counters[e] = 0;
  if (gid == 0 && clear_err) *errflag = 0;
}

// T of H(1)...H(nq) = I - Y T Y^T for every block (DLARFT 'F','C',
// rht2ht.f90:1599-1610); blockIdx.y: 0 for U, 1 for V.  One wave per block:
// G = Y^T Y, then the recurrence T(0:j-1, j) = T(0:j-1, 0:j-1) (-tau_j G(0:j-1, j)),
// lane i owning row i of T.
__global__ __launch_bounds__(64) void k_hh_larft(const double* __restrict__ U,
                                                 const double* __restrict__ V,
                                                 const double* __restrict__ tauU,
                                                 const double* __restrict__ tauV, double* TU,
                                                 double* TV, int ldu, int nq) {
  __shared__ double Ys[kHhMaxLdu * kHhMaxNq];
  __shared__ double G[kHhMaxNq * (kHhMaxNq + 1)];
  __shared__ double Tl[kHhMaxNq * (kHhMaxNq + 1)];
  const int ib = blockIdx.x, lane = threadIdx.x;
  const bool isV = blockIdx.y == 1;
  const double* Y = (isV ? V : U) + (size_t)ib * ldu * nq;
  const double* tau = (isV ? tauV : tauU) + (size_t)ib * nq;
  double* T = (isV ? TV : TU) + (size_t)ib * nq * nq;
  for (int e = lane; e < ldu * nq; e += 64) Ys[e] = Y[e];
  __syncthreads();
  constexpr int LG = kHhMaxNq + 1;
  if (lane < nq) {
    for (int j = 0; j < nq; ++j) {
      double g = 0.0;
      for (int l = 0; l < ldu; ++l) g = fma(Ys[lane * ldu + l], Ys[j * ldu + l], g);
      G[lane * LG + j] = g;
    }
  }
  __syncthreads();
  if (lane < nq) {
    for (int j = 0; j < nq; ++j) {
      const double tj = tau[j];
      double tij = 0.0;
      if (tj != 0.0 && lane < j) {
        for (int m = lane; m < j; ++m) tij = fma(Tl[lane * LG + m], G[m * LG + j], tij);
        tij *= -tj;
      }
      Tl[lane * LG + j] = (lane == j) ? tj : (lane < j ? tij : 0.0);
    }
    for (int j = 0; j < nq; ++j) T[(size_t)j * nq + lane] = Tl[lane * LG + j];
  }
}

// ---------------------------------------------------------------------------
// Far updates on FP64 MFMA.  X <- (I - Y T^T Y^T) X over the windows of the
// sweep's blocks, bottom-up, where X is
//   * COLS = false: the transpose of rows of A, B (above the windows), Q or Z
//     (right reflectors, C <- C (I - Y T Y^T), rht2ht.f90:1644-1765,
//     1818-1951, 1976-2018);
//   * COLS = true: columns of A and B right of the windows (left reflectors,
//     C <- (I - Y T Y^T)^T C).
// A wave owns 16 of the "other" indices (rows, or columns for COLS) and holds
// the block's window of R+NQ indices in registers: lane l has other index
// l%16 and window indices 4j + l/16 (chunk j), which is exactly the operand
// and accumulator layout of v_mfma_f64_16x16x4_f64 for
//   W  = Y^T X   (NQ x 16),  W2 = T^T W,  X -= Y W2.
// Consecutive blocks' windows overlap by NQ indices: they stay in registers
// (shifted by R/4 chunks) and only R new indices are loaded per block, one
// block ahead.  Per other index, the sequences of a block that apply form a
// range [lo, hi] (the phase's edge rules); W and W2 are masked to it, which
// makes the product exactly H(lo) ... H(hi).  Only the window indices the
// applied reflectors reach are stored, since the rest may belong to the next
// sweep's generation running beside this kernel.
// ---------------------------------------------------------------------------
struct HhFarArgs {
  double* M[2];
  int ld[2];
  const double* Y[2];
  const double* T[2];
  int isA[2];  // COLS: A's stop columns (1) or B's (0)
  int n, r, nq, s, nblocks, ldu, phase;
};

__device__ __forceinline__ int floordiv(int a, int b) { return a >= 0 ? a / b : -((-a + b - 1) / b); }

// Sequences [lo, hi] (1-based) of block ib that row `row` (0-based) of A or B
// receives in `phase`, or all of them (kHhFull: Q and Z).
__device__ __forceinline__ void hh_row_range(int row, int ib, int phase, int s, int r, int nq,
                                             int& lo, int& hi) {
  lo = 1;
  hi = nq;
  if (phase == kHhFull) return;
  const int i = row + 1;
  if (i <= s) {
    if (phase == kHhLook) hi = 0;
    return;
  }
  const int base = s + 1 + (ib - 1) * r;  // the block's first window row
  if (i >= base) {
    hi = 0;
    return;
  }
  const int d = floordiv(i - base, r);
  if (phase == kHhAll) {
    lo = imax(1, d + nq + 1);
  } else if (phase == kHhLook) {
    lo = imax(1, d + nq + 1);
    hi = imin(nq, d + 2 * nq);
  } else if (d > -2 * nq) {
    lo = d + 2 * nq + 1;  // kHhFar: the edge rows
  }
}

// Whether column `col` (0-based) of A or B receives block ib's left
// reflectors in `phase` (rht2ht.f90:1690-1704, 1752-1765, 1869-1884, 1941-1951).
__device__ __forceinline__ bool hh_col_active(int col, int ib, int phase, bool isA, int n, int s,
                                              int r, int nq) {
  const int j = col + 1;
  const int istop = imin(n, isA ? ((ib == 1) ? s + nq - 1 : s + (ib - 2) * r + nq)
                                : s + (ib - 1) * r + r + nq - 1);
  const int istop2 = imin(n, s + (ib - 1) * r + r + nq + r * nq - 1);
  if (phase == kHhLook) return j > istop && j <= istop2;
  if (phase == kHhFar) return j > istop2;
  return j > istop;
}

template <int R, int NQ, int WAVES, bool COLS>
__global__ __launch_bounds__(64 * WAVES) void k_hh_far(HhFarArgs a) {
  static_assert(R % 4 == 0 && NQ % 4 == 0, "chunks of four window indices");
  constexpr int W = R + NQ;          // window indices (= the tape's ldu)
  constexpr int WT = (W + 15) / 16;  // 16-index output tiles
  constexpr int WC = W / 4;          // chunks of four
  constexpr int RC = R / 4, NC = NQ / 4;
  constexpr int KT = (NQ + 15) / 16;  // 16-sequence tiles
  constexpr int SK = 16 * KT + 1;     // LDS row stride of Y and T
  constexpr int NT = 64 * WAVES;
  constexpr int YS = WT * 16 * SK, TS = 16 * KT * SK;
  extern __shared__ double lds[];
  // two buffers of Y (YS) and T (TS) staged per block; buffer b at b*(YS+TS)
  int* rng = (int*)(lds + 2 * (YS + TS));
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, t = threadIdx.x;
  const int l15 = lane & 15, cs = lane >> 4;
  const int y = blockIdx.y;
  double* __restrict__ M = a.M[y];
  const int ld = a.ld[y];
  const double* __restrict__ Yt = a.Y[y];
  const double* __restrict__ Tt = a.T[y];
  const bool isA = a.isA[y] != 0;
  const int n = a.n, r = a.r, nq = a.nq, s = a.s;
  const int other = (blockIdx.x * WAVES + wave) * 16 + l15;  // row (col for COLS)

  auto range = [&](int ib, int& lo, int& hi) {
    if (other >= n) {
      lo = 1;
      hi = 0;
    } else if (COLS) {
      lo = 1;
      hi = hh_col_active(other, ib, a.phase, isA, n, s, r, nq) ? nq : 0;
    } else {
      hh_row_range(other, ib, a.phase, s, r, nq, lo, hi);
    }
  };
  // the workgroup's blocks: the union of its rows' (columns') ranges, contiguous
  int bhi = 0, blo = a.nblocks + 1;
  for (int ib = a.nblocks; ib >= 1; --ib) {
    int lo, hi;
    range(ib, lo, hi);
    if (__any(lo <= hi)) {
      bhi = imax(bhi, ib);
      blo = imin(blo, ib);
    }
  }
  if (lane == 0) {
    rng[2 * wave] = bhi;
    rng[2 * wave + 1] = blo;
  }
  for (int e = t; e < 2 * (YS + TS); e += NT) lds[e] = 0.0;
  __syncthreads();
#pragma unroll
  for (int w = 0; w < WAVES; ++w) {
    bhi = imax(bhi, rng[2 * w]);
    blo = imin(blo, rng[2 * w + 1]);
  }
  if (bhi < blo) return;

  auto elem = [&](int win0, int w) -> size_t {  // address of (window index, other)
    return COLS ? (size_t)(win0 + w) + (size_t)other * ld : (size_t)other + (size_t)(win0 + w) * ld;
  };
  auto stage = [&](int ib, int buf) {
    const double* yb = Yt + (size_t)(ib - 1) * W * NQ;
    const double* tb = Tt + (size_t)(ib - 1) * NQ * NQ;
    double* ys = lds + buf * (YS + TS);
    double* ts = ys + YS;
    for (int e = t; e < W * NQ; e += NT) ys[(e % W) * SK + e / W] = yb[e];
    for (int e = t; e < NQ * NQ; e += NT) ts[(e % NQ) * SK + e / NQ] = tb[e];
  };

  double X[WC];
  double Xn[RC];
  int dlo = 1 << 30, dhi = -1;
  // first block: the whole window
  {
    const int w0 = s + (bhi - 1) * r;
#pragma unroll
    for (int j = 0; j < WC; ++j) {
      const int w = 4 * j + cs;
      X[j] = (other < n && w0 + w < n) ? M[elem(w0, w)] : 0.0;
    }
  }
  stage(bhi, 0);
  __syncthreads();
  for (int ib = bhi, buf = 0; ib >= blo; --ib, buf ^= 1) {
    const int w0 = s + (ib - 1) * r;  // first window index (0-based)
    const bool more = ib > blo;
    // prefetch: the next block's R new indices and its factors
    if (more) {
      const int wn = w0 - r;
#pragma unroll
      for (int j = 0; j < RC; ++j) {
        const int w = 4 * j + cs;
        Xn[j] = (other < n && wn + w >= 0 && wn + w < n) ? M[elem(wn, w)] : 0.0;
      }
      stage(ib - 1, buf ^ 1);
    }
    int lo, hi;
    range(ib, lo, hi);
    const bool act = lo <= hi;
    if (act) {
      // the applied reflectors' reach: local indices lo-1 .. hi+r-2
      dlo = imin(dlo, w0 + lo - 1);
      dhi = imax(dhi, w0 + hi + r - 2);
    }
    const double* Yb = lds + buf * (YS + TS);
    const double* Tb = Yb + YS;
    // W = Y^T X
    dbl4 Wt[KT];
#pragma unroll
    for (int kt = 0; kt < KT; ++kt) {
      dbl4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int j = 0; j < WC; ++j) acc = mfma16(Yb[(4 * j + cs) * SK + 16 * kt + l15], X[j], acc);
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int q = 16 * kt + 4 * i + cs + 1;  // 1-based sequence
        if (!(act && q >= lo && q <= hi)) acc[i] = 0.0;
      }
      Wt[kt] = acc;
    }
    // W2 = T^T W (T upper triangular)
    dbl4 W2[KT];
#pragma unroll
    for (int ko = 0; ko < KT; ++ko) {
      dbl4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int kt = 0; kt <= ko; ++kt)
#pragma unroll
        for (int i = 0; i < 4; ++i)
          acc = mfma16(Tb[(16 * kt + 4 * i + cs) * SK + 16 * ko + l15], Wt[kt][i], acc);
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int q = 16 * ko + 4 * i + cs + 1;
        if (!(act && q >= lo && q <= hi)) acc[i] = 0.0;
      }
      W2[ko] = acc;
    }
    // X -= Y W2
#pragma unroll
    for (int wt = 0; wt < WT; ++wt) {
      dbl4 acc;
#pragma unroll
      for (int i = 0; i < 4; ++i) acc[i] = (4 * wt + i < WC) ? X[4 * wt + i] : 0.0;
#pragma unroll
      for (int kt = 0; kt < KT; ++kt)
#pragma unroll
        for (int i = 0; i < 4; ++i)
          acc = mfma16(-Yb[(16 * wt + l15) * SK + 16 * kt + 4 * i + cs], W2[kt][i], acc);
#pragma unroll
      for (int i = 0; i < 4; ++i)
        if (4 * wt + i < WC) X[4 * wt + i] = acc[i];
    }
    // store what leaves the window (all of it after the last block)
#pragma unroll
    for (int j = 0; j < WC; ++j) {
      if (more && j < NC) continue;
      const int w = 4 * j + cs, g = w0 + w;
      if (other < n && g < n && g >= dlo && g <= dhi) M[elem(w0, w)] = X[j];
    }
    if (more) {
#pragma unroll
      for (int j = NC - 1; j >= 0; --j) X[j + RC] = X[j];
#pragma unroll
      for (int j = 0; j < RC; ++j) X[j] = Xn[j];
    }
    __syncthreads();  // the next block's factors are staged; this buffer is free
  }
}

// The same far updates for any (r, nq), one thread per other index applying
// the reflectors H(lo), ..., H(hi) of each block one at a time from the tape
// (no MFMA; used where the MFMA kernels have no instantiation).
template <bool COLS>
__global__ __launch_bounds__(128) void k_hh_far_generic(HhFarArgs a, const double* tau0,
                                                        const double* tau1) {
  const int y = blockIdx.y;
  double* __restrict__ M = a.M[y];
  const int ld = a.ld[y];
  const double* __restrict__ Yt = a.Y[y];
  const double* __restrict__ tau = y ? tau1 : tau0;
  const bool isA = a.isA[y] != 0;
  const int n = a.n, r = a.r, nq = a.nq, s = a.s, W = a.ldu;
  const int other = blockIdx.x * blockDim.x + threadIdx.x;
  if (other >= n) return;
  for (int ib = a.nblocks; ib >= 1; --ib) {
    int lo = 1, hi = nq;
    if (COLS)
      hi = hh_col_active(other, ib, a.phase, isA, n, s, r, nq) ? nq : 0;
    else
      hh_row_range(other, ib, a.phase, s, r, nq, lo, hi);
    const int w0 = s + (ib - 1) * r;
    for (int q = lo; q <= hi; ++q) {
      const double tq = tau[(size_t)(ib - 1) * nq + q - 1];
      if (tq == 0.0) continue;
      const double* yq = Yt + (size_t)(ib - 1) * W * nq + (size_t)(q - 1) * W;
      const int w1 = imin(W, n - w0);
      double d = 0.0;
      for (int w = q - 1; w < w1; ++w) {
        const size_t e = COLS ? (size_t)(w0 + w) + (size_t)other * ld
                              : (size_t)other + (size_t)(w0 + w) * ld;
        d = fma(yq[w], M[e], d);
      }
      d *= tq;
      for (int w = q - 1; w < w1; ++w) {
        const double v = yq[w];
        if (v == 0.0) continue;
        const size_t e = COLS ? (size_t)(w0 + w) + (size_t)other * ld
                              : (size_t)other + (size_t)(w0 + w) * ld;
        M[e] = fma(-d, v, M[e]);
      }
    }
  }
}

// ---------------------------------------------------------------------------
// launchers
size_t hh_sweep_lds(int r) {
  const int rm = r <= 16 ? 16 : (r <= 32 ? 32 : 64);
  if (rm == 16) return sizeof(double) * HhLds<16>::kSize;
  if (rm == 32) return sizeof(double) * HhLds<32>::kSize;
  return sizeof(double) * HhLds<64>::kSize;
}

bool hh_supported(int r, int nq) {
  return r >= 1 && r <= 64 && nq >= 1 && nq <= kHhMaxNq && r + nq <= kHhMaxLdu;
}

hipError_t hh_prepare(hipStream_t st, double* tape, size_t nd, int* counters, int nc,
                      int* errflag, int clear_err) {
  hipLaunchKernelGGL(k_hh_prep, dim3(256), dim3(256), 0, st, tape, nd, counters, nc, errflag,
                     clear_err);
  return hipGetLastError();
}

hipError_t hh_generate(hipStream_t st, double* A, int lda, double* B, int ldb, int n, int r,
                       int nq, int s, double* U, double* V, double* tauU, double* tauV,
                       int* progress, int* hprog, int* errflag, int spin, int nhelp) {
  if (!hh_supported(r, nq)) return hipErrorInvalidValue;
  HhSweepArgs a{A, B, lda, ldb, n, r, nq, s, r + nq, U, V, tauU, tauV, progress, hprog, errflag,
                spin, nhelp, (n - 1) / r + ((n - 1) % r != 0)};
  const dim3 g(nq * (1 + nhelp)), b(kHT);
  const size_t lds = hh_sweep_lds(r);
#define PHT_HH(RM)                                                                     \
  do {                                                                                 \
    static bool attr = false;                                                          \
    if (!attr) {                                                                       \
      hipError_t e = hipFuncSetAttribute((const void*)k_hh_sweep<RM>,                  \
                                         hipFuncAttributeMaxDynamicSharedMemorySize,   \
                                         (int)(sizeof(double) * HhLds<RM>::kSize));    \
      if (e != hipSuccess) return e;                                                   \
      attr = true;                                                                     \
    }                                                                                  \
    hipLaunchKernelGGL(k_hh_sweep<RM>, g, b, lds, st, a);                              \
  } while (0)
  if (r <= 16)
    PHT_HH(16);
  else if (r <= 32)
    PHT_HH(32);
  else
    PHT_HH(64);
#undef PHT_HH
  return hipGetLastError();
}

hipError_t hh_factors(hipStream_t st, const double* U, const double* V, const double* tauU,
                      const double* tauV, double* TU, double* TV, int ldu, int nq, int nblocks) {
  if (nblocks <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_hh_larft, dim3(nblocks, 2), dim3(64), 0, st, U, V, tauU, tauV, TU, TV,
                     ldu, nq);
  return hipGetLastError();
}

namespace {
template <int R, int NQ, bool COLS>
hipError_t far_launch(hipStream_t st, const HhFarArgs& a, int nmat) {
  constexpr int WAVES = 2;
  constexpr int W = R + NQ, WT = (W + 15) / 16, KT = (NQ + 15) / 16, SK = 16 * KT + 1;
  const size_t lds = sizeof(double) * 2 * (WT * 16 * SK + 16 * KT * SK) + 2 * WAVES * sizeof(int);
  static bool attr = false;
  if (!attr) {
    hipError_t e = hipFuncSetAttribute((const void*)k_hh_far<R, NQ, WAVES, COLS>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    attr = true;
  }
  const dim3 g((a.n + 16 * WAVES - 1) / (16 * WAVES), nmat), b(64 * WAVES);
  hipLaunchKernelGGL((k_hh_far<R, NQ, WAVES, COLS>), g, b, lds, st, a);
  return hipGetLastError();
}
}  // namespace

#define PHT_HH_PAIRS(X) \
  X(8, 4) X(8, 8) X(16, 4) X(16, 8) X(16, 16) X(32, 8) X(32, 16) X(32, 32) X(64, 8) X(64, 16) X(64, 32)

bool hh_far_supported(int r, int nq) {
#define PHT_HAS(RR, NN) \
  if (r == RR && nq == NN) return true;
  PHT_HH_PAIRS(PHT_HAS)
#undef PHT_HAS
  return false;
}

hipError_t hh_far(hipStream_t st, bool cols, double* M0, int ld0, double* M1, int ld1,
                  const double* Y0, const double* T0, const double* tau0, const double* Y1,
                  const double* T1, const double* tau1, int isA0, int isA1, int n, int r, int nq,
                  int s, int nblocks, int phase, bool generic) {
  if (!hh_far_supported(r, nq)) generic = true;
  HhFarArgs a;
  a.M[0] = M0;
  a.M[1] = M1;
  a.ld[0] = ld0;
  a.ld[1] = ld1;
  a.Y[0] = Y0;
  a.Y[1] = Y1;
  a.T[0] = T0;
  a.T[1] = T1;
  a.isA[0] = isA0;
  a.isA[1] = isA1;
  a.n = n;
  a.r = r;
  a.nq = nq;
  a.s = s;
  a.nblocks = nblocks;
  a.ldu = r + nq;
  a.phase = phase;
  const int nmat = M1 ? 2 : 1;
  if (nblocks <= 0) return hipSuccess;
  if (!generic) {
#define PHT_FAR(RR, NN)                                                      \
  if (r == RR && nq == NN)                                                   \
    return cols ? far_launch<RR, NN, true>(st, a, nmat) : far_launch<RR, NN, false>(st, a, nmat);
    PHT_HH_PAIRS(PHT_FAR)
#undef PHT_FAR
  }
  const dim3 g((n + 127) / 128, nmat), b(128);
  if (cols)
    hipLaunchKernelGGL(k_hh_far_generic<true>, g, b, 0, st, a, tau0, tau1);
  else
    hipLaunchKernelGGL(k_hh_far_generic<false>, g, b, 0, st, a, tau0, tau1);
  return hipGetLastError();
}

}  // namespace pht
