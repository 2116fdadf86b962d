// FP64 per-sample evaluation for large surface counts (A <= 16, padded to 16), one wave per
// sample.
//
// Same estimator as k_eval_fp64 (pibronic/pimc/pimc.py:1087-1213 fused; block loops
// :1341-1462), mapped for matrices that do not fit one lane's registers:
//
//  * Beads are processed four at a time. Lane (g, j) = (lane / 16, lane % 16) owns column j
//    of the coupling matrix of bead 4c + g: it assembles V[:, j] from the linear and
//    quadratic couplings, forms the O-factor of surface j for every temperature set, and
//    takes part in the per-bead scaling maximum over its 16-lane group.
//  * Eigen-decomposition: one-sided (Hestenes) Jacobi on G = V + sigma I, sigma = 1.5 |V|_F,
//    which is positive definite with condition number <= 5. Its singular vectors are the
//    eigenvectors of V and lambda_j = |G u_j| - sigma. A sweep is the 15 perfect matchings
//    {j, j xor k}, k = 1..15 (every pair once); partners swap columns with cross-lane moves,
//    so no rotation parameters have to be broadcast. The eigenvectors are the normalised
//    columns of the rotated G (Demmel & Veselic, "Jacobi's method is more accurate than QR",
//    SIAM J. Matrix Anal. Appl. 13 (1992), for the one-sided scheme).
//  * M = U diag(exp(-tau lambda)) U^T and the ring product run on the matrix cores: 16x16x4
//    FP64 MFMA tiles. The FP64 accumulator layout is: register r of lane l holds
//    D[l/16 + 4r][l%16] (measured on gfx950). Permuting the sum index of each tile step
//    (k = lane/16 + 4*step) makes that layout directly the operand layout of the next
//    product, so the chain never leaves registers.
//    The chain is carried transposed, R^T <- diag(O_p / S_p) M_p R^T, whose trace equals the
//    reference's tr(prod_p M_p O_p) because M_p is symmetric and O_p diagonal.
//
// Padding: columns j >= A hold sigma e_j, are never rotated (their inner products with real
// columns are exactly zero) and carry O = 0, so they drop out of every trace.
#include "pib_common.hpp"
#include "pib_kernels.hpp"

namespace pib {

typedef double d4 __attribute__((ext_vector_type(4)));

constexpr int kW = 16;                 // padded matrix dimension = lanes per bead
constexpr int kBeadsPerChunk = kWave / kW;
constexpr int kWideWaves = 2;          // waves (samples) per work-group
constexpr int kURow = kW + 1;          // LDS row stride of the eigenvector tiles (bank spread)

struct WideLds {
  double U[kBeadsPerChunk][kW * kURow];  // U[b][i * kURow + m] = eigenvector m, component i
  double D[kBeadsPerChunk][kW];          // exp(-tau lambda_m)
  double O[3][kBeadsPerChunk][kW];       // scaled O-factors (vib, vib+, vib-)
};

// hardware 1/x and 1/sqrt(x) (~2^-29) with Newton steps (see rcp_nr in the per-lane kernel):
// two steps where orthogonality or an output depends on it, one inside the rotation angle
template <int STEPS = 2>
__device__ __forceinline__ double rcp_nr_w(double x) {
  double y = __builtin_amdgcn_rcp(x);
#pragma unroll
  for (int it = 0; it < STEPS; ++it) {
    const double e = fma(-x, y, 1.0);
    y = fma(y, e, y);
  }
  return y;
}

template <int STEPS = 2>
__device__ __forceinline__ double rsq_nr_w(double x) {
  double y = __builtin_amdgcn_rsq(x);
#pragma unroll
  for (int it = 0; it < STEPS; ++it) {
    const double e = fma(-(x * y), y, 1.0);
    y = fma(0.5 * y, e, y);
  }
  return y;
}

__device__ __forceinline__ double oterm_w(double coth, double csch2, double x1, double x2) {
  return coth * (x1 * x1 + x2 * x2) - (csch2 * x1) * x2;
}

// Wave-local ordering of LDS traffic: every access of the wave before this point is complete
// and visible to every lane of the same wave after it.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// sum / max over the 16 lanes of a group; the xor butterfly gives every lane the same bits
// (a + b == b + a exactly)
__device__ __forceinline__ double group_sum(double x) {
#pragma unroll
  for (int m = 1; m < kW; m <<= 1) x += shfl_xor_d(x, m);
  return x;
}
__device__ __forceinline__ double group_max(double x) {
#pragma unroll
  for (int m = 1; m < kW; m <<= 1) x = fmax(x, shfl_xor_d(x, m));
  return x;
}

// BF16x3 ring products: FP32 values split into bf16 hi + lo, products on the bf16 matrix cores
// (v_mfma_f32_16x16x16_bf16: lane l supplies A[l%16][4(l/16)+e] and B[4(l/16)+e][l%16],
// result register r = D[4(l/16)+r][l%16]), FP32 accumulation.
typedef float f4w __attribute__((ext_vector_type(4)));
typedef short s4w __attribute__((ext_vector_type(4)));

__device__ __forceinline__ unsigned short f2bf_w(float x) {
  unsigned u = __float_as_uint(x);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}
__device__ __forceinline__ void split_bf(float x, short& h, short& l) {
  const unsigned short hb = f2bf_w(x);
  h = (short)hb;
  l = (short)f2bf_w(x - __uint_as_float((unsigned)hb << 16));
}
__device__ __forceinline__ f4w mfma_bf16x3(const s4w& ah, const s4w& al, const s4w& bh, const s4w& bl,
                                           f4w c) {
  c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(al, bh, c, 0, 0, 0);
  c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah, bl, c, 0, 0, 0);
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah, bh, c, 0, 0, 0);
}
// scale a 16x16 FP32 matrix (one wave, 4 values per lane) by a power of two so that its largest
// magnitude lies in [0.5, 1); the exponent (the same on every lane) goes to e
__device__ __forceinline__ void renorm_wave(f4w& R, int& e) {
  float m = fmaxf(fmaxf(fabsf(R[0]), fabsf(R[1])), fmaxf(fabsf(R[2]), fabsf(R[3])));
#pragma unroll
  for (int s = 1; s < kWave; s <<= 1) m = fmaxf(m, __shfl_xor(m, s, kWave));
  int ex = 0;
  (void)frexpf(m, &ex);
  if (m > 0.0f && ex != 0) {
    const float f = ldexpf(1.0f, -ex);
#pragma unroll
    for (int r = 0; r < 4; ++r) R[r] *= f;
    e += ex;
  }
}

template <int K>
struct HiBit {
  static constexpr int v = (K >= 8) ? 8 : (K >= 4) ? 4 : (K >= 2) ? 2 : 1;
};

// One step of the sweep: lane j and lane j ^ K orthogonalise their columns.
template <int K>
__device__ __forceinline__ void hestenes_step(double (&v)[kW], double& al, int j, double tol2,
                                              bool& busy) {
  double pc[kW];
#pragma unroll
  for (int i = 0; i < kW; ++i) pc[i] = shfl_xor_d(v[i], K);
  const double be = shfl_xor_d(al, K);
  double gam = 0.0;
#pragma unroll
  for (int i = 0; i < kW; ++i) gam = fma(v[i], pc[i], gam);
  const bool low = (j & HiBit<K>::v) == 0;  // this lane holds the lower column index p
  const double app = low ? al : be, aqq = low ? be : al;
  // Gram block [[app, gam], [gam, aqq]]: t = tan of the rotation that zeroes gam (smaller
  // root, |t| <= 1); gam == 0 gives t = 0, an exact no-op (see jacobi_rotate)
  const double d = aqq - app, h = 2.0 * gam;
  const double x = fma(d, d, h * h);
  const double r = x * rsq_nr_w<1>(x > 0.0 ? x : 1.0);
  const double den = fabs(d) + r;
  double t = h * rcp_nr_w<1>(den > 0.0 ? den : 1.0);
  t = d < 0.0 ? -t : t;
  const double c = rsq_nr_w<2>(fma(t, t, 1.0));
  const double s = t * c;
  busy |= (gam * gam > tol2 * (app * aqq));
  const double sg = low ? -s : s;  // p: c g_p - s g_q ; q: s g_p + c g_q
#pragma unroll
  for (int i = 0; i < kW; ++i) v[i] = fma(c, v[i], sg * pc[i]);
  // the rotation maps the Gram block to diag(app - t gam, aqq + t gam) (as a_pp - t a_pq in
  // the two-sided method); the norms are recomputed exactly at every sweep start
  al = low ? fma(-t, gam, al) : fma(t, gam, al);
}

// Round r of a 12-column round-robin (circle method) sweep, for A <= 12: rows and columns
// 12..15 of G hold only padding (zero rows under the real columns, sigma e_j columns whose
// rotations are no-ops), so each round pairs the twelve real-or-padding columns below 12 --
// lane 11 with lane r, lane j with lane (2r - j) mod 11 otherwise -- and touches twelve rows.
// 11 rounds of 6 disjoint pairs replace the 15 xor matchings of the 16-column sweep. Lanes
// 12..15 pair with themselves (no-op).
__device__ __forceinline__ void hestenes_round12(double (&v)[kW], double& al, int j, int r, int group,
                                                 double tol2, bool& busy) {
  constexpr int kR = 12;
  const int pj = (j >= kR) ? j : (j == kR - 1) ? r : (j == r) ? kR - 1 : (2 * r - j + 2 * (kR - 1)) % (kR - 1);
  const int src = group * kW + pj;
  double pc[kR];
#pragma unroll
  for (int i = 0; i < kR; ++i) pc[i] = __shfl(v[i], src, kWave);
  const double be = __shfl(al, src, kWave);
  double gam = 0.0;
#pragma unroll
  for (int i = 0; i < kR; ++i) gam = fma(v[i], pc[i], gam);
  const bool self = pj == j;
  const bool low = j < pj;  // this lane holds the lower column index p
  const double app = low ? al : be, aqq = low ? be : al;
  const double d = aqq - app, h = self ? 0.0 : 2.0 * gam;
  const double x = fma(d, d, h * h);
  const double rr = x * rsq_nr_w<1>(x > 0.0 ? x : 1.0);
  const double den = fabs(d) + rr;
  double t = h * rcp_nr_w<1>(den > 0.0 ? den : 1.0);
  t = d < 0.0 ? -t : t;
  const double c = rsq_nr_w<2>(fma(t, t, 1.0));
  const double sn = t * c;
  busy |= !self && (gam * gam > tol2 * (app * aqq));
  const double sg = low ? -sn : sn;  // p: c g_p - s g_q ; q: s g_p + c g_q
#pragma unroll
  for (int i = 0; i < kR; ++i) v[i] = fma(c, v[i], sg * pc[i]);
  al = self ? al : (low ? fma(-t, gam, al) : fma(t, gam, al));
}

__device__ __forceinline__ double column_norm2_12(const double (&v)[kW]) {
  double a2 = 0.0;
#pragma unroll
  for (int i = 0; i < 12; ++i) a2 = fma(v[i], v[i], a2);
  return a2;
}

__device__ __forceinline__ double column_norm2(const double (&v)[kW]) {
  double a2 = 0.0;
#pragma unroll
  for (int i = 0; i < kW; ++i) a2 = fma(v[i], v[i], a2);
  return a2;
}

// One sample per wave; L is the wave's private LDS tile set. BF: ring products (and M) as
// BF16x3 on the bf16 matrix cores with FP32 accumulation, renormalised every bead.
template <bool PM, bool DBG, bool BF = false, bool C12 = false>
__device__ __forceinline__ void eval_wide_sample(const EvalArgs& args, int64_t sample, WideLds& L) {
  const int lane = threadIdx.x & (kWave - 1);
  const int g = lane / kW, j = lane % kW;
  const int A = args.A, AR = args.AR, N = args.N, P = args.P;
  const bool sep = AR != A;
  const bool jr = j < AR, jv = j < A;
  const double* __restrict__ qs = args.q + sample * (int64_t)N * P;
  const OSet sr = args.set[SET_RHO], sv = args.set[SET_VIB];
  const OSet sp = args.set[SET_VIB_PLUS], sm = args.set[SET_VIB_MINUS];
  const double2* __restrict__ LQ = args.wide_lq;
  const double* __restrict__ Ew = args.wide_e;
  const int* __restrict__ pairs = args.pair;
  const double* __restrict__ Qp = args.wide_qpair;
  const double tol2 = 1e-26;  // |g_p . g_q| <= 1e-13 |g_p| |g_q| for every pair ends the sweeps
  constexpr int kMaxSweeps = 16;

  // transposed ring products in accumulator layout: reg r of lane l = R^T[l/16 + 4r][l%16]
  d4 Rg, Rp, Rm;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    Rg[r] = (g + 4 * r == j) ? 1.0 : 0.0;
    Rp[r] = Rg[r];
    Rm[r] = Rg[r];
  }
  // BF16x3 ring states in the FP32 accumulator layout: reg r of lane l = R^T[4(l/16)+r][l%16]
  f4w Fg, Fp, Fm;
  int xg = 0, xp = 0, xm = 0;  // power-of-two exponents of the BF16x3 ring states
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    Fg[r] = (4 * g + r == j) ? 1.0f : 0.0f;
    Fp[r] = Fg[r];
    Fm[r] = Fg[r];
  }
  double rp = 1.0;  // rho: product over this group's beads of O_rho,j / S
  double smant = 1.0;
  int sexp = 0;

  for (int p0 = 0; p0 < P; p0 += kBeadsPerChunk) {
    const int p = p0 + g;
    const bool bead = p < P;  // uniform over the 16-lane group
    const int pc_ = bead ? p : P - 1;
    const int pn = (pc_ + 1 == P) ? 0 : pc_ + 1;

    double er = 0.0, ev = 0.0, ep = 0.0, em = 0.0;
    double v[kW];
#pragma unroll
    for (int i = 0; i < kW; ++i) v[i] = 0.0;
    for (int d = 0; d < N; ++d) {
      const double x0 = qs[d * P + pc_], xn = qs[d * P + pn];
      double q1 = x0, q1n = xn;
      if (jr) {  // sampling model: q -= d_rho; O_rho; q += d_rho (pimc.py:1091-1127)
        const double dd = sr.shift[j * N + d];
        const double x1 = x0 - dd, x2 = xn - dd;
        er += oterm_w(sr.coth[d], 2.0 * sr.csch[d], x1, x2);
        if (!sep) { q1 = x1 + dd; q1n = x2 + dd; }
      }
      if (jv) {  // coupled model at tau (and tau +- dbeta/P)
        const double dd = sv.shift[j * N + d];
        const double y1 = q1 - dd, y2 = q1n - dd;
        ev += oterm_w(sv.coth[d], 2.0 * sv.csch[d], y1, y2);
        if (PM) {
          const double z1 = (y1 + dd) - dd, z2 = (y2 + dd) - dd;
          ep += oterm_w(sp.coth[d], 2.0 * sp.csch[d], z1, z2);
          const double w1 = (z1 + dd) - dd, w2 = (z2 + dd) - dd;
          em += oterm_w(sm.coth[d], 2.0 * sm.csch[d], w1, w2);
        }
      }
      // column j of V (lower triangle mirrored): sum_d x_d (L + Qh x_d)  (pimc.py:1147-1160)
      const double2* lq = LQ + (size_t)d * kW * kW + j;
#pragma unroll
      for (int i = 0; i < kW; ++i) {
        if (i < A) {
          const double2 c = lq[i * kW];
          v[i] = fma(x0, fma(c.y, x0, c.x), v[i]);
        }
      }
    }
    for (int t = 0; t < args.npair; ++t) {  // off-diagonal quadratic mode pairs
      const int jj = pairs[2 * t], kk = pairs[2 * t + 1];
      const double xj = qs[jj * P + pc_], xk = qs[kk * P + pc_];
      const double* qp = Qp + (size_t)t * kW * kW + j;
#pragma unroll
      for (int i = 0; i < kW; ++i)
        if (i < A) v[i] = fma(xk * qp[i * kW], xj, v[i]);
    }
#pragma unroll
    for (int i = 0; i < kW; ++i)
      if (i < A) v[i] += Ew[i * kW + j];

    // O-factors and the per-bead scaling S_p = max over both O's (pimc.py:1076-1084)
    const double Or = jr ? exp(-0.5 * er) * sr.pref[j] : 0.0;
    const double Ov = jv ? exp(-0.5 * ev) * sv.pref[j] : 0.0;
    double Opl = 0.0, Omi = 0.0;
    if (PM && jv) {
      Opl = exp(-0.5 * ep) * sp.pref[j];
      Omi = exp(-0.5 * em) * sm.pref[j];
    }
    double S = group_max(fmax(Or, Ov));
    if (args.noscale) S = 1.0;

    if (DBG && bead) {
      const int rec = eval_debug_record(A, AR);
      double* dr = args.dbg + ((int64_t)sample * P + p) * rec;
      if (jr) dr[j] = Or;
      if (jv) dr[AR + j] = Ov;
      if (j == 0) dr[AR + A] = S;
    }

    // G = V + sigma I, one-sided Jacobi
    double f = 0.0;
#pragma unroll
    for (int i = 0; i < kW; ++i) f = fma(v[i], v[i], f);
    f = group_sum(f);
    const double sig = 1.5 * (f * rsq_nr_w(f > 0.0 ? f : 1.0)) + 1e-100;
#pragma unroll
    for (int i = 0; i < kW; ++i) v[i] += (i == j) ? sig : 0.0;
    double al = 0.0;
    if (C12) {  // A <= 12
      for (int sweep = 0; sweep < kMaxSweeps; ++sweep) {
        al = column_norm2_12(v);
        bool busy = false;
        for (int r = 0; r < 11; ++r) hestenes_round12(v, al, j, r, g, tol2, busy);
        if (!__any(busy)) break;
      }
    } else for (int sweep = 0; sweep < kMaxSweeps; ++sweep) {
      al = column_norm2(v);
      bool busy = false;
      hestenes_step<1>(v, al, j, tol2, busy);
      hestenes_step<2>(v, al, j, tol2, busy);
      hestenes_step<3>(v, al, j, tol2, busy);
      hestenes_step<4>(v, al, j, tol2, busy);
      hestenes_step<5>(v, al, j, tol2, busy);
      hestenes_step<6>(v, al, j, tol2, busy);
      hestenes_step<7>(v, al, j, tol2, busy);
      hestenes_step<8>(v, al, j, tol2, busy);
      hestenes_step<9>(v, al, j, tol2, busy);
      hestenes_step<10>(v, al, j, tol2, busy);
      hestenes_step<11>(v, al, j, tol2, busy);
      hestenes_step<12>(v, al, j, tol2, busy);
      hestenes_step<13>(v, al, j, tol2, busy);
      hestenes_step<14>(v, al, j, tol2, busy);
      hestenes_step<15>(v, al, j, tol2, busy);
      if (!__any(busy)) break;
    }
    al = column_norm2(v);
    const double rs = rsq_nr_w(al);
    const double lam = al * rs - sig;
    const double inv = 1.0 / S;

    wave_sync();  // the previous chunk's ring steps have read the tiles
#pragma unroll
    for (int i = 0; i < kW; ++i) L.U[g][i * kURow + j] = v[i] * rs;
    L.D[g][j] = exp(-args.tau * lam);
    L.O[0][g][j] = Ov * inv;
    if (PM) {
      L.O[1][g][j] = Opl * inv;
      L.O[2][g][j] = Omi * inv;
    }
    if (bead) {
      rp *= Or * inv;
      int e;
      smant = frexp(smant * S, &e);
      sexp += e;
    }
    wave_sync();

    // ring steps, beads in order (pimc.py:1198-1206)
    const int nb = min(kBeadsPerChunk, P - p0);
    for (int b = 0; b < nb; ++b) {
      if constexpr (BF) {
        // M = (U diag(d)) U^T: A[i][k] = U[i][k] d_k, B[k][i'] = U[i'][k], k = 4(l/16) + e
        s4w ah, al, bh, bl;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const int k = 4 * g + e;
          const double u = L.U[b][j * kURow + k];
          short h, l;
          split_bf((float)(u * L.D[b][k]), h, l);
          ah[e] = h;
          al[e] = l;
          split_bf((float)u, h, l);
          bh[e] = h;
          bl[e] = l;
        }
        const f4w M = mfma_bf16x3(ah, al, bh, bl, f4w{0.f, 0.f, 0.f, 0.f});
        if (DBG) {
          const int rec = eval_debug_record(A, AR);
          double* dr = args.dbg + ((int64_t)sample * P + p0 + b) * rec + AR + A + 1;
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int row = 4 * g + r;
            if (row < A && j < A) dr[row * A + j] = (double)M[r];
          }
        }
        // Z = M R^T: the A operand M[j][4(l/16)+e] is M's own register e (M symmetric), the
        // B operand is the ring state's register e
        s4w mh, ml;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          short h, l;
          split_bf(M[e], h, l);
          mh[e] = h;
          ml[e] = l;
        }
#pragma unroll
        for (int w = 0; w < (PM ? 3 : 1); ++w) {
          f4w& F = (w == 0) ? Fg : (w == 1) ? Fp : Fm;
          int& fe = (w == 0) ? xg : (w == 1) ? xp : xm;
          s4w rh, rl;
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            short h, l;
            split_bf(F[e], h, l);
            rh[e] = h;
            rl[e] = l;
          }
          const f4w Z = mfma_bf16x3(mh, ml, rh, rl, f4w{0.f, 0.f, 0.f, 0.f});
#pragma unroll
          for (int r = 0; r < 4; ++r) F[r] = Z[r] * (float)L.O[w][b][4 * g + r];
          renorm_wave(F, fe);
        }
      } else {
      d4 M = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int kb = 0; kb < 4; ++kb) {
        const int m = g + 4 * kb;
        const double u = L.U[b][j * kURow + m];
        M = __builtin_amdgcn_mfma_f64_16x16x4f64(u * L.D[b][m], u, M, 0, 0, 0);
      }
      if (DBG) {
        const int rec = eval_debug_record(A, AR);
        double* dr = args.dbg + ((int64_t)sample * P + p0 + b) * rec + AR + A + 1;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int row = g + 4 * r;
          if (row < A && j < A) dr[row * A + j] = M[r];
        }
      }
      d4 Z = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int kb = 0; kb < 4; ++kb) Z = __builtin_amdgcn_mfma_f64_16x16x4f64(M[kb], Rg[kb], Z, 0, 0, 0);
#pragma unroll
      for (int r = 0; r < 4; ++r) Rg[r] = Z[r] * L.O[0][b][g + 4 * r];
      if (PM) {
        d4 Zp = {0.0, 0.0, 0.0, 0.0}, Zm = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int kb = 0; kb < 4; ++kb) {
          Zp = __builtin_amdgcn_mfma_f64_16x16x4f64(M[kb], Rp[kb], Zp, 0, 0, 0);
          Zm = __builtin_amdgcn_mfma_f64_16x16x4f64(M[kb], Rm[kb], Zm, 0, 0, 0);
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          Rp[r] = Zp[r] * L.O[1][b][g + 4 * r];
          Rm[r] = Zm[r] * L.O[2][b][g + 4 * r];
        }
      }
      }
    }
  }

  // traces: FP64 layout: lane l holds (j, j) iff j % 4 == l / 16, in register j / 4;
  // FP32 layout: iff j / 4 == l / 16, in register j % 4
  double tg = 0.0, tp = 0.0, tm = 0.0;
  if constexpr (BF) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      if ((j >> 2) == g && (j & 3) == r) {
        tg = (double)Fg[r];
        tp = (double)Fp[r];
        tm = (double)Fm[r];
      }
    }
  } else {
    const bool diag = (j & 3) == g;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      if (diag && (j >> 2) == r) {
        tg = Rg[r];
        tp = Rp[r];
        tm = Rm[r];
      }
    }
  }
  double ls = log(smant) + (double)sexp * 0.69314718055994530942;
  // rho: product over the four bead groups, then sum over surfaces
  rp *= shfl_xor_d(rp, 16);
  rp *= shfl_xor_d(rp, 32);
  double rho = group_sum(jr ? rp : 0.0);
  ls += shfl_xor_d(ls, 16);
  ls += shfl_xor_d(ls, 32);
#pragma unroll
  for (int m = 1; m < kWave; m <<= 1) {
    tg += shfl_xor_d(tg, m);
    if (PM) {
      tp += shfl_xor_d(tp, m);
      tm += shfl_xor_d(tm, m);
    }
  }
  if (BF) {  // exponents are uniform over the wave
    tg = ldexp(tg, xg);
    tp = ldexp(tp, xp);
    tm = ldexp(tm, xm);
  }
  if (lane == 0) {
    const int64_t st = args.out_stride;
    args.out[OUT_RHO * st + sample] = rho;
    args.out[OUT_G * st + sample] = tg;
    if (PM) {
      args.out[OUT_GP * st + sample] = tp;
      args.out[OUT_GM * st + sample] = tm;
    }
    args.out[OUT_LOGS * st + sample] = ls;
  }
}

template <bool PM, bool DBG, bool BF, bool C12>
__global__ void __launch_bounds__(kWave * kWideWaves) k_eval_wide(const EvalArgs args) {
  __shared__ WideLds lds_all[kWideWaves];
  const int wv = threadIdx.x / kWave;
  const int64_t sample = (int64_t)blockIdx.x * kWideWaves + wv;
  if (sample >= args.B) return;  // whole wave; no work-group barrier is used
  eval_wide_sample<PM, DBG, BF, C12>(args, sample, lds_all[wv]);
}

// Persistent batched launch (see k_eval_batched): a tile is a range of samples ("waves") of
// one problem; problems may differ in A, A_rho, N and P (all runtime here).
template <bool PM>
__global__ void __launch_bounds__(kWave * kWideWaves)
k_eval_wide_batched(const EvalArgs* __restrict__ probs, const BatchTile* __restrict__ tiles, int ntiles,
                    int* counter) {
  __shared__ WideLds lds_all[kWideWaves];
  const int lane = threadIdx.x & (kWave - 1);
  WideLds& L = lds_all[threadIdx.x / kWave];
  for (;;) {
    int t = 0;
    if (lane == 0) t = atomicAdd(counter, 1);
    t = __builtin_amdgcn_readfirstlane(__shfl(t, 0, kWave));
    if (t >= ntiles) break;
    const BatchTile tile = tiles[t];
    const EvalArgs& a = probs[tile.prob];
    for (int64_t smp = tile.wave0; smp < tile.wave0 + tile.nwaves; ++smp) {
      if (smp < a.B) {
        if (a.A <= 12)
          eval_wide_sample<PM, false, false, true>(a, smp, L);
        else
          eval_wide_sample<PM, false, false, false>(a, smp, L);
      }
      wave_sync();
    }
  }
}

hipError_t launch_eval_wide_batched(int pm, const EvalArgs* probs, const BatchTile* tiles, int ntiles,
                                    int* counter, int waves, hipStream_t s) {
  const dim3 grid((unsigned)((waves + kWideWaves - 1) / kWideWaves));
  if (pm)
    hipLaunchKernelGGL((k_eval_wide_batched<true>), grid, dim3(kWave * kWideWaves), 0, s, probs, tiles,
                       ntiles, counter);
  else
    hipLaunchKernelGGL((k_eval_wide_batched<false>), grid, dim3(kWave * kWideWaves), 0, s, probs, tiles,
                       ntiles, counter);
  return hipGetLastError();
}

hipError_t launch_eval_wide(int pm, const EvalArgs& a, hipStream_t s, bool bf16x3) {
  const dim3 grid((unsigned)((a.B + kWideWaves - 1) / kWideWaves));
  const dim3 block(kWave * kWideWaves);
  const bool dbg = a.dbg != nullptr;
  const bool c12 = a.A <= 12;
#define PIB_WIDE(pm_, dbg_, bf_) \
  if ((pm != 0) == pm_ && dbg == dbg_ && bf16x3 == bf_) { \
    if (c12) hipLaunchKernelGGL((k_eval_wide<pm_, dbg_, bf_, true>), grid, block, 0, s, a); \
    else hipLaunchKernelGGL((k_eval_wide<pm_, dbg_, bf_, false>), grid, block, 0, s, a); \
    return hipGetLastError(); \
  }
  PIB_WIDE(false, false, false) PIB_WIDE(true, false, false) PIB_WIDE(false, true, false)
  PIB_WIDE(true, true, false) PIB_WIDE(false, false, true) PIB_WIDE(true, false, true)
  PIB_WIDE(false, true, true) PIB_WIDE(true, true, true)
#undef PIB_WIDE
  return hipErrorInvalidValue;
}

}  // namespace pib
