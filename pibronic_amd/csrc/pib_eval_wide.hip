// FP64 per-sample evaluation for large surface counts (A <= 16, padded to 16), one wave per
// sample.
//
// Same estimator as k_eval_fp64 (pibronic/pimc/pimc.py:1087-1213 fused; block loops
// :1341-1462), mapped for matrices that do not fit one lane's registers:
//
//  * Beads are processed four at a time. For the O-factors lane (g, j) = (lane / 16, lane % 16)
//    takes surface j of bead 4c + g: it forms the O-factor of surface j for every temperature
//    set and takes part in the per-bead scaling maximum over its 16-lane group.
//  * The coupling matrices V of the four beads are assembled directly in the FP64 matrix-core
//    accumulator layout: register r of lane l holds X[l/16 + 4r][l%16] (measured on gfx950).
//    For a symmetric X that layout is also the operand layout of a product X Y: with the sum
//    index of tile step k taken as lane/16 + 4k, lane l supplies X[l%16][l/16 + 4k] =
//    X[l/16 + 4k][l%16] (its own register k) as the A operand and Y[l/16 + 4k][l%16] (register
//    k of Y) as the B operand. Every product below chains in registers that way.
//  * M = exp(-tau V) directly, without eigenvectors: the reference forms
//    M = U diag(exp(-tau lambda)) U^T from eigh(V) (pimc.py:1167-1187), and M is the only
//    consumer of the decomposition. X = -tau V is scaled by 2^-s, a degree-16 polynomial equal
//    to the Taylor polynomial through X^15 is evaluated in four products (expm_sym) and
//    squared s times. The four beads of a chunk go through it in lockstep (one s, the largest
//    they need), so four independent product chains fill the matrix-core pipeline. V is
//    symmetric, so every power and polynomial of it is, and the accumulator layout of each
//    product feeds the next one.
//  * The ring product runs on the same FP64 matrix cores, carried transposed,
//    R^T <- diag(O_p / S_p) M_p R^T, whose trace equals the reference's tr(prod_p M_p O_p)
//    because M_p is symmetric and O_p diagonal.
//
// Padding: rows and columns j >= A of V are zero, the identity of the Taylor polynomial covers
// only the leading A x A block, so M is zero outside it; padded surfaces carry O = 0 and drop
// out of every trace.
#include "pib_common.hpp"
#include "pib_kernels.hpp"

namespace pib {

typedef double d4 __attribute__((ext_vector_type(4)));

constexpr int kW = 16;                 // padded matrix dimension = lanes per bead
constexpr int kBeadsPerChunk = kWave / kW;
constexpr int kWideWaves = 2;          // waves (samples) per work-group

struct WideLds {
  double X[kBeadsPerChunk][4][kWave];  // -tau V of the chunk's beads, register r of lane l
  double O[3][kBeadsPerChunk][kW];     // scaled O-factors (vib, vib+, vib-)
  float Mf[kW][kW + 1];                // M in FP32 for the BF16x3 ring (layout change)
};

// Wave-local ordering of LDS traffic: every access of the wave before this point is complete
// and visible to every lane of the same wave after it.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// sum / max over the 16 lanes of a group, and sum over the wave; the xor butterfly gives every
// lane the same bits (a + b == b + a exactly)
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
__device__ __forceinline__ double wave_sum(double x) {
#pragma unroll
  for (int m = 1; m < kWave; m <<= 1) x += shfl_xor_d(x, m);
  return x;
}

// Sum over the wave with the same bits on every lane, without LDS: within each row of 16
// lanes the quad permutations [1,0,3,2] and [2,3,0,1], the half-row mirror and the row mirror
// give every lane its row's sum (the two lanes of each exchange add the same two values); the
// four row sums are read from lanes 0, 16, 32 and 48 and added in a fixed order.
template <int CTRL>
__device__ __forceinline__ double dpp_move_d(double x) {
  const long long u = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)u, CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(u >> 32), CTRL, 0xF, 0xF, false);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__device__ __forceinline__ double readlane_d(double x, int l) {
  const long long u = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_readlane((int)u, l);
  const int hi = __builtin_amdgcn_readlane((int)(u >> 32), l);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__device__ __forceinline__ double wave_sum_dpp(double x) {
  x += dpp_move_d<0xB1>(x);   // quad_perm [1,0,3,2]
  x += dpp_move_d<0x4E>(x);   // quad_perm [2,3,0,1]
  x += dpp_move_d<0x141>(x);  // row_half_mirror
  x += dpp_move_d<0x140>(x);  // row_mirror
  return (readlane_d(x, 0) + readlane_d(x, 16)) + (readlane_d(x, 32) + readlane_d(x, 48));
}

// Z + X Y for a symmetric X, all in the FP64 accumulator layout. Tile step k sums over the
// indices 4k..4k+3; with NR < 4 the indices from 4 NR on are zero padding of both operands and
// their steps are skipped.
template <int NR = 4>
__device__ __forceinline__ d4 sym_mul(const d4& X, const d4& Y, d4 Z) {
#pragma unroll
  for (int k = 0; k < NR; ++k) Z = __builtin_amdgcn_mfma_f64_16x16x4f64(X[k], Y[k], Z, 0, 0, 0);
  return Z;
}
template <int NR = 4>
__device__ __forceinline__ d4 sym_mul(const d4& X, const d4& Y) {
  return sym_mul<NR>(X, Y, d4{0.0, 0.0, 0.0, 0.0});
}

constexpr double kLog2Theta15 = -0.5640051;  // log2(0.67642)

// exp(X) for a symmetric X in the FP64 accumulator layout, zero outside its leading n x n
// block, n <= 4 NR: scaling and squaring of the degree-16 polynomial that is the Taylor
// polynomial through X^15, in four products (the per-lane kernel's form, after Sastre 2018):
//   y02 = X2 (c1 X2 + c2 X),  y12 = (y02 + c3 X2 + c4 X)(y02 + c5 X2) + c6 y02 + c7 X2,
//   T = (y12 + c8 X2 + c9 X)(y12 + c10 y02 + c11 X) + c12 y12 + c13 y02 + c14 X2 + X + I,
// with y02 held as G = y02 + c5 X2. alpha = |X^2|_F^(1/2) >= the spectral radius of a symmetric
// X; alpha 2^-s <= theta_15+ = 0.6764, where the form's relative backward error is below the
// FP64 unit roundoff. On the A = 12 workload matrices a chunk of four beads squares 2.1 times
// instead of 3.1 with degree-12 Paterson-Stockmeyer in X^4: 6.1 products per bead instead of
// 8.1. Every factor is a polynomial in X, hence symmetric, as sym_mul needs of its first
// operand; only the identity term touches the diagonal, and only in the leading block, so the
// result is zero outside it.
template <int NR>
__device__ __forceinline__ d4 expm_sym(const d4& X, int n, int g, int j) {
  constexpr double c1 = -4.01876161020103546293293287596e-4, c2 = -2.94553144027968298047549126357e-3;
  constexpr double c3 = 8.70906657683767596765346131523e-3, c4 = -4.01756844067356780152362494573e-1;
  constexpr double c5 = -3.23076288812231208197540935626e-2, c6 = 2.33731940471157926649827413344e-2;
  constexpr double c7 = 2.61492797729811710806033238185e-1, c8 = -2.38107037387098722469262858218e-1;
  constexpr double c9 = -4.1302763659297829110328428344e-2, c10 = -5.79236170707326052179805145605e0;
  constexpr double c11 = 2.22420917249637356119032703879e0, c12 = 1.04080173523135436461532671084e1;
  constexpr double c13 = 3.03012340073871230596564458406e0, c14 = -2.12975559049643578453357460382e0;
  constexpr double d3 = c3 - c5, e7 = c7 - c6 * c5, f10 = -c10 * c5, g14 = c14 - c13 * c5;
  const bool dl = ((j & 3) == g) && (j < n);  // register j / 4 of this lane holds X[j][j]
  const int dr = j >> 2;
  const d4 X2 = sym_mul<NR>(X, X);
  double n2 = 0.0;
#pragma unroll
  for (int r = 0; r < 4; ++r) n2 = fma(X2[r], X2[r], n2);
  n2 = wave_sum_dpp(n2);
  // smallest s >= 0 with |X^2|_F^(1/2) 2^-s <= theta_15+; a zero matrix gives -inf, a NaN s = 0,
  // the count is capped at 64
  const double sd = 0.25 * log2_coarse(n2) - kLog2Theta15;
  int s = 0;
  if (sd > 0.0) s = sd < 64.0 ? (int)ceil(sd) : 64;
  s = __builtin_amdgcn_readfirstlane(s);  // identical on every lane
  d4 x1, x2, H;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    x1[r] = ldexp(X[r], -s);
    x2[r] = ldexp(X2[r], -2 * s);
    H[r] = fma(c1, x2[r], c2 * x1[r]);
  }
  const d4 G = sym_mul<NR>(x2, H, c5 * x2);  // y02 + c5 X2
  d4 F, acc;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    F[r] = fma(c4, x1[r], fma(d3, x2[r], G[r]));  // y02 + c3 X2 + c4 X
    acc[r] = fma(c6, G[r], e7 * x2[r]);
  }
  const d4 Y = sym_mul<NR>(F, G, acc);  // y12
  d4 F1, F2, R;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const double I = (dl && r == dr) ? 1.0 : 0.0;
    F1[r] = fma(c8, x2[r], fma(c9, x1[r], Y[r]));
    F2[r] = fma(c10, G[r], fma(c11, x1[r], fma(f10, x2[r], Y[r])));
    R[r] = fma(c12, Y[r], fma(c13, G[r], fma(g14, x2[r], x1[r]))) + I;
  }
  d4 T = sym_mul<NR>(F1, F2, R);
  for (int k = 0; k < s; ++k) T = sym_mul<NR>(T, T);
  return T;
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
// the three partial products go to separate zero accumulators and are added on the VALU: a
// chain through the accumulator lets the register allocator give a product a destination that
// partly overlaps its accumulator input, which the matrix cores do not compute correctly
__device__ __forceinline__ f4w mfma_bf16x3(const s4w& ah, const s4w& al, const s4w& bh, const s4w& bl) {
  const f4w z = {0.f, 0.f, 0.f, 0.f};
  const f4w lo1 = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(al, bh, z, 0, 0, 0);
  const f4w lo2 = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah, bl, z, 0, 0, 0);
  const f4w hi = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah, bh, z, 0, 0, 0);
  return (lo1 + lo2) + hi;
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

// One sample per wave; L is the wave's private LDS tile set. BF: ring products as BF16x3 on
// the bf16 matrix cores with FP32 accumulation, renormalised every bead (M stays FP64).
// One sample per wave; L is the wave's private LDS tile set. BF: ring products as BF16x3 on
// the bf16 matrix cores with FP32 accumulation, renormalised every bead (M stays FP64). NR:
// rows g + 4r, r < NR, of V are assembled (NR = ceil(A / 4); rows >= A are zero padding).
template <bool PM, bool DBG, bool BF, int NR>
__device__ __forceinline__ void eval_wide_sample(const EvalArgs& args, int64_t sample, WideLds& L) {
  const int lane = threadIdx.x & (kWave - 1);
  const int g = lane / kW, j = lane % kW;
  const int A = args.A, AR = args.AR, N = args.N, P = args.P;
  const bool jr = j < AR, jv = j < A;
  const int jrc = jr ? j : AR - 1, jvc = jv ? j : A - 1;  // in-bounds surface of this lane
  const double* __restrict__ qs = args.q + sample * (int64_t)N * P;
  const OSet sr = args.set[SET_RHO], sv = args.set[SET_VIB];
  const OSet sp = args.set[SET_VIB_PLUS], sm = args.set[SET_VIB_MINUS];
  const ModeRec R(A, AR);
  const int rl = R.len();
  const double* __restrict__ tab = args.mode_tab;
  const double2* __restrict__ LQ = args.wide_lq;
  const double* __restrict__ Ew = args.wide_e;
  const int* __restrict__ pairs = args.pair;
  const double* __restrict__ Qp = args.wide_qpair;
  const double ntau = -args.tau;

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
    const int nb = min(kBeadsPerChunk, P - p0);
    const int p = p0 + g;
    const bool bead = p < P;  // uniform over the 16-lane group
    const int pc_ = bead ? p : P - 1;
    const int pn = (pc_ + 1 == P) ? 0 : pc_ + 1;
    int pb[kBeadsPerChunk];  // the chunk's beads (clamped; results past P are not used)
#pragma unroll
    for (int b = 0; b < kBeadsPerChunk; ++b) pb[b] = min(p0 + b, P - 1);

    // O-exponents in the expanded form of the lanes kernel (see OSet): surface-independent
    // parts t (identical on the lanes of a group) and this lane's surface's linear part e
    double tr = 0.0, tv = 0.0, tp = 0.0, tm = 0.0;
    double er = 0.0, ev = 0.0, ep = 0.0, em = 0.0;
    d4 v[kBeadsPerChunk];  // V of the chunk's beads, accumulator layout
#pragma unroll
    for (int b = 0; b < kBeadsPerChunk; ++b) v[b] = d4{0.0, 0.0, 0.0, 0.0};
    // everything mode d needs, loaded while mode d - 1's arithmetic runs
    struct ModeIn {
      double x0, xn;                  // this group's bead and its ring neighbour
      double xb[kBeadsPerChunk];      // the chunk's beads (uniform)
      double c0, c1, c2, c3;          // th, csch of the rho and vib sets (uniform)
      double p0, p1, p2, p3;          // th, csch of the vib+ and vib- sets (uniform)
      double lr, lv, lp, lm;          // this lane's surface: lin of rho, vib, vib+, vib-
      double2 c[NR];                  // {L, Qh} at rows g + 4r, column j
    };
    auto load_mode = [&](int d) {
      ModeIn m;
      const double* __restrict__ t = tab + d * rl;
      const double* qd = qs + d * P;
      m.x0 = qd[pc_];
      m.xn = qd[pn];
#pragma unroll
      for (int b = 0; b < kBeadsPerChunk; ++b) m.xb[b] = qd[pb[b]];
      m.c0 = t[0];
      m.c1 = t[1];
      m.c2 = t[2];
      m.c3 = t[3];
      m.lr = t[R.lin_rho() + jrc];
      m.lv = t[R.lin_vib() + jvc];
      if (PM) {
        m.p0 = t[R.pm() + 0];
        m.p1 = t[R.pm() + 1];
        m.p2 = t[R.pm() + 2];
        m.p3 = t[R.pm() + 3];
        m.lp = t[R.lin_plus() + jvc];
        m.lm = t[R.lin_minus() + jvc];
      }
#pragma unroll
      for (int r = 0; r < NR; ++r) m.c[r] = LQ[((size_t)d * kW + g + 4 * r) * kW + j];
      return m;
    };
    ModeIn cur = load_mode(0);
    for (int d = 0; d < N; ++d) {
      const ModeIn nxt = load_mode(d + 1 < N ? d + 1 : d);
      const double sx = cur.x0 + cur.xn, dx = cur.x0 - cur.xn;
      const double q2s = fma(cur.x0, cur.x0, cur.xn * cur.xn), d2 = dx * dx;
      tr = fma(cur.c1, d2, fma(cur.c0, q2s, tr));
      tv = fma(cur.c3, d2, fma(cur.c2, q2s, tv));
      er = fma(cur.lr, sx, er);
      ev = fma(cur.lv, sx, ev);
      if (PM) {
        tp = fma(cur.p1, d2, fma(cur.p0, q2s, tp));
        tm = fma(cur.p3, d2, fma(cur.p2, q2s, tm));
        ep = fma(cur.lp, sx, ep);
        em = fma(cur.lm, sx, em);
      }
      // V[i][j] += x_d (L + Qh x_d), i = g + 4r (pimc.py:1147-1160)
#pragma unroll
      for (int r = 0; r < NR; ++r) {
#pragma unroll
        for (int b = 0; b < kBeadsPerChunk; ++b)
          v[b][r] = fma(cur.xb[b], fma(cur.c[r].y, cur.xb[b], cur.c[r].x), v[b][r]);
      }
      cur = nxt;
    }
    for (int t = 0; t < args.npair; ++t) {  // off-diagonal quadratic mode pairs
      const int jj = pairs[2 * t], kk = pairs[2 * t + 1];
      double xj[kBeadsPerChunk], xk[kBeadsPerChunk];
#pragma unroll
      for (int b = 0; b < kBeadsPerChunk; ++b) {
        xj[b] = qs[jj * P + pb[b]];
        xk[b] = qs[kk * P + pb[b]];
      }
      const double* qp = Qp + (size_t)t * kW * kW + g * kW + j;
#pragma unroll
      for (int r = 0; r < NR; ++r) {
        const double cq = qp[4 * r * kW];
#pragma unroll
        for (int b = 0; b < kBeadsPerChunk; ++b) v[b][r] = fma(xk[b] * cq, xj[b], v[b][r]);
      }
    }
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      const double e = Ew[(g + 4 * r) * kW + j];
#pragma unroll
      for (int b = 0; b < kBeadsPerChunk; ++b) v[b][r] += e;
    }

    // O-factors and the per-bead scaling S_p = max over both O's (pimc.py:1076-1084)
    const double Or = jr ? exp(-0.5 * ((tr + er) + sr.k2[jrc])) * sr.pref[jrc] : 0.0;
    const double Ov = jv ? exp(-0.5 * ((tv + ev) + sv.k2[jvc])) * sv.pref[jvc] : 0.0;
    double Opl = 0.0, Omi = 0.0;
    if (PM && jv) {
      Opl = exp(-0.5 * ((tp + ep) + sp.k2[jvc])) * sp.pref[jvc];
      Omi = exp(-0.5 * ((tm + em) + sm.k2[jvc])) * sm.pref[jvc];
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

    const double inv = 1.0 / S;
    wave_sync();  // the previous chunk's ring steps have read the X and O tiles
#pragma unroll
    for (int b = 0; b < kBeadsPerChunk; ++b) {
#pragma unroll
      for (int r = 0; r < 4; ++r) L.X[b][r][lane] = ntau * v[b][r];
    }
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

    // M_p = exp(-tau V_p) and the ring steps, beads in order (pimc.py:1181-1206); one bead's
    // exponential at a time keeps the register count (and so the waves per SIMD) in check
#pragma clang loop unroll(disable)
    for (int b = 0; b < nb; ++b) {
      {
        d4 X;
#pragma unroll
        for (int r = 0; r < 4; ++r) X[r] = L.X[b][r][lane];
        const d4 M = expm_sym<NR>(X, A, g, j);
        if (DBG) {
          const int rec = eval_debug_record(A, AR);
          double* dr = args.dbg + ((int64_t)sample * P + p0 + b) * rec + AR + A + 1;
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int row = g + 4 * r;
            if (row < A && j < A) dr[row * A + j] = M[r];
          }
        }
        if constexpr (BF) {
          // M into the FP32 layout (reg e = M[4(l/16)+e][l%16] = M[l%16][4(l/16)+e]), split
#pragma unroll
          for (int r = 0; r < 4; ++r) L.Mf[g + 4 * r][j] = (float)M[r];
          wave_sync();
          s4w mh, ml;
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            short h, l;
            split_bf(L.Mf[4 * g + e][j], h, l);
            mh[e] = h;
            ml[e] = l;
          }
          wave_sync();  // every lane has read Mf before the next bead writes it
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
            const f4w Z = mfma_bf16x3(mh, ml, rh, rl);
#pragma unroll
            for (int r = 0; r < 4; ++r) F[r] = Z[r] * (float)L.O[w][b][4 * g + r];
            renorm_wave(F, fe);
          }
        } else {
          // sum index >= 4 NR: M's padding columns are zero
          const d4 Z = sym_mul<NR>(M, Rg);
#pragma unroll
          for (int r = 0; r < 4; ++r) Rg[r] = Z[r] * L.O[0][b][g + 4 * r];
          if (PM) {
            const d4 Zp = sym_mul<NR>(M, Rp), Zm = sym_mul<NR>(M, Rm);
#pragma unroll
            for (int r = 0; r < 4; ++r) {
              Rp[r] = Zp[r] * L.O[1][b][g + 4 * r];
              Rm[r] = Zm[r] * L.O[2][b][g + 4 * r];
            }
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
  tg = wave_sum(tg);
  if (PM) {
    tp = wave_sum(tp);
    tm = wave_sum(tm);
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

// Occupancy: the plain kernel for A <= 12 fits 168 VGPRs (three waves per SIMD; 7.3 against
// 9.3 ms at the compiler's default register budget, A=12, N=24, P=128, measured on MI355X);
// the +-dbeta ring states and the other variants keep two waves, where they do not spill.
template <bool PM, bool DBG, bool BF, int NR>
__global__ void __launch_bounds__(kWave * kWideWaves)
__attribute__((amdgpu_waves_per_eu((PM || DBG || BF || NR == 4) ? 2 : 3)))
k_eval_wide(const EvalArgs args) {
  __shared__ WideLds lds_all[kWideWaves];
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
  const int64_t sample = (int64_t)blockIdx.x * kWideWaves + wv;
  if (sample >= args.B) return;  // whole wave; no work-group barrier is used
  eval_wide_sample<PM, DBG, BF, NR>(args, sample, lds_all[wv]);
}

// Persistent batched launch (see k_eval_batched): a tile is a range of samples ("waves") of
// one problem; problems may differ in A, A_rho, N and P (all runtime here).
template <bool PM>
__global__ void __launch_bounds__(kWave * kWideWaves) __attribute__((amdgpu_waves_per_eu(2)))
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
      if (smp < a.B) eval_wide_sample<PM, false, false, 4>(a, smp, L);
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
  // rows of V assembled: ceil(A / 4) in the plain and +-dbeta kernels, all 16 otherwise
  const int nr = (a.A + 3) / 4;
  if (!dbg && !bf16x3) {
#define PIB_WIDE_NR(pm_, nr_) \
  if ((pm != 0) == pm_ && nr == nr_) { \
    hipLaunchKernelGGL((k_eval_wide<pm_, false, false, nr_>), grid, block, 0, s, a); \
    return hipGetLastError(); \
  }
    PIB_WIDE_NR(false, 1) PIB_WIDE_NR(false, 2) PIB_WIDE_NR(false, 3) PIB_WIDE_NR(false, 4)
    PIB_WIDE_NR(true, 1) PIB_WIDE_NR(true, 2) PIB_WIDE_NR(true, 3) PIB_WIDE_NR(true, 4)
#undef PIB_WIDE_NR
    return hipErrorInvalidValue;
  }
#define PIB_WIDE(pm_, dbg_, bf_) \
  if ((pm != 0) == pm_ && dbg == dbg_ && bf16x3 == bf_) { \
    hipLaunchKernelGGL((k_eval_wide<pm_, dbg_, bf_, 4>), grid, block, 0, s, a); \
    return hipGetLastError(); \
  }
  PIB_WIDE(false, true, false) PIB_WIDE(true, true, false) PIB_WIDE(false, false, true) PIB_WIDE(true, false, true)
  PIB_WIDE(false, true, true) PIB_WIDE(true, true, true)
#undef PIB_WIDE
  return hipErrorInvalidValue;
}

}  // namespace pib
