// FP64 per-sample evaluation of the PIMC estimator (one launch per block of samples).
//
// Replaces the body of block_compute / block_compute_pm after the coordinate transform
// (pibronic/pimc/pimc.py:1366-1381, 1421-1449): build_o_matrix (:1087-1129),
// build_scaling_factors + scale_o_matrices (:1062-1084), build_denominator (:1132-1136),
// diagonalize_coupling_matrix (:1139-1174) and build_numerator (:1177-1213), fused.
//
// Mapping (gfx950, wave64): G lanes cooperate on one sample, 64/G samples per wave. Lane g
// owns the contiguous bead chunk [g*K, min((g+1)*K, P)). Per bead it forms the diagonal
// O-factors of every temperature set, the residual coupling V in registers, M = exp(-tau V)
// directly (a scaled and squared Taylor polynomial, per lane, A x A in registers; the
// reference's eigendecomposition serves only to form M) and multiplies M * diag(O_vib / S_p)
// into its running ring product. The G partial products are
// then combined by a log2(G)-level ordered tree of A x A products over cross-lane moves
// (chunk 0 * chunk 1 * ... keeps the reference's left-to-right order).
//
// Numerics: the O-matrix exponent sum_j coth (x1^2 + x2^2) - 2 csch x1 x2 (x1, x2 the shifted
// coordinates of beads p, p+1) is evaluated as
//   sum_j th (x^2 + x'^2) + csch (x - x')^2 + lin_a (x + x') + k2_a,   th = coth - csch,
// which needs one FMA per surface and mode instead of nine operations. The reference's
// expression cancels when tau*omega is small (coth*x^2 ~ 1e3 against a result ~1e1); this
// form does not, so it sits about 100x closer to the exact value of the same rounded
// constants, and the O-factors differ from the reference's by the reference's own rounding
// (<= 1e-11 relative at A=12, N=24, P=128; ~1e-13 at the small workloads). V is built from
// the unshifted coordinates (the reference's shift/unshift leaves an ulp of drift in q,
// which moves V by ~1e-16). Linear algebra (V, M, ring) uses explicit fma().
#include <type_traits>

#include "pib_common.hpp"
#include "pib_kernels.hpp"

namespace pib {

// ---------------------------------------------------------------------------------------
// M = exp(X) for a symmetric A x A matrix held in one lane's registers, X = -tau V.
//
// The reference forms M = U diag(exp(-tau lambda)) U^T from eigh(V) (pimc.py:1167-1187); M is
// the only consumer of the decomposition, so it is computed directly: scaling and squaring of
// a Taylor polynomial, in one of three evaluation schemes (expm_lane picks one per kernel). The
// first is the degree-12 Taylor polynomial T_12, evaluated by Paterson-Stockmeyer in X^3:
//   T_12 = ((B3 X^3 + B2) X^3 + B1) X^3 + B0,  B_n = c_3n I + c_3n+1 X + c_3n+2 X^2 (n < 3),
//   B3 = c9 I + c10 X + c11 X^2 + c12 X^3,
// five products (X^2, X^3, three Horner steps) against seven for Horner's rule at degree 8.
// X is scaled by 2^-s so that alpha 2^-s <= theta_12 = 0.2996 (Sastre et al. 2015: the backward
// error of T_12 below the unit roundoff), alpha = min(|X|_F, |X^3|_F^(1/3)) >= the spectral
// radius (X^2 and X^3 are formed before scaling; scaling by powers of two is exact). For a
// symmetric X the 2-norm error of a polynomial in X is its largest error over the eigenvalues.
// T_12 is then squared s times: on the workload matrices 2.0 squarings per wave at A = 4
// instead of 5.0 at degree 8 with the Frobenius bound (7.0 instead of 12.0 products per bead).
// Every intermediate is a polynomial in X, hence symmetric, so products form the upper
// triangle only. s differs between lanes; the squaring loop runs the wave's largest count.
// Z = X Y + d I for commuting symmetric X, Y (Z symmetric; Z must not alias X or Y)
template <int A>
__device__ __forceinline__ void sym_prod_lane(const double (&X)[A][A], const double (&Y)[A][A], double d,
                                              double (&Z)[A][A]) {
#pragma unroll
  for (int i = 0; i < A; ++i)
#pragma unroll
    for (int k = i; k < A; ++k) {
      double s = (i == k) ? fma(X[i][0], Y[0][k], d) : X[i][0] * Y[0][k];
#pragma unroll
      for (int c = 1; c < A; ++c) s = fma(X[i][c], Y[c][k], s);
      Z[i][k] = s;
      Z[k][i] = s;
    }
}

// Z = T Y + c0 I + c1 X + c2 X2 for commuting symmetric T, Y, X, X2 (Z symmetric, no aliasing)
template <int A>
__device__ __forceinline__ void sym_prod_add(const double (&T)[A][A], const double (&Y)[A][A], double c0,
                                             double c1, const double (&X)[A][A], double c2,
                                             const double (&X2)[A][A], double (&Z)[A][A]) {
#pragma unroll
  for (int i = 0; i < A; ++i)
#pragma unroll
    for (int k = i; k < A; ++k) {
      double s = fma(c2, X2[i][k], c1 * X[i][k]);
      if (i == k) s += c0;
#pragma unroll
      for (int c = 0; c < A; ++c) s = fma(T[i][c], Y[c][k], s);
      Z[i][k] = s;
      Z[k][i] = s;
    }
}

// T <- T^(2^s) for a symmetric T. Two squarings per iteration (T -> Z -> T), so that no
// register copy follows each one; an odd count ends with one squaring and a copy. s differs
// between lanes; the loop runs the wave's largest count.
template <int A>
__device__ __forceinline__ void square_times(double (&T)[A][A], int s) {
  double Z[A][A];
  for (int k = 0; k + 1 < s; k += 2) {
    sym_prod_lane<A>(T, T, 0.0, Z);
    sym_prod_lane<A>(Z, Z, 0.0, T);
  }
  if (s & 1) {
    sym_prod_lane<A>(T, T, 0.0, Z);
#pragma unroll
    for (int i = 0; i < A; ++i)
#pragma unroll
      for (int j = 0; j < A; ++j) T[i][j] = Z[i][j];
  }
}

constexpr double kLog2Theta12 = -1.738865;  // log2(0.2996)

template <int A>
__device__ __forceinline__ void expm_lane_ps12(double (&X)[A][A], double (&T)[A][A]) {
  double X2[A][A], X3[A][A];
  sym_prod_lane<A>(X, X, 0.0, X2);
  sym_prod_lane<A>(X2, X, 0.0, X3);
  double nf = 0.0, n3 = 0.0;
#pragma unroll
  for (int i = 0; i < A; ++i) {
    nf = fma(X[i][i], X[i][i], nf);
    n3 = fma(X3[i][i], X3[i][i], n3);
#pragma unroll
    for (int k = i + 1; k < A; ++k) {
      nf = fma(2.0 * X[i][k], X[i][k], nf);
      n3 = fma(2.0 * X3[i][k], X3[i][k], n3);
    }
  }
  const double la = fmin(0.5 * log2_coarse(nf), log2_coarse(n3) / 6.0);
  const double sd = la - kLog2Theta12;
  int s = 0;
  if (sd > 0.0) s = sd < 64.0 ? (int)ceil(sd) : 64;
#pragma unroll
  for (int i = 0; i < A; ++i)
#pragma unroll
    for (int k = i; k < A; ++k) {
      X[i][k] = ldexp(X[i][k], -s);
      X[k][i] = X[i][k];
      X2[i][k] = ldexp(X2[i][k], -2 * s);
      X2[k][i] = X2[i][k];
      X3[i][k] = ldexp(X3[i][k], -3 * s);
      X3[k][i] = X3[i][k];
    }
  constexpr double c[13] = {1.0, 1.0, 1.0 / 2, 1.0 / 6, 1.0 / 24, 1.0 / 120, 1.0 / 720, 1.0 / 5040,
                            1.0 / 40320, 1.0 / 362880, 1.0 / 3628800, 1.0 / 39916800, 1.0 / 479001600};
#pragma unroll
  for (int i = 0; i < A; ++i)
#pragma unroll
    for (int k = 0; k < A; ++k) {
      double v = fma(c[12], X3[i][k], fma(c[11], X2[i][k], c[10] * X[i][k]));
      T[i][k] = (i == k) ? v + c[9] : v;
    }
#pragma unroll
  for (int n = 2; n >= 0; --n) {
    double Z[A][A];
    sym_prod_add<A>(T, X3, c[3 * n], c[3 * n + 1], X, c[3 * n + 2], X2, Z);
#pragma unroll
    for (int i = 0; i < A; ++i)
#pragma unroll
      for (int k = 0; k < A; ++k) T[i][k] = Z[i][k];
  }
  square_times<A>(T, s);
}

// theta_8 = 0.049: below the root of |y|^8 / 9! = u (the degree-8 polynomial's truncation)
constexpr double kLog2Theta8 = -4.3511;  // log2(0.049)

// The degree-8 Taylor polynomial with three matrix products (Horner's rule takes seven), in the
// evaluation form of Sastre (J. Sastre, "Efficient evaluation of matrix polynomials", Linear
// Algebra Appl. 539 (2018) 229-250):
//   Y  = X2 (c1 X2 + c2 X),
//   T8 = (Y + c3 X2 + c4 X)(Y + c5 X2) + c6 Y + X2/2 + X + I,
// whose coefficients solve the six equations that make the expansion's X^3..X^8 terms 1/k!
// (c1^2 = 1/8!, 2 c1 c2 = 1/7!, ...; of the two real solution pairs, the one with the smaller
// c6). Y's product is taken as two dot products per entry (X2 X2 and X2 X), so that no fourth
// matrix is live beside X, X2 and Y; the three factors of the last product then replace X, X2
// and Y entry by entry. alpha = |X^2|_F^(1/2), a tighter bound on the spectral radius of a
// symmetric X than |X|_F (the root of the largest of sum lambda^4), from X2, which the
// polynomial needs anyway. For |X| <= theta_8 the terms beyond I + X + X^2/2 are below 1e-5 of
// the result, so the form rounds like Horner's rule.
template <int A>
__device__ __forceinline__ void expm_lane_s8(double (&X)[A][A], double (&T)[A][A]) {
  constexpr double c1 = 4.98011920555997349987007158e-3, c2 = 1.99204768222398939994802863e-2;
  constexpr double c3 = 7.66526532111914669031909419e-2, c4 = 8.76500980178555335977132598e-1;
  constexpr double c5 = 1.22552115011207473091611921e-1, c6 = 2.97430720484762666375079572e0;
  double X2[A][A];
  sym_prod_lane<A>(X, X, 0.0, X2);
  double n2 = 0.0;
#pragma unroll
  for (int i = 0; i < A; ++i) {
    n2 = fma(X2[i][i], X2[i][i], n2);
#pragma unroll
    for (int k = i + 1; k < A; ++k) n2 = fma(2.0 * X2[i][k], X2[i][k], n2);
  }
  // smallest s >= 0 with |X^2|_F^(1/2) 2^-s <= theta_8; a zero matrix gives -inf, NaN s = 0
  const double sd = 0.25 * log2_coarse(n2) - kLog2Theta8;
  int s = 0;
  if (sd > 0.0) s = sd < 64.0 ? (int)ceil(sd) : 64;
#pragma unroll
  for (int i = 0; i < A; ++i)
#pragma unroll
    for (int k = i; k < A; ++k) {
      X[i][k] = ldexp(X[i][k], -s);
      X[k][i] = X[i][k];
      X2[i][k] = ldexp(X2[i][k], -2 * s);
      X2[k][i] = X2[i][k];
    }
  // T <- Y = c1 X2 X2 + c2 X2 X (symmetric: polynomials in X commute)
#pragma unroll
  for (int i = 0; i < A; ++i)
#pragma unroll
    for (int k = i; k < A; ++k) {
      double a = X2[i][0] * X2[0][k], b = X2[i][0] * X[0][k];
#pragma unroll
      for (int c = 1; c < A; ++c) {
        a = fma(X2[i][c], X2[c][k], a);
        b = fma(X2[i][c], X[c][k], b);
      }
      T[i][k] = fma(c1, a, c2 * b);
      T[k][i] = T[i][k];
    }
  // factors in place: X <- Y + c3 X2 + c4 X, X2 <- Y + c5 X2, T <- c6 Y + X2/2 + X + I
#pragma unroll
  for (int i = 0; i < A; ++i)
#pragma unroll
    for (int k = i; k < A; ++k) {
      const double x = X[i][k], x2 = X2[i][k], y = T[i][k];
      double w = fma(c6, y, fma(0.5, x2, x));
      if (i == k) w += 1.0;
      X[i][k] = fma(c3, x2, fma(c4, x, y));
      X[k][i] = X[i][k];
      X2[i][k] = fma(c5, x2, y);
      X2[k][i] = X2[i][k];
      T[i][k] = w;
      T[k][i] = w;
    }
  // T <- X X2 + T: each output entry reads only its own entry of T
#pragma unroll
  for (int i = 0; i < A; ++i)
#pragma unroll
    for (int k = i; k < A; ++k) {
      double acc = T[i][k];
#pragma unroll
      for (int c = 0; c < A; ++c) acc = fma(X[i][c], X2[c][k], acc);
      T[i][k] = acc;
      T[k][i] = acc;
    }
  square_times<A>(T, s);
}

// theta_15+ = 0.6764: the largest |X| at which the form below has a relative backward error
// sum_{k>=16} |h_k| |X|^(k-1) <= 2^-53, h(x) = log(e^-x p(x)) (h_16 = -2.17e-14, h_17 = 1.89e-14)
constexpr double kLog2Theta15 = -0.5640051;  // log2(0.67642)

// Degree 16 in four products, exact Taylor coefficients through x^15 (x^16 at 0.546 / 16!), in
// the evaluation form of Sastre (2018) one level up from the degree-8 form:
//   y02 = X2 (c1 X2 + c2 X),
//   y12 = (y02 + c3 X2 + c4 X)(y02 + c5 X2) + c6 y02 + c7 X2,
//   T   = (y12 + c8 X2 + c9 X)(y12 + c10 y02 + c11 X) + c12 y12 + c13 y02 + c14 X2 + X + I.
// The fourteen coefficients solve the fourteen equations that make the x^2..x^15 terms 1/k! (of
// the real solutions, one with the smallest coefficients). To keep four matrices live at most, y02
// is held as G = y02 + c5 X2 (the second factor of y12's product; the later terms in y02 are
// rewritten in G and X2), and y12's first factor is formed one row at a time. alpha = |X^2|_F^(1/2)
// as in the degree-8 form; on the A = 8 workload matrices a wave squares 2 times instead of 5-6.
template <int A>
__device__ __forceinline__ void expm_lane_s15(double (&X)[A][A], double (&T)[A][A]) {
  constexpr double c1 = -4.01876161020103546293293287596e-4, c2 = -2.94553144027968298047549126357e-3;
  constexpr double c3 = 8.70906657683767596765346131523e-3, c4 = -4.01756844067356780152362494573e-1;
  constexpr double c5 = -3.23076288812231208197540935626e-2, c6 = 2.33731940471157926649827413344e-2;
  constexpr double c7 = 2.61492797729811710806033238185e-1, c8 = -2.38107037387098722469262858218e-1;
  constexpr double c9 = -4.1302763659297829110328428344e-2, c10 = -5.79236170707326052179805145605e0;
  constexpr double c11 = 2.22420917249637356119032703879e0, c12 = 1.04080173523135436461532671084e1;
  constexpr double c13 = 3.03012340073871230596564458406e0, c14 = -2.12975559049643578453357460382e0;
  // the same polynomial in G = y02 + c5 X2
  constexpr double d3 = c3 - c5, e7 = c7 - c6 * c5, f10 = -c10 * c5, g14 = c14 - c13 * c5;
  double X2[A][A], Gm[A][A];
  sym_prod_lane<A>(X, X, 0.0, X2);
  double n2 = 0.0;
#pragma unroll
  for (int i = 0; i < A; ++i) {
    n2 = fma(X2[i][i], X2[i][i], n2);
#pragma unroll
    for (int k = i + 1; k < A; ++k) n2 = fma(2.0 * X2[i][k], X2[i][k], n2);
  }
  const double sd = 0.25 * log2_coarse(n2) - kLog2Theta15;
  int s = 0;
  if (sd > 0.0) s = sd < 64.0 ? (int)ceil(sd) : 64;
#pragma unroll
  for (int i = 0; i < A; ++i)
#pragma unroll
    for (int k = i; k < A; ++k) {
      X[i][k] = ldexp(X[i][k], -s);
      X[k][i] = X[i][k];
      X2[i][k] = ldexp(X2[i][k], -2 * s);
      X2[k][i] = X2[i][k];
    }
  // T <- c1 X2 + c2 X; G <- X2 T + c5 X2
#pragma unroll
  for (int i = 0; i < A; ++i)
#pragma unroll
    for (int k = i; k < A; ++k) {
      T[i][k] = fma(c1, X2[i][k], c2 * X[i][k]);
      T[k][i] = T[i][k];
    }
#pragma unroll
  for (int i = 0; i < A; ++i)
#pragma unroll
    for (int k = i; k < A; ++k) {
      double acc = c5 * X2[i][k];
#pragma unroll
      for (int c = 0; c < A; ++c) acc = fma(X2[i][c], T[c][k], acc);
      Gm[i][k] = acc;
      Gm[k][i] = acc;
    }
  // T <- y12 = F G + c6 G + e7 X2, F = G + d3 X2 + c4 X taken one row at a time
#pragma unroll
  for (int i = 0; i < A; ++i) {
    double f[A];
#pragma unroll
    for (int c = 0; c < A; ++c) f[c] = fma(c4, X[i][c], fma(d3, X2[i][c], Gm[i][c]));
#pragma unroll
    for (int k = i; k < A; ++k) {
      double acc = fma(c6, Gm[i][k], e7 * X2[i][k]);
#pragma unroll
      for (int c = 0; c < A; ++c) acc = fma(f[c], Gm[c][k], acc);
      T[i][k] = acc;
      T[k][i] = acc;
    }
  }
  // factors in place: X <- y12 + c8 X2 + c9 X, X2 <- y12 + c10 G + c11 X + f10 X2,
  // T <- c12 y12 + c13 G + g14 X2 + X + I
#pragma unroll
  for (int i = 0; i < A; ++i)
#pragma unroll
    for (int k = i; k < A; ++k) {
      const double x = X[i][k], x2 = X2[i][k], g = Gm[i][k], y = T[i][k];
      double r = fma(c12, y, fma(c13, g, fma(g14, x2, x)));
      if (i == k) r += 1.0;
      X[i][k] = fma(c8, x2, fma(c9, x, y));
      X[k][i] = X[i][k];
      X2[i][k] = fma(c10, g, fma(c11, x, fma(f10, x2, y)));
      X2[k][i] = X2[i][k];
      T[i][k] = r;
      T[k][i] = r;
    }
  // T <- X X2 + T: each output entry reads only its own entry of T
#pragma unroll
  for (int i = 0; i < A; ++i)
#pragma unroll
    for (int k = i; k < A; ++k) {
      double acc = T[i][k];
#pragma unroll
      for (int c = 0; c < A; ++c) acc = fma(X[i][c], X2[c][k], acc);
      T[i][k] = acc;
      T[k][i] = acc;
    }
  square_times<A>(T, s);
}

// Which scheme each kernel uses: the four-product degree-15+ form without +-dbeta (its four
// live matrices fit beside one ring state), except at A = 6, whose kernel holds two waves per
// SIMD in 256 VGPRs and spills with four live matrices; else Paterson-Stockmeyer degree 12 for
// A <= 4 and the three-product degree-8 form above (the +-dbeta kernels carry three ring states).
// Measured on MI355X (eval ms, degree 15+ against the previous scheme, 1e5 samples, same GPU):
// A=4 N=8 P=64 0.308 / 0.315 (PS-12), A=5 N=10 P=64 0.472 / 0.508, A=6 N=12 P=128 1.758 / 1.427
// (spilling), A=7 N=14 P=64 1.173 / 1.260, A=8 N=16 P=64 1.723 / 1.846 (4.05 / 4.43 at 262,144).
template <int A, bool PM>
__device__ __forceinline__ void expm_lane(double (&X)[A][A], double (&T)[A][A]) {
  if constexpr (!PM && A != 6)
    expm_lane_s15<A>(X, T);
  else if constexpr (A <= 4)
    expm_lane_ps12<A>(X, T);
  else
    expm_lane_s8<A>(X, T);
}

template <int A, int AR, bool PM, bool QD, class Q>
__device__ __forceinline__ void mode_terms(const_as<double>* tab, const Q& qs, int N,
                                           int P, int p, int pn, double& tr, double& tv, double& tp,
                                           double& tm, double (&er)[AR], double (&ev)[A], double (&ep)[A],
                                           double (&em)[A], double (&v)[A][A]) {
  constexpr ModeRec R(A, AR);
  // one mode's terms from its coordinates x0 (bead p) and xn (bead pn)
  auto mode = [&](int j, double x0, double xn) {
    const_as<double>* t = tab + j * R.len();
    const double sx = x0 + xn, dx = x0 - xn;
    const double q2s = fma(x0, x0, xn * xn), d2 = dx * dx;
    tr = fma(t[1], d2, fma(t[0], q2s, tr));
    tv = fma(t[3], d2, fma(t[2], q2s, tv));
    if (PM) {
      tp = fma(t[R.pm() + 1], d2, fma(t[R.pm() + 0], q2s, tp));
      tm = fma(t[R.pm() + 3], d2, fma(t[R.pm() + 2], q2s, tm));
    }
#pragma unroll
    for (int a = 0; a < AR; ++a) er[a] = fma(t[R.lin_rho() + a], sx, er[a]);
#pragma unroll
    for (int a = 0; a < A; ++a) {
      ev[a] = fma(t[R.lin_vib() + a], sx, ev[a]);
      if (PM) {
        ep[a] = fma(t[R.lin_plus() + a], sx, ep[a]);
        em[a] = fma(t[R.lin_minus() + a], sx, em[a]);
      }
    }
    // one scalar operand per FMA (a VALU op reads at most one SGPR pair)
    const double xx = x0 * x0;
#pragma unroll
    for (int b = 0, e = 0; b < A; ++b)
#pragma unroll
      for (int c = 0; c <= b; ++c, ++e) {
        v[b][c] = fma(t[R.lin_q() + e], x0, v[b][c]);
        if (QD) v[b][c] = fma(t[R.quad() + e], xx, v[b][c]);
      }
  };
  // Eight modes per iteration in two halves of four: the coordinates of one half are loaded
  // while the other half's arithmetic runs, four modes ahead of their use, into registers that
  // are free by then (a copy of a value still in flight would wait for its load). The modes are
  // taken in groups of four: modes past N read the zero records the table is padded with (and
  // the last mode's coordinates again). The loop body is straight-line code, which keeps the
  // compiler's wait counts exact (with a branch per mode it waited for every load in flight at
  // the top of each iteration).
  // (the loader's raw values are carried between iterations; the coordinate is formed at use)
  typedef typename Q::raw_t raw_t;
  auto ld4 = [&](int j, raw_t (&x)[4][2]) {
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int jj = j + k < N ? j + k : N - 1;
      x[k][0] = qs.load(jj, P, p);
      x[k][1] = qs.load(jj, P, pn);
    }
  };
  // modes past N read the centre of mode N - 1 (offset loader) with a zero record: nothing added
  auto use = [&](int j, const raw_t (&x)[2]) {
    const int jj = j < N ? j : N - 1;
    mode(j, qs.value(jj, x[0]), qs.value(jj, x[1]));
  };
  const int N4 = (N + 3) & ~3;
  raw_t xa[4][2], xb[4][2];
  ld4(0, xa);
  int j = 0;
  for (; j + 8 <= N4; j += 8) {
    ld4(j + 4, xb);
#pragma unroll
    for (int k = 0; k < 4; ++k) use(j + k, xa[k]);
    ld4(j + 8, xa);
#pragma unroll
    for (int k = 0; k < 4; ++k) use(j + 4 + k, xb[k]);
  }
  if (j < N4) {
#pragma unroll
    for (int k = 0; k < 4; ++k) use(j + k, xa[k]);
  }
}

// Occupancy: for A <= 4 the VGPR budget is capped at 256 (two waves per SIMD): the plain
// kernel needs 128 (four waves), the +-dbeta kernel's three ring states need ~200, and a cap
// of 168 made it spill (0.78 against 0.59 ms at A = 4, N = 8, P = 64, 1e5 samples, measured on
// MI355X). A = 5, 6 without +-dbeta are held to 256 as well (the persistent batched kernel
// otherwise takes 268 at A = 6 and drops to one wave per SIMD). Larger A, and the +-dbeta
// kernels of A > 4, keep the default budget (512 with the accumulation registers).
template <int A, bool PM>
struct EvalOccupancy {
  static constexpr int waves = (A <= 4 || (A <= 6 && !PM)) ? 2 : 1;
};

// The samples of one wave: wave w evaluates samples [w * 64/G, (w + 1) * 64/G) of `args`.
// QF: coordinates as float offsets from a per-mode centre (args.qf, args.qcentre).
// CH (with PM): 0 accumulates the three ring chains of g, g+ and g- together; 1, 2, 3 only the
// chain of g (writing rho, g and log S), of g+ or of g- (writing that output alone). Every pass
// forms the same M and scaling as the three-chain form, so three passes give its outputs bit
// for bit, with a third of the accumulators live (at A = 8 the three chains do not fit in the
// registers of one lane).
template <int A, int AR, bool PM, bool DBG, bool QF = false, int CH = 0>
__device__ __forceinline__ void eval_wave(const EvalArgs& args, int64_t wave) {
  static_assert(CH == 0 || (PM && !DBG && !QF), "single-chain passes are +-dbeta passes without debug records");
  constexpr bool ALL = PM && CH == 0;  // the g+ and g- chains beside g's
  const int lane = threadIdx.x & (kWave - 1);
  const int G = args.G, K = args.K, N = args.N, P = args.P;
  const int gl = lane & (G - 1);
  const int spw = kWave / G;
  const int64_t sample = wave * spw + lane / G;
  const bool live = sample < args.B;
  const int64_t NP = (int64_t)N * P;
  typedef typename std::conditional<QF, CoordOffsetBuf, CoordBuf>::type Loader;
  Loader qb;
  if constexpr (QF)
    qb = CoordOffsetBuf::wave_samples(args.qf, args.qcentre, args.B, NP, wave * spw, spw, live ? sample : -1);
  else
    qb = CoordBuf::wave_samples(args.q, args.B, NP, wave * spw, spw, live ? sample : -1);

  // read-only constants through the constant address space: scalar loads
  struct RO {
    const_as<double>* k2;
    const_as<double>* lpref;
  };
  const RO sr{as_const(args.set[SET_RHO].k2), as_const(args.set[SET_RHO].lpref)};
  const RO sv{as_const(args.set[SET_VIB].k2), as_const(args.set[SET_VIB].lpref)};
  const RO sp{as_const(args.set[SET_VIB_PLUS].k2), as_const(args.set[SET_VIB_PLUS].lpref)};
  const RO sm{as_const(args.set[SET_VIB_MINUS].k2), as_const(args.set[SET_VIB_MINUS].lpref)};
  const_as<double>* Ev = as_const(args.energy);
  const double* __restrict__ Qd = args.qdiag;
  const_as<int>* pairs = as_const(args.pair);
  const_as<double>* qpair = as_const(args.qpair);

  double Ng[A][A], Np[A][A], Nm[A][A];
#pragma unroll
  for (int i = 0; i < A; ++i)
#pragma unroll
    for (int k = 0; k < A; ++k) {
      Ng[i][k] = (i == k) ? 1.0 : 0.0;
      if (ALL) { Np[i][k] = Ng[i][k]; Nm[i][k] = Ng[i][k]; }
    }
  // log-domain products: lr[a] = sum_p log(O_rho,a(p) / S_p), ls = sum_p log S_p
  double lr[AR];
#pragma unroll
  for (int a = 0; a < AR; ++a) lr[a] = 0.0;
  double ls = 0.0;

  const int pbeg = gl * K;
  const int pend = live ? min(pbeg + K, P) : pbeg;
  for (int p = pbeg; p < pend; ++p) {
    const int pn = (p + 1 == P) ? 0 : p + 1;
    // O-matrix exponents (pimc.py:1087-1129) in the expanded cancellation-free form of OSet:
    // the surface-independent parts t* once per mode, one FMA per surface and mode
    double er[AR], ev[A], ep[A], em[A], v[A][A];
    double tr = 0.0, tv = 0.0, tp = 0.0, tm = 0.0;
#pragma unroll
    for (int a = 0; a < AR; ++a) er[a] = 0.0;
#pragma unroll
    for (int a = 0; a < A; ++a) {
      ev[a] = 0.0;
      ep[a] = 0.0;
      em[a] = 0.0;
#pragma unroll
      for (int c = 0; c <= a; ++c) v[a][c] = 0.0;
    }
    if (Qd)
      mode_terms<A, AR, PM, true>(as_const(args.mode_tab), qb, N, P, p, pn, tr, tv, tp, tm, er, ev, ep, em, v);
    else
      mode_terms<A, AR, PM, false>(as_const(args.mode_tab), qb, N, P, p, pn, tr, tv, tp, tm, er, ev, ep, em, v);
    // off-diagonal quadratic mode pairs (j != k)
    for (int t = 0; t < args.npair; ++t) {
      const int jj = pairs[2 * t], kk = pairs[2 * t + 1];
      const double xj = qb.value(jj, qb.load(jj, P, p)), xk = qb.value(kk, qb.load(kk, P, p));
      const_as<double>* Qh = qpair + (int64_t)t * A * A;
#pragma unroll
      for (int b = 0; b < A; ++b)
#pragma unroll
        for (int c = 0; c <= b; ++c) v[b][c] = fma(xk * Qh[b * A + c], xj, v[b][c]);
    }

    // O-factors and the per-bead scaling S_p = max over both O's (pimc.py:1076-1084), in the
    // log domain: log O_a = log F_a - e_a / 2, log S_p = max; rho's factors never leave it,
    // so per bead only the vib O-factors (and M's eigenvalues) need an exponential
    double lOr[AR], lOv[A], lOp[A], lOm[A];
    double lS = -INFINITY;
#pragma unroll
    for (int a = 0; a < AR; ++a) {
      lOr[a] = fma(-0.5, (tr + er[a]) + sr.k2[a], sr.lpref[a]);
      lS = fmax(lS, lOr[a]);
    }
#pragma unroll
    for (int a = 0; a < A; ++a) {
      lOv[a] = fma(-0.5, (tv + ev[a]) + sv.k2[a], sv.lpref[a]);
      lS = fmax(lS, lOv[a]);
      if (ALL || CH == 2) lOp[a] = fma(-0.5, (tp + ep[a]) + sp.k2[a], sp.lpref[a]);
      if (ALL || CH == 3) lOm[a] = fma(-0.5, (tm + em[a]) + sm.k2[a], sm.lpref[a]);
    }
    if (args.noscale) lS = 0.0;

    // coupling matrix -> M = exp(-tau V) (pimc.py:1167-1187)
    double M[A][A];
    {
      double X[A][A];
#pragma unroll
      for (int b = 0; b < A; ++b)
#pragma unroll
        for (int c = 0; c <= b; ++c) {
          X[b][c] = -args.tau * (v[b][c] + Ev[b * A + c]);
          X[c][b] = X[b][c];
        }
      expm_lane<A, PM>(X, M);
    }

    if (DBG) {
      const int rec = eval_debug_record(A, AR);
      double* dr = args.dbg + ((int64_t)sample * P + p) * rec;
#pragma unroll
      for (int a = 0; a < AR; ++a) dr[a] = exp(lOr[a]);
#pragma unroll
      for (int a = 0; a < A; ++a) dr[AR + a] = exp(lOv[a]);
      dr[AR + A] = exp(lS);
#pragma unroll
      for (int c = 0; c < A; ++c)
#pragma unroll
        for (int e = 0; e < A; ++e) dr[AR + A + 1 + c * A + e] = M[c][e];
    }

    // scale (pimc.py:1062-1066, 1445, 1448), then accumulate rho and the ring products
    double Ov[A], Op[A], Om[A];
#pragma unroll
    for (int a = 0; a < AR; ++a) lr[a] += lOr[a] - lS;
#pragma unroll
    for (int a = 0; a < A; ++a) {
      // in a single-chain pass Ov holds the factors of that pass's chain
      Ov[a] = exp((CH == 2 ? lOp[a] : (CH == 3 ? lOm[a] : lOv[a])) - lS);
      if (ALL) {
        Op[a] = exp(lOp[a] - lS);
        Om[a] = exp(lOm[a] - lS);
      }
    }
    ls += lS;
#pragma unroll
    for (int i = 0; i < A; ++i) {
      double rg[A], rpl[A], rmi[A];
#pragma unroll
      for (int c = 0; c < A; ++c) {
        rg[c] = Ng[i][c];
        if (ALL) { rpl[c] = Np[i][c]; rmi[c] = Nm[i][c]; }
      }
#pragma unroll
      for (int k = 0; k < A; ++k) {
        double s = rg[0] * M[0][k];
#pragma unroll
        for (int c = 1; c < A; ++c) s = fma(rg[c], M[c][k], s);
        Ng[i][k] = s * Ov[k];
        if (ALL) {
          double sp_ = rpl[0] * M[0][k], sm_ = rmi[0] * M[0][k];
#pragma unroll
          for (int c = 1; c < A; ++c) {
            sp_ = fma(rpl[c], M[c][k], sp_);
            sm_ = fma(rmi[c], M[c][k], sm_);
          }
          Np[i][k] = sp_ * Op[k];
          Nm[i][k] = sm_ * Om[k];
        }
      }
    }
  }


  // ordered tree over the G lanes of a sample: lane g (g % 2s == 0) absorbs lane g + s
  for (int s = 1; s < G; s <<= 1) {
    const bool leader = (gl & (2 * s - 1)) == 0;
    double o[A][A];
#pragma unroll
    for (int i = 0; i < A; ++i)
#pragma unroll
      for (int k = 0; k < A; ++k) o[i][k] = shfl_down_d(Ng[i][k], s);
    if (leader) {
#pragma unroll
      for (int i = 0; i < A; ++i) {
        double r[A];
#pragma unroll
        for (int c = 0; c < A; ++c) r[c] = Ng[i][c];
#pragma unroll
        for (int k = 0; k < A; ++k) {
          double acc = r[0] * o[0][k];
#pragma unroll
          for (int c = 1; c < A; ++c) acc = fma(r[c], o[c][k], acc);
          Ng[i][k] = acc;
        }
      }
    }
    if (ALL) {
#pragma unroll
      for (int w = 0; w < 2; ++w) {
        double (&X)[A][A] = (w == 0) ? Np : Nm;
#pragma unroll
        for (int i = 0; i < A; ++i)
#pragma unroll
          for (int k = 0; k < A; ++k) o[i][k] = shfl_down_d(X[i][k], s);
        if (leader) {
#pragma unroll
          for (int i = 0; i < A; ++i) {
            double r[A];
#pragma unroll
            for (int c = 0; c < A; ++c) r[c] = X[i][c];
#pragma unroll
            for (int k = 0; k < A; ++k) {
              double acc = r[0] * o[0][k];
#pragma unroll
              for (int c = 1; c < A; ++c) acc = fma(r[c], o[c][k], acc);
              X[i][k] = acc;
            }
          }
        }
      }
    }
#pragma unroll
    for (int a = 0; a < AR; ++a) {
      const double t = shfl_down_d(lr[a], s);
      if (leader) lr[a] += t;
    }
    const double tl = shfl_down_d(ls, s);
    if (leader) ls += tl;
  }

  if (live && gl == 0 && CH >= 2) {  // the g+ or g- pass: the trace of its chain alone
    double t = Ng[0][0];
#pragma unroll
    for (int i = 1; i < A; ++i) t += Ng[i][i];
    args.out[(CH == 2 ? OUT_GP : OUT_GM) * args.out_stride + sample] = t;
  }
  if (live && gl == 0 && CH < 2) {
    const int64_t st = args.out_stride;
    double rho = exp(lr[0]);
#pragma unroll
    for (int a = 1; a < AR; ++a) rho += exp(lr[a]);
    double g = Ng[0][0];
#pragma unroll
    for (int i = 1; i < A; ++i) g += Ng[i][i];
    args.out[OUT_RHO * st + sample] = rho;
    args.out[OUT_G * st + sample] = g;
    if (ALL) {
      double gp = Np[0][0], gm = Nm[0][0];
#pragma unroll
      for (int i = 1; i < A; ++i) { gp += Np[i][i]; gm += Nm[i][i]; }
      args.out[OUT_GP * st + sample] = gp;
      args.out[OUT_GM * st + sample] = gm;
    }
    args.out[OUT_LOGS * st + sample] = ls;
  }
}

template <int A, int AR, bool PM, bool DBG>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(EvalOccupancy<A, PM>::waves)))
k_eval_fp64(const EvalArgs args) {
  eval_wave<A, AR, PM, DBG>(args, ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave);
}

// one ring chain of the +-dbeta evaluation (CH 1: g with rho and log S, 2: g+, 3: g-)
template <int A, int AR, int CH>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(EvalOccupancy<A, true>::waves)))
k_eval_fp64_chain(const EvalArgs args) {
  eval_wave<A, AR, true, false, false, CH>(args, ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave);
}

// the same on float offsets from a per-mode centre (pib_eval_block_flow_offsets)
template <int A>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(EvalOccupancy<A, false>::waves)))
k_eval_fp64_offsets(const EvalArgs args) {
  eval_wave<A, A, false, false, true>(args, ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave);
}

LaneProb lane_prob(const EvalArgs& a) {
  LaneProb p{};
  p.q = a.q;
  p.out = a.out;
  p.out_stride = a.out_stride;
  p.B = a.B;
  p.N = a.N;
  p.P = a.P;
  p.G = a.G;
  p.K = a.K;
  p.npair = a.npair;
  p.noscale = a.noscale;
  p.tau = a.tau;
  p.mode_tab = a.mode_tab;
  p.energy = a.energy;
  p.qdiag = a.qdiag;
  p.pair = a.pair;
  p.qpair = a.qpair;
  for (int k = 0; k < NUM_SETS; ++k) {
    p.k2[k] = a.set[k].k2;
    p.lpref[k] = a.set[k].lpref;
  }
  return p;
}

// Persistent batched launch: many blocks of samples (temperatures with different P, models
// with the same surface counts) in one grid. Each wave repeatedly takes the next tile
// (problem, first wave, wave count) from a device counter until the tile list is exhausted;
// the host orders tiles by decreasing cost, so the long ones start first. Every wave reaches
// the exit once the counter passes the tile count. The problems' arguments come by value in
// the kernel arguments (scalar registers), as in the stand-alone launch.
template <int A, int AR, bool PM>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(EvalOccupancy<A, PM>::waves)))
k_eval_batched(const LaneProbs probs, const BatchTile* __restrict__ tiles, int ntiles, int* counter) {
  const int lane = threadIdx.x & (kWave - 1);
  for (;;) {
    int t = 0;
    if (lane == 0) t = atomicAdd(counter, 1);
    t = __builtin_amdgcn_readfirstlane(__shfl(t, 0, kWave));
    if (t >= ntiles) break;
    const int pi = __builtin_amdgcn_readfirstlane(tiles[t].prob);
    const int64_t w0 = __builtin_amdgcn_readfirstlane(tiles[t].wave0);
    const int nw = __builtin_amdgcn_readfirstlane(tiles[t].nwaves);
    const LaneProb& d = probs.p[pi];
    EvalArgs a{};
    a.q = d.q;
    a.out = d.out;
    a.out_stride = d.out_stride;
    a.B = d.B;
    a.N = d.N;
    a.P = d.P;
    a.G = d.G;
    a.K = d.K;
    a.npair = d.npair;
    a.noscale = d.noscale;
    a.tau = d.tau;
    a.mode_tab = d.mode_tab;
    a.energy = d.energy;
    a.qdiag = d.qdiag;
    a.pair = d.pair;
    a.qpair = d.qpair;
    a.dbg = nullptr;
#pragma unroll
    for (int k = 0; k < NUM_SETS; ++k) {
      a.set[k].k2 = d.k2[k];
      a.set[k].lpref = d.lpref[k];
    }
    for (int64_t w = w0; w < w0 + nw; ++w) eval_wave<A, AR, PM, false>(a, w);
  }
}

// ---------------------------------------------------------------------------------------
template <int A, int AR>
static hipError_t launch_a(int pm, const EvalArgs& a, hipStream_t s) {
  const int spw = kWave / a.G;
  const int64_t waves = (a.B + spw - 1) / spw;
  const int threads = 256;
  const dim3 grid((unsigned)((waves * kWave + threads - 1) / threads));
  const bool dbg = a.dbg != nullptr;
  // q is read straight from global memory. Staging each wave's samples through LDS (one
  // coalesced copy) removes the ~2.5x re-fetch of q seen in FETCH_SIZE, but measured 3-13%
  // slower on MI355X: the copy prologue adds latency and the LDS it needs costs occupancy,
  // while this kernel is bound by dependent FP64 arithmetic, not by memory.
  const EvalArgs& b = a;
  const size_t lds = 0;
  if constexpr (A == 8 && AR == 8) {
    // +-dbeta at A = 8: the three ring chains in three passes of one chain each (the three
    // together spill 1.5 KB per lane to scratch: 5.78 ms against 3.55 ms for the passes at
    // 65,536 samples, N = 16, P = 64, timed alone; in a block loop after evaluations without
    // +-dbeta both forms have been measured at ~1.6x their cycles, the passes never slower; at
    // A = 7 one pass is still faster, 2.00 ms against about three times the plain kernel's
    // 0.77); the debug records keep the one-pass form
    if (pm && !dbg) {
      hipLaunchKernelGGL((k_eval_fp64_chain<A, AR, 1>), grid, dim3(threads), lds, s, b);
      hipLaunchKernelGGL((k_eval_fp64_chain<A, AR, 2>), grid, dim3(threads), lds, s, b);
      hipLaunchKernelGGL((k_eval_fp64_chain<A, AR, 3>), grid, dim3(threads), lds, s, b);
      return hipGetLastError();
    }
  }
  if (pm && dbg)
    hipLaunchKernelGGL((k_eval_fp64<A, AR, true, true>), grid, dim3(threads), lds, s, b);
  else if (pm)
    hipLaunchKernelGGL((k_eval_fp64<A, AR, true, false>), grid, dim3(threads), lds, s, b);
  else if (dbg)
    hipLaunchKernelGGL((k_eval_fp64<A, AR, false, true>), grid, dim3(threads), lds, s, b);
  else
    hipLaunchKernelGGL((k_eval_fp64<A, AR, false, false>), grid, dim3(threads), lds, s, b);
  return hipGetLastError();
}

template <int A, int AR>
static hipError_t launch_batched_a(int pm, const LaneProbs& probs, const BatchTile* tiles, int ntiles,
                                   int* counter, int waves, hipStream_t s) {
  const dim3 grid((unsigned)((waves + 3) / 4));
  if (pm)
    hipLaunchKernelGGL((k_eval_batched<A, AR, true>), grid, dim3(256), 0, s, probs, tiles, ntiles, counter);
  else
    hipLaunchKernelGGL((k_eval_batched<A, AR, false>), grid, dim3(256), 0, s, probs, tiles, ntiles, counter);
  return hipGetLastError();
}

#define PIB_EVAL_CASES(X) \
  X(1, 1) X(2, 2) X(3, 3) X(4, 4) X(5, 5) X(6, 6) X(7, 7) X(8, 8) \
  X(2, 1) X(2, 4) X(2, 8) X(4, 2) X(4, 8)

bool eval_supported(int A, int AR) {
#define PIB_CASE(a_, ar_) if (A == a_ && AR == ar_) return true;
  PIB_EVAL_CASES(PIB_CASE)
#undef PIB_CASE
  return false;
}

hipError_t launch_eval_fp64(int A, int AR, int pm, const EvalArgs& a, hipStream_t s) {
#define PIB_CASE(a_, ar_) if (A == a_ && AR == ar_) return launch_a<a_, ar_>(pm, a, s);
  PIB_EVAL_CASES(PIB_CASE)
#undef PIB_CASE
  return hipErrorInvalidValue;
}

template <int A>
static hipError_t launch_offsets_a(const EvalArgs& a, hipStream_t s) {
  const int spw = kWave / a.G;
  const int64_t waves = (a.B + spw - 1) / spw;
  const int threads = 256;
  const dim3 grid((unsigned)((waves * kWave + threads - 1) / threads));
  hipLaunchKernelGGL((k_eval_fp64_offsets<A>), grid, dim3(threads), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_eval_fp64_offsets(int A, const EvalArgs& a, hipStream_t s) {
  switch (A) {
    case 1: return launch_offsets_a<1>(a, s);
    case 2: return launch_offsets_a<2>(a, s);
    case 3: return launch_offsets_a<3>(a, s);
    case 4: return launch_offsets_a<4>(a, s);
    case 5: return launch_offsets_a<5>(a, s);
    case 6: return launch_offsets_a<6>(a, s);
    case 7: return launch_offsets_a<7>(a, s);
    case 8: return launch_offsets_a<8>(a, s);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_eval_batched(int A, int AR, int pm, const LaneProbs& probs, const BatchTile* tiles,
                               int ntiles, int* counter, int waves, hipStream_t s) {
#define PIB_CASE(a_, ar_) \
  if (A == a_ && AR == ar_) return launch_batched_a<a_, ar_>(pm, probs, tiles, ntiles, counter, waves, s);
  PIB_EVAL_CASES(PIB_CASE)
#undef PIB_CASE
  return hipErrorInvalidValue;
}

}  // namespace pib
