// Reduced-precision variant of k_eval_fp64 (same mapping, same outputs, A <= 8).
//
// What runs in FP32: the O-matrix exponents, the coupling matrix V, M = exp(-tau V) and the
// ring products. What stays FP64: the bead coordinates as read (rounded to float once per
// mode), the per-sample log-scale sum_p log S_p and the per-surface log-products of rho, and
// the final assembly of s_g, s_rho.
//
// Range: the O-factors are formed in the log domain, log O_a = log F_a - e_a / 2, so that
// S_p = max_a O_a and O_a / S_p = exp(log O_a - log S_p) never under/overflow before scaling.
// The ring products are renormalised by a power of two after every bead (exact in binary
// floating point) and the exponents are carried as integers; s_g = trace * 2^exp in FP64.
//
// The O exponent coth (x1^2 + x2^2) - 2 csch x1 x2 cancels when tau w is small; here it is
// evaluated, as in the FP64 kernel, in the expanded form
// sum_j th (x^2 + x'^2) + csch (x - x')^2 + lin_a (x + x') + k2_a (th = coth - csch), whose
// FP32 rounding is no larger than that of the shifted form th ((x-d)^2 + (x'-d)^2) +
// csch (x - x')^2 evaluated with x - d in FP64 (measured on the workload models: exponent
// error <= 5e-5 absolute for |e| ~ 10-20).
#include "pib_common.hpp"
#include "pib_kernels.hpp"

namespace pib {

namespace {

// M = exp(X) for a symmetric A x A matrix in one lane's registers, X = -tau V, in FP32: the
// FP64 kernel's scaled and squared degree-8 Taylor polynomial in three products (see
// expm_lane_s8 there: Y = X2 (c1 X2 + c2 X), T8 = (Y + c3 X2 + c4 X)(Y + c5 X2) + c6 Y + X2/2 +
// X + I), with |X^2|_F^(1/2) 2^-s <= theta_8 = 0.619, below the root of |y|^8 / 9! = 2^-24.
// Registers are half as dear as in FP64, so Y's right factor is formed as a matrix (three
// products). Against Horner's rule at degree 5 (theta_5 = 0.13 on |X|_F, four products) the
// squaring loop runs 2.1 instead of 5.0 times per wave at A = 8, N = 16, P = 64 and 0.8 instead
// of 3.0 at A = 4, N = 8, P = 64 (products per bead 5.1 against 9.0 and 3.8 against 7.0).
constexpr float kLog2Theta8f = -0.69197f;  // log2(0.619)

// Z = X Y + d I for commuting symmetric X, Y (Z symmetric; Z must not alias X or Y)
template <int A>
__device__ __forceinline__ void sym_prod_lane_f(const float (&X)[A][A], const float (&Y)[A][A], float d,
                                                float (&Z)[A][A]) {
#pragma unroll
  for (int i = 0; i < A; ++i)
#pragma unroll
    for (int k = i; k < A; ++k) {
      float s = (i == k) ? fmaf(X[i][0], Y[0][k], d) : X[i][0] * Y[0][k];
#pragma unroll
      for (int c = 1; c < A; ++c) s = fmaf(X[i][c], Y[c][k], s);
      Z[i][k] = s;
      Z[k][i] = s;
    }
}

template <int A>
__device__ __forceinline__ void expm_lane_f(float (&X)[A][A], float (&T)[A][A]) {
  constexpr float c1 = 4.98011920555997e-3f, c2 = 1.99204768222399e-2f, c3 = 7.66526532111915e-2f;
  constexpr float c4 = 8.76500980178555e-1f, c5 = 1.22552115011207e-1f, c6 = 2.97430720484763e0f;
  float X2[A][A];
  sym_prod_lane_f<A>(X, X, 0.0f, X2);
  float n2 = 0.0f;
#pragma unroll
  for (int i = 0; i < A; ++i) {
    n2 = fmaf(X2[i][i], X2[i][i], n2);
#pragma unroll
    for (int k = i + 1; k < A; ++k) n2 = fmaf(2.0f * X2[i][k], X2[i][k], n2);
  }
  // smallest s >= 0 with |X^2|_F^(1/2) 2^-s <= theta_8; a zero matrix gives -inf, NaN s = 0
  const float sd = 0.25f * __log2f(n2) - kLog2Theta8f;
  int s = 0;
  if (sd > 0.0f) s = sd < 64.0f ? (int)ceilf(sd) : 64;
  float Y[A][A];
  {
    float W[A][A];
#pragma unroll
    for (int i = 0; i < A; ++i)
#pragma unroll
      for (int k = i; k < A; ++k) {
        X[i][k] = ldexpf(X[i][k], -s);
        X[k][i] = X[i][k];
        X2[i][k] = ldexpf(X2[i][k], -2 * s);
        X2[k][i] = X2[i][k];
        W[i][k] = fmaf(c1, X2[i][k], c2 * X[i][k]);
        W[k][i] = W[i][k];
      }
    sym_prod_lane_f<A>(X2, W, 0.0f, Y);
  }
  // factors in place: X <- Y + c3 X2 + c4 X, X2 <- Y + c5 X2, T <- c6 Y + X2/2 + X + I
#pragma unroll
  for (int i = 0; i < A; ++i)
#pragma unroll
    for (int k = i; k < A; ++k) {
      const float x = X[i][k], x2 = X2[i][k], y = Y[i][k];
      float w = fmaf(c6, y, fmaf(0.5f, x2, x));
      if (i == k) w += 1.0f;
      X[i][k] = fmaf(c3, x2, fmaf(c4, x, y));
      X[k][i] = X[i][k];
      X2[i][k] = fmaf(c5, x2, y);
      X2[k][i] = X2[i][k];
      T[i][k] = w;
      T[k][i] = w;
    }
  // T <- X X2 + T: each output entry reads only its own entry of T
#pragma unroll
  for (int i = 0; i < A; ++i)
#pragma unroll
    for (int k = i; k < A; ++k) {
      float acc = T[i][k];
#pragma unroll
      for (int c = 0; c < A; ++c) acc = fmaf(X[i][c], X2[c][k], acc);
      T[i][k] = acc;
      T[k][i] = acc;
    }
  // two squarings per iteration (T -> Z -> T): no register copy after each one
  {
    float Z[A][A];
    for (int k = 0; k + 1 < s; k += 2) {
      sym_prod_lane_f<A>(T, T, 0.0f, Z);
      sym_prod_lane_f<A>(Z, Z, 0.0f, T);
    }
    if (s & 1) {
      sym_prod_lane_f<A>(T, T, 0.0f, Z);
#pragma unroll
      for (int i = 0; i < A; ++i)
#pragma unroll
        for (int j = 0; j < A; ++j) T[i][j] = Z[i][j];
    }
  }
}

template <int A, int AR, bool PM, bool QD>
__device__ __forceinline__ void mode_terms_f(const float* __restrict__ tab, const CoordBuf& qs, int N, int P,
                                             int p, int pn, float& tr, float& tv, float& tp, float& tm,
                                             float (&er)[AR], float (&ev)[A], float (&ep)[A], float (&em)[A],
                                             float (&v)[A][A]) {
  constexpr ModeRec R(A, AR);
  auto mode = [&](int j, double x0d, double xnd) {
    const float* __restrict__ t = tab + j * R.len_f();
    const float x0 = (float)x0d, xn = (float)xnd;
    const float sx = x0 + xn, dx = x0 - xn;
    const float q2s = fmaf(x0, x0, xn * xn), d2 = dx * dx;
    tr = fmaf(t[1], d2, fmaf(t[0], q2s, tr));
    tv = fmaf(t[3], d2, fmaf(t[2], q2s, tv));
    if (PM) {
      tp = fmaf(t[R.pm() + 1], d2, fmaf(t[R.pm() + 0], q2s, tp));
      tm = fmaf(t[R.pm() + 3], d2, fmaf(t[R.pm() + 2], q2s, tm));
    }
#pragma unroll
    for (int a = 0; a < AR; ++a) er[a] = fmaf(t[R.lin_rho() + a], sx, er[a]);
#pragma unroll
    for (int a = 0; a < A; ++a) {
      ev[a] = fmaf(t[R.lin_vib() + a], sx, ev[a]);
      if (PM) {
        ep[a] = fmaf(t[R.lin_plus() + a], sx, ep[a]);
        em[a] = fmaf(t[R.lin_minus() + a], sx, em[a]);
      }
    }
    const float xx = x0 * x0;
#pragma unroll
    for (int b = 0, e = 0; b < A; ++b)
#pragma unroll
      for (int c = 0; c <= b; ++c, ++e) {
        v[b][c] = fmaf(t[R.lin_q() + e], x0, v[b][c]);
        if (QD) v[b][c] = fmaf(t[R.quad() + e], xx, v[b][c]);
      }
  };
  // the FP64 kernel's schedule (see mode_terms there): eight modes per iteration in two halves
  // of four, one half's coordinates loaded while the other half's arithmetic runs; modes past N
  // read the zero records the table is padded with
  auto ld4 = [&](int j, double (&x)[4][2]) {
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int jj = j + k < N ? j + k : N - 1;
      x[k][0] = qs[jj * P + p];
      x[k][1] = qs[jj * P + pn];
    }
  };
  const int N4 = (N + 3) & ~3;
  double xa[4][2], xb[4][2];
  ld4(0, xa);
  int j = 0;
  for (; j + 8 <= N4; j += 8) {
    ld4(j + 4, xb);
#pragma unroll
    for (int k = 0; k < 4; ++k) mode(j + k, xa[k][0], xa[k][1]);
    ld4(j + 8, xa);
#pragma unroll
    for (int k = 0; k < 4; ++k) mode(j + 4 + k, xb[k][0], xb[k][1]);
  }
  if (j < N4) {
#pragma unroll
    for (int k = 0; k < 4; ++k) mode(j + k, xa[k][0], xa[k][1]);
  }
}

// A x A matrix scaled by a power of two so that its largest magnitude lies in [0.5, 1);
// the exponent goes to e (exact; zero matrices are left alone)
template <int A>
__device__ __forceinline__ void renorm(float (&m)[A][A], int& e) {
  float mx = 0.0f;
#pragma unroll
  for (int i = 0; i < A; ++i)
#pragma unroll
    for (int k = 0; k < A; ++k) mx = fmaxf(mx, fabsf(m[i][k]));
  int ex = 0;
  (void)frexpf(mx, &ex);
  if (mx > 0.0f && ex != 0) {
    const float f = ldexpf(1.0f, -ex);
#pragma unroll
    for (int i = 0; i < A; ++i)
#pragma unroll
      for (int k = 0; k < A; ++k) m[i][k] *= f;
    e += ex;
  }
}

template <int A>
__device__ __forceinline__ void matmul_left(float (&X)[A][A], const float (&Y)[A][A]) {
#pragma unroll
  for (int i = 0; i < A; ++i) {
    float r[A];
#pragma unroll
    for (int c = 0; c < A; ++c) r[c] = X[i][c];
#pragma unroll
    for (int k = 0; k < A; ++k) {
      float acc = r[0] * Y[0][k];
#pragma unroll
      for (int c = 1; c < A; ++c) acc = fmaf(r[c], Y[c][k], acc);
      X[i][k] = acc;
    }
  }
}

__device__ __forceinline__ float shfl_down_f(float v, int d) { return __shfl_down(v, (unsigned)d, kWave); }

}  // namespace

template <int A, int AR, bool PM>
__global__ void __launch_bounds__(256) k_eval_fp32(const EvalArgs args) {
  const int lane = threadIdx.x & (kWave - 1);
  const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  const int G = args.G, K = args.K, N = args.N, P = args.P;
  const int gl = lane & (G - 1);
  const int64_t sample = wave * (kWave / G) + lane / G;
  const bool live = sample < args.B;
  const double* __restrict__ qs = args.q + (live ? sample : 0) * (int64_t)N * P;
  const CoordBuf qb = CoordBuf::wave_samples(args.q, args.B, (int64_t)N * P, wave * (kWave / G), kWave / G,
                                             live ? sample : -1);
  const OSetF fr = args.setf[SET_RHO], fv = args.setf[SET_VIB];
  const OSetF fp = args.setf[SET_VIB_PLUS], fm = args.setf[SET_VIB_MINUS];
  const float* __restrict__ Ev = args.energy_f;
  const bool has_q = args.qdiag_f != nullptr;
  const float tau = (float)args.tau;

  float Ng[A][A], Np[A][A], Nm[A][A];
  int eg = 0, ep_ = 0, em_ = 0;
#pragma unroll
  for (int i = 0; i < A; ++i)
#pragma unroll
    for (int k = 0; k < A; ++k) {
      Ng[i][k] = (i == k) ? 1.0f : 0.0f;
      if (PM) { Np[i][k] = Ng[i][k]; Nm[i][k] = Ng[i][k]; }
    }
  double lr[AR];  // sum_p log(O_rho,a / S_p)
#pragma unroll
  for (int a = 0; a < AR; ++a) lr[a] = 0.0;
  double ls = 0.0;  // sum_p log S_p

  const int pbeg = gl * K;
  const int pend = live ? min(pbeg + K, P) : pbeg;
  for (int p = pbeg; p < pend; ++p) {
    const int pn = (p + 1 == P) ? 0 : p + 1;
    float er[AR], ev[A], epl[A], emi[A], v[A][A];
    float tr = 0.f, tv = 0.f, tpl = 0.f, tmi = 0.f;
#pragma unroll
    for (int a = 0; a < AR; ++a) er[a] = 0.0f;
#pragma unroll
    for (int a = 0; a < A; ++a) {
      ev[a] = 0.0f;
      epl[a] = 0.0f;
      emi[a] = 0.0f;
#pragma unroll
      for (int c = 0; c <= a; ++c) v[a][c] = 0.0f;
    }
    if (has_q)
      mode_terms_f<A, AR, PM, true>(args.mode_tab_f, qb, N, P, p, pn, tr, tv, tpl, tmi, er, ev, epl, emi, v);
    else
      mode_terms_f<A, AR, PM, false>(args.mode_tab_f, qb, N, P, p, pn, tr, tv, tpl, tmi, er, ev, epl, emi, v);
    for (int t = 0; t < args.npair; ++t) {
      const int jj = args.pair[2 * t], kk = args.pair[2 * t + 1];
      const float xj = (float)qs[jj * P + p], xk = (float)qs[kk * P + p];
      const float* Qh = args.qpair_f + (int64_t)t * A * A;
#pragma unroll
      for (int b = 0; b < A; ++b)
#pragma unroll
        for (int c = 0; c <= b; ++c) v[b][c] = fmaf(xk * Qh[b * A + c], xj, v[b][c]);
    }

    // log O-factors and the scaling S_p = max over both O's
    float lOr[AR], lOv[A], lOp[A], lOm[A];
    float lS = -INFINITY;
#pragma unroll
    for (int a = 0; a < AR; ++a) {
      lOr[a] = fmaf(-0.5f, (tr + er[a]) + fr.k2[a], fr.lpref[a]);
      lS = fmaxf(lS, lOr[a]);
    }
#pragma unroll
    for (int a = 0; a < A; ++a) {
      lOv[a] = fmaf(-0.5f, (tv + ev[a]) + fv.k2[a], fv.lpref[a]);
      lS = fmaxf(lS, lOv[a]);
      if (PM) {
        lOp[a] = fmaf(-0.5f, (tpl + epl[a]) + fp.k2[a], fp.lpref[a]);
        lOm[a] = fmaf(-0.5f, (tmi + emi[a]) + fm.k2[a], fm.lpref[a]);
      }
    }
    if (args.noscale) lS = 0.0f;
    ls += (double)lS;
#pragma unroll
    for (int a = 0; a < AR; ++a) lr[a] += (double)(lOr[a] - lS);

    float M[A][A];
    {
      float X[A][A];
#pragma unroll
      for (int b = 0; b < A; ++b)
#pragma unroll
        for (int c = 0; c <= b; ++c) {
          X[b][c] = -tau * (v[b][c] + Ev[b * A + c]);
          X[c][b] = X[b][c];
        }
      expm_lane_f<A>(X, M);
    }
    float Ov[A], Opl[A], Omi[A];
#pragma unroll
    for (int a = 0; a < A; ++a) {
      Ov[a] = expf(lOv[a] - lS);
      if (PM) {
        Opl[a] = expf(lOp[a] - lS);
        Omi[a] = expf(lOm[a] - lS);
      }
    }
    // ring: N <- N M diag(O / S), renormalised
#pragma unroll
    for (int i = 0; i < A; ++i) {
      float rg[A], rp_[A], rm_[A];
#pragma unroll
      for (int c = 0; c < A; ++c) {
        rg[c] = Ng[i][c];
        if (PM) { rp_[c] = Np[i][c]; rm_[c] = Nm[i][c]; }
      }
#pragma unroll
      for (int k = 0; k < A; ++k) {
        float s = rg[0] * M[0][k];
#pragma unroll
        for (int c = 1; c < A; ++c) s = fmaf(rg[c], M[c][k], s);
        Ng[i][k] = s * Ov[k];
        if (PM) {
          float sp = rp_[0] * M[0][k], sm = rm_[0] * M[0][k];
#pragma unroll
          for (int c = 1; c < A; ++c) {
            sp = fmaf(rp_[c], M[c][k], sp);
            sm = fmaf(rm_[c], M[c][k], sm);
          }
          Np[i][k] = sp * Opl[k];
          Nm[i][k] = sm * Omi[k];
        }
      }
    }
    renorm<A>(Ng, eg);
    if (PM) {
      renorm<A>(Np, ep_);
      renorm<A>(Nm, em_);
    }
  }

  // ordered tree over the G lanes of a sample (chunk 0 * chunk 1 * ...)
  for (int s = 1; s < G; s <<= 1) {
    const bool leader = (gl & (2 * s - 1)) == 0;
#pragma unroll
    for (int w = 0; w < (PM ? 3 : 1); ++w) {
      float (&X)[A][A] = (w == 0) ? Ng : (w == 1) ? Np : Nm;
      int& xe = (w == 0) ? eg : (w == 1) ? ep_ : em_;
      float o[A][A];
#pragma unroll
      for (int i = 0; i < A; ++i)
#pragma unroll
        for (int k = 0; k < A; ++k) o[i][k] = shfl_down_f(X[i][k], s);
      const int oe = __shfl_down(xe, (unsigned)s, kWave);
      if (leader) {
        matmul_left<A>(X, o);
        xe += oe;
        renorm<A>(X, xe);
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

  if (live && gl == 0) {
    const int64_t st = args.out_stride;
    double lmax = lr[0];
#pragma unroll
    for (int a = 1; a < AR; ++a) lmax = fmax(lmax, lr[a]);
    double rho = 0.0;
#pragma unroll
    for (int a = 0; a < AR; ++a) rho += exp(lr[a] - lmax);
    rho *= exp(lmax);
    float tg = Ng[0][0], tp = PM ? Np[0][0] : 0.f, tm = PM ? Nm[0][0] : 0.f;
#pragma unroll
    for (int i = 1; i < A; ++i) {
      tg += Ng[i][i];
      if (PM) { tp += Np[i][i]; tm += Nm[i][i]; }
    }
    args.out[OUT_RHO * st + sample] = rho;
    args.out[OUT_G * st + sample] = ldexp((double)tg, eg);
    if (PM) {
      args.out[OUT_GP * st + sample] = ldexp((double)tp, ep_);
      args.out[OUT_GM * st + sample] = ldexp((double)tm, em_);
    }
    args.out[OUT_LOGS * st + sample] = ls;
  }
}

template <int A, int AR>
static hipError_t launch_f(int pm, const EvalArgs& a, hipStream_t s) {
  const int spw = kWave / a.G;
  const int64_t waves = (a.B + spw - 1) / spw;
  const dim3 grid((unsigned)((waves * kWave + 255) / 256));
  if (pm)
    hipLaunchKernelGGL((k_eval_fp32<A, AR, true>), grid, dim3(256), 0, s, a);
  else
    hipLaunchKernelGGL((k_eval_fp32<A, AR, false>), grid, dim3(256), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_eval_fp32(int A, int AR, int pm, const EvalArgs& a, hipStream_t s) {
#define PIB_F32_CASES(X) \
  X(1, 1) X(2, 2) X(3, 3) X(4, 4) X(5, 5) X(6, 6) X(7, 7) X(8, 8) X(2, 1) X(2, 4) X(2, 8) X(4, 2) X(4, 8)
#define PIB_CASE(a_, ar_) if (A == a_ && AR == ar_) return launch_f<a_, ar_>(pm, a, s);
  PIB_F32_CASES(PIB_CASE)
#undef PIB_CASE
#undef PIB_F32_CASES
  return hipErrorInvalidValue;
}

}  // namespace pib
