// C-ABI layer: context, constant upload, argument checks and kernel dispatch.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/pibronic_hip.h"
#include "pib_common.hpp"
#include "pib_ctx.hpp"
#include "pib_flow.hpp"
#include "pib_kernels.hpp"

using namespace pib;

namespace pib {

// Per-mode constant records of the lanes kernel (layout: ModeRec). Rebuilt by the coupling
// and O-matrix set setters once the coupling and the rho and vib sets agree in shape; the
// +-dbeta entries stay zero until those sets are given.
int rebuild_mode_table(pib_ctx* c) {
  c->mode_tab_ok = false;
  const int A = c->A, N = c->N, AR = c->set_A[SET_RHO];
  if (!c->coupling_set || !AR || c->set_A[SET_VIB] != A || c->set_N[SET_RHO] != N || c->set_N[SET_VIB] != N)
    return PIB_OK;
  const ModeRec R(A, AR);
  const size_t AA = (size_t)A * A;
  // zero records after the last mode up to a multiple of four (the lanes kernel takes modes in
  // groups of four; a zero record adds exactly nothing)
  const int N4 = (N + 3) & ~3;
  std::vector<double> tab((size_t)N4 * R.len(), 0.0);
  const bool pm = c->set_A[SET_VIB_PLUS] == A && c->set_A[SET_VIB_MINUS] == A &&
                  c->set_N[SET_VIB_PLUS] == N && c->set_N[SET_VIB_MINUS] == N;
  for (int j = 0; j < N; ++j) {
    double* t = tab.data() + (size_t)j * R.len();
    t[0] = c->h_th[SET_RHO][j];
    t[1] = c->h_cs[SET_RHO][j];
    t[2] = c->h_th[SET_VIB][j];
    t[3] = c->h_cs[SET_VIB][j];
    for (int a = 0; a < AR; ++a) t[R.lin_rho() + a] = c->h_lin[SET_RHO][(size_t)a * N + j];
    for (int a = 0; a < A; ++a) t[R.lin_vib() + a] = c->h_lin[SET_VIB][(size_t)a * N + j];
    int e = 0;
    for (int b = 0; b < A; ++b)
      for (int d = 0; d <= b; ++d, ++e) {
        t[R.lin_q() + e] = c->h_linear[(size_t)j * AA + b * A + d];
        t[R.quad() + e] = c->h_qdiag.empty() ? 0.0 : c->h_qdiag[(size_t)j * AA + b * A + d];
      }
    if (pm) {
      t[R.pm() + 0] = c->h_th[SET_VIB_PLUS][j];
      t[R.pm() + 1] = c->h_cs[SET_VIB_PLUS][j];
      t[R.pm() + 2] = c->h_th[SET_VIB_MINUS][j];
      t[R.pm() + 3] = c->h_cs[SET_VIB_MINUS][j];
      for (int a = 0; a < A; ++a) {
        t[R.lin_plus() + a] = c->h_lin[SET_VIB_PLUS][(size_t)a * N + j];
        t[R.lin_minus() + a] = c->h_lin[SET_VIB_MINUS][(size_t)a * N + j];
      }
    }
  }
  // float copy, records padded to 16 bytes (and zero records up to N4, as above)
  std::vector<float> tabf((size_t)N4 * R.len_f(), 0.f);
  for (int j = 0; j < N; ++j)
    for (int i = 0; i < R.lin_minus() + A; ++i)
      tabf[(size_t)j * R.len_f() + i] = (float)tab[(size_t)j * R.len() + i];
  int r;
  if ((r = upload(c, c->mode_tab, tab.data(), tab.size() * sizeof(double)))) return r;
  if ((r = upload(c, c->mode_tab_f, tabf.data(), tabf.size() * sizeof(float)))) return r;
  c->mode_tab_ok = true;
  return PIB_OK;
}

int make_eval_args(pib_ctx* c, int flags, int64_t B, EvalArgs* out_args, bool* wide_out) {
  const bool pm = (flags & PIB_EVAL_PM) != 0;
  if (flags & ~(PIB_EVAL_PM | PIB_EVAL_NOSCALE | PIB_EVAL_WIDE | PIB_EVAL_FP32 | PIB_EVAL_BF16X3))
    return fail(c, PIB_ERR_ARG, "eval_block: unknown flag");
  if ((flags & PIB_EVAL_FP32) && (flags & PIB_EVAL_BF16X3))
    return fail(c, PIB_ERR_ARG, "eval_block: PIB_EVAL_FP32 and PIB_EVAL_BF16X3 are different kernels");
  if ((flags & PIB_EVAL_FP32) && (flags & PIB_EVAL_WIDE))
    return fail(c, PIB_ERR_UNSUPPORTED, "eval_block: the FP32 kernel is the lanes-per-sample kernel (no PIB_EVAL_WIDE)");
  if (!c->coupling_set || c->P < 3) return fail(c, PIB_ERR_STATE, "eval_block: coupling or beads not set");
  const int need = pm ? 4 : 2;
  for (int s = 0; s < need; ++s)
    if (!c->set_A[s]) return fail(c, PIB_ERR_STATE, "eval_block: O-matrix set " + std::to_string(s) + " not set");
  for (int s = 1; s < need; ++s)
    if (c->set_A[s] != c->A || c->set_N[s] != c->N)
      return fail(c, PIB_ERR_ARG, "eval_block: coupled-model O-matrix set has the wrong surface/mode count");
  if (c->set_N[0] != c->N) return fail(c, PIB_ERR_ARG, "eval_block: sampling model has a different number of modes");
  const int AR = c->set_A[SET_RHO];
  const bool fits_wide = c->A <= kWideMaxStates && AR <= kWideMaxStates;
  const bool fp32 = (flags & PIB_EVAL_FP32) != 0;
  const bool bf = (flags & PIB_EVAL_BF16X3) != 0;
  if (bf && !fits_wide)
    return fail(c, PIB_ERR_UNSUPPORTED, "eval_block: BF16x3 ring products are in the wide kernel (A, A_rho <= 16)");
  const bool wide = !fp32 && fits_wide && (bf || (flags & PIB_EVAL_WIDE) || !eval_supported(c->A, AR));
  if (!wide)
    for (int s = 0; s < need; ++s)
      if (!c->pref_positive[s])
        return fail(c, PIB_ERR_UNSUPPORTED, std::string("eval_block: the ") + (fp32 ? "FP32" : "FP64") +
                                                " lanes kernel works with log F_a and needs every O-matrix prefactor > 0");
  if (!wide && !eval_supported(c->A, AR))
    return fail(c, PIB_ERR_UNSUPPORTED, std::string("eval_block: no ") + (fp32 ? "FP32" : "FP64") +
                                            " kernel for A=" + std::to_string(c->A) +
                                            ", A_rho=" + std::to_string(AR));
  if (B < 0) return fail(c, PIB_ERR_ARG, "eval_block: B < 0 or out_stride < B");
  if (!c->mode_tab_ok) return fail(c, PIB_ERR_STATE, "eval_block: per-mode constants not built");
  EvalArgs a{};
  a.B = B;
  a.N = c->N;
  a.P = c->P;
  a.G = c->lanes ? c->lanes : auto_lanes(c, B, fp32);
  a.K = (c->P + a.G - 1) / a.G;
  a.noscale = (flags & PIB_EVAL_NOSCALE) ? 1 : 0;
  a.tau = c->tau;
  for (int s = 0; s < NUM_SETS; ++s) {
    a.set[s].coth = (const double*)c->coth[s].p;
    a.set[s].csch = (const double*)c->csch[s].p;
    a.set[s].pref = (const double*)c->pref[s].p;
    a.set[s].shift = (const double*)c->shift[s].p;
    a.set[s].th = (const double*)c->th[s].p;
    a.set[s].lin = (const double*)c->lin[s].p;
    a.set[s].k2 = (const double*)c->k2[s].p;
    a.set[s].lpref = (const double*)c->lpref[s].p;
    if (s >= need) a.set[s] = a.set[SET_VIB];
  }
  a.energy = (const double*)c->energy.p;
  a.linear = (const double*)c->linear.p;
  a.qdiag = c->has_qdiag ? (const double*)c->qdiag.p : nullptr;
  a.npair = c->npair;
  a.pair = (const int*)c->pair.p;
  a.qpair = (const double*)c->qpair.p;
  a.mode_tab = (const double*)c->mode_tab.p;
  a.mode_tab_f = (const float*)c->mode_tab_f.p;
  a.A = c->A;
  a.AR = AR;
  a.wide_lq = (const double2*)c->w_lq.p;
  a.wide_e = (const double*)c->w_e.p;
  a.wide_qpair = (const double*)c->w_qpair.p;
  for (int s = 0; s < NUM_SETS; ++s) {
    const int src = (s >= need) ? SET_VIB : s;
    a.setf[s].tc = (const float*)c->tc_f[src].p;
    a.setf[s].cs = (const float*)c->cs_f[src].p;
    a.setf[s].lpref = (const float*)c->lpref_f[src].p;
    a.setf[s].k2 = (const float*)c->k2_f[src].p;
  }
  a.energy_f = (const float*)c->energy_f.p;
  a.linear_f = (const float*)c->linear_f.p;
  a.qdiag_f = c->has_qdiag ? (const float*)c->qdiag_f.p : nullptr;
  a.qpair_f = (const float*)c->qpair_f.p;
  *out_args = a;
  *wide_out = wide;
  return PIB_OK;
}

}  // namespace pib

extern "C" {

int pib_abi_version(void) { return PIB_ABI_VERSION; }

int pib_ctx_create(int device, pib_ctx** out) {
  if (!out) return PIB_ERR_ARG;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) return PIB_ERR_HIP;
  pib_ctx* c = new pib_ctx();
  c->device = device;
  if (hipSetDevice(device) != hipSuccess) {
    delete c;
    return PIB_ERR_HIP;
  }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess && prop.multiProcessorCount > 0)
    c->num_cu = prop.multiProcessorCount;
  // the allocation-free block sums' partial rows: resolved here, with the device current
  if (sum_scratch_address(device, &c->sum_scratch) != hipSuccess) {
    delete c;
    return PIB_ERR_HIP;
  }
  *out = c;
  return PIB_OK;
}

int pib_ctx_destroy(pib_ctx* c) {
  if (!c) return PIB_OK;
  (void)hipSetDevice(c->device);
  (void)hipDeviceSynchronize();
  release(c->energy);
  release(c->linear);
  release(c->qdiag);
  release(c->pair);
  release(c->qpair);
  release(c->energy_f);
  release(c->linear_f);
  release(c->qdiag_f);
  release(c->qpair_f);
  release(c->w_lq);
  release(c->w_e);
  release(c->w_qpair);
  for (int s = 0; s < NUM_SETS; ++s) {
    release(c->coth[s]);
    release(c->csch[s]);
    release(c->pref[s]);
    release(c->shift[s]);
    release(c->tc_f[s]);
    release(c->cs_f[s]);
    release(c->lpref_f[s]);
    release(c->th[s]);
    release(c->lin[s]);
    release(c->k2[s]);
    release(c->lpref[s]);
    release(c->k2_f[s]);
  }
  release(c->mode_tab);
  release(c->fb_mu);
  release(c->fb_c);
  release(c->fb_qidx);
  release(c->fb_qlist);
  release(c->mode_tab_f);
  release(c->s_cdf);
  release(c->s_chol);
  release(c->s_shift);
  release(c->partials);
  release(c->work);
  release(c->f_U);
  release(c->f_sigma);
  release(c->f_isigma);
  release(c->f_m);
  delete c;
  return PIB_OK;
}

const char* pib_last_error(const pib_ctx* c) { return c ? c->err.c_str() : "null context"; }

int pib_set_coupling(pib_ctx* c, int A, int N, const double* E, const double* L,
                     const double* Q) {
  if (!c) return PIB_ERR_ARG;
  ++c->gen;
  if (A < 1 || N < 1 || !E || !L) return fail(c, PIB_ERR_ARG, "set_coupling: bad A/N or null array");
  const size_t AA = (size_t)A * A;
  if (!finite_all(E, AA) || !finite_all(L, (size_t)N * AA) || (Q && !finite_all(Q, (size_t)N * N * AA)))
    return fail(c, PIB_ERR_ARG, "set_coupling: non-finite model parameter");
  // V must be symmetric in surfaces (pibronic/pimc/pimc.py:1169)
  for (int j = 0; j <= N; ++j) {
    const double* X = (j == N) ? E : L + (size_t)j * AA;
    for (int b = 0; b < A; ++b)
      for (int d = 0; d < b; ++d)
        if (std::fabs(X[b * A + d] - X[d * A + b]) > 1e-8 + 1e-5 * std::fabs(X[d * A + b]))
          return fail(c, PIB_ERR_ARG, "set_coupling: coupling matrix is not symmetric in surfaces");
  }
  int r;
  if ((r = upload(c, c->energy, E, AA * sizeof(double)))) return r;
  if ((r = upload(c, c->linear, L, (size_t)N * AA * sizeof(double)))) return r;
  std::vector<double> qd((size_t)N * AA, 0.0), qp;
  std::vector<int> pairs;
  bool any_diag = false;
  if (Q) {
    for (int j = 0; j < N; ++j)
      for (int k = 0; k < N; ++k) {
        const double* blk = Q + ((size_t)j * N + k) * AA;
        bool nz = false;
        for (size_t i = 0; i < AA; ++i) nz |= (blk[i] != 0.0);
        if (!nz) continue;
        if (j == k) {
          any_diag = true;
          for (size_t i = 0; i < AA; ++i) qd[(size_t)j * AA + i] = 0.5 * blk[i];
        } else {
          pairs.push_back(j);
          pairs.push_back(k);
          for (size_t i = 0; i < AA; ++i) qp.push_back(0.5 * blk[i]);
        }
      }
  }
  c->has_qdiag = any_diag;
  if (any_diag && (r = upload(c, c->qdiag, qd.data(), qd.size() * sizeof(double)))) return r;
  c->h_linear.assign(L, L + (size_t)N * AA);
  if (any_diag) c->h_qdiag = qd; else c->h_qdiag.clear();
  c->npair = (int)(pairs.size() / 2);
  if (c->npair) {
    if ((r = upload(c, c->pair, pairs.data(), pairs.size() * sizeof(int)))) return r;
    if ((r = upload(c, c->qpair, qp.data(), qp.size() * sizeof(double)))) return r;
  }
  {  // FP32 copies
    auto tof = [](const double* x, size_t n) { return std::vector<float>(x, x + n); };
    std::vector<float> ef = tof(E, AA), lf = tof(L, (size_t)N * AA), qdf = tof(qd.data(), qd.size());
    std::vector<float> qpf = tof(qp.data(), qp.size());
    if (qpf.empty()) qpf.push_back(0.f);
    if ((r = upload(c, c->energy_f, ef.data(), ef.size() * sizeof(float)))) return r;
    if ((r = upload(c, c->linear_f, lf.data(), lf.size() * sizeof(float)))) return r;
    if ((r = upload(c, c->qdiag_f, qdf.data(), qdf.size() * sizeof(float)))) return r;
    if ((r = upload(c, c->qpair_f, qpf.data(), qpf.size() * sizeof(float)))) return r;
  }
  if (A <= kWideMaxStates) {
    // 16 x 16 padded copies for the wide kernel; element (b, c) of every matrix is taken
    // from the lower triangle (max(b,c), min(b,c)), as eigh(UPLO='L') reads V
    constexpr int W = kWideMaxStates;
    auto lo = [A](int b, int cc) { return b >= cc ? b * A + cc : cc * A + b; };
    std::vector<double> lq((size_t)N * W * W * 2, 0.0), ew((size_t)W * W, 0.0);
    std::vector<double> qw((size_t)std::max(c->npair, 1) * W * W, 0.0);
    for (int b = 0; b < A; ++b)
      for (int cc = 0; cc < A; ++cc) {
        ew[b * W + cc] = E[lo(b, cc)];
        for (int j = 0; j < N; ++j) {
          lq[(((size_t)j * W + b) * W + cc) * 2] = L[(size_t)j * AA + lo(b, cc)];
          lq[(((size_t)j * W + b) * W + cc) * 2 + 1] = qd[(size_t)j * AA + lo(b, cc)];
        }
        for (int t = 0; t < c->npair; ++t) qw[((size_t)t * W + b) * W + cc] = qp[(size_t)t * AA + lo(b, cc)];
      }
    if ((r = upload(c, c->w_lq, lq.data(), lq.size() * sizeof(double)))) return r;
    if ((r = upload(c, c->w_e, ew.data(), ew.size() * sizeof(double)))) return r;
    if ((r = upload(c, c->w_qpair, qw.data(), qw.size() * sizeof(double)))) return r;
  }
  c->A = A;
  c->N = N;
  c->coupling_set = true;
  return rebuild_mode_table(c);
}

int pib_set_omatrix_set(pib_ctx* c, int slot, int A_set, int N, const double* coth,
                        const double* csch, const double* pref, const double* shift) {
  if (!c) return PIB_ERR_ARG;
  ++c->gen;
  if (slot < 0 || slot >= NUM_SETS) return fail(c, PIB_ERR_ARG, "set_omatrix_set: bad slot");
  if (A_set < 1 || N < 1 || !coth || !csch || !pref || !shift)
    return fail(c, PIB_ERR_ARG, "set_omatrix_set: bad size or null array");
  if (!finite_all(coth, N) || !finite_all(csch, N) || !finite_all(pref, A_set) ||
      !finite_all(shift, (size_t)A_set * N))
    return fail(c, PIB_ERR_ARG, "set_omatrix_set: non-finite constant");
  int r;
  if ((r = upload(c, c->coth[slot], coth, N * sizeof(double)))) return r;
  if ((r = upload(c, c->csch[slot], csch, N * sizeof(double)))) return r;
  if ((r = upload(c, c->pref[slot], pref, A_set * sizeof(double)))) return r;
  if ((r = upload(c, c->shift[slot], shift, (size_t)A_set * N * sizeof(double)))) return r;
  {  // expanded exponent constants (FP64 lanes kernel)
    std::vector<double> th(N), lin((size_t)A_set * N), k2(A_set);
    for (int j = 0; j < N; ++j) th[j] = coth[j] - csch[j];
    for (int a = 0; a < A_set; ++a) {
      double acc = 0.0;
      for (int j = 0; j < N; ++j) {
        const double d = shift[(size_t)a * N + j];
        lin[(size_t)a * N + j] = -2.0 * th[j] * d;
        acc += th[j] * d * d;
      }
      k2[a] = 2.0 * acc;
    }
    if ((r = upload(c, c->th[slot], th.data(), N * sizeof(double)))) return r;
    if ((r = upload(c, c->lin[slot], lin.data(), (size_t)A_set * N * sizeof(double)))) return r;
    if ((r = upload(c, c->k2[slot], k2.data(), A_set * sizeof(double)))) return r;
    std::vector<double> lp(A_set);
    for (int q = 0; q < A_set; ++q) lp[q] = pref[q] > 0.0 ? std::log(pref[q]) : 0.0;
    if ((r = upload(c, c->lpref[slot], lp.data(), A_set * sizeof(double)))) return r;
    std::vector<float> k2f(k2.begin(), k2.end());
    if ((r = upload(c, c->k2_f[slot], k2f.data(), A_set * sizeof(float)))) return r;
    c->h_th[slot] = th;
    c->h_cs[slot].assign(csch, csch + N);
    c->h_lin[slot] = lin;
  }
  {  // FP32 kernel constants
    std::vector<float> tc(N), cs(N), lp(A_set);
    for (int j = 0; j < N; ++j) {
      tc[j] = (float)(coth[j] - csch[j]);
      cs[j] = (float)csch[j];
    }
    bool pos = true;
    for (int a = 0; a < A_set; ++a) {
      pos = pos && pref[a] > 0.0;
      lp[a] = pref[a] > 0.0 ? (float)std::log(pref[a]) : 0.f;
    }
    c->pref_positive[slot] = pos;
    if ((r = upload(c, c->tc_f[slot], tc.data(), N * sizeof(float)))) return r;
    if ((r = upload(c, c->cs_f[slot], cs.data(), N * sizeof(float)))) return r;
    if ((r = upload(c, c->lpref_f[slot], lp.data(), A_set * sizeof(float)))) return r;
  }
  c->set_A[slot] = A_set;
  c->set_N[slot] = N;
  return rebuild_mode_table(c);
}

int pib_set_beads(pib_ctx* c, int P, double tau) {
  if (!c) return PIB_ERR_ARG;
  ++c->gen;
  if (P < 3) return fail(c, PIB_ERR_ARG, "circulant matrix requires 3 or more beads");
  if (!std::isfinite(tau) || tau <= 0.0) return fail(c, PIB_ERR_ARG, "set_beads: tau must be > 0");
  c->P = P;
  c->tau = tau;
  return PIB_OK;
}

int pib_set_lanes_per_sample(pib_ctx* c, int lanes) {
  if (!c) return PIB_ERR_ARG;
  if (lanes != 0 && (lanes < 1 || lanes > 64 || (lanes & (lanes - 1))))
    return fail(c, PIB_ERR_ARG, "lanes per sample must be 0 or a power of two <= 64");
  c->lanes = lanes;
  return PIB_OK;
}

int pib_get_lanes_per_sample(pib_ctx* c, int64_t B) {
  if (!c) return PIB_ERR_ARG;
  if (c->lanes) return c->lanes;
  if (c->P < 3) return fail(c, PIB_ERR_STATE, "beads not set");
  return auto_lanes(c, B, false);
}

int pib_eval_block(pib_ctx* c, const double* q, int64_t B, int flags, double* out,
                   int64_t out_stride, double* dbg, void* stream) {
  if (!c) return PIB_ERR_ARG;
  EvalArgs a{};
  bool wide = false;
  int r = make_eval_args(c, flags, B, &a, &wide);
  if (r) return r;
  if (out_stride < B) return fail(c, PIB_ERR_ARG, "eval_block: B < 0 or out_stride < B");
  if (B == 0) return PIB_OK;
  if (!q || !out) return fail(c, PIB_ERR_ARG, "eval_block: null q or out");
  a.q = q;
  a.out = out;
  a.out_stride = out_stride;
  a.dbg = dbg;
  PIB_HIP(c, hipSetDevice(c->device));
  const int pm = (flags & PIB_EVAL_PM) ? 1 : 0;
  hipError_t e = (flags & PIB_EVAL_FP32) ? launch_eval_fp32(c->A, a.AR, pm, a, (hipStream_t)stream)
                 : wide ? launch_eval_wide(pm, a, (hipStream_t)stream, (flags & PIB_EVAL_BF16X3) != 0)
                                         : launch_eval_fp64(c->A, a.AR, pm, a, (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(c, e, "eval_block launch");
  return PIB_OK;
}

int pib_eval_block_flow_offsets(pib_ctx* c, const float* dq, int64_t B, int flags, double* out,
                                int64_t out_stride, void* stream) {
  if (!c) return PIB_ERR_ARG;
  if (flags & ~PIB_EVAL_NOSCALE)
    return fail(c, PIB_ERR_UNSUPPORTED, "eval_block_flow_offsets: flags other than PIB_EVAL_NOSCALE");
  if (!c->f_P || !c->f_m.p) return fail(c, PIB_ERR_STATE, "eval_block_flow_offsets: flow basis not set");
  if (c->f_N != c->N || c->f_P != c->P)
    return fail(c, PIB_ERR_STATE, "eval_block_flow_offsets: flow basis and beads disagree in N or P");
  EvalArgs a{};
  bool wide = false;
  int r = make_eval_args(c, flags, B, &a, &wide);
  if (r) return r;
  if (wide || a.AR != c->A || c->A > 8)
    return fail(c, PIB_ERR_UNSUPPORTED, "eval_block_flow_offsets: the FP64 lanes kernel with A_rho = A <= 8 only");
  if (out_stride < B) return fail(c, PIB_ERR_ARG, "eval_block_flow_offsets: out_stride < B");
  if (B == 0) return PIB_OK;
  if (!dq || !out) return fail(c, PIB_ERR_ARG, "eval_block_flow_offsets: null dq or out");
  a.qf = dq;
  a.qcentre = (const double*)c->f_m.p;
  a.out = out;
  a.out_stride = out_stride;
  PIB_HIP(c, hipSetDevice(c->device));
  hipError_t e = launch_eval_fp64_offsets(c->A, a, (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(c, e, "eval_block_flow_offsets launch");
  return PIB_OK;
}

int pib_set_sampler(pib_ctx* c, int A_rho, int N, int P, const double* w, const double* coth,
                    const double* csch, const double* shift) {
  if (!c) return PIB_ERR_ARG;
  ++c->gen;
  if (A_rho < 1 || N < 1 || P < 3 || !w || !coth || !csch || !shift)
    return fail(c, PIB_ERR_ARG, "set_sampler: bad size or null array");
  if (P > 4096) return fail(c, PIB_ERR_ARG, "set_sampler: P > 4096 not supported");
  if (!finite_all(w, A_rho) || !finite_all(coth, N) || !finite_all(csch, N) ||
      !finite_all(shift, (size_t)A_rho * N))
    return fail(c, PIB_ERR_ARG, "set_sampler: non-finite constant");
  std::vector<double> cdf(A_rho);
  double tot = 0.0;
  for (int a = 0; a < A_rho; ++a) {
    if (w[a] < 0.0) return fail(c, PIB_ERR_ARG, "set_sampler: negative surface weight");
    tot += w[a];
  }
  if (!(tot > 0.0)) return fail(c, PIB_ERR_ARG, "set_sampler: surface weights sum to zero");
  double run = 0.0;
  for (int a = 0; a < A_rho; ++a) {
    run += w[a];
    cdf[a] = run / tot;
  }
  cdf[A_rho - 1] = 1.0;
  // Cholesky factor of the cyclic tridiagonal precision matrix of each mode:
  // Lambda = alpha I + beta C, alpha = 2 coth, beta = -csch, C = ring adjacency (ones at
  // +-1 mod P). Non-zeros of L: diagonal, subdiagonal (k+1, k) for k <= P-3, last row.
  std::vector<double> chol((size_t)N * 3 * P, 0.0);
  for (int j = 0; j < N; ++j) {
    const double alpha = 2.0 * coth[j], beta = -csch[j];
    double* dinv = chol.data() + (size_t)j * 3 * P;
    double* sub = dinv + P;
    double* last = sub + P;
    std::vector<double> diag(P);
    double lastsq = 0.0;
    for (int k = 0; k < P - 1; ++k) {
      const double ps = k >= 1 ? sub[k - 1] : 0.0;
      const double dd = alpha - ps * ps;
      if (!(dd > 0.0)) return fail(c, PIB_ERR_ARG, "set_sampler: precision matrix is not positive definite");
      diag[k] = std::sqrt(dd);
      if (k <= P - 3) sub[k] = beta / diag[k];
      const double lam = (k == 0 || k == P - 2) ? beta : 0.0;
      const double pl = k >= 1 ? last[k - 1] * sub[k - 1] : 0.0;
      last[k] = (lam - pl) / diag[k];
      lastsq += last[k] * last[k];
    }
    const double dl = alpha - lastsq;
    if (!(dl > 0.0)) return fail(c, PIB_ERR_ARG, "set_sampler: precision matrix is not positive definite");
    diag[P - 1] = std::sqrt(dl);
    for (int k = 0; k < P; ++k) dinv[k] = 1.0 / diag[k];
  }
  int r;
  if ((r = upload(c, c->s_cdf, cdf.data(), cdf.size() * sizeof(double)))) return r;
  if ((r = upload(c, c->s_chol, chol.data(), chol.size() * sizeof(double)))) return r;
  if ((r = upload(c, c->s_shift, shift, (size_t)A_rho * N * sizeof(double)))) return r;
  c->s_A = A_rho;
  c->s_N = N;
  c->s_P = P;
  c->sampler_set = true;
  return PIB_OK;
}

int pib_sample_block(pib_ctx* c, uint64_t seed, int64_t offset, int64_t B, double* q,
                     int32_t* src, void* stream) {
  if (!c) return PIB_ERR_ARG;
  if (!c->sampler_set) return fail(c, PIB_ERR_STATE, "sample_block: sampler not set");
  if (B < 0 || offset < 0) return fail(c, PIB_ERR_ARG, "sample_block: negative B or offset");
  if (B == 0) return PIB_OK;
  if (!q) return fail(c, PIB_ERR_ARG, "sample_block: null q");
  SampleArgs a{};
  a.B = B;
  a.N = c->s_N;
  a.P = c->s_P;
  a.A_rho = c->s_A;
  a.offset = offset;
  a.key0 = (uint32_t)seed;
  a.key1 = (uint32_t)(seed >> 32);
  a.cdf = (const double*)c->s_cdf.p;
  a.chol = (const double*)c->s_chol.p;
  a.shift = (const double*)c->s_shift.p;
  a.q = q;
  a.src = src;
  PIB_HIP(c, hipSetDevice(c->device));
  hipError_t e = launch_sample_fp64(a, (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(c, e, "sample_block launch");
  return PIB_OK;
}

static int block_sums_impl(pib_ctx* c, const double* out, int64_t out_stride, int64_t B, int pm,
                           double delta_beta, const double* logq, double* sums, void* stream);

int pib_block_sums(pib_ctx* c, const double* out, int64_t out_stride, int64_t B, int pm,
                   double delta_beta, double* sums, void* stream) {
  return block_sums_impl(c, out, out_stride, B, pm, delta_beta, nullptr, sums, stream);
}

int pib_block_sums_q(pib_ctx* c, const double* out, int64_t out_stride, int64_t B, int pm,
                     double delta_beta, const double* logq, double* sums, void* stream) {
  if (!c) return PIB_ERR_ARG;
  if (!logq) return fail(c, PIB_ERR_ARG, "block_sums_q: null logq");
  return block_sums_impl(c, out, out_stride, B, pm, delta_beta, logq, sums, stream);
}

static int block_sums_impl(pib_ctx* c, const double* out, int64_t out_stride, int64_t B, int pm,
                           double delta_beta, const double* logq, double* sums, void* stream) {
  if (!c) return PIB_ERR_ARG;
  if (!out || !sums || B < 1 || out_stride < B)
    return fail(c, PIB_ERR_ARG, "block_sums: null pointer, B < 1 or out_stride < B");
  if (pm && !(delta_beta > 0.0)) return fail(c, PIB_ERR_ARG, "block_sums: delta_beta must be > 0");
  const int nparts = sum_parts(B);
  if (c->nparts < nparts) {
    std::vector<double> zero((size_t)nparts * NUM_SUMS, 0.0);
    int r = upload(c, c->partials, zero.data(), zero.size() * sizeof(double));
    if (r) return r;
    c->nparts = nparts;
  }
  SumArgs a{};
  a.out = out;
  a.out_stride = out_stride;
  a.B = B;
  a.pm = pm ? 1 : 0;
  a.delta_beta = pm ? delta_beta : 1.0;
  a.logq = logq;
  PIB_HIP(c, hipSetDevice(c->device));
  hipError_t e = launch_block_sums(a, (double*)c->partials.p, nparts, sums, (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(c, e, "block_sums launch");
  return PIB_OK;
}

int pib_block_sums_noalloc(pib_ctx* c, const double* out, int64_t out_stride, int64_t B, int pm,
                           double delta_beta, const double* logq, double* sums, void* stream) {
  if (!c) return PIB_ERR_ARG;
  if (!out || !sums || B < 1 || out_stride < B)
    return fail(c, PIB_ERR_ARG, "block_sums_noalloc: null pointer, B < 1 or out_stride < B");
  if (pm && !(delta_beta > 0.0)) return fail(c, PIB_ERR_ARG, "block_sums_noalloc: delta_beta must be > 0");
  const int nparts = sum_parts(B);
  if (nparts > kSumMaxParts)
    return fail(c, PIB_ERR_ARG, "block_sums_noalloc: more partial rows than the static scratch holds");
  if (!c->sum_scratch) return fail(c, PIB_ERR_STATE, "block_sums_noalloc: static scratch not resolved");
  SumArgs a{};
  a.out = out;
  a.out_stride = out_stride;
  a.B = B;
  a.pm = pm ? 1 : 0;
  a.delta_beta = pm ? delta_beta : 1.0;
  a.logq = logq;
  PIB_HIP(c, hipSetDevice(c->device));
  hipError_t e = launch_block_sums(a, c->sum_scratch, nparts, sums, (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(c, e, "block_sums_noalloc launch");
  return PIB_OK;
}

static int gemm_impl(pib_ctx* c, int bf16x3, int ta, int tb, int64_t M, int64_t N, int64_t K, const float* A,
                 int64_t lda, const float* B, int64_t ldb, float* C, int64_t ldc,
                 const float* bias, int relu, const float* mask, int64_t ldmask, int accumulate,
                 void* stream) {
  if (!c) return PIB_ERR_ARG;
  GemmArgs g{A, B, C, bias, mask, M, N, K, lda, ldb, ldc, ldmask, relu ? 1 : 0, accumulate ? 1 : 0,
             0, nullptr, bf16x3};
  if (!A || !B || !C) return fail(c, PIB_ERR_ARG, "gemm_f32: null operand");
  if (!gemm_shape_ok(g))
    return fail(c, PIB_ERR_ARG, bf16x3 ? "gemm_bf16x3: M and N must be multiples of 64, K of 32"
                                       : "gemm_f32: M and N must be multiples of 64, K of 16");
  const int64_t rowsA = ta ? K : M, colsA = ta ? M : K, rowsB = tb ? N : K, colsB = tb ? K : N;
  (void)rowsA; (void)rowsB;
  if (lda < colsA || ldb < colsB || ldc < N || (mask && ldmask < N))
    return fail(c, PIB_ERR_ARG, "gemm_f32: leading dimension smaller than the row length");
  PIB_HIP(c, hipSetDevice(c->device));
  const int splits = gemm_splits(g, c->num_cu);
  if (splits > 1) {
    int r;
    if ((r = reserve(c, c->work, (size_t)splits * M * N * sizeof(float)))) return r;
    g.kchunk = K / splits;
    g.partial = (float*)c->work.p;
  }
  hipError_t e = launch_gemm_f32(ta != 0, tb != 0, g, (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(c, e, "gemm_f32 launch");
  return PIB_OK;
}

int pib_gemm_f32(pib_ctx* c, int ta, int tb, int64_t M, int64_t N, int64_t K, const float* A,
                 int64_t lda, const float* B, int64_t ldb, float* C, int64_t ldc,
                 const float* bias, int relu, const float* mask, int64_t ldmask, int accumulate,
                 void* stream) {
  return gemm_impl(c, 0, ta, tb, M, N, K, A, lda, B, ldb, C, ldc, bias, relu, mask, ldmask, accumulate, stream);
}

int pib_gemm_bf16x3(pib_ctx* c, int ta, int tb, int64_t M, int64_t N, int64_t K, const float* A,
                    int64_t lda, const float* B, int64_t ldb, float* C, int64_t ldc,
                    const float* bias, int relu, const float* mask, int64_t ldmask, int accumulate,
                    void* stream) {
  return gemm_impl(c, 1, ta, tb, M, N, K, A, lda, B, ldb, C, ldc, bias, relu, mask, ldmask, accumulate, stream);
}

int pib_gemm_bf16x3_presplit(pib_ctx* c, int64_t M, int64_t N, int64_t K, const uint16_t* A_hi,
                             const uint16_t* A_lo, int64_t lda, const uint16_t* Bt_hi,
                             const uint16_t* Bt_lo, int64_t ldb, float* C, int64_t ldc,
                             uint16_t* O_hi, uint16_t* O_lo, int64_t ldo, const float* bias, int relu,
                             void* stream) {
  if (!c) return PIB_ERR_ARG;
  GemmPreArgs g{A_hi, A_lo, Bt_hi, Bt_lo, C, O_hi, O_lo, bias, M, N, K, lda, ldb, ldc, ldo, relu ? 1 : 0};
  if (!A_hi || !A_lo || !Bt_hi || !Bt_lo) return fail(c, PIB_ERR_ARG, "gemm_bf16x3_presplit: null operand");
  if (lda < K || ldb < K || (C && ldc < N) || (O_hi && ldo < N))
    return fail(c, PIB_ERR_ARG, "gemm_bf16x3_presplit: leading dimension smaller than the row length");
  if (!gemm_pre_shape_ok(g))
    return fail(c, PIB_ERR_ARG, "gemm_bf16x3_presplit: M, N multiples of 64, K of 32, lda/ldb multiples of 8, "
                                "16-byte aligned operands, an FP32 or a split output (hi and lo together)");
  PIB_HIP(c, hipSetDevice(c->device));
  hipError_t e = launch_gemm_bf16x3_pre(g, (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(c, e, "gemm_bf16x3_presplit launch");
  return PIB_OK;
}

int pib_split_bf16(pib_ctx* c, int64_t rows, int64_t cols, const float* X, int64_t ldx, int mode,
                   uint16_t* hi, uint16_t* lo, int64_t ldo, void* stream) {
  if (!c) return PIB_ERR_ARG;
  if (mode & ~(PIB_SPLIT_TRANSPOSE | PIB_SPLIT_PERM32 | PIB_SPLIT_FRAG32))
    return fail(c, PIB_ERR_ARG, "split_bf16: unknown mode bit");
  if ((mode & PIB_SPLIT_PERM32) && (!(mode & PIB_SPLIT_TRANSPOSE) || rows % 32))
    return fail(c, PIB_ERR_ARG, "split_bf16: PIB_SPLIT_PERM32 needs PIB_SPLIT_TRANSPOSE and rows a multiple of 32");
  if ((mode & PIB_SPLIT_FRAG32) && (mode != (PIB_SPLIT_TRANSPOSE | PIB_SPLIT_FRAG32) || cols % 32 || rows % 16))
    return fail(c, PIB_ERR_ARG, "split_bf16: PIB_SPLIT_FRAG32 needs PIB_SPLIT_TRANSPOSE alone beside it, "
                                "cols a multiple of 32 and rows of 16");
  const bool tr = (mode & PIB_SPLIT_TRANSPOSE) != 0;
  if (rows < 1 || cols < 1 || !X || !hi || ldx < cols || ldo < (tr ? rows : cols))
    return fail(c, PIB_ERR_ARG, "split_bf16: bad size, leading dimension or null pointer");
  PIB_HIP(c, hipSetDevice(c->device));
  hipError_t e = launch_split_bf16(rows, cols, X, ldx, mode, hi, lo, ldo, (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(c, e, "split_bf16 launch");
  return PIB_OK;
}

int pib_flow_layer_fused(pib_ctx* c, int inverse, int64_t B, int D, int hidden, int a_off, int b_off,
                         float smax, float* z, double* logdet, const uint16_t* W1_hi,
                         const uint16_t* W1_lo, const uint16_t* W2_hi, const uint16_t* W2_lo,
                         const uint16_t* W3_hi, const uint16_t* W3_lo, const float* b1, const float* b2,
                         const float* b3, void* stream) {
  if (!c) return PIB_ERR_ARG;
  if (!z || !logdet || !W1_hi || !W2_hi || !W3_hi || !b1 || !b2 || !b3)
    return fail(c, PIB_ERR_ARG, "flow_layer_fused: null pointer");
  if ((W1_lo == nullptr) != (W2_lo == nullptr) || (W1_lo == nullptr) != (W3_lo == nullptr))
    return fail(c, PIB_ERR_ARG, "flow_layer_fused: the lo parts must be all given or all null");
  if (!fused_layer_supported(B, D, hidden, W1_lo != nullptr))
    return fail(c, PIB_ERR_UNSUPPORTED, "flow_layer_fused: needs hidden 256, D = 512 (or 1024 with single "
                                        "bf16 products) and B a multiple of 128");
  const int H = D / 2;
  if (!((a_off == 0 && b_off == H) || (a_off == H && b_off == 0)))
    return fail(c, PIB_ERR_ARG, "flow_layer_fused: the halves must be (0, D/2) or (D/2, 0)");
  if (!(smax > 0.f)) return fail(c, PIB_ERR_ARG, "flow_layer_fused: smax must be > 0");
  FusedLayerArgs a{B, D, hidden, a_off, b_off, inverse ? 1 : 0, smax, z, logdet,
                   W1_hi, W1_lo, W2_hi, W2_lo, W3_hi, W3_lo, b1, b2, b3};
  PIB_HIP(c, hipSetDevice(c->device));
  hipError_t e = launch_flow_layer(a, (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(c, e, "flow_layer_fused launch");
  return PIB_OK;
}

int pib_flow_layer_fused_frag(pib_ctx* c, int inverse, int64_t B, int D, int hidden, int a_off, int b_off,
                              float smax, float* z, double* logdet, const uint16_t* W1_frag,
                              const uint16_t* W2_frag, const uint16_t* W3_frag, const float* b1,
                              const float* b2, const float* b3, void* stream) {
  if (!c) return PIB_ERR_ARG;
  if (!z || !logdet || !W1_frag || !W2_frag || !W3_frag || !b1 || !b2 || !b3)
    return fail(c, PIB_ERR_ARG, "flow_layer_fused_frag: null pointer");
  if (!frag_layer_supported(B, D, hidden))
    return fail(c, PIB_ERR_UNSUPPORTED, "flow_layer_fused_frag: needs hidden 256, D = 512 or 1024 and B a multiple of 128");
  const int H = D / 2;
  if (!((a_off == 0 && b_off == H) || (a_off == H && b_off == 0)))
    return fail(c, PIB_ERR_ARG, "flow_layer_fused_frag: the halves must be (0, D/2) or (D/2, 0)");
  if (!(smax > 0.f)) return fail(c, PIB_ERR_ARG, "flow_layer_fused_frag: smax must be > 0");
  if (((uintptr_t)W1_frag | (uintptr_t)W2_frag | (uintptr_t)W3_frag | (uintptr_t)z) % 16)
    return fail(c, PIB_ERR_ARG, "flow_layer_fused_frag: weights and z must be 16-byte aligned");
  FusedLayerArgs a{B, D, hidden, a_off, b_off, inverse ? 1 : 0, smax, z, logdet,
                   W1_frag, nullptr, W2_frag, nullptr, W3_frag, nullptr, b1, b2, b3};
  PIB_HIP(c, hipSetDevice(c->device));
  hipError_t e = launch_flow_layer_frag(a, (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(c, e, "flow_layer_fused_frag launch");
  return PIB_OK;
}

int pib_flow_set_basis(pib_ctx* c, int N, int P, const double* U, const double* sigma,
                       const double* m) {
  if (!c) return PIB_ERR_ARG;
  if (N < 1 || P < 64 || (P % 64) || !U || !sigma || !m)
    return fail(c, PIB_ERR_ARG, "flow_set_basis: need N >= 1, P a multiple of 64 and non-null arrays");
  if (!finite_all(U, (size_t)P * P) || !finite_all(sigma, (size_t)N * P) || !finite_all(m, N))
    return fail(c, PIB_ERR_ARG, "flow_set_basis: non-finite value");
  std::vector<float> Uf((size_t)P * P), sf((size_t)N * P), isf((size_t)N * P);
  for (size_t i = 0; i < Uf.size(); ++i) Uf[i] = (float)U[i];
  for (size_t i = 0; i < sf.size(); ++i) {
    if (!(sigma[i] > 0.0)) return fail(c, PIB_ERR_ARG, "flow_set_basis: sigma must be > 0");
    sf[i] = (float)sigma[i];
    isf[i] = (float)(1.0 / sigma[i]);
  }
  int r;
  if ((r = upload(c, c->f_U, Uf.data(), Uf.size() * sizeof(float)))) return r;
  if ((r = upload(c, c->f_sigma, sf.data(), sf.size() * sizeof(float)))) return r;
  if ((r = upload(c, c->f_isigma, isf.data(), isf.size() * sizeof(float)))) return r;
  if ((r = upload(c, c->f_m, m, (size_t)N * sizeof(double)))) return r;
  c->f_N = N;
  c->f_P = P;
  return PIB_OK;
}

int pib_flow_whiten(pib_ctx* c, int mode, int64_t B, const void* in, float* work, void* out,
                    void* stream) {
  if (!c) return PIB_ERR_ARG;
  if (!c->f_P) return fail(c, PIB_ERR_STATE, "flow_whiten: basis not set");
  if (mode < PIB_FLOW_WHITEN || mode > PIB_FLOW_UNWHITEN_OFFSETS) return fail(c, PIB_ERR_ARG, "flow_whiten: bad mode");
  if (B < 1 || !in || !work || !out) return fail(c, PIB_ERR_ARG, "flow_whiten: B < 1 or null pointer");
  const int N = c->f_N, P = c->f_P;
  if ((B * N) % 64 != 0 || P % 16 != 0)
    return fail(c, PIB_ERR_ARG, "flow_whiten: B*N must be a multiple of 64 and P of 16");
  BasisArgs ba{B, N, P, (const float*)c->f_sigma.p, (const float*)c->f_isigma.p, (const double*)c->f_m.p};
  float* X = work;
  float* Y = work + B * (int64_t)N * P;
  hipStream_t s = (hipStream_t)stream;
  PIB_HIP(c, hipSetDevice(c->device));
  const float* U = (const float*)c->f_U.p;
  hipError_t e;
  if (mode == PIB_FLOW_UNWHITEN_OFFSETS && !whiten_fused_supported(P))
    return fail(c, PIB_ERR_UNSUPPORTED, "flow_whiten: offsets need P = 64 or 128");
  if (whiten_fused_supported(P)) {
    const int m = mode == PIB_FLOW_WHITEN     ? FLOW_WHITEN
                  : mode == PIB_FLOW_UNWHITEN ? FLOW_UNWHITEN
                  : mode == PIB_FLOW_GRAD     ? FLOW_GRAD
                                              : FLOW_UNWHITEN_OFFSETS;
    if ((e = launch_whiten_fused(m, ba, in, out, U, c->num_cu, s)) != hipSuccess) return hip_fail(c, e, "flow whiten");
    return PIB_OK;
  }
  if (mode == PIB_FLOW_UNWHITEN) {
    if ((e = launch_flow_pre(FLOW_UNWHITEN, ba, in, Y, s)) != hipSuccess) return hip_fail(c, e, "flow pre");
    // X = Y U^T : B operand U transposed (B(k, p) = U[p][k])
    GemmArgs g{Y, U, X, nullptr, nullptr, B * N, P, P, P, P, P, 0, 0, 0};
    if ((e = launch_gemm_f32(false, true, g, s)) != hipSuccess) return hip_fail(c, e, "flow gemm");
    if ((e = launch_flow_post(FLOW_UNWHITEN, ba, X, out, s)) != hipSuccess) return hip_fail(c, e, "flow post");
  } else {
    const int pre = mode == PIB_FLOW_WHITEN ? FLOW_WHITEN : FLOW_GRAD;
    if ((e = launch_flow_pre(pre, ba, in, X, s)) != hipSuccess) return hip_fail(c, e, "flow pre");
    GemmArgs g{X, U, Y, nullptr, nullptr, B * N, P, P, P, P, P, 0, 0, 0};
    if ((e = launch_gemm_f32(false, false, g, s)) != hipSuccess) return hip_fail(c, e, "flow gemm");
    if ((e = launch_flow_post(pre, ba, Y, out, s)) != hipSuccess) return hip_fail(c, e, "flow post");
  }
  return PIB_OK;
}

int pib_flow_coupling(pib_ctx* c, int inverse, int64_t B, int D, int a_off, int b_off,
                      float smax, float* z, const float* o, double* logdet, void* stream) {
  if (!c) return PIB_ERR_ARG;
  if (B < 0 || D < 2 || (D & 1) || !z || !o || !logdet || !(smax > 0.0f))
    return fail(c, PIB_ERR_ARG, "flow_coupling: bad size, smax or null pointer");
  if (!((a_off == 0 && b_off == D / 2) || (a_off == D / 2 && b_off == 0)))
    return fail(c, PIB_ERR_ARG, "flow_coupling: halves must be (0, D/2) or (D/2, 0)");
  CouplingArgs a{};
  a.B = B; a.D = D; a.a_off = a_off; a.b_off = b_off; a.smax = smax;
  a.z = z; a.o = o; a.logdet = logdet;
  PIB_HIP(c, hipSetDevice(c->device));
  hipError_t e = launch_coupling(inverse ? 1 : 0, a, (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(c, e, "flow_coupling launch");
  return PIB_OK;
}

int pib_flow_coupling_inverse_backward(pib_ctx* c, int64_t B, int D, int a_off, int b_off,
                                       float smax, const float* y, const float* o, float* dz,
                                       float* dout, const double* cb, void* stream) {
  if (!c) return PIB_ERR_ARG;
  if (B < 0 || D < 2 || (D & 1) || !y || !o || !dz || !dout || !cb || !(smax > 0.0f))
    return fail(c, PIB_ERR_ARG, "flow_coupling_inverse_backward: bad size, smax or null pointer");
  if (!((a_off == 0 && b_off == D / 2) || (a_off == D / 2 && b_off == 0)))
    return fail(c, PIB_ERR_ARG, "flow_coupling_inverse_backward: halves must be (0, D/2) or (D/2, 0)");
  CouplingArgs a{};
  a.B = B; a.D = D; a.a_off = a_off; a.b_off = b_off; a.smax = smax;
  a.o = o; a.y = y; a.dz = dz; a.dout = dout; a.c = cb;
  PIB_HIP(c, hipSetDevice(c->device));
  hipError_t e = launch_coupling_inverse_backward(a, (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(c, e, "flow_coupling_inverse_backward launch");
  return PIB_OK;
}

int pib_flow_set_base(pib_ctx* c, int A_rho, int D, const double* mu, const double* cc) {
  if (!c) return PIB_ERR_ARG;
  if (A_rho < 1 || A_rho > 16 || D < 2 || (D & 1) || !mu || !cc)
    return fail(c, PIB_ERR_ARG, "flow_set_base: need 1 <= A_rho <= 16, D even and non-null arrays");
  if (!finite_all(mu, (size_t)A_rho * D) || !finite_all(cc, A_rho))
    return fail(c, PIB_ERR_ARG, "flow_set_base: non-finite value");
  if (D % 4) return fail(c, PIB_ERR_ARG, "flow_set_base: D must be a multiple of 4");
  // only the feature quads where some mu_a is non-zero are kept (the displacements live in the
  // constant normal mode of each bead mode: N of the D / 4 quads); the draw adds and dots
  // exactly those
  std::vector<int> qidx(D / 4, -1), qlist;
  std::vector<double> mq;
  int nq = 0;
  for (int m = 0; m < D / 4; ++m) {
    bool nz = false;
    for (int b = 0; b < A_rho && !nz; ++b)
      for (int e = 0; e < 4; ++e) nz = nz || mu[(size_t)b * D + 4 * m + e] != 0.0;
    if (!nz) continue;
    qidx[m] = nq++;
    qlist.push_back(m);
    for (int b = 0; b < A_rho; ++b)
      for (int e = 0; e < 4; ++e) mq.push_back(mu[(size_t)b * D + 4 * m + e]);
  }
  if (mq.empty()) mq.assign(4 * A_rho, 0.0);  // keep valid buffers
  if (qlist.empty()) qlist.push_back(0);
  int r;
  if ((r = upload(c, c->fb_mu, mq.data(), mq.size() * sizeof(double)))) return r;
  if ((r = upload(c, c->fb_qidx, qidx.data(), qidx.size() * sizeof(int)))) return r;
  if ((r = upload(c, c->fb_qlist, qlist.data(), qlist.size() * sizeof(int)))) return r;
  if ((r = upload(c, c->fb_c, cc, (size_t)A_rho * sizeof(double)))) return r;
  c->fb_A = A_rho;
  c->fb_D = D;
  c->fb_nq = nq;
  return PIB_OK;
}

int pib_flow_draw_base(pib_ctx* c, uint64_t seed, int64_t offset, int64_t B, float* z, double* logrho,
                       void* stream) {
  if (!c) return PIB_ERR_ARG;
  if (!c->fb_A) return fail(c, PIB_ERR_STATE, "flow_draw_base: base distribution not set");
  if (!c->sampler_set || c->s_A != c->fb_A)
    return fail(c, PIB_ERR_STATE, "flow_draw_base: sampler surface weights not set for this A_rho");
  if (B < 0 || offset < 0) return fail(c, PIB_ERR_ARG, "flow_draw_base: negative B or offset");
  if (B == 0) return PIB_OK;
  if (!z || !logrho) return fail(c, PIB_ERR_ARG, "flow_draw_base: null output");
  BaseArgs a{};
  a.B = B;
  a.offset = offset;
  a.D = c->fb_D;
  a.A_rho = c->fb_A;
  a.key0 = (uint32_t)seed;
  a.key1 = (uint32_t)(seed >> 32);
  a.cdf = (const double*)c->s_cdf.p;
  a.mu = (const double*)c->fb_mu.p;
  a.qidx = (const int*)c->fb_qidx.p;
  a.qlist = (const int*)c->fb_qlist.p;
  a.nq = c->fb_nq;
  a.c = (const double*)c->fb_c.p;
  a.z = z;
  a.logrho = logrho;
  PIB_HIP(c, hipSetDevice(c->device));
  hipError_t e = launch_flow_base(a, (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(c, e, "flow_draw_base launch");
  return PIB_OK;
}

int pib_flow_logrho(pib_ctx* c, int64_t B, const double* R, const double* cb, float* dR,
                    double* logrho, void* stream) {
  if (!c) return PIB_ERR_ARG;
  if (!c->set_A[SET_RHO] || c->P < 3) return fail(c, PIB_ERR_STATE, "flow_logrho: sampling constants or beads not set");
  if (B < 0 || !R || (!dR && !logrho)) return fail(c, PIB_ERR_ARG, "flow_logrho: B < 0 or nothing to compute");
  LogRhoArgs a{};
  a.B = B; a.N = c->set_N[SET_RHO]; a.P = c->P; a.A = c->set_A[SET_RHO];
  a.coth = (const double*)c->coth[SET_RHO].p;
  a.csch = (const double*)c->csch[SET_RHO].p;
  a.pref = (const double*)c->pref[SET_RHO].p;
  a.shift = (const double*)c->shift[SET_RHO].p;
  a.R = R; a.c = cb; a.dR = dR; a.logrho = logrho;
  PIB_HIP(c, hipSetDevice(c->device));
  hipError_t e = launch_logrho(a, (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(c, e, "flow_logrho launch (more than 16 surfaces?)");
  return PIB_OK;
}

int pib_colsum_f32(pib_ctx* c, int64_t M, int64_t N, const float* X, int64_t ldx, float* out,
                   void* stream) {
  if (!c) return PIB_ERR_ARG;
  if (M < 1 || N < 1 || ldx < N || !X || !out) return fail(c, PIB_ERR_ARG, "colsum_f32: bad size or null pointer");
  int r;
  if ((r = reserve(c, c->work, (size_t)colsum_chunks(M) * N * sizeof(double)))) return r;
  PIB_HIP(c, hipSetDevice(c->device));
  hipError_t e = launch_colsum_f32(M, N, X, ldx, out, (double*)c->work.p, (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(c, e, "colsum_f32 launch");
  return PIB_OK;
}

int pib_adam_f32(pib_ctx* c, int64_t n, float* p, const float* g, float* m, float* v, float lr,
                 float b1, float b2, float eps, int step, void* stream) {
  if (!c) return PIB_ERR_ARG;
  if (n < 1 || !p || !g || !m || !v || step < 1) return fail(c, PIB_ERR_ARG, "adam_f32: bad size, step or null pointer");
  PIB_HIP(c, hipSetDevice(c->device));
  hipError_t e = launch_adam_f32(n, p, g, m, v, lr, b1, b2, eps, step, (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(c, e, "adam_f32 launch");
  return PIB_OK;
}

}  // extern "C"
