// Learned proposal: elementwise and reduction kernels of the affine-coupling flow.
//
// The flow acts on whitened normal-mode coordinates of the bead ring. With U the circulant
// eigenvectors (the reference's normal-mode basis, pibronic/pimc/pimc.py:684-689) and
// sigma_jk the harmonic standard deviations (:401-404), Z = (R - m) U / sigma; a single
// harmonic surface gives Z ~ N(0, I). Coupling layers (RealNVP, Dinh et al. 2017) alternate
// between the even-k and odd-k normal modes; the state is kept permuted so that each half is
// one contiguous row slice: zp[b][h*D/2 + j*P/2 + kk] = Z[b][j][2 kk + h].
#include <math.h>

#include "pib_common.hpp"
#include "pib_flow.hpp"

#include <algorithm>

namespace pib {

// ------------------------------------------------------------------ whitening pre/post
// Elementwise whitening steps on a 2-D grid: blockIdx.y strides over samples, x over the
// N*P row, so the mode and bead indices need only 32-bit arithmetic.
__global__ void k_flow_pre(int mode, BasisArgs b, const void* in, float* out) {
  const int NP = b.N * b.P, Ph = b.P / 2, H = NP / 2;
  for (int64_t s = blockIdx.y; s < b.B; s += gridDim.y) {
    for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < NP; r += gridDim.x * blockDim.x) {
      const int64_t e = s * NP + r;
      const int j = r / b.P, k = r - j * b.P;
      if (mode == FLOW_WHITEN) {
        out[e] = (float)(((const double*)in)[e] - b.m[j]);
      } else if (mode == FLOW_GRAD) {
        out[e] = ((const float*)in)[e];
      } else {  // FLOW_UNWHITEN: out (row-major B*N x P) = unperm(zp) * sigma
        const int h = k & 1, kk = k >> 1;
        out[e] = ((const float*)in)[s * NP + h * H + j * Ph + kk] * b.sigma[r];
      }
    }
  }
}

__global__ void k_flow_post(int mode, BasisArgs b, const float* in, void* out) {
  const int NP = b.N * b.P, Ph = b.P / 2, H = NP / 2;
  for (int64_t s = blockIdx.y; s < b.B; s += gridDim.y) {
    for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < NP; r += gridDim.x * blockDim.x) {
      const int64_t e = s * NP + r;
      const int j = r / b.P, k = r - j * b.P;
      if (mode == FLOW_UNWHITEN) {
        ((double*)out)[e] = (double)in[e] + b.m[j];
      } else {
        const float f = (mode == FLOW_WHITEN) ? b.inv_sigma[r] : b.sigma[r];
        const int h = k & 1, kk = k >> 1;
        ((float*)out)[s * NP + h * H + j * Ph + kk] = in[e] * f;
      }
    }
  }
}

static dim3 grid_rows(int64_t rows, int64_t cols, int threads) {
  const int64_t gx = std::min<int64_t>((cols + threads - 1) / threads, 64);
  const int64_t gy = std::min<int64_t>(rows, 65535);
  return dim3((unsigned)std::max<int64_t>(gx, 1), (unsigned)std::max<int64_t>(gy, 1));
}

static dim3 grid_for(int64_t n, int threads) {
  int64_t g = (n + threads - 1) / threads;
  if (g > 65536) g = 65536;
  if (g < 1) g = 1;
  return dim3((unsigned)g);
}

hipError_t launch_flow_pre(int mode, const BasisArgs& b, const void* in, float* out, hipStream_t s) {
  hipLaunchKernelGGL(k_flow_pre, grid_rows(b.B, (int64_t)b.N * b.P, 256), dim3(256), 0, s, mode, b, in, out);
  return hipGetLastError();
}

hipError_t launch_flow_post(int mode, const BasisArgs& b, const float* in, void* out, hipStream_t s) {
  hipLaunchKernelGGL(k_flow_post, grid_rows(b.B, (int64_t)b.N * b.P, 256), dim3(256), 0, s, mode, b, in, out);
  return hipGetLastError();
}

// ------------------------------------------------------------------ fused whitening
// One kernel for each direction of the whitening transform: a work-group stages the P x P basis
// (U, or U^T for the inverse, padded rows) in LDS once; then each wave walks blocks of 16 rows
// (sample, mode) of its own: their input is gathered into the wave's LDS rows (R - m_j, or the
// un-permuted zp * sigma), the product with the basis runs on the FP32 matrix cores
// (v_mfma_f32_16x16x4f32, 16 rows x P columns), and the result goes back through LDS to be
// written with row-contiguous stores (the permuted even/odd halves, or R = X + m_j in double). Same arithmetic as pre + GEMM +
// post up to the FP32 summation order.
typedef float f32x4w __attribute__((ext_vector_type(4)));

// LDS traffic of one wave ordered: every access before is complete and visible to every lane of
// the same wave after it
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int P>
struct WhitenLds {
  static constexpr int ldu = P + 16;  // basis rows 16 floats apart mod 64 banks
  static constexpr int ldx = P + 4;   // row tiles 4 floats apart
  // the basis, the four waves' input rows and (all modes but the offsets) their output rows
  static constexpr size_t bytes(int mode) {
    return sizeof(float) * ((size_t)P * ldu + (mode == 3 ? 1 : 2) * 64 * (size_t)ldx);
  }
};

template <int P, int MODE>
__global__ void __launch_bounds__(256) k_whiten(BasisArgs b, const void* in, void* out, const float* U) {
  using L = WhitenLds<P>;
  extern __shared__ __align__(16) float wsm[];
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);  // uniform: row indices in scalars
  float* Us = wsm;                                  // [P][ldu]: B operand rows k, columns c
  float* Xs = wsm + P * L::ldu + wave * 16 * L::ldx;  // this wave's 16 input rows
  float* Ys = Xs + 64 * L::ldx;                     // this wave's 16 output rows
  const int N = b.N, NP = N * P, Ph = P / 2, H = NP / 2;
  constexpr bool inv = MODE == FLOW_UNWHITEN || MODE == FLOW_UNWHITEN_OFFSETS;
  for (int e = tid; e < P * P; e += 256) {
    const int r = e / P, c = e - r * P;
    if (inv) Us[c * L::ldu + r] = U[e];  // B[k][c] = U[c][k]
    else Us[r * L::ldu + c] = U[e];      // B[k][c] = U[k][c]
  }
  __syncthreads();
  // each wave owns 16 rows (sample s, mode j) per step: it stages, multiplies and writes only its
  // own rows, so the loop needs no work-group barrier; (s, j) advance without divisions
  const int64_t rows = b.B * (int64_t)N;
  // Everything a step reads from memory -- its 16 rows and their per-row factors (sigma or m
  // on the way in, m, 1/sigma or sigma on the way out) -- is loaded one step ahead, before the
  // current step's products and write-out, and combined only when used: vector loads complete
  // in order, so a load waited for at once would hold the wave for a memory round trip (per
  // row, if the factors were read where they are applied).
  constexpr int NO = P / 64;
  struct Rows {
    float x[16][NO];     // input values (whiten: R as float after subtracting m happens at use)
    double xd[16][NO];   // whiten: R (double)
    float fin[16][NO];   // unwhiten: sigma of the input element
    float fout[16][NO];  // whiten: 1/sigma, grad: sigma of the output element
    double m[16];        // whiten: subtracted on the way in; unwhiten: added on the way out
  };
  auto load_rows = [&](int64_t rw, Rows& r) {
    const int64_t gr0 = rw * 16, s0 = gr0 / N;
    int64_t s = s0;
    int j = (int)(gr0 - s0 * N);
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      const int64_t gr = gr0 + i;
      if (MODE == FLOW_WHITEN || MODE == FLOW_UNWHITEN) r.m[i] = b.m[j];
#pragma unroll
      for (int t = 0; t < NO; ++t) {
        const int o = lane + 64 * t;
        const int h = o / Ph, kk = o - h * Ph, k = 2 * kk + h;  // permuted position o <-> mode k
        if (MODE == FLOW_WHITEN) {
          r.xd[i][t] = ((const double*)in)[gr * P + o];
          r.fout[i][t] = b.inv_sigma[j * P + k];
        } else if (MODE == FLOW_GRAD) {
          r.x[i][t] = ((const float*)in)[gr * P + o];
          r.fout[i][t] = b.sigma[j * P + k];
        } else {
          r.x[i][t] = ((const float*)in)[s * NP + h * H + j * Ph + kk];
          r.fin[i][t] = b.sigma[j * P + k];
        }
      }
      if (++j == N) {
        j = 0;
        ++s;
      }
    }
  };
  const int64_t rstep = (int64_t)gridDim.x * 4;
  Rows nxt;
  if (((int64_t)blockIdx.x * 4 + wave) * 16 < rows) load_rows((int64_t)blockIdx.x * 4 + wave, nxt);
  for (int64_t rw = (int64_t)blockIdx.x * 4 + wave; rw * 16 < rows; rw += rstep) {
    const int64_t gr0 = rw * 16, s0 = gr0 / N;
    const int j0 = (int)(gr0 - s0 * N);
    const Rows cur = nxt;
#pragma unroll
    for (int i = 0; i < 16; ++i)
#pragma unroll
      for (int t = 0; t < NO; ++t) {
        const int o = lane + 64 * t;
        if (inv) {
          const int h = o / Ph, kk = o - h * Ph;
          Xs[i * L::ldx + 2 * kk + h] = cur.x[i][t] * cur.fin[i][t];
        } else if (MODE == FLOW_WHITEN) {
          Xs[i * L::ldx + o] = (float)(cur.xd[i][t] - cur.m[i]);
        } else {
          Xs[i * L::ldx + o] = cur.x[i][t];
        }
      }
    wave_lds_sync();
    if ((rw + rstep) * 16 < rows) load_rows(rw + rstep, nxt);
    // one column tile at a time, each a single in-place accumulator chain written to Ys at
    // once (keeping all P/16 tiles live let the compiler rename accumulators between unrolled
    // steps with a destination that partly overlaps the C operand, which gave wrong sums). The
    // offsets go from the accumulators straight to memory (register r of lane l holds row
    // 4(l/16) + r, column 16c + l%16: 64-byte row pieces), with no output rows in LDS: 38
    // instead of 55 KB of LDS per work-group, four work-groups per CU instead of two.
#pragma unroll 1
    for (int c = 0; c < P / 16; ++c) {
      f32x4w acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int k0 = 0; k0 < P; k0 += 4) {
        const float a = Xs[(lane & 15) * L::ldx + k0 + (lane >> 4)];
        const float bb = Us[(k0 + (lane >> 4)) * L::ldu + 16 * c + (lane & 15)];
        acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a, bb, acc, 0, 0, 0);
      }
      if constexpr (MODE == FLOW_UNWHITEN_OFFSETS) {
#pragma unroll
        for (int r = 0; r < 4; ++r)
          ((float*)out)[(gr0 + 4 * (lane >> 4) + r) * P + 16 * c + (lane & 15)] = acc[r];
      } else {
#pragma unroll
        for (int r = 0; r < 4; ++r) Ys[(4 * (lane >> 4) + r) * L::ldx + 16 * c + (lane & 15)] = acc[r];
      }
    }
    wave_lds_sync();
    int64_t s = s0;
    int j = j0;
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      if constexpr (MODE == FLOW_UNWHITEN_OFFSETS) break;  // stored from the accumulators
      const int64_t gr = gr0 + i;
#pragma unroll
      for (int t = 0; t < NO; ++t) {
        const int o = lane + 64 * t;
        if (inv) {
          ((double*)out)[gr * P + o] = (double)Ys[i * L::ldx + o] + cur.m[i];
        } else {
          const int h = o / Ph, kk = o - h * Ph, k = 2 * kk + h;
          ((float*)out)[s * NP + h * H + j * Ph + kk] = Ys[i * L::ldx + k] * cur.fout[i][t];
        }
      }
      if (++j == N) {
        j = 0;
        ++s;
      }
    }
    wave_lds_sync();  // the next step overwrites this wave's rows
  }
}

bool whiten_fused_supported(int P) { return P == 64 || P == 128; }

template <int P, int MODE>
static hipError_t launch_whiten_pm(const BasisArgs& b, const void* in, void* out, const float* U, int num_cu,
                                   hipStream_t s) {
  const size_t lds = WhitenLds<P>::bytes(MODE);
  static unsigned long long lds_set = 0;  // more than the default 64 KB of dynamic LDS when P = 128
  if (hipError_t e = ensure_dynamic_lds((const void*)k_whiten<P, MODE>, lds, lds_set)) return e;
  const int64_t blocks = (b.B * (int64_t)b.N + 63) / 64;  // 4 waves x 16 rows
  const int64_t grid = std::max<int64_t>(1, std::min<int64_t>(blocks, (int64_t)4 * num_cu));
  hipLaunchKernelGGL((k_whiten<P, MODE>), dim3((unsigned)grid), dim3(256), lds, s, b, in, out, U);
  return hipGetLastError();
}

template <int P>
static hipError_t launch_whiten_p(int mode, const BasisArgs& b, const void* in, void* out, const float* U,
                                  int num_cu, hipStream_t s) {
  if (mode == FLOW_WHITEN) return launch_whiten_pm<P, FLOW_WHITEN>(b, in, out, U, num_cu, s);
  if (mode == FLOW_UNWHITEN) return launch_whiten_pm<P, FLOW_UNWHITEN>(b, in, out, U, num_cu, s);
  if (mode == FLOW_UNWHITEN_OFFSETS) return launch_whiten_pm<P, FLOW_UNWHITEN_OFFSETS>(b, in, out, U, num_cu, s);
  return launch_whiten_pm<P, FLOW_GRAD>(b, in, out, U, num_cu, s);
}

hipError_t launch_whiten_fused(int mode, const BasisArgs& b, const void* in, void* out, const float* U,
                               int num_cu, hipStream_t s) {
  if ((b.B * (int64_t)b.N) % 16) return hipErrorInvalidValue;
  if (b.P == 64) return launch_whiten_p<64>(mode, b, in, out, U, num_cu, s);
  if (b.P == 128) return launch_whiten_p<128>(mode, b, in, out, U, num_cu, s);
  return hipErrorInvalidValue;
}

// ------------------------------------------------------------------ flow base draws
// Sixteen lanes per sample, four samples per wave: the surface a from the cumulative weights
// (the sampler's counter (s, 0, 1)), features 4m..4m+3 from the Philox counter (s, m, 2) by
// Box-Muller (refinement of a small radius uniform from (s, m, 3)), z = eps + mu_a rounded to
// float (the value the flow receives; its density is evaluated at that value), and
// log rho = logsumexp_b (c_b + z . mu_b) - |z|^2 / 2 with fixed-order sums over the sample's
// sixteen lanes.
constexpr int kBaseMaxA = 16;
constexpr int kBaseLanes = 16;

__device__ __forceinline__ double group16_sum(double v) {
#pragma unroll
  for (int o = 1; o < kBaseLanes; o <<= 1) v += shfl_xor_d(v, o);
  return v;
}

// The four normals of feature quad m of global sample gs: two Box-Muller pairs from one Philox
// call (32-bit uniforms, the radius one refined to 64 bits from a second call when it is below
// 2^-20, so that the tails reach ~9.4 sigma; a 32-bit radius uniform alone stops at 6.7 sigma),
// the transform in FP32 with the hardware log2, sqrt and sin/cos (z is float)
__device__ __forceinline__ void base_normals(int64_t gs, int m, uint32_t k0, uint32_t k1, float (&zo)[4]) {
  const u32x4 r = philox4x32_10(u32x4{(uint32_t)gs, (uint32_t)(gs >> 32), (uint32_t)m, 2u}, k0, k1);
  float l2a = __log2f(((float)r.x + 0.5f) * 0x1.0p-32f);
  float l2b = __log2f(((float)r.z + 0.5f) * 0x1.0p-32f);
  if (r.x < (1u << 12) || r.z < (1u << 12)) {  // rare: P = 2^-19 per call
    const u32x4 t = philox4x32_10(u32x4{(uint32_t)gs, (uint32_t)(gs >> 32), (uint32_t)m, 3u}, k0, k1);
    if (r.x < (1u << 12)) l2a = (float)log2(((double)r.x * 0x1.0p32 + (double)t.x + 0.5) * 0x1.0p-64);
    if (r.z < (1u << 12)) l2b = (float)log2(((double)r.z * 0x1.0p32 + (double)t.z + 0.5) * 0x1.0p-64);
  }
  const float ra = __fsqrt_rn(-1.3862943611198906f * l2a), rb = __fsqrt_rn(-1.3862943611198906f * l2b);
  const float ua = ((float)r.y + 0.5f) * 0x1.0p-32f, ub = ((float)r.w + 0.5f) * 0x1.0p-32f;
  // v_cos_f32 / v_sin_f32 take their argument in revolutions
  zo[0] = ra * __builtin_amdgcn_cosf(ua);
  zo[1] = ra * __builtin_amdgcn_sinf(ua);
  zo[2] = rb * __builtin_amdgcn_cosf(ub);
  zo[3] = rb * __builtin_amdgcn_sinf(ub);
}

// The displacements mu are non-zero only in a few feature quads (the constant normal mode of
// each bead mode, N of the D / 4 quads). qidx maps a quad to its row of the compact mu
// (nq * A_rho * 4) or -1, qlist a row to its quad. The lanes first draw every other quad
// (z = eps; no displacement, no contribution to z . mu_b), then the displaced quads, one per
// lane and pass (z = eps + mu_src, and z . mu_b): with the displaced quads inside the main loop,
// the lane that met one (lane 0 of each sample for P = 64, every step) held its whole wave in a
// branch of FP64 dot products. mu, qidx and qlist are staged in LDS once per work-group when
// they fit in 64 KB (MU_LDS), the work-groups then looping over groups of sixteen samples.
// MAXA >= A_rho bounds the per-surface sums.
template <bool MU_LDS, int MAXA>
__global__ void __launch_bounds__(256) k_flow_base(BaseArgs a) {
  extern __shared__ double lds[];
  const double* __restrict__ mu = a.mu;
  const int* __restrict__ qidx = a.qidx;
  const int* __restrict__ qlist = a.qlist;
  if constexpr (MU_LDS) {
    const int nmu = a.nq * a.A_rho * 4;
    for (int i = threadIdx.x; i < nmu; i += 256) lds[i] = a.mu[i];
    int* qi_lds = (int*)(lds + nmu);
    for (int i = threadIdx.x; i < a.D / 4; i += 256) qi_lds[i] = a.qidx[i];
    int* ql_lds = qi_lds + a.D / 4;
    for (int i = threadIdx.x; i < a.nq; i += 256) ql_lds[i] = a.qlist[i];
    __syncthreads();
    mu = lds;
    qidx = qi_lds;
    qlist = ql_lds;
  }
  const int AR = a.A_rho;
  const int sl = threadIdx.x & (kBaseLanes - 1);
  const int64_t groups = (a.B + 256 / kBaseLanes - 1) / (256 / kBaseLanes);
  for (int64_t grp = blockIdx.x; grp < groups; grp += gridDim.x) {  // uniform over the block
    const int64_t s = grp * (256 / kBaseLanes) + threadIdx.x / kBaseLanes;
    const bool live = s < a.B;  // uniform over the sixteen lanes of a sample
    const int64_t gs = a.offset + (live ? s : 0);
    int src = 0;
    if (AR > 1) {
      const u32x4 r = philox4x32_10(u32x4{(uint32_t)gs, (uint32_t)(gs >> 32), 0u, 1u}, a.key0, a.key1);
      const double u = u53(r.x, r.y);
      while (src < AR - 1 && u >= a.cdf[src]) ++src;
    }
    float* zr = a.z + (live ? s : 0) * a.D;
    double zz = 0.0, dot[MAXA];
#pragma unroll
    for (int b = 0; b < MAXA; ++b) dot[b] = 0.0;
    if (live) {
      for (int m = sl; 4 * m < a.D; m += kBaseLanes) {
        float zo[4];
        base_normals(gs, m, a.key0, a.key1, zo);
        if (qidx[m] < 0) {
          *(float4*)(zr + 4 * m) = float4{zo[0], zo[1], zo[2], zo[3]};
#pragma unroll
          for (int e = 0; e < 4; ++e) zz = fma((double)zo[e], (double)zo[e], zz);
        }
      }
      for (int t = sl; t < a.nq; t += kBaseLanes) {
        const int m = qlist[t];
        float zo[4];
        base_normals(gs, m, a.key0, a.key1, zo);
        const double* __restrict__ mq = mu + (int64_t)t * AR * 4;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          zo[e] = (float)((double)zo[e] + mq[src * 4 + e]);
          const double d = zo[e];
          zz = fma(d, d, zz);
#pragma unroll
          for (int b = 0; b < MAXA; ++b)
            if (b < AR) dot[b] = fma(d, mq[b * 4 + e], dot[b]);
        }
        *(float4*)(zr + 4 * m) = float4{zo[0], zo[1], zo[2], zo[3]};
      }
    }
    zz = group16_sum(zz);
    double mx = -INFINITY, t[MAXA];
#pragma unroll
    for (int b = 0; b < MAXA; ++b) {
      if (b < AR) {
        t[b] = a.c[b] + group16_sum(dot[b]);
        mx = fmax(mx, t[b]);
      }
    }
    if (live && sl == 0) {
      double acc = 0.0;
#pragma unroll
      for (int b = 0; b < MAXA; ++b)
        if (b < AR) acc += exp(t[b] - mx);
      a.logrho[s] = mx + log(acc) - 0.5 * zz;
    }
  }
}

template <int MAXA>
static hipError_t launch_flow_base_a(const BaseArgs& a, hipStream_t s) {
  const int64_t groups = (a.B + 256 / kBaseLanes - 1) / (256 / kBaseLanes);
  const size_t mu_bytes = (size_t)a.nq * a.A_rho * 4 * sizeof(double) + (size_t)(a.D / 4 + a.nq) * sizeof(int);
  if (mu_bytes <= 65536) {
    // a few work-groups per CU, each staging mu once
    const int64_t grid = std::min<int64_t>(groups, 2048);
    hipLaunchKernelGGL((k_flow_base<true, MAXA>), dim3((unsigned)grid), dim3(256), mu_bytes, s, a);
  } else {
    if (groups > 0x7fffffff) return hipErrorInvalidValue;
    hipLaunchKernelGGL((k_flow_base<false, MAXA>), dim3((unsigned)groups), dim3(256), 0, s, a);
  }
  return hipGetLastError();
}

hipError_t launch_flow_base(const BaseArgs& a, hipStream_t s) {
  if (a.B <= 0 || a.D <= 0 || (a.D & 3) || a.A_rho < 1 || a.A_rho > kBaseMaxA || a.nq < 0) return hipErrorInvalidValue;
  if (a.A_rho <= 4) return launch_flow_base_a<4>(a, s);
  if (a.A_rho <= 8) return launch_flow_base_a<8>(a, s);
  return launch_flow_base_a<16>(a, s);
}

// ------------------------------------------------------------------ block reduction helper

// ------------------------------------------------------------------ coupling layers
// forward:  y_B = z_B * exp(s) + t,   logdet += sum s
// inverse:  z_B = (y_B - t) * exp(-s), logdet -= sum s
// with s = smax * tanh(raw / smax) (bounded scales). One wave per sample row, four rows per
// work-group; the logdet sum is a fixed-order xor butterfly over the wave.
constexpr int kRowsPerBlock = 4;

__device__ __forceinline__ double wave_sum_d(double v) {
#pragma unroll
  for (int m = 1; m < 64; m <<= 1) v += __shfl_xor(v, m, 64);
  return v;
}

__global__ void __launch_bounds__(64 * kRowsPerBlock) k_coupling(int inverse, CouplingArgs a) {
  const int lane = threadIdx.x & 63;
  const int64_t b = (int64_t)blockIdx.x * kRowsPerBlock + threadIdx.x / 64;
  if (b >= a.B) return;  // whole wave
  const int H = a.D / 2;
  const float* o = a.o + b * a.D;
  float* z = a.z + b * a.D + a.b_off;
  double part = 0.0;
  for (int i = lane; i < H; i += 64) {
    const float s = a.smax * tanhf(o[i] / a.smax);
    const float t = o[H + i];
    z[i] = inverse ? (z[i] - t) * expf(-s) : fmaf(z[i], expf(s), t);
    part += (double)s;
  }
  const double tot = wave_sum_d(part);
  if (lane == 0) a.logdet[b] += inverse ? -tot : tot;
}

hipError_t launch_coupling(int inverse, const CouplingArgs& a, hipStream_t s) {
  if (a.B < 1) return hipSuccess;
  hipLaunchKernelGGL(k_coupling, dim3((unsigned)((a.B + kRowsPerBlock - 1) / kRowsPerBlock)),
                     dim3(64 * kRowsPerBlock), 0, s, inverse, a);
  return hipGetLastError();
}

// Backward of one inverse coupling layer (input y, output x; x_A = y_A,
// x_B = (y_B - t) e^{-s}, logdet_inv -= sum s) for the loss
// L = -sum_b c_b (log rho(R0_b) + logdet_inv_b). Given dz = dL/dx:
//   dL/dy_B = dz_B e^{-s};  dL/dt = -dz_B e^{-s};  dL/ds = -dz_B x_B + c_b;
//   dL/draw = dL/ds (1 - (s/smax)^2).
// dz's B-half is overwritten with dL/dy_B; dz's A-half (dL/dx_A, which passes through) is left
// for the MLP backward to accumulate into.
__global__ void __launch_bounds__(256) k_coupling_inv_bwd(CouplingArgs a) {
  const int64_t b = blockIdx.x;
  const int H = a.D / 2;
  const float* o = a.o + b * a.D;
  const float* yb = a.y + b * a.D + a.b_off;
  float* dzb = a.dz + b * a.D + a.b_off;
  float* dout = a.dout + b * a.D;
  const float cb = (float)a.c[b];
  for (int i = threadIdx.x; i < H; i += blockDim.x) {
    const float th = tanhf(o[i] / a.smax);
    const float s = a.smax * th;
    const float t = o[H + i];
    const float e = expf(-s);
    const float xb = (yb[i] - t) * e;
    const float g = dzb[i];
    const float ds = fmaf(-g, xb, cb);
    dout[i] = ds * (1.0f - th * th);
    dout[H + i] = -g * e;
    dzb[i] = g * e;
  }
}

hipError_t launch_coupling_inverse_backward(const CouplingArgs& a, hipStream_t s) {
  if (a.B < 1) return hipSuccess;
  hipLaunchKernelGGL(k_coupling_inv_bwd, dim3((unsigned)a.B), dim3(256), 0, s, a);
  return hipGetLastError();
}

// ------------------------------------------------------------------ log rho and its gradient
// rho(R) = sum_a prod_p O^rho_a(p) of the sampling model (pibronic/pimc/pimc.py:1087-1136,
// unscaled), log_a = P log F_a - 1/2 sum_jp [coth_j (x_p^2 + x_{p+1}^2) - 2 csch_j x_p x_{p+1}],
// x = R - d_a. d log_a / dx_p = -(2 coth_j x_p - csch_j (x_{p-1} + x_{p+1})). The gradient of
// log rho is the responsibility-weighted sum over surfaces. One work-group per sample.
constexpr int kMaxRhoSurfaces = 16;

__global__ void __launch_bounds__(64 * kRowsPerBlock) k_logrho(LogRhoArgs a) {
  const int lane = threadIdx.x & 63;
  const int64_t b = (int64_t)blockIdx.x * kRowsPerBlock + threadIdx.x / 64;
  if (b >= a.B) return;  // whole wave
  const int N = a.N, P = a.P, NP = N * P;
  const double* R = a.R + b * (int64_t)NP;
  double part[kMaxRhoSurfaces];
#pragma unroll
  for (int s = 0; s < kMaxRhoSurfaces; ++s) part[s] = 0.0;
  for (int j = 0; j < N; ++j) {  // mode-uniform constants: scalar loads, no index division
    const double cj = a.coth[j], sj2 = 2.0 * a.csch[j];
    for (int p = lane; p < P; p += 64) {
      const int pn = (p + 1 == P) ? 0 : p + 1;
      const double r0 = R[j * P + p], r1 = R[j * P + pn];
#pragma unroll
      for (int s = 0; s < kMaxRhoSurfaces; ++s) {
        if (s < a.A) {
          const double d = a.shift[s * N + j];
          const double x1 = r0 - d, x2 = r1 - d;
          part[s] += cj * (x1 * x1 + x2 * x2) - (sj2 * x1) * x2;
        }
      }
    }
  }
  // every lane ends with the same sums (xor butterfly), so every lane holds la, lmax, tot
  double lmax = -INFINITY;
  double la[kMaxRhoSurfaces];
#pragma unroll
  for (int s = 0; s < kMaxRhoSurfaces; ++s) {
    la[s] = -INFINITY;
    if (s < a.A) {
      la[s] = P * log(a.pref[s]) - 0.5 * wave_sum_d(part[s]);
      lmax = fmax(lmax, la[s]);
    }
  }
  double tot = 0.0;
#pragma unroll
  for (int s = 0; s < kMaxRhoSurfaces; ++s)
    if (s < a.A) tot += exp(la[s] - lmax);
  if (lane == 0 && a.logrho) a.logrho[b] = lmax + log(tot);
  if (!a.dR) return;
  double resp[kMaxRhoSurfaces];
#pragma unroll
  for (int s = 0; s < kMaxRhoSurfaces; ++s) resp[s] = (s < a.A) ? exp(la[s] - lmax) / tot : 0.0;
  const double cb = a.c ? a.c[b] : 1.0;
  for (int j = 0; j < N; ++j) {
    const double cj2 = 2.0 * a.coth[j], sj = a.csch[j];
    for (int p = lane; p < P; p += 64) {
      const int pn = (p + 1 == P) ? 0 : p + 1, pp = (p == 0) ? P - 1 : p - 1;
      const double r0 = R[j * P + p], rn = R[j * P + pn], rp = R[j * P + pp];
      double g = 0.0;
#pragma unroll
      for (int s = 0; s < kMaxRhoSurfaces; ++s) {
        if (s < a.A) {
          const double d = a.shift[s * N + j];
          g -= resp[s] * (cj2 * (r0 - d) - sj * ((rn - d) + (rp - d)));
        }
      }
      a.dR[b * (int64_t)NP + j * P + p] = (float)(-cb * g);
    }
  }
}

hipError_t launch_logrho(const LogRhoArgs& a, hipStream_t s) {
  if (a.A > kMaxRhoSurfaces) return hipErrorInvalidValue;
  if (a.B < 1) return hipSuccess;
  hipLaunchKernelGGL(k_logrho, dim3((unsigned)((a.B + kRowsPerBlock - 1) / kRowsPerBlock)),
                     dim3(64 * kRowsPerBlock), 0, s, a);
  return hipGetLastError();
}

// ------------------------------------------------------------------ column sums, Adam
// Pass 1: work-group (column tile of 64, row chunk of kColsumRows) -> partial[chunk][col], the
// 256 threads as 64 columns x 4 row phases combined through LDS. Pass 2: per column, the
// chunks in order.
constexpr int kColsumRows = 64;

int64_t colsum_chunks(int64_t M) { return (M + kColsumRows - 1) / kColsumRows; }

__global__ void __launch_bounds__(256) k_colsum_partial(int64_t M, int64_t N, const float* X,
                                                        int64_t ldx, double* partial) {
  __shared__ double sh[4][64];
  const int c = threadIdx.x & 63, ph = threadIdx.x >> 6;
  const int64_t col = (int64_t)blockIdx.x * 64 + c;
  const int64_t r0 = (int64_t)blockIdx.y * kColsumRows;
  double acc = 0.0;
  if (col < N)
    for (int64_t i = r0 + ph; i < min(r0 + kColsumRows, M); i += 4) acc += X[i * ldx + col];
  sh[ph][c] = acc;
  __syncthreads();
  if (ph == 0 && col < N)
    partial[(int64_t)blockIdx.y * N + col] = ((sh[0][c] + sh[1][c]) + sh[2][c]) + sh[3][c];
}

__global__ void k_colsum_final(int64_t chunks, int64_t N, const double* partial, float* out) {
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < N;
       j += (int64_t)gridDim.x * blockDim.x) {
    double acc = 0.0;
    for (int64_t k = 0; k < chunks; ++k) acc += partial[k * N + j];
    out[j] = (float)acc;
  }
}

hipError_t launch_colsum_f32(int64_t M, int64_t N, const float* X, int64_t ldx, float* out,
                             double* partial, hipStream_t s) {
  const int64_t chunks = colsum_chunks(M);
  hipLaunchKernelGGL(k_colsum_partial, dim3((unsigned)((N + 63) / 64), (unsigned)chunks),
                     dim3(256), 0, s, M, N, X, ldx, partial);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_colsum_final, grid_for(N, 256), dim3(256), 0, s, chunks, N, partial, out);
  return hipGetLastError();
}

// Adam (Kingma & Ba 2015) with bias correction; step counts from 1.
__global__ void k_adam(int64_t n, float* p, const float* g, float* m, float* v, float lr, float b1,
                       float b2, float eps, float c1, float c2) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const float gi = g[i];
    const float mi = fmaf(b1, m[i], (1.0f - b1) * gi);
    const float vi = fmaf(b2, v[i], (1.0f - b2) * gi * gi);
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  }
}

hipError_t launch_adam_f32(int64_t n, float* p, const float* g, float* m, float* v, float lr,
                           float b1, float b2, float eps, int step, hipStream_t s) {
  const float c1 = 1.0f - powf(b1, (float)step), c2 = 1.0f - powf(b2, (float)step);
  hipLaunchKernelGGL(k_adam, grid_for(n, 256), dim3(256), 0, s, n, p, g, m, v, lr, b1, b2, eps,
                     c1, c2);
  return hipGetLastError();
}

}  // namespace pib
