// One affine-coupling layer in one kernel (flow sampling and density directions), with the
// activations resident in registers.
//
// The three products of the layer's MLP (H -> h -> h -> 2H, ReLU, ReLU) are computed
// transposed, Y^T = W^T X^T, on the bf16 matrix cores as BF16x3 (hi/lo splits, FP32
// accumulation), so that samples run along the MFMA N dimension and features along M. With
// v_mfma_f32_16x16x32_bf16 (lane l: A[l%16][8(l/16)+e], B[8(l/16)+e][l%16],
// D[4(l/16)+r][l%16]) the accumulator of an output tile gives lane l the features
// 4(l/16)+r of sample l%16 -- the same sample the lane supplies as B operand. Two such tiles
// (features 32c..32c+15 and 32c+16..32c+31) therefore form the B operand of input block c of
// the next product directly in registers, provided that product's weights list their inputs
// in the matching order (pib_split_bf16 with PIB_SPLIT_PERM32 stores input 32c + w at
// 32c + 8((w%16)/4) + 4(w/16) + w%4). The hidden activations never leave the registers.
//
// Work-group: 8 waves, 16 samples each (128 samples). Weight tiles (16 output rows x K, hi
// and lo) are staged through a double-buffered LDS ring shared by the 8 waves; the next tile
// is fetched into registers while the current one is multiplied. The last product is taken
// tile pair by tile pair (raw scales s and shifts t of the same 16 features), and its
// epilogue applies the coupling -- s = smax tanh(raw / smax), y = z e^s + t (or
// z = (y - t) e^-s) -- to the lane's four features of its sample and accumulates sum s.
#include <hip/hip_runtime.h>

#include "pib_common.hpp"
#include "pib_flow.hpp"

namespace pib {

namespace {

typedef float f32x4f __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8f __attribute__((ext_vector_type(8)));
typedef unsigned uvec4f __attribute__((ext_vector_type(4)));

constexpr int kWaves = 8;               // waves per work-group
constexpr int kSamplesPerWave = 16;     // one MFMA N tile
constexpr int kRows = kWaves * kSamplesPerWave;
constexpr int kThreads = kWaves * 64;
constexpr int kPadF = 8;                // LDS row padding (bf16 elements): rows 16 bytes apart mod 256

__device__ __forceinline__ unsigned short f2bf_f(float x) {
  unsigned u = __float_as_uint(x);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf2f_f(unsigned short h) { return __uint_as_float((unsigned)h << 16); }

// 8 bf16 packed in 4 dwords (element e in the low half of dword e/2 when e is even)
struct Op8 {
  unsigned w[4];
};
__device__ __forceinline__ void put(Op8& o, int e, unsigned short v) {
  o.w[e >> 1] = (e & 1) ? ((o.w[e >> 1] & 0xffffu) | ((unsigned)v << 16)) : ((o.w[e >> 1] & 0xffff0000u) | v);
}
__device__ __forceinline__ bf16x8f as_bf(const Op8& o) {
  return __builtin_bit_cast(bf16x8f, uvec4f{o.w[0], o.w[1], o.w[2], o.w[3]});
}
__device__ __forceinline__ void split4(float v, int e, Op8& h, Op8& l) {
  const unsigned short hb = f2bf_f(v);
  put(h, e, hb);
  put(l, e, f2bf_f(v - bf2f_f(hb)));
}

__device__ __forceinline__ f32x4f mfma3(const bf16x8f& ah, const bf16x8f& al, const bf16x8f& bh,
                                        const bf16x8f& bl, f32x4f c) {
  c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh, c, 0, 0, 0);
  c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl, c, 0, 0, 0);
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh, c, 0, 0, 0);
}

// The tile sequence of one layer: the HID/16 row tiles of W1^T (K = H), the HID/16 row tiles
// of W2^T (K = HID), then for q = 0..H/16-1 the rows 16q.. (scales) and H + 16q.. (shifts)
// of W3^T (K = HID).
template <int H, int HID>
struct Tiles {
  static constexpr int n1 = HID / 16, n2 = HID / 16, n3 = 2 * (H / 16);
  static constexpr int total = n1 + n2 + n3;
  static constexpr int kmax = H > HID ? H : HID;
  static constexpr int ld = kmax + kPadF;      // LDS row stride (elements)
  static constexpr int tile = 16 * ld;         // elements of one hi (or lo) tile
  static constexpr int units = 2 * 16 * kmax / 8 / kThreads;  // 16-byte loads per thread per tile
  __device__ static constexpr int K(int t) { return t < n1 ? H : HID; }
};

template <int H, int HID>
__device__ __forceinline__ void tile_source(const FusedLayerArgs& a, int t, const unsigned short*& wh,
                                            const unsigned short*& wl, int& row0) {
  using T = Tiles<H, HID>;
  if (t < T::n1) {
    wh = a.W1h;
    wl = a.W1l;
    row0 = 16 * t;
  } else if (t < T::n1 + T::n2) {
    wh = a.W2h;
    wl = a.W2l;
    row0 = 16 * (t - T::n1);
  } else {
    const int u = t - T::n1 - T::n2;
    wh = a.W3h;
    wl = a.W3l;
    row0 = (u & 1) ? H + 16 * (u >> 1) : 16 * (u >> 1);
  }
}

// Fetch tile t into registers: unit i of this thread is 16-byte unit tid + kThreads * i of the
// tile (hi units first, then lo); units beyond the tile's K are not loaded.
template <int H, int HID>
__device__ __forceinline__ void fetch(const FusedLayerArgs& a, int t, int tid, uvec4f (&pre)[Tiles<H, HID>::units]) {
  using T = Tiles<H, HID>;
  const unsigned short *wh, *wl;
  int row0;
  tile_source<H, HID>(a, t, wh, wl, row0);
  const int K = T::K(t), per = 16 * K / 8;  // units per array
#pragma unroll
  for (int i = 0; i < T::units; ++i) {
    const int u = tid + kThreads * i;
    if (u < 2 * per) {
      const int arr = u >= per, r = u - arr * per;
      const int row = r / (K / 8), c8 = r - row * (K / 8);
      const unsigned short* src = (arr ? wl : wh) + (int64_t)(row0 + row) * K + 8 * c8;
      pre[i] = *(const uvec4f*)src;
    }
  }
}

template <int H, int HID>
__device__ __forceinline__ void store(int t, int tid, const uvec4f (&pre)[Tiles<H, HID>::units], unsigned short* buf) {
  using T = Tiles<H, HID>;
  const int K = T::K(t), per = 16 * K / 8;
#pragma unroll
  for (int i = 0; i < T::units; ++i) {
    const int u = tid + kThreads * i;
    if (u < 2 * per) {
      const int arr = u >= per, r = u - arr * per;
      const int row = r / (K / 8), c8 = r - row * (K / 8);
      *(uvec4f*)(buf + arr * T::tile + row * T::ld + 8 * c8) = pre[i];
    }
  }
}

// acc += W^T tile (LDS, 16 rows x K) * X^T (operands in registers, K/32 blocks)
template <int NB, int LD>
__device__ __forceinline__ f32x4f tile_product(const unsigned short* buf, const Op8 (&bh)[NB],
                                               const Op8 (&bl)[NB], int li, int g, int tile_elems) {
  f32x4f acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int c = 0; c < NB; ++c) {
    const unsigned short* p = buf + li * LD + 32 * c + 8 * g;
    const bf16x8f ah = __builtin_bit_cast(bf16x8f, *(const uvec4f*)p);
    const bf16x8f al = __builtin_bit_cast(bf16x8f, *(const uvec4f*)(p + tile_elems));
    acc = mfma3(ah, al, as_bf(bh[c]), as_bf(bl[c]), acc);
  }
  return acc;
}

template <int H, int HID>
__global__ void __launch_bounds__(kThreads) k_flow_layer(FusedLayerArgs a) {
  using T = Tiles<H, HID>;
  constexpr int NBX = H / 32, NBH = HID / 32;
  extern __shared__ __align__(16) unsigned short lds[];  // [2 buffers][hi, lo][16][ld]
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int li = lane & 15, g = lane >> 4;
  const int D = a.D;
  const int64_t n = (int64_t)blockIdx.x * kRows + wave * kSamplesPerWave + li;  // this lane's sample
  float* zrow = a.z + n * D;

  // conditioning half as the first product's B operand: block c, elements 32c + 8g + e
  Op8 xh[NBX], xl[NBX];
#pragma unroll
  for (int c = 0; c < NBX; ++c) {
    const float4* src = (const float4*)(zrow + a.a_off + 32 * c + 8 * g);
    const float4 v0 = src[0], v1 = src[1];
    const float v[8] = {v0.x, v0.y, v0.z, v0.w, v1.x, v1.y, v1.z, v1.w};
#pragma unroll
    for (int e = 0; e < 8; ++e) split4(v[e], e, xh[c], xl[c]);
  }

  Op8 h1h[NBH], h1l[NBH], h2h[NBH], h2l[NBH];
  double part = 0.0;

  uvec4f pre[T::units];
  fetch<H, HID>(a, 0, tid, pre);
  store<H, HID>(0, tid, pre, lds);
  // pipeline step around tile t: prefetch t+1 into registers, wait for tile t in LDS, then
  // (after the caller's product) write t+1 into the other buffer
#define PIB_TILE_BEGIN(t)                                         \
  if ((t) + 1 < T::total) fetch<H, HID>(a, (t) + 1, tid, pre);    \
  __syncthreads();                                                \
  const unsigned short* buf = lds + ((t) & 1) * 2 * T::tile;
#define PIB_TILE_END(t) \
  if ((t) + 1 < T::total) store<H, HID>((t) + 1, tid, pre, lds + (((t) + 1) & 1) * 2 * T::tile);

#pragma clang loop unroll(full)
  for (int t = 0; t < T::n1; ++t) {
    PIB_TILE_BEGIN(t)
    const f32x4f acc = tile_product<NBX, T::ld>(buf, xh, xl, li, g, T::tile);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const float v = acc[r] + a.b1[16 * t + 4 * g + r];
      split4(v > 0.f ? v : 0.f, 4 * (t & 1) + r, h1h[t >> 1], h1l[t >> 1]);
    }
    PIB_TILE_END(t)
  }
#pragma clang loop unroll(full)
  for (int u = 0; u < T::n2; ++u) {
    const int t = T::n1 + u;
    PIB_TILE_BEGIN(t)
    const f32x4f acc = tile_product<NBH, T::ld>(buf, h1h, h1l, li, g, T::tile);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const float v = acc[r] + a.b2[16 * u + 4 * g + r];
      split4(v > 0.f ? v : 0.f, 4 * (u & 1) + r, h2h[u >> 1], h2l[u >> 1]);
    }
    PIB_TILE_END(t)
  }
  for (int q = 0; q < H / 16; ++q) {  // the operands h2 are indexed by constants only here
    f32x4f sraw;
    {
      const int t = T::n1 + T::n2 + 2 * q;
      PIB_TILE_BEGIN(t)
      sraw = tile_product<NBH, T::ld>(buf, h2h, h2l, li, g, T::tile);
      PIB_TILE_END(t)
    }
    {
      const int t = T::n1 + T::n2 + 2 * q + 1;
      PIB_TILE_BEGIN(t)
      const f32x4f acc = tile_product<NBH, T::ld>(buf, h2h, h2l, li, g, T::tile);
      const int f0 = 16 * q + 4 * g;  // this lane's four features of the transformed half
      float4* zp = (float4*)(zrow + a.b_off + f0);
      const float4 zv = *zp;
      float zz[4] = {zv.x, zv.y, zv.z, zv.w};
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const float s = a.smax * tanhf((sraw[r] + a.b3[f0 + r]) / a.smax);
        const float sh = acc[r] + a.b3[H + f0 + r];
        zz[r] = a.inverse ? (zz[r] - sh) * expf(-s) : fmaf(zz[r], expf(s), sh);
        part += (double)s;
      }
      *zp = float4{zz[0], zz[1], zz[2], zz[3]};
      PIB_TILE_END(t)
    }
  }
#undef PIB_TILE_BEGIN
#undef PIB_TILE_END
  // sum s over the lane's features, then over the four lanes that hold the same sample
  part += shfl_xor_d(part, 16);
  part += shfl_xor_d(part, 32);
  if (g == 0) a.logdet[n] += a.inverse ? -part : part;
}

template <int H, int HID>
hipError_t launch_hh(const FusedLayerArgs& a, hipStream_t s) {
  using T = Tiles<H, HID>;
  const size_t lds = (size_t)2 * 2 * T::tile * sizeof(unsigned short);
  static bool attr_set = false;  // more than the default 64 KB of dynamic LDS when H = 512
  if (!attr_set) {
    hipError_t e = hipFuncSetAttribute((const void*)k_flow_layer<H, HID>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    attr_set = true;
  }
  hipLaunchKernelGGL((k_flow_layer<H, HID>), dim3((unsigned)(a.B / kRows)), dim3(kThreads), lds, s, a);
  return hipGetLastError();
}

}  // namespace

bool fused_layer_supported(int64_t B, int D, int hidden) {
  const int H = D / 2;
  return B > 0 && B % kRows == 0 && hidden == 256 && D == 2 * H && (H == 256 || H == 512);
}

hipError_t launch_flow_layer(const FusedLayerArgs& a, hipStream_t s) {
  if (!fused_layer_supported(a.B, a.D, a.hidden)) return hipErrorInvalidValue;
  if ((a.a_off != 0 && a.a_off != a.D / 2) || (a.b_off != 0 && a.b_off != a.D / 2) || a.a_off == a.b_off)
    return hipErrorInvalidValue;
  if (a.D / 2 == 256) return launch_hh<256, 256>(a, s);
  return launch_hh<512, 256>(a, s);
}

}  // namespace pib
