// One affine-coupling layer in one kernel (flow sampling and density directions). Two kernels:
// k_flow_layer (below: pib_flow_layer_fused, BF16x3 or single bf16), activations resident in
// registers and weights staged through LDS; and k_flow_layer_ws (further down:
// pib_flow_layer_fused_frag, single bf16 at D = 512), fragment-ordered weights streamed into
// registers per wave and activations staged in LDS.
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
// Work-group: 128 samples; each wave holds kNT MFMA N tiles of 16 samples. kNT = 1 (eight waves,
// two per SIMD) measured faster than kNT = 2 (four waves with the whole register file each:
// 2.79 vs 2.09 ms per draw of 65536 samples at D = 512): the second wave per SIMD hides more
// LDS and L2 latency than the doubled reuse of each weight read saves. Weight tiles (16 output rows x K, hi
// and lo) are staged two at a time through a double-buffered LDS ring shared by the 8 waves;
// the next pair is fetched into registers while the current one is multiplied. The last
// product is taken tile pair by tile pair (raw scales s and shifts t of the same 16
// features), and its
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

constexpr int kNT = 1;                 // MFMA N tiles (16 samples each) per wave
constexpr int kWaves = 8 / kNT;         // waves per work-group
constexpr int kRows = kWaves * 16 * kNT;  // samples per work-group
constexpr int kThreads = kWaves * 64;
constexpr int kPadF = 8;                // LDS row padding (bf16 elements): rows 16 bytes apart mod 256

__device__ __forceinline__ unsigned short f2bf_f(float x) {
  unsigned u = __float_as_uint(x);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf2f_f(unsigned short h) { return __uint_as_float((unsigned)h << 16); }

typedef __bf16 bf16x2f __attribute__((ext_vector_type(2)));
typedef float f32x2f __attribute__((ext_vector_type(2)));
// two floats to packed bf16 (a in the low half), round to nearest even: one v_cvt_pk_bf16_f32
__device__ __forceinline__ unsigned pack_bf16(float a, float b) {
  return __builtin_bit_cast(unsigned, __builtin_convertvector(f32x2f{a, b}, bf16x2f));
}

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

// The stage sequence of one layer; a stage is two 16-row weight tiles (hi and, with LO, lo),
// staged in LDS together behind one barrier: the HID/32 row pairs of W1^T (K = H), the HID/32
// row pairs of W2^T (K = HID), then for q = 0..H/16-1 the rows 16q.. (scales) with H + 16q..
// (shifts) of W3^T (K = HID). Without LO (single bf16 products) only the hi parts are staged.
template <int H, int HID, bool LO>
struct Stages {
  static constexpr int n1 = HID / 32, n2 = HID / 32, n3 = H / 16;
  static constexpr int total = n1 + n2 + n3;
  static constexpr int kmax = H > HID ? H : HID;
  static constexpr int parts = LO ? 2 : 1;     // hi (and lo) per tile
  static constexpr int ld = kmax + kPadF;      // LDS row stride (elements)
  static constexpr int tile = 16 * ld;         // elements of one hi (or lo) tile
  static constexpr int stage = 2 * parts * tile;  // [tile 0 hi, (tile 0 lo), tile 1 hi, (tile 1 lo)]
  static constexpr int units = parts * 2 * 16 * kmax / 8 / kThreads;  // 16-byte loads per thread
  __device__ static constexpr int K(int s) { return s < n1 ? H : HID; }
};

// source of stage s: weight arrays and the first rows of its two tiles
template <int H, int HID, bool LO>
__device__ __forceinline__ void stage_source(const FusedLayerArgs& a, int s, const unsigned short*& wh,
                                             const unsigned short*& wl, int& rowA, int& rowB) {
  using S = Stages<H, HID, LO>;
  if (s < S::n1) {
    wh = a.W1h;
    wl = a.W1l;
    rowA = 32 * s;
    rowB = rowA + 16;
  } else if (s < S::n1 + S::n2) {
    wh = a.W2h;
    wl = a.W2l;
    rowA = 32 * (s - S::n1);
    rowB = rowA + 16;
  } else {
    const int q = s - S::n1 - S::n2;
    wh = a.W3h;
    wl = a.W3l;
    rowA = 16 * q;
    rowB = H + 16 * q;
  }
}

// Unit i of this thread is 16-byte unit tid + kThreads * i of the stage, ordered
// (tile, hi/lo, row, 8-element column group); units beyond the stage's K are skipped.
template <int H, int HID, bool LO>
__device__ __forceinline__ void fetch(const FusedLayerArgs& a, int s, int tid,
                                      uvec4f (&pre)[Stages<H, HID, LO>::units]) {
  using S = Stages<H, HID, LO>;
  const unsigned short *wh, *wl;
  int rowA, rowB;
  stage_source<H, HID, LO>(a, s, wh, wl, rowA, rowB);
  const int K = S::K(s), per = 16 * K / 8;  // units per (tile, hi/lo)
#pragma unroll
  for (int i = 0; i < S::units; ++i) {
    const int u = tid + kThreads * i;
    if (u < 2 * S::parts * per) {
      const int blk = u / per, r = u - blk * per;  // blk = parts * tile + (lo ? 1 : 0)
      const int row = r / (K / 8), c8 = r - row * (K / 8);
      const bool lo = LO && (blk & 1);
      const int t = LO ? (blk >> 1) : blk;
      const unsigned short* src = (lo ? wl : wh) + (int64_t)((t ? rowB : rowA) + row) * K + 8 * c8;
      pre[i] = *(const uvec4f*)src;
    }
  }
}

template <int H, int HID, bool LO>
__device__ __forceinline__ void store(int s, int tid, const uvec4f (&pre)[Stages<H, HID, LO>::units],
                                      unsigned short* buf) {
  using S = Stages<H, HID, LO>;
  const int K = S::K(s), per = 16 * K / 8;
#pragma unroll
  for (int i = 0; i < S::units; ++i) {
    const int u = tid + kThreads * i;
    if (u < 2 * S::parts * per) {
      const int blk = u / per, r = u - blk * per;
      const int row = r / (K / 8), c8 = r - row * (K / 8);
      *(uvec4f*)(buf + blk * S::tile + row * S::ld + 8 * c8) = pre[i];
    }
  }
}

// W^T tile (LDS, 16 rows x K, lo part tile_elems after hi) times X^T for each of the wave's
// kNT sample tiles (operands in registers): one LDS read of a weight block feeds 3 kNT MFMAs
// (LO: BF16x3) or kNT (single bf16 products of the hi parts)
template <int NB, int LD, bool LO>
__device__ __forceinline__ void tile_product(const unsigned short* buf, const Op8 (&bh)[kNT][NB],
                                             const Op8 (&bl)[kNT][NB], int li, int g, int tile_elems,
                                             f32x4f (&acc)[kNT]) {
  // single bf16 products: even and odd blocks in two accumulator chains, added at the end
  // (with one chain of NB the register allocator emitted products whose destination partly
  // overlaps their accumulator input)
  f32x4f odd[kNT];
#pragma unroll
  for (int j = 0; j < kNT; ++j) {
    acc[j] = f32x4f{0.f, 0.f, 0.f, 0.f};
    odd[j] = acc[j];
  }
#pragma unroll
  for (int c = 0; c < NB; ++c) {
    const unsigned short* p = buf + li * LD + 32 * c + 8 * g;
    const bf16x8f ah = __builtin_bit_cast(bf16x8f, *(const uvec4f*)p);
    if constexpr (LO) {
      const bf16x8f al = __builtin_bit_cast(bf16x8f, *(const uvec4f*)(p + tile_elems));
#pragma unroll
      for (int j = 0; j < kNT; ++j) acc[j] = mfma3(ah, al, as_bf(bh[j][c]), as_bf(bl[j][c]), acc[j]);
    } else {
#pragma unroll
      for (int j = 0; j < kNT; ++j) {
        if (c & 1)
          odd[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, as_bf(bh[j][c]), odd[j], 0, 0, 0);
        else
          acc[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, as_bf(bh[j][c]), acc[j], 0, 0, 0);
      }
    }
  }
  if constexpr (!LO) {
#pragma unroll
    for (int j = 0; j < kNT; ++j) acc[j] += odd[j];
  }
}

// operand element e of a block: bf16 hi (and with LO the lo remainder)
template <bool LO>
__device__ __forceinline__ void split_op(float v, int e, Op8& h, Op8& l) {
  if constexpr (LO) {
    split4(v, e, h, l);
  } else {
    put(h, e, f2bf_f(v));
  }
}

template <int H, int HID, bool LO>
__global__ void __launch_bounds__(kThreads) k_flow_layer(FusedLayerArgs a) {
  using S = Stages<H, HID, LO>;
  constexpr int NBX = H / 32, NBH = HID / 32;
  extern __shared__ __align__(16) unsigned short lds[];  // [2 buffers][stage]
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int li = lane & 15, g = lane >> 4;
  const int D = a.D;
  // this lane's samples: one per sample tile j
  const int64_t n0 = (int64_t)blockIdx.x * kRows + wave * 16 * kNT + li;
  float* zrow[kNT];
#pragma unroll
  for (int j = 0; j < kNT; ++j) zrow[j] = a.z + (n0 + 16 * j) * D;

  // conditioning half as the first product's B operand: block c, elements 32c + 8g + e
  Op8 xh[kNT][NBX], xl[kNT][NBX];
#pragma unroll
  for (int j = 0; j < kNT; ++j)
#pragma unroll
    for (int c = 0; c < NBX; ++c) {
      const float4* src = (const float4*)(zrow[j] + a.a_off + 32 * c + 8 * g);
      const float4 v0 = src[0], v1 = src[1];
      const float v[8] = {v0.x, v0.y, v0.z, v0.w, v1.x, v1.y, v1.z, v1.w};
#pragma unroll
      for (int e = 0; e < 8; ++e) split_op<LO>(v[e], e, xh[j][c], xl[j][c]);
    }

  Op8 h1h[kNT][NBH], h1l[kNT][NBH], h2h[kNT][NBH], h2l[kNT][NBH];
  const float ismax = 1.0f / a.smax;
  double part[kNT];
#pragma unroll
  for (int j = 0; j < kNT; ++j) part[j] = 0.0;

  uvec4f pre[S::units];
  fetch<H, HID, LO>(a, 0, tid, pre);
  store<H, HID, LO>(0, tid, pre, lds);
  // pipeline step around stage s: prefetch s+1 into registers, wait for stage s in LDS, then
  // (after the caller's products) write s+1 into the other buffer
#define PIB_STAGE_BEGIN(s)                                        \
  if ((s) + 1 < S::total) fetch<H, HID, LO>(a, (s) + 1, tid, pre); \
  __syncthreads();                                                \
  const unsigned short* buf = lds + ((s) & 1) * S::stage;
#define PIB_STAGE_END(s) \
  if ((s) + 1 < S::total) store<H, HID, LO>((s) + 1, tid, pre, lds + (((s) + 1) & 1) * S::stage);

  // first product: stage c gives output tiles 2c, 2c+1 = operand block c of the second
#pragma clang loop unroll(full)
  for (int c = 0; c < S::n1; ++c) {
    PIB_STAGE_BEGIN(c)
    f32x4f acc0[kNT], acc1[kNT];
    tile_product<NBX, S::ld, LO>(buf, xh, xl, li, g, S::tile, acc0);
    tile_product<NBX, S::ld, LO>(buf + S::parts * S::tile, xh, xl, li, g, S::tile, acc1);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const float b0 = a.b1[32 * c + 4 * g + r], b1 = a.b1[32 * c + 16 + 4 * g + r];
#pragma unroll
      for (int j = 0; j < kNT; ++j) {
        const float v0 = acc0[j][r] + b0, v1 = acc1[j][r] + b1;
        split_op<LO>(v0 > 0.f ? v0 : 0.f, r, h1h[j][c], h1l[j][c]);
        split_op<LO>(v1 > 0.f ? v1 : 0.f, 4 + r, h1h[j][c], h1l[j][c]);
      }
    }
    PIB_STAGE_END(c)
  }
#pragma clang loop unroll(full)
  for (int c = 0; c < S::n2; ++c) {
    const int s = S::n1 + c;
    PIB_STAGE_BEGIN(s)
    f32x4f acc0[kNT], acc1[kNT];
    tile_product<NBH, S::ld, LO>(buf, h1h, h1l, li, g, S::tile, acc0);
    tile_product<NBH, S::ld, LO>(buf + S::parts * S::tile, h1h, h1l, li, g, S::tile, acc1);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const float b0 = a.b2[32 * c + 4 * g + r], b1 = a.b2[32 * c + 16 + 4 * g + r];
#pragma unroll
      for (int j = 0; j < kNT; ++j) {
        const float v0 = acc0[j][r] + b0, v1 = acc1[j][r] + b1;
        split_op<LO>(v0 > 0.f ? v0 : 0.f, r, h2h[j][c], h2l[j][c]);
        split_op<LO>(v1 > 0.f ? v1 : 0.f, 4 + r, h2h[j][c], h2l[j][c]);
      }
    }
    PIB_STAGE_END(s)
  }
  // third product and coupling; the transformed half's z (four features per lane per q and
  // sample tile) is loaded two q-steps ahead
  float4 zq0[kNT], zq1[kNT];
#pragma unroll
  for (int j = 0; j < kNT; ++j) {
    zq0[j] = *(const float4*)(zrow[j] + a.b_off + 4 * g);
    zq1[j] = *(const float4*)(zrow[j] + a.b_off + 16 + 4 * g);
  }
  for (int q = 0; q < S::n3; ++q) {  // the operands h2 are indexed by constants only here
    const int s = S::n1 + S::n2 + q;
    float4 zv[kNT];
#pragma unroll
    for (int j = 0; j < kNT; ++j) {
      zv[j] = zq0[j];
      zq0[j] = zq1[j];
      if (q + 2 < S::n3) zq1[j] = *(const float4*)(zrow[j] + a.b_off + 16 * (q + 2) + 4 * g);
    }
    PIB_STAGE_BEGIN(s)
    f32x4f sraw[kNT], acc[kNT];
    tile_product<NBH, S::ld, LO>(buf, h2h, h2l, li, g, S::tile, sraw);
    tile_product<NBH, S::ld, LO>(buf + S::parts * S::tile, h2h, h2l, li, g, S::tile, acc);
    const int f0 = 16 * q + 4 * g;  // this lane's four features of the transformed half
#pragma unroll
    for (int j = 0; j < kNT; ++j) {
      float zz[4] = {zv[j].x, zv[j].y, zv[j].z, zv[j].w};
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const float sh = acc[j][r] + a.b3[H + f0 + r];
        float sc, es;
        if constexpr (LO) {  // the layer-by-layer path's arithmetic (pib_flow_coupling)
          sc = a.smax * tanhf((sraw[j][r] + a.b3[f0 + r]) / a.smax);
          es = expf(a.inverse ? -sc : sc);
        } else {
          // single-bf16 draws: tanh from the hardware exp2 and reciprocal (it only shapes the
          // network's output), e^{+-s} from the hardware exp2 (~2e-7 relative; the log-det
          // adds the same s the map exponentiates)
          const float x = (sraw[j][r] + a.b3[f0 + r]) * ismax;
          const float e2 = __builtin_amdgcn_exp2f(2.8853900817779268f * x);  // e^{2x}
          sc = a.smax * fmaf(-2.0f, __builtin_amdgcn_rcpf(e2 + 1.0f), 1.0f);
          es = __builtin_amdgcn_exp2f((a.inverse ? -1.4426950408889634f : 1.4426950408889634f) * sc);
        }
        zz[r] = a.inverse ? (zz[r] - sh) * es : fmaf(zz[r], es, sh);
        part[j] += (double)sc;
      }
      *(float4*)(zrow[j] + a.b_off + f0) = float4{zz[0], zz[1], zz[2], zz[3]};
    }
    PIB_STAGE_END(s)
  }
#undef PIB_STAGE_BEGIN
#undef PIB_STAGE_END
  // sum s over the lane's features, then over the four lanes that hold the same sample
#pragma unroll
  for (int j = 0; j < kNT; ++j) {
    double t = part[j];
    t += shfl_xor_d(t, 16);
    t += shfl_xor_d(t, 32);
    if (g == 0) a.logdet[n0 + 16 * j] += a.inverse ? -t : t;
  }
}

// ---------------------------------------------------------------------------------------
// k_flow_layer_ws: the single-bf16 layer at D = 512 and 1024 on fragment-ordered weights
// (pib_flow_layer_fused_frag), weights streamed per wave (v_mfma_f32_32x32x16_bf16: lane l
// supplies A[l%32][8(l/32)+e] and B[8(l/32)+e][l%32] and holds D[8(r/4)+4(l/32)+r%4][l%32] in
// register r). Work-group: 8 waves and SB = 128 samples. The layer is a sequence of weight
// blocks of 16 inputs: the first product's, the second's, then those of H/256 passes of the
// last. In the hidden products wave w owns output rows 32w..32w+31 (one M tile); in pass p of
// the last it owns the scale rows 32q.. and shift rows H + 32q.. of features 32q..32q+31, q =
// (H/256) w + p (two M tiles). The weights are stored in the MFMA's A-operand order, so a
// wave's block of a tile is one contiguous 1 KB load (with rows of W^T, each lane's 16 bytes
// would come from a different 32-byte piece of 32 rows, and the address and L1 units, not the
// matrix cores, would bound the layer). The blocks stream from L2 into a register ring
// kRing - 1 blocks ahead of use; the biases sit in LDS and the coupling's z is loaded ahead of
// its epilogue, because vector loads complete in order: a load issued behind the ring and
// waited for at once would wait for the whole ring. z moves in whole 128-byte row segments,
// transposed through per-wave LDS slots. The log-det partials are added to the result by a
// no-return atomic (a load at the end would wait for every z store of the wave to complete).
// The activations of all SB samples are
// staged in LDS as bf16 rows in natural order -- the conditioning half (at D = 1024 in two
// parts of 256 inputs, the second behind a barrier pair halfway through the first product),
// then h1, then h2 over the conditioning half -- and every wave reads them as B operands one block ahead, each
// 16-byte read feeding one 32x32x16 MFMA (two in the last product) for 32 samples. Against
// k_flow_layer (each wave reading every staged weight tile for its own 16 samples, a barrier
// per weight stage) this halves the LDS bytes per flop, removes its bank conflicts and takes
// four barriers per layer (six at D = 1024) instead of 32. The 16-register MFMA results are early-clobber for
// the compiler, so a destination never partly overlaps its accumulator input.
// ---------------------------------------------------------------------------------------
typedef float f32x16f __attribute__((ext_vector_type(16)));

template <int H, int SB>
struct WsGeo {
  static constexpr int HID = 256;
  static constexpr int NT = SB / 32;  // MFMA N tiles (32 samples each)
  static constexpr int P3 = H / 256;  // passes of the last product per wave
  static constexpr int KB1 = H / 16, KB2 = HID / 16;
  static constexpr int NBLK = KB1 + KB2 + P3 * KB2;  // weight blocks per wave and layer
  // weight blocks in registers: kRing - 1 in flight ahead of the one in use (fewer at
  // D = 1024, whose longer unrolled first product needs the registers)
  static constexpr int kRing = H == 256 ? 7 : 5;
  static constexpr int XH = 256;      // conditioning inputs staged at a time (H / XH stagings)
  static constexpr int NXH = H / XH;
  static constexpr int ldx = XH + kPadF, ldh = HID + kPadF;
  static constexpr int elemsA = SB * ldx, elemsB = SB * ldh;
  static constexpr int nbias = 2 * HID + 2 * H;  // b1, b2, b3
  static constexpr size_t lds_bytes =
      (size_t)(elemsA + elemsB) * 2 + (size_t)8 * SB * sizeof(double) + (size_t)nbias * sizeof(float);
};

template <int H, int SB>
__global__ void __launch_bounds__(512) k_flow_layer_ws(FusedLayerArgs a) {
  using G = WsGeo<H, SB>;
  constexpr int NT = G::NT, KB1 = G::KB1, KB2 = G::KB2, HID = G::HID, kRing = G::kRing;
  extern __shared__ __align__(16) unsigned short lds[];
  unsigned short* bufA = lds;              // conditioning half, later h2
  unsigned short* bufB = lds + G::elemsA;  // h1
  double* red = (double*)(lds + G::elemsA + G::elemsB);  // [8 waves][SB] log-det parts
  // biases: in LDS, so that the epilogues wait on no global load (vector loads complete in
  // order: waiting for one issued behind the weight ring would wait for the whole ring)
  float* bl1 = (float*)(red + 8 * SB);
  float* bl2 = bl1 + HID;
  float* bl3 = bl2 + HID;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int l32 = lane & 31, h = lane >> 5;
  const int64_t s0 = (int64_t)blockIdx.x * SB;
  const int D = a.D;

  // weight block b of this wave: the A operands of its M tiles (the second in the last
  // product only), 1 KB contiguous per tile in the fragment order of pib_split_bf16's
  // PIB_SPLIT_FRAG32: tile m, block c of a K-input matrix at ((m K/16 + c) 64 + lane) 8
  uvec4f ring[kRing][2];
  auto fetch_block = [&](int b, uvec4f(&dst)[2]) {
    if (b < KB1 + KB2) {
      const bool first = b < KB1;
      const int nb = first ? KB1 : KB2, c = first ? b : b - KB1;
      dst[0] = *(const uvec4f*)((first ? a.W1h : a.W2h) + ((int64_t)(w * nb + c) * 64 + lane) * 8);
    } else {
      const int t = b - KB1 - KB2, q = G::P3 * w + t / KB2, c = t % KB2;
      dst[0] = *(const uvec4f*)(a.W3h + ((int64_t)(q * KB2 + c) * 64 + lane) * 8);
      dst[1] = *(const uvec4f*)(a.W3h + ((int64_t)((H / 32 + q) * KB2 + c) * 64 + lane) * 8);
    }
  };
#pragma unroll
  for (int b = 0; b < kRing - 1; ++b) fetch_block(b, ring[b]);

  for (int i = tid; i < G::nbias; i += 512) bl1[i] = i < HID ? a.b1[i] : (i < 2 * HID ? a.b2[i - HID] : a.b3[i - 2 * HID]);
  // inputs XH x part.. of the conditioning half of the SB samples -> bf16 rows of bufA (at
  // D = 1024 in two parts, the second staged halfway through the first product)
  auto stage_x = [&](int part) {
#pragma unroll 8
    for (int i = 0; i < SB * G::XH / 4 / 512; ++i) {
      const int u = tid + 512 * i, row = u / (G::XH / 4), c4 = u - row * (G::XH / 4);
      const float4 v = *(const float4*)(a.z + (s0 + row) * D + a.a_off + G::XH * part + 4 * c4);
      uint2 pk;
      pk.x = pack_bf16(v.x, v.y);
      pk.y = pack_bf16(v.z, v.w);
      *(uint2*)(bufA + row * G::ldx + 4 * c4) = pk;
    }
  };
  stage_x(0);
  __syncthreads();

  const float ismax = 1.0f / a.smax;
  f32x16f acc[2][NT];
  double part[NT];
#pragma unroll
  for (int n = 0; n < NT; ++n) part[n] = 0.0;
  // B operands of block b (cur) and b + 1 (nxt, read one block ahead within a step)
  auto read_b = [&](int b, uvec4f(&dst)[NT]) {
    const int step = b < KB1 ? 0 : (b < KB1 + KB2 ? 1 : 2);
    const int c = step == 0 ? b : (step == 1 ? b - KB1 : (b - KB1 - KB2) % KB2);
    const unsigned short* src = step == 1 ? bufB : bufA;
    const int ld = step == 1 ? G::ldh : G::ldx;
    const int cc = step == 0 ? c % (G::XH / 16) : c;  // block within the staged part
#pragma unroll
    for (int n = 0; n < NT; ++n) dst[n] = *(const uvec4f*)(src + (32 * n + l32) * ld + 16 * cc + 8 * h);
  };
  uvec4f bcur[NT], bnxt[NT];
  read_b(0, bcur);
  // z of the transformed half for the coupling, one tile (32 samples x the 32 features of block
  // q) at a time: loaded in whole 128-byte row segments (lane L: row 8i + L / 8, 16-byte chunk
  // (L % 8) ^ (L / 8), i = 0..3) kZAhead blocks before the end of a pass (tile 0) or while the
  // previous tile is processed, then transposed through a 4 KB LDS slot of the wave (bufB, free
  // once step 1 is done; row r at 32 r floats, chunk k at position k ^ (r % 8): conflict-free
  // reads of one chunk by 32 rows) into the accumulator layout. The coupled tile goes back into
  // its slot and out as whole row segments through a buffer resource (scalar row offsets).
  constexpr int kZAhead = 4;
  float* zslot = (float*)bufB + w * 2048;
  const int wu = __builtin_amdgcn_readfirstlane(w);  // the wave index in a scalar register
  const int zr = lane >> 3, zk = (lane & 7) ^ zr;
  const unsigned zoff = (unsigned)(zr * D + 4 * zk) * 4u;
  const __amdgpu_buffer_rsrc_t zrsrc =
      __builtin_amdgcn_make_buffer_rsrc((void*)(a.z + s0 * D), (short)0, SB * D * 4, 0x00020000);
  auto zsoff = [&](int n, int q, int i) { return ((32 * n + 8 * i) * D + a.b_off + 32 * q) * 4; };
  uvec4f zq[4];
  auto load_z = [&](int n, int q, uvec4f(&dst)[4]) {
#pragma unroll
    for (int i = 0; i < 4; ++i) dst[i] = __builtin_amdgcn_raw_buffer_load_b128(zrsrc, (int)zoff, zsoff(n, q, i), 0);
  };
  // The slot accesses of one wave are ordered by compiler fences (no barrier separates them): at
  // D = 512 the scheduler moved the stores' slot reads above the coupled values' write-back.
#define PIB_FENCE() do { asm volatile("" ::: "memory"); __builtin_amdgcn_sched_barrier(0); } while (0)
  auto put_z = [&](int t, const uvec4f(&src)[4]) {
#pragma unroll
    for (int i = 0; i < 4; ++i) *(float4*)(zslot + 1024 * t + 256 * i + 4 * lane) = __builtin_bit_cast(float4, src[i]);
  };
  auto z_store = [&](int n, int q, int t) {
#pragma unroll
    for (int i = 0; i < 4; ++i)
      __builtin_amdgcn_raw_buffer_store_b128(
          __builtin_bit_cast(uvec4f, *(const float4*)(zslot + 1024 * t + 256 * i + 4 * lane)), zrsrc, (int)zoff,
          zsoff(n, q, i), 0);
  };
#pragma clang loop unroll(full)
  for (int b = 0; b < G::NBLK; ++b) {
    if (b + kRing - 1 < G::NBLK) fetch_block(b + kRing - 1, ring[(b + kRing - 1) % kRing]);
    const int step = b < KB1 ? 0 : (b < KB1 + KB2 ? 1 : 2);
    const int t3 = b - KB1 - KB2;
    const int c = step == 0 ? b : (step == 1 ? b - KB1 : t3 % KB2);
    const int last = (step == 0 ? KB1 : KB2) - 1;
    // within a step, and across the passes of the last product (h2 stays put), the next
    // block's operands are ready now; after a hidden product they are read after its barrier
    const bool restage = step == 0 && c % (G::XH / 16) == G::XH / 16 - 1 && c != last;
    const bool ahead = b + 1 < G::NBLK && (c != last || step == 2) && !restage;
    if (ahead) read_b(b + 1, bnxt);
    if (step == 2 && c == last - kZAhead) load_z(0, G::P3 * wu + t3 / KB2, zq);
    // keep the loads here: the scheduler would otherwise sink them next to their use, leaving
    // one weight block in flight and the LDS latency exposed
    __builtin_amdgcn_sched_barrier(0);
    const bf16x8f a0 = __builtin_bit_cast(bf16x8f, ring[b % kRing][0]);
    const bf16x8f a1 = __builtin_bit_cast(bf16x8f, ring[b % kRing][1]);
#pragma unroll
    for (int n = 0; n < NT; ++n) {
      const bf16x8f bo = __builtin_bit_cast(bf16x8f, bcur[n]);
      if (c == 0) {
        acc[0][n] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a0, bo, f32x16f{}, 0, 0, 0);
        if (step == 2) acc[1][n] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a1, bo, f32x16f{}, 0, 0, 0);
      } else {
        acc[0][n] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a0, bo, acc[0][n], 0, 0, 0);
        if (step == 2) acc[1][n] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a1, bo, acc[1][n], 0, 0, 0);
      }
    }
    if (ahead) {
#pragma unroll
      for (int n = 0; n < NT; ++n) bcur[n] = bnxt[n];
    }
    if (restage) {  // the next part of the conditioning half, once every wave is done with this one
      __syncthreads();
      stage_x(c / (G::XH / 16) + 1);
      __syncthreads();
      read_b(b + 1, bcur);
    }
    if (c != last) continue;
    if (step < 2) {
      // bias, ReLU, bf16 into the next product's input rows (natural order)
      const float* bias = step == 0 ? bl1 : bl2;
      unsigned short* dst = step == 0 ? bufB : bufA;
      const int ldd = step == 0 ? G::ldh : G::ldx;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        float bv[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) bv[i] = bias[32 * w + 8 * j + 4 * h + i];
#pragma unroll
        for (int n = 0; n < NT; ++n) {
          float hv[4];
#pragma unroll
          for (int i = 0; i < 4; ++i) hv[i] = fmaxf(acc[0][n][4 * j + i] + bv[i], 0.f);
          uint2 pk;
          pk.x = pack_bf16(hv[0], hv[1]);
          pk.y = pack_bf16(hv[2], hv[3]);
          *(uint2*)(dst + (32 * n + l32) * ldd + 32 * w + 8 * j + 4 * h) = pk;
        }
      }
      __syncthreads();
      read_b(b + 1, bcur);
    } else {
      // coupling of features 32q + 8j + 4h + 0..3 of sample 32n + l32 (the arithmetic of
      // k_flow_layer's single-bf16 epilogue)
      const int q = G::P3 * wu + t3 / KB2;
#pragma unroll
      for (int n = 0; n < NT; ++n) {
        const int t = n & 1;
        put_z(t, zq);
        PIB_FENCE();
        if (n + 1 < NT) load_z(n + 1, q, zq);
        float* zs = zslot + 1024 * t + 32 * l32;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int f0 = 32 * q + 8 * j + 4 * h;
          float* zsj = zs + 4 * ((2 * j + h) ^ (l32 & 7));
          const float4 zc = *(const float4*)zsj;
          float zz[4] = {zc.x, zc.y, zc.z, zc.w};
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            const float sh = acc[1][n][4 * j + i] + bl3[H + f0 + i];
            const float x = (acc[0][n][4 * j + i] + bl3[f0 + i]) * ismax;
            const float e2 = __builtin_amdgcn_exp2f(2.8853900817779268f * x);  // e^{2x}
            const float sc = a.smax * fmaf(-2.0f, __builtin_amdgcn_rcpf(e2 + 1.0f), 1.0f);
            const float es = __builtin_amdgcn_exp2f((a.inverse ? -1.4426950408889634f : 1.4426950408889634f) * sc);
            zz[i] = a.inverse ? (zz[i] - sh) * es : fmaf(zz[i], es, sh);
            part[n] += (double)sc;
          }
          *(float4*)zsj = float4{zz[0], zz[1], zz[2], zz[3]};
        }
        PIB_FENCE();
        z_store(n, q, t);
        // the log-det partial materialised here: left to the scheduler, the FP64 additions sank to
        // the kernel's end and the 64 scales per lane they need were spilled
        asm volatile("" : "+v"(part[n]));
      }
    }
  }
  // log-det: over the lane's features, the two lanes of a sample, then the eight waves
#pragma unroll
  for (int n = 0; n < NT; ++n) {
    double t = part[n];
    t += shfl_xor_d(t, 32);
    if (h == 0) red[w * SB + 32 * n + l32] = t;
  }
  __syncthreads();
  if (tid < SB) {
    double t = 0.0;
#pragma unroll
    for (int v = 0; v < 8; ++v) t += red[v * SB + tid];
    // a no-return atomic add: a load here would wait for every z store of the wave to complete
    unsafeAtomicAdd(a.logdet + s0 + tid, a.inverse ? -t : t);
  }
}
#undef PIB_FENCE

template <int H, int SB>
hipError_t launch_ws(const FusedLayerArgs& a, hipStream_t s) {
  using G = WsGeo<H, SB>;
  static unsigned long long lds_set = 0;  // 144 KB of dynamic LDS: above the default 64 KB
  if (hipError_t e = ensure_dynamic_lds((const void*)k_flow_layer_ws<H, SB>, G::lds_bytes, lds_set)) return e;
  hipLaunchKernelGGL((k_flow_layer_ws<H, SB>), dim3((unsigned)(a.B / SB)), dim3(512), G::lds_bytes, s, a);
  return hipGetLastError();
}

template <int H, int HID, bool LO>
hipError_t launch_hh(const FusedLayerArgs& a, hipStream_t s) {
  using S = Stages<H, HID, LO>;
  const size_t lds = (size_t)2 * S::stage * sizeof(unsigned short);
  static unsigned long long lds_set = 0;  // 67.6 KB of dynamic LDS with LO: above the default 64 KB
  if (hipError_t e = ensure_dynamic_lds((const void*)k_flow_layer<H, HID, LO>, lds, lds_set)) return e;
  hipLaunchKernelGGL((k_flow_layer<H, HID, LO>), dim3((unsigned)(a.B / kRows)), dim3(kThreads), lds, s, a);
  return hipGetLastError();
}

}  // namespace

// BF16x3 at D/2 = 256 only: at D/2 = 512 the first product's hi and lo operands (128
// registers) push the kernel past 256 VGPRs, and the compiler then assigns some chained MFMAs
// a destination that partly overlaps their accumulator input, which the matrix cores do not
// compute correctly. Single bf16 products (no lo operands) also run at D/2 = 512.
bool fused_layer_supported(int64_t B, int D, int hidden, bool lo) {
  return B > 0 && B % kRows == 0 && hidden == 256 && (D == 512 || (!lo && D == 1024));
}

hipError_t launch_flow_layer(const FusedLayerArgs& a, hipStream_t s) {
  const bool lo = a.W1l != nullptr;
  if (lo != (a.W2l != nullptr) || lo != (a.W3l != nullptr)) return hipErrorInvalidValue;
  if (!fused_layer_supported(a.B, a.D, a.hidden, lo)) return hipErrorInvalidValue;
  if ((a.a_off != 0 && a.a_off != a.D / 2) || (a.b_off != 0 && a.b_off != a.D / 2) || a.a_off == a.b_off)
    return hipErrorInvalidValue;
  if (a.D == 1024) return launch_hh<512, 256, false>(a, s);
  return lo ? launch_hh<256, 256, true>(a, s) : launch_hh<256, 256, false>(a, s);
}

bool frag_layer_supported(int64_t B, int D, int hidden) {
  return B > 0 && B % 128 == 0 && hidden == 256 && (D == 512 || D == 1024);
}

hipError_t launch_flow_layer_frag(const FusedLayerArgs& a, hipStream_t s) {
  if (!frag_layer_supported(a.B, a.D, a.hidden)) return hipErrorInvalidValue;
  if ((a.a_off != 0 && a.a_off != a.D / 2) || (a.b_off != 0 && a.b_off != a.D / 2) || a.a_off == a.b_off)
    return hipErrorInvalidValue;
  return a.D == 1024 ? launch_ws<512, 128>(a, s) : launch_ws<256, 128>(a, s);
}

}  // namespace pib
