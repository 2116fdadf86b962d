// Shared device helpers for the PIMC kernels (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pib {

constexpr int kWave = 64;

// ---------------------------------------------------------------------------------------
// Philox4x32-10 counter-based generator (Salmon, Moraes, Dror, Shaw, "Parallel random
// numbers: as easy as 1, 2, 3", SC'11). Written from the published round function and
// constants; 10 rounds, key bumped by the Weyl constants between rounds.
// ---------------------------------------------------------------------------------------
struct u32x4 {
  uint32_t x, y, z, w;
};

__device__ __forceinline__ void mulhilo(uint32_t a, uint32_t b, uint32_t& hi, uint32_t& lo) {
  uint64_t p = (uint64_t)a * (uint64_t)b;
  hi = (uint32_t)(p >> 32);
  lo = (uint32_t)p;
}

__device__ __forceinline__ u32x4 philox4x32_10(u32x4 ctr, uint32_t k0, uint32_t k1) {
  const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
  const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    uint32_t hi0, lo0, hi1, lo1;
    mulhilo(M0, ctr.x, hi0, lo0);
    mulhilo(M1, ctr.z, hi1, lo1);
    u32x4 n;
    n.x = hi1 ^ ctr.y ^ k0;
    n.y = lo1;
    n.z = hi0 ^ ctr.w ^ k1;
    n.w = lo0;
    ctr = n;
    k0 += W0;
    k1 += W1;
  }
  return ctr;
}

// Read-only tables through the constant address space: a wave-uniform address then becomes a
// scalar load even where the kernel also stores to global memory (the compiler cannot prove
// that a store does not clobber a generic pointer's target, and a persistent loop stores
// between reads). Only for data no kernel writes during the launch.
template <class T>
using const_as = const T __attribute__((address_space(4)));
template <class T>
__device__ __forceinline__ const_as<T>* as_const(const T* p) {
  return (const_as<T>*)p;
}

// One wave's bead coordinates (samples [s0, s0 + n) of a (B, N*P) double array) through a
// buffer resource: base and size wave-uniform, loads past the end return 0. The evaluation
// kernels' mode loops load each mode's coordinates ahead of their use; with plain loads the
// optimizer folds the loop's phi of two loads (the prologue's and the previous iteration's)
// into one load of a phi of addresses at the top of the iteration, right before the use, so a
// wave waited a full memory latency per mode. The buffer-load intrinsic is a call, which that
// fold leaves alone.
struct CoordBuf {
  __amdgpu_buffer_rsrc_t rsrc;
  int base;  // this lane's sample: index of its first coordinate in the resource (doubles)
  typedef double raw_t;
  __device__ __forceinline__ double operator[](int i) const {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(rsrc, (base + i) * 8, 0, 0));
  }
  // coordinate of mode j, bead p (P beads per mode), in two steps: the load, issued ahead,
  // and the value, formed where it is used
  __device__ __forceinline__ raw_t load(int j, int P, int p) const { return (*this)[j * P + p]; }
  __device__ __forceinline__ double value(int, raw_t r) const { return r; }
  // the wave's samples [first, first + per_wave) (first made wave-uniform explicitly); `sample`
  // is this lane's sample, or -1 for a lane past the block (which reads the wave's first)
  __device__ __forceinline__ static CoordBuf wave_samples(const double* q, int64_t B, int64_t NP, int64_t first,
                                                          int per_wave, int64_t sample) {
    // gfx9 buffer descriptor word 3: 32-bit data format, raw (stride 0) addressing
    constexpr int kWord3 = 0x00020000;
    const int64_t s0 = ((int64_t)__builtin_amdgcn_readfirstlane((int)((uint64_t)first >> 32)) << 32) |
                       (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)first);
    const int64_t n = B - s0 < per_wave ? (B - s0 > 0 ? B - s0 : 0) : per_wave;
    return CoordBuf{__builtin_amdgcn_make_buffer_rsrc((void*)(q + s0 * NP), (short)0, (int)(n * NP * 8), kWord3),
                    sample >= 0 ? (int)((sample - s0) * NP) : 0};
  }
};

// The same for coordinates given as float offsets from a per-mode centre (FP64, read through
// the constant address space): x(j, p) = centre[j] + (double)dq, the doubles an unwhitening
// that adds the centre in FP64 would have written, at half the bytes.
struct CoordOffsetBuf {
  __amdgpu_buffer_rsrc_t rsrc;
  int base;  // this lane's sample: index of its first offset in the resource (floats)
  const_as<double>* centre;
  typedef float raw_t;
  __device__ __forceinline__ raw_t load(int j, int P, int p) const {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rsrc, (base + j * P + p) * 4, 0, 0));
  }
  __device__ __forceinline__ double value(int j, raw_t r) const { return centre[j] + (double)r; }
  __device__ __forceinline__ static CoordOffsetBuf wave_samples(const float* q, const double* c, int64_t B,
                                                                int64_t NP, int64_t first, int per_wave,
                                                                int64_t sample) {
    constexpr int kWord3 = 0x00020000;  // as CoordBuf
    const int64_t s0 = ((int64_t)__builtin_amdgcn_readfirstlane((int)((uint64_t)first >> 32)) << 32) |
                       (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)first);
    const int64_t n = B - s0 < per_wave ? (B - s0 > 0 ? B - s0 : 0) : per_wave;
    return CoordOffsetBuf{
        __builtin_amdgcn_make_buffer_rsrc((void*)(q + s0 * NP), (short)0, (int)(n * NP * 4), kWord3),
        sample >= 0 ? (int)((sample - s0) * NP) : 0, as_const(c)};
  }
};

// 53-bit uniform in [0, 1) from two 32-bit words.
__device__ __forceinline__ double u53(uint32_t hi, uint32_t lo) {
  const uint64_t m = ((uint64_t)(hi >> 5) << 26) | (uint64_t)(lo >> 6);
  return (double)m * 0x1.0p-53;
}

// ---------------------------------------------------------------------------------------
// Cross-lane moves of doubles within a wave (ds_bpermute based; width 64).
// ---------------------------------------------------------------------------------------
__device__ __forceinline__ double shfl_down_d(double v, int delta) {
  return __shfl_down(v, (unsigned int)delta, kWave);
}

__device__ __forceinline__ double shfl_xor_d(double v, int mask) {
  return __shfl_xor(v, mask, kWave);
}

// log2(x) to ~1e-7 (frexp and the hardware log2 of the mantissa); 0 gives -inf, NaN NaN. Used
// to pick squaring counts, where only the integer part matters.
__device__ __forceinline__ double log2_coarse(double x) {
  int e = 0;
  const double m = frexp(x, &e);
  return (double)e + (double)__log2f((float)m);
}

// Allow `bytes` of dynamic LDS for kernel `fn` on the current device (needed above 64 KB).
// The attribute is per device; it is set on the first launch on each device (a bit per device
// in `done`, which the caller keeps per kernel; setting it twice is harmless).
inline hipError_t ensure_dynamic_lds(const void* fn, size_t bytes, unsigned long long& done) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  const unsigned long long bit = (dev >= 0 && dev < 64) ? (1ull << dev) : 0ull;
  if (bit && (__atomic_load_n(&done, __ATOMIC_ACQUIRE) & bit)) return hipSuccess;
  e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
  if (e == hipSuccess && bit) __atomic_fetch_or(&done, bit, __ATOMIC_RELEASE);
  return e;
}

}  // namespace pib
