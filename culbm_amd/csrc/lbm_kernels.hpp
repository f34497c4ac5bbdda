// Kernel-side data structures shared by the step, dump and layout kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "lbm_physics.hpp"

namespace culbm {

// Unsigned 32-bit division by a run-time invariant (round-up multiplier):
// q = (mulhi(m, n) + n) >> l, evaluated in 64 bits; exact for all n < 2^32.
struct FastDiv {
  uint32_t d, m, l;
  __host__ static FastDiv make(uint32_t d) {
    FastDiv f;
    f.d = d;
    uint32_t l = 0;
    while ((uint64_t(1) << l) < d) ++l;
    f.l = l;
    f.m = uint32_t(((((uint64_t(1) << l) - d) << 32) / d) + 1);
    return f;
  }
  __host__ __device__ uint32_t div(uint32_t n) const {
#if defined(__HIP_DEVICE_COMPILE__)
    const uint64_t t = uint64_t(__umulhi(m, n)) + n;
#else
    const uint64_t t = ((uint64_t(m) * n) >> 32) + n;
#endif
    return uint32_t(t >> l);
  }
};

// Storage formats of the populations.  All three hold the shifted value
// f_i - w_i; the 16-bit ones hold it multiplied by a power-of-two scale, so
// encode = round-to-nearest-even(scale * f) and decode = h * (1/scale) are
// exact apart from the single rounding.
enum Storage : int32_t { kFP32 = 0, kFP16 = 1, kBF16 = 2 };

// Stochastic rounding of the 16-bit formats (opt-in).  Every value a step
// stores gets one random byte b from a hash of (cell, direction, number of the
// state the step produces), the same in every kernel and layout, so all
// variants stay bit-identical; b decides the rounding direction with
// probability equal to the value's position between its two neighbours in the
// format (to 1/256).
__host__ __device__ inline uint32_t mix32(uint32_t x) {  // two multiply-xorshift rounds
  x ^= x >> 16;
  x *= 0x7feb352du;
  x ^= x >> 15;
  x *= 0x846ca68bu;
  x ^= x >> 16;
  return x;
}

struct SrHash {
  uint32_t h[7];  // 4 random bytes each, one per direction
  __device__ __forceinline__ void make(uint32_t cell, uint32_t state) {
    const uint32_t key = mix32((cell * 0x9e3779b1u) ^ mix32(state * 0x85ebca77u + 0x632be5abu));
#pragma unroll
    for (int j = 0; j < 7; ++j) h[j] = mix32(key + uint32_t(j) * 0x27d4eb2fu);
  }
  template <int I>
  __device__ __forceinline__ uint32_t byte() const {
    return (h[I / 4] >> (8 * (I % 4))) & 0xffu;
  }
};

struct StoreF32 {
  using T = float;
  static constexpr int kBytes = 4;
  __device__ __forceinline__ static float dec(T v, float) { return v; }
  __device__ __forceinline__ static float dec_fast(T v, float, float) { return v; }
  __device__ __forceinline__ static T enc(float v, float) { return v; }
  __device__ __forceinline__ static T enc_sr(float v, float, uint32_t) { return v; }
};

struct StoreF16 {
  using T = _Float16;
  static constexpr int kBytes = 2;
  __device__ __forceinline__ static float dec(T v, float inv_s) { return float(v) * inv_s; }
  // The same value as one mixed-precision FMA (v_fma_mix_f32): the product of a
  // binary16 value and a power of two is exact, and a -0 addend leaves every
  // product unchanged, the sign of a zero included.  nz must be -0.0f; it comes
  // in as a kernel argument, since a -0 constant lets the compiler fold the FMA
  // back into a conversion and a multiply.
  __device__ __forceinline__ static float dec_fast(T v, float inv_s, float nz) { return fmaf(float(v), inv_s, nz); }
  // round-to-nearest-even(v * s + 0) in one mixed-precision FMA with a binary16
  // result (v_fma_mixlo_f16): v * s is exact, so the single rounding equals the
  // conversion of v * s; the +0 addend maps -0 to +0.
  __device__ __forceinline__ static T enc(float v, float s) { return T(fmaf(v, s, 0.0f)); }
  // stochastic: add b * 2^5 + 2^4 below the 10 kept mantissa bits of the FP32
  // value, clear the 13 dropped bits, convert (exact in the binary16 normal
  // range; below it the conversion rounds to the subnormal grid)
  __device__ __forceinline__ static T enc_sr(float v, float s, uint32_t b) {
    const uint32_t u = (__float_as_uint(v * s) + (b << 5) + 0x10u) & ~0x1fffu;
    return T(__uint_as_float(u));
  }
};

struct StoreBF16 {
  using T = uint16_t;
  static constexpr int kBytes = 2;
  __device__ __forceinline__ static float dec(T v, float inv_s) {
    return __uint_as_float(uint32_t(v) << 16) * inv_s;
  }
  __device__ __forceinline__ static float dec_fast(T v, float inv_s, float) { return dec(v, inv_s); }
  // round to nearest even by the hardware conversion (v_cvt_pk_bf16_f32), the
  // same bits as adding 0x7fff + (bit 16) to the FP32 pattern and truncating
  __device__ __forceinline__ static T enc(float v, float s) {
    return __builtin_bit_cast(uint16_t, static_cast<__bf16>(v * s));
  }
  // stochastic: add b * 2^8 + 2^7 to the FP32 bits, then truncate
  __device__ __forceinline__ static T enc_sr(float v, float s, uint32_t b) {
    const uint32_t u = __float_as_uint(v * s) + (b << 8) + 0x80u;
    return T(u >> 16);
  }
};

// Workgroup orders of a cell launch (StepArgs::order).
constexpr uint32_t kOrderDispatch = 0;  // blockIdx as dispatched, round robin over the XCDs
constexpr uint32_t kOrderSlabs = 1;     // one contiguous slab of workgroups per XCD
constexpr uint32_t kOrderRuns = 2;      // runs of consecutive workgroups per XCD, dealt round robin

// Arguments of one step launch over the z-plane range [z0, z1).
//
// Domain arrays (flags, rho/u) are dense [nz][ny][nx].  Population storage is
// [27][pz][py][px] with direction stride `dstride`; the pull layout uses
// px=nx, py=ny, pz=nz, the in-place layout pads every axis by one cell so the
// twisted store at x+1 of the last cell has a home.
struct StepArgs {
  int32_t nx, ny, nz;
  int32_t px, py;        // layout pitches (cells per row, rows per plane)
  uint32_t gcol;         // in-place rows packed at pitch nx: first element of the ghost column
                         // (the c_x = +1 stores of x = nx-1) in each direction array; 0: rows
                         // padded, the ghost column is x = nx
  int32_t z0, z1;        // plane range of this launch
  uint32_t ncells;       // (z1-z0)*nx*ny
  int64_t dstride;       // elements between consecutive directions
  FastDiv divx, divy;    // divide by nx, by ny
  uint32_t nblk;         // workgroups in the launch (for the XCD remap)
  uint32_t order;         // workgroup order: kOrderDispatch / kOrderSlabs / kOrderRuns (cell_block)
  const uint8_t* sflag;  // per-cell step bits (kStepBits of the flag word)
  const float* rho;
  const float* vx;
  const float* vy;
  const float* vz;
  const void* src;
  void* dst;  // pull: destination buffer; in-place: same as src
  float inv_tau, omt;
  float scale, inv_scale;
  float neg_zero;        // -0.0f: the addend of the one-instruction 16-bit decode (dec_fast)
  uint32_t state;        // number of the state this launch produces (stochastic rounding key)
  // fused dump (the DUMP step kernels): the full flag word for the DUMP bits
  // and the rho/u buffers it writes (the same buffers as the EQU inputs above)
  const uint32_t* flag;
  float* mrho;
  float* mvx;
  float* mvy;
  float* mvz;
};

}  // namespace culbm
