// D3Q27 BGK physics on shifted populations (stored value = f_i - w_i).
//
// Device-side restatement of the reference update, written for gfx950 with an
// explicit FP32 operation order so that every stream/blocking variant (and
// the host FP32 restatement) produce identical bits:
//   equilibrium  src/lbm_core/device_function.hpp:305-398   (calcEqu<27>)
//   moments      src/lbm_core/device_function.hpp:439-463   (calcRhoU<27>)
//   collision    src/lbm_core/device_function.hpp:676-754   (collision2<27>)
// The library is compiled with -ffp-contract=off: only the fmaf() calls below
// fuse, exactly where the reference writes fma().
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace culbm {

// Per-cell flag bits, same values as the reference
// (src/simulator/single_dev_expt/local_config.hpp:12-22).
enum : uint32_t {
  kLoad = 1u << 0,
  kEqu = 1u << 1,
  kCollide = 1u << 2,
  kBounceBack = 1u << 3,
  kStore = 1u << 4,
  kDumpRho = 1u << 27,
  kDumpVx = 1u << 28,
  kDumpVy = 1u << 29,
  kDumpVz = 1u << 30,
  kCorrect = 1u << 31,
};
// The step kernels read a compact 1-byte copy holding bits 0..4.
constexpr uint32_t kStepBits = 0x1fu;

// Direction i = (cx+1) + 3(cy+1) + 9(cz+1); opposite 26-i; rest 13
// (src/basic/velocity_set.hpp:189-238).
__host__ __device__ constexpr int cx_of(int i) { return i % 3 - 1; }
__host__ __device__ constexpr int cy_of(int i) { return (i / 3) % 3 - 1; }
__host__ __device__ constexpr int cz_of(int i) { return i / 9 - 1; }

struct Moments {
  float rho, ux, uy, uz;
};

// calcRhoU<27>: rho = (1+f0) + (f1+f2) + ... + (f25+f26) left to right;
// u = (sum of c-signed pairs) * (1/rho).
__device__ __forceinline__ Moments moments(const float* f) {
  Moments m;
  float rho = 1.0f + f[0];
#pragma unroll
  for (int k = 1; k < 27; k += 2) rho = rho + (f[k] + f[k + 1]);
  const float inv = 1.0f / rho;
  float ux = f[2] - f[0];
  ux = ux + (f[5] - f[3]);
  ux = ux + (f[8] - f[6]);
  ux = ux + (f[11] - f[9]);
  ux = ux + (f[14] - f[12]);
  ux = ux + (f[17] - f[15]);
  ux = ux + (f[20] - f[18]);
  ux = ux + (f[23] - f[21]);
  ux = ux + (f[26] - f[24]);
  float uy = f[6] - f[0];
  uy = uy + (f[7] - f[1]);
  uy = uy + (f[8] - f[2]);
  uy = uy + (f[15] - f[9]);
  uy = uy + (f[16] - f[10]);
  uy = uy + (f[17] - f[11]);
  uy = uy + (f[24] - f[18]);
  uy = uy + (f[25] - f[19]);
  uy = uy + (f[26] - f[20]);
  float uz = f[18] - f[0];
  uz = uz + (f[19] - f[1]);
  uz = uz + (f[20] - f[2]);
  uz = uz + (f[21] - f[3]);
  uz = uz + (f[22] - f[4]);
  uz = uz + (f[23] - f[5]);
  uz = uz + (f[24] - f[6]);
  uz = uz + (f[25] - f[7]);
  uz = uz + (f[26] - f[8]);
  m.rho = rho;
  m.ux = ux * inv;
  m.uy = uy * inv;
  m.uz = uz * inv;
  return m;
}

constexpr float kW0 = 8.0f / 27.0f;
constexpr float kWS = 2.0f / 27.0f;
constexpr float kWE = 1.0f / 54.0f;
constexpr float kWC = 1.0f / 216.0f;

// Terms shared by the equilibrium of all 27 directions.
struct EqTerms {
  float s[27];  // signed projection c_i . (3u), built from the reference's pair sums
  float c3, rho, rhom1;
  float rws, rwe, rwc, mws, mwe, mwc;
};

__device__ __forceinline__ void eq_terms(float rho, float vx, float vy, float vz, EqTerms& e) {
  e.rho = rho;
  e.rhom1 = rho - 1.0f;
  e.c3 = -3.0f * (vx * vx + vy * vy + vz * vz);
  vx *= 3.0f;
  vy *= 3.0f;
  vz *= 3.0f;
  const float v0 = vx + vy, v1 = vx + vz, v2 = vy + vz, v3 = vx - vy, v4 = vx - vz,
              v5 = vy - vz, v6 = vx + vy + vz, v7 = vx + vy - vz, v8 = vx - vy + vz,
              v9 = vy - vx + vz;
  e.s[0] = -v6;  e.s[1] = -v2;  e.s[2] = -v9;
  e.s[3] = -v1;  e.s[4] = -vz;  e.s[5] = v4;
  e.s[6] = -v8;  e.s[7] = v5;   e.s[8] = v7;
  e.s[9] = -v0;  e.s[10] = -vy; e.s[11] = v3;
  e.s[12] = -vx; e.s[13] = 0.0f; e.s[14] = vx;
  e.s[15] = -v3; e.s[16] = vy;  e.s[17] = v0;
  e.s[18] = -v7; e.s[19] = -v5; e.s[20] = v8;
  e.s[21] = -v4; e.s[22] = vz;  e.s[23] = v1;
  e.s[24] = v9;  e.s[25] = v2;  e.s[26] = v6;
  e.rws = kWS * rho;
  e.rwe = kWE * rho;
  e.rwc = kWC * rho;
  e.mws = kWS * e.rhom1;
  e.mwe = kWE * e.rhom1;
  e.mwc = kWC * e.rhom1;
}

template <int I>
__device__ __forceinline__ float feq_dir(const EqTerms& e) {
  constexpr int k = (cx_of(I) != 0) + (cy_of(I) != 0) + (cz_of(I) != 0);
  const float rw = k == 1 ? e.rws : (k == 2 ? e.rwe : e.rwc);
  const float mw = k == 1 ? e.mws : (k == 2 ? e.mwe : e.mwc);
  const float s = e.s[I];
  return __builtin_fmaf(rw, __builtin_fmaf(0.5f, __builtin_fmaf(s, s, e.c3), s), mw);
}

template <int I>
__device__ __forceinline__ void equ_one(const EqTerms& e, float* f) {
  if constexpr (I == 13) {
    f[13] = kW0 * __builtin_fmaf(e.rho, 0.5f * e.c3, e.rhom1);
  } else {
    f[I] = feq_dir<I>(e);
  }
  if constexpr (I < 26) equ_one<I + 1>(e, f);
}

// calcEqu<27>
__device__ __forceinline__ void equilibrium(float rho, float vx, float vy, float vz, float* f) {
  EqTerms e;
  eq_terms(rho, vx, vy, vz, e);
  equ_one<0>(e, f);
}

template <int I>
__device__ __forceinline__ void relax_one(const EqTerms& e, float inv_tau, float omt, float* f) {
  if constexpr (I == 13) {
    f[13] = inv_tau * kW0 * __builtin_fmaf(e.rho, 0.5f * e.c3, e.rhom1) + omt * f[13];
  } else {
    f[I] = inv_tau * feq_dir<I>(e) + omt * f[I];
  }
  if constexpr (I < 26) relax_one<I + 1>(e, inv_tau, omt, f);
}

// collision2<27>: f <- invTau*feq(rho(f), u(f)) + (1-invTau)*f
__device__ __forceinline__ void collide(float inv_tau, float omt, float* f) {
  const Moments m = moments(f);
  EqTerms e;
  eq_terms(m.rho, m.ux, m.uy, m.uz, e);
  relax_one<0>(e, inv_tau, omt, f);
}

__device__ __forceinline__ void bounce_back(float* f) {
#pragma unroll
  for (int i = 0; i < 13; ++i) {
    const float t = f[i];
    f[i] = f[26 - i];
    f[26 - i] = t;
  }
}

}  // namespace culbm
