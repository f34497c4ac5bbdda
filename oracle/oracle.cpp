// CPU oracle for the D3Q27 BGK collide+stream path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker: tests, the package's
// smoke() and bench.py's CPU-baseline leg may load it (as liboracle.so);
// the product library never does.
//
// It restates, in plain host code, the reference update of cuLBM's
// single-device engine:
//   - velocity set and weights           src/basic/velocity_set.hpp:189-255
//   - shifted equilibrium calcEqu<27>    src/lbm_core/device_function.hpp:305-398
//   - moments calcRhoU<27>               src/lbm_core/device_function.hpp:439-463
//   - BGK collision2<27>                 src/lbm_core/device_function.hpp:676-754
//   - flag bits                          src/simulator/single_dev_expt/local_config.hpp:12-22
//   - pull step with face clamp          src/simulator/single_dev_expt/kernel.cu:2489-2632
//   - dump of rho/u                      src/simulator/single_dev_expt/kernel.cu:2634-2735
//   - halo-blocking axis math            src/blocking_core/blocking_alogrithm.hpp:8-41
//
// Two arithmetic modes:
//   * FP64: the same algebra evaluated in double.  This is the accuracy oracle.
//   * FP32: the reference's exact single-precision operation order, compiled
//     with -ffp-contract=off, so results are reproducible bit for bit.  The
//     FP32 lattice step is OpenMP-parallel over (z, y) rows and doubles as
//     the CPU baseline.
// 16-bit shifted storage (FP16 / BF16 with a power-of-two scale), its
// stochastic rounding and the discrete adjoint have no reference
// counterpart; the functions here define them.
//
// What pins each part (the tests under tests/):
//   - equilibrium / moments / collision: golden vectors from the reference
//     headers compiled on the host (bit-exact), the closed forms and
//     conservation laws in FP64, feq as a fixed point;
//   - pull step, clamp, bounce-back, dump: a brute-force pure-Python step on
//     tiny grids, a uniform equilibrium that stays steady under the face
//     clamp, and cell-id tags streamed through every GPU policy (the step's
//     own source, kernel.cu, does not build here);
//   - blocking math: exhaustive brute force over small (N, B, I) and the
//     reference header's functions through oracle/_ref;
//   - 16-bit codecs: round-to-nearest-even against numpy's float16 and the
//     bfloat16 bit truncation rule; stochastic rounding: each result is one
//     of the two neighbouring representable values, its mean converges to
//     the input (unbiased), and it is a function of (cell, direction, state);
//   - adjoint: central finite differences of the same FP64 forward run
//     (rho/u of EQU cells, invTau, f^0), and the checkpointed parallel form
//     against the serial one bit for bit.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <omp.h>

#include <algorithm>
#include <vector>

namespace {

constexpr uint32_t F_LOAD = 1u << 0;
constexpr uint32_t F_EQU = 1u << 1;
constexpr uint32_t F_COLLIDE = 1u << 2;
constexpr uint32_t F_BB = 1u << 3;
constexpr uint32_t F_STORE = 1u << 4;
constexpr uint32_t F_DUMP_RHO = 1u << 27;
constexpr uint32_t F_DUMP_VX = 1u << 28;
constexpr uint32_t F_DUMP_VY = 1u << 29;
constexpr uint32_t F_DUMP_VZ = 1u << 30;

// Direction i = (cx+1) + 3(cy+1) + 9(cz+1); opposite = 26 - i; rest = 13.
inline int cx_of(int i) { return i % 3 - 1; }
inline int cy_of(int i) { return (i / 3) % 3 - 1; }
inline int cz_of(int i) { return i / 9 - 1; }

// Equilibrium deviation feq_i - w_i for one cell (calcEqu<27>).
// u' = 3u, c3 = -3|u|^2; non-rest: fma(w rho, fma(1/2, fma(s, s, c3), s), w (rho-1))
// with s = c_i . u' formed from the same pairwise sums as the reference;
// rest: w0 * fma(rho, c3/2, rho-1).
template <typename R>
struct Eq {
  R s[27];    // signed projection c_i . u' (s[13] unused)
  R c3, rho, rhom1;
};

template <typename R>
inline void eq_terms(R rho, R vx, R vy, R vz, Eq<R>& e) {
  e.rho = rho;
  e.rhom1 = rho - R(1);
  e.c3 = R(-3) * (vx * vx + vy * vy + vz * vz);
  vx *= R(3);
  vy *= R(3);
  vz *= R(3);
  const R v0 = vx + vy, v1 = vx + vz, v2 = vy + vz, v3 = vx - vy, v4 = vx - vz,
          v5 = vy - vz, v6 = vx + vy + vz, v7 = vx + vy - vz, v8 = vx - vy + vz,
          v9 = vy - vx + vz;
  R* s = e.s;
  s[0] = -v6;  s[1] = -v2;  s[2] = -v9;
  s[3] = -v1;  s[4] = -vz;  s[5] = v4;
  s[6] = -v8;  s[7] = v5;   s[8] = v7;
  s[9] = -v0;  s[10] = -vy; s[11] = v3;
  s[12] = -vx; s[13] = R(0); s[14] = vx;
  s[15] = -v3; s[16] = vy;  s[17] = v0;
  s[18] = -v7; s[19] = -v5; s[20] = v8;
  s[21] = -v4; s[22] = vz;  s[23] = v1;
  s[24] = v9;  s[25] = v2;  s[26] = v6;
}

template <typename R>
struct W {
  static constexpr R W0 = R(8) / R(27);
  static constexpr R WS = R(2) / R(27);
  static constexpr R WE = R(1) / R(54);
  static constexpr R WC = R(1) / R(216);
};

// weight class per direction: 0 rest, 1 face (straight), 2 edge, 3 corner
inline int wclass(int i) {
  const int n = (cx_of(i) != 0) + (cy_of(i) != 0) + (cz_of(i) != 0);
  return n;
}

template <typename R>
inline void feq_from(const Eq<R>& e, R* feq) {
  const R rhos = W<R>::WS * e.rho, rhoe = W<R>::WE * e.rho, rhoc = W<R>::WC * e.rho;
  const R rhom1s = W<R>::WS * e.rhom1, rhom1e = W<R>::WE * e.rhom1,
          rhom1c = W<R>::WC * e.rhom1;
#pragma GCC unroll 27
  for (int i = 0; i < 27; ++i) {
    if (i == 13) continue;
    const int k = wclass(i);
    const R rw = k == 1 ? rhos : (k == 2 ? rhoe : rhoc);
    const R mw = k == 1 ? rhom1s : (k == 2 ? rhom1e : rhom1c);
    const R s = e.s[i];
    feq[i] = std::fma(rw, std::fma(R(0.5), std::fma(s, s, e.c3), s), mw);
  }
}

template <typename R>
inline void equ(R rho, R vx, R vy, R vz, R* f) {
  Eq<R> e;
  eq_terms(rho, vx, vy, vz, e);
  feq_from(e, f);
  f[13] = W<R>::W0 * std::fma(e.rho, R(0.5) * e.c3, e.rhom1);
}

// calcRhoU<27>: fixed pairwise grouping, summed left to right.
template <typename R>
inline void moments(const R* f, R& rho, R& vx, R& vy, R& vz) {
  rho = (R(1) + f[0]);
#pragma GCC unroll 13
  for (int k = 1; k < 27; k += 2) rho = rho + (f[k] + f[k + 1]);
  const R inv = R(1) / rho;
  static const int px[9] = {2, 5, 8, 11, 14, 17, 20, 23, 26};
  static const int mx[9] = {0, 3, 6, 9, 12, 15, 18, 21, 24};
  static const int py[9] = {6, 7, 8, 15, 16, 17, 24, 25, 26};
  static const int my[9] = {0, 1, 2, 9, 10, 11, 18, 19, 20};
  static const int pz[9] = {18, 19, 20, 21, 22, 23, 24, 25, 26};
  static const int mz[9] = {0, 1, 2, 3, 4, 5, 6, 7, 8};
  vx = f[px[0]] - f[mx[0]];
  vy = f[py[0]] - f[my[0]];
  vz = f[pz[0]] - f[mz[0]];
#pragma GCC unroll 8
  for (int k = 1; k < 9; ++k) {
    vx = vx + (f[px[k]] - f[mx[k]]);
    vy = vy + (f[py[k]] - f[my[k]]);
    vz = vz + (f[pz[k]] - f[mz[k]]);
  }
  vx = vx * inv;
  vy = vy * inv;
  vz = vz * inv;
}

// collision2<27>: f_i <- invTau*feq_i + (1-invTau)*f_i  (no contraction);
// the rest direction multiplies (invTau*W0) first, as the reference does.
template <typename R>
inline void collide(R inv_tau, R* f) {
  R rho, vx, vy, vz;
  moments(f, rho, vx, vy, vz);
  Eq<R> e;
  eq_terms(rho, vx, vy, vz, e);
  R feq[27];
  feq_from(e, feq);
  const R omt = R(1) - inv_tau;
#pragma GCC unroll 27
  for (int i = 0; i < 27; ++i) {
    if (i == 13) continue;
    f[i] = inv_tau * feq[i] + omt * f[i];
  }
  f[13] = inv_tau * W<R>::W0 * std::fma(e.rho, R(0.5) * e.c3, e.rhom1) + omt * f[13];
}

inline void bounce_back(float* f) {
  for (int i = 0; i < 13; ++i) { const float t = f[i]; f[i] = f[26 - i]; f[26 - i] = t; }
}
inline void bounce_back(double* f) {
  for (int i = 0; i < 13; ++i) { const double t = f[i]; f[i] = f[26 - i]; f[26 - i] = t; }
}

// ---- 16-bit storage codecs (round to nearest even) -------------------------
inline uint32_t fbits(float f) { uint32_t u; std::memcpy(&u, &f, 4); return u; }
inline float bitsf(uint32_t u) { float f; std::memcpy(&f, &u, 4); return f; }

inline uint16_t f32_to_f16(float f) {
  const uint32_t u = fbits(f);
  const uint32_t sign = (u >> 16) & 0x8000u;
  const uint32_t a = u & 0x7fffffffu;
  if (a >= 0x7f800000u) return uint16_t(sign | 0x7c00u | (a > 0x7f800000u ? 0x200u : 0u));
  if (a >= 0x477ff000u) return uint16_t(sign | 0x7c00u);  // rounds to >= 65520 -> inf
  if (a < 0x38800000u) {                                   // result subnormal or zero
    // value = a_float; subnormal step is 2^-24
    const float af = bitsf(a);
    const float q = af * 16777216.0f;  // exact scaling by 2^24
    // round half to even to an integer
    float r = std::nearbyint(q);
    return uint16_t(sign | uint32_t(r));
  }
  const uint32_t exp = ((a >> 23) - 112u) << 10;  // rebias 127 -> 15
  const uint32_t mant = a & 0x7fffffu;
  uint32_t h = exp | (mant >> 13);
  const uint32_t rem = mant & 0x1fffu;
  if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) h += 1u;
  return uint16_t(sign | h);
}

inline float f16_to_f32(uint16_t h) {
  const uint32_t sign = uint32_t(h & 0x8000u) << 16;
  const uint32_t e = (h >> 10) & 0x1fu;
  const uint32_t m = h & 0x3ffu;
  if (e == 0) {
    const float v = float(m) * (1.0f / 16777216.0f);
    return sign ? -v : v;
  }
  if (e == 31) return bitsf(sign | 0x7f800000u | (m << 13));
  return bitsf(sign | ((e + 112u) << 23) | (m << 13));
}

inline uint16_t f32_to_bf16(float f) {
  uint32_t u = fbits(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return uint16_t((u >> 16) | 0x40u);
  u += 0x7fffu + ((u >> 16) & 1u);
  return uint16_t(u >> 16);
}

// ---- stochastic rounding of the 16-bit formats -------------------------------
// Each value a step stores gets one random byte b from a hash of (cell, direction,
// number of the state produced); the value is rounded up with probability equal
// to its position between the two neighbouring 16-bit values (to 1/256):
//  * binary16: add b * 2^5 + 2^4 to the FP32 bits, clear the 13 bits binary16
//    drops, convert (exact in the binary16 normal range);
//  * bfloat16: add b * 2^8 + 2^7 to the FP32 bits, then truncate.
inline uint32_t mix32(uint32_t x) {  // a 32-bit integer finaliser (two multiply-xorshift rounds)
  x ^= x >> 16;
  x *= 0x7feb352du;
  x ^= x >> 15;
  x *= 0x846ca68bu;
  x ^= x >> 16;
  return x;
}
inline uint32_t sr_key(uint32_t cell, uint32_t state) {
  return mix32((cell * 0x9e3779b1u) ^ mix32(state * 0x85ebca77u + 0x632be5abu));
}
inline uint32_t sr_byte(uint32_t key, int i) {
  return (mix32(key + uint32_t(i / 4) * 0x27d4eb2fu) >> (8 * (i % 4))) & 0xffu;
}
inline uint16_t f32_to_f16_sr(float t, uint32_t b) {
  return f32_to_f16(bitsf((fbits(t) + (b << 5) + 0x10u) & ~0x1fffu));
}
inline uint16_t f32_to_bf16_sr(float t, uint32_t b) {
  const uint32_t u = fbits(t) + (b << 8) + 0x80u;
  return uint16_t(u >> 16);
}

inline float bf16_to_f32(uint16_t h) { return bitsf(uint32_t(h) << 16); }

// ---- lattice helpers -------------------------------------------------------
struct Dom {
  int nx, ny, nz;
  int64_t n;
};

// Pull gather with the reference's face clamp: any offset component that
// would leave the domain is replaced by 0 (the cell pulls from itself on that
// axis).  kernel.cu:2504-2569.
inline int64_t pull_src(const Dom& d, int x, int y, int z, int i) {
  const int cx = cx_of(i), cy = cy_of(i), cz = cz_of(i);
  int sx = x - cx, sy = y - cy, sz = z - cz;
  if (sx < 0 || sx >= d.nx) sx = x;
  if (sy < 0 || sy >= d.ny) sy = y;
  if (sz < 0 || sz >= d.nz) sz = z;
  return (int64_t(sz) * d.ny + sy) * d.nx + sx;
}

template <typename R, typename LoadF, typename StoreF>
void step_generic(const Dom& d, const uint32_t* flag, const float* rho_b, const float* vx_b,
                  const float* vy_b, const float* vz_b, R inv_tau, LoadF load, StoreF store,
                  int nthreads) {
#pragma omp parallel for collapse(2) schedule(static) num_threads(nthreads)
  for (int z = 0; z < d.nz; ++z)
    for (int y = 0; y < d.ny; ++y)
      for (int x = 0; x < d.nx; ++x) {
        const int64_t c = (int64_t(z) * d.ny + y) * d.nx + x;
        const uint32_t fl = flag[c];
        R f[27];
        for (int i = 0; i < 27; ++i) f[i] = R(0);
        if (fl & F_LOAD)
          for (int i = 0; i < 27; ++i) f[i] = load(i, pull_src(d, x, y, z, i));
        if (fl & F_EQU) equ<R>(R(rho_b[c]), R(vx_b[c]), R(vy_b[c]), R(vz_b[c]), f);
        if (fl & F_BB) bounce_back(f);
        if (fl & F_COLLIDE) collide<R>(inv_tau, f);
        if (fl & F_STORE)
          for (int i = 0; i < 27; ++i) store(i, c, f[i]);
      }
}

template <typename R, typename LoadF>
void dump_generic(const Dom& d, const uint32_t* flag, LoadF load, float* rho_b, float* vx_b,
                  float* vy_b, float* vz_b, int nthreads) {
#pragma omp parallel for collapse(2) schedule(static) num_threads(nthreads)
  for (int z = 0; z < d.nz; ++z)
    for (int y = 0; y < d.ny; ++y)
      for (int x = 0; x < d.nx; ++x) {
        const int64_t c = (int64_t(z) * d.ny + y) * d.nx + x;
        const uint32_t fl = flag[c];
        R f[27];
        for (int i = 0; i < 27; ++i) f[i] = load(i, pull_src(d, x, y, z, i));
        R rho, vx, vy, vz;
        moments(f, rho, vx, vy, vz);
        if (fl & F_DUMP_RHO) rho_b[c] = float(rho);
        if (fl & F_DUMP_VX) vx_b[c] = float(vx);
        if (fl & F_DUMP_VY) vy_b[c] = float(vy);
        if (fl & F_DUMP_VZ) vz_b[c] = float(vz);
      }
}

// FP64 dump keeps double moments (the accuracy oracle keeps its own buffers).
void dump_f64_impl(const Dom& d, const uint32_t* flag, const double* src, double* rho_b,
                   double* vx_b, double* vy_b, double* vz_b, int nthreads) {
#pragma omp parallel for collapse(2) schedule(static) num_threads(nthreads)
  for (int z = 0; z < d.nz; ++z)
    for (int y = 0; y < d.ny; ++y)
      for (int x = 0; x < d.nx; ++x) {
        const int64_t c = (int64_t(z) * d.ny + y) * d.nx + x;
        const uint32_t fl = flag[c];
        double f[27];
        for (int i = 0; i < 27; ++i) f[i] = src[int64_t(i) * d.n + pull_src(d, x, y, z, i)];
        double rho, vx, vy, vz;
        moments(f, rho, vx, vy, vz);
        if (fl & F_DUMP_RHO) rho_b[c] = rho;
        if (fl & F_DUMP_VX) vx_b[c] = vx;
        if (fl & F_DUMP_VY) vy_b[c] = vy;
        if (fl & F_DUMP_VZ) vz_b[c] = vz;
      }
}

}  // namespace

extern "C" {

// ---- velocity set (velocity_set.hpp:189-255) -------------------------------
void orc_velset(int32_t* cx, int32_t* cy, int32_t* cz, double* w) {
  for (int i = 0; i < 27; ++i) {
    cx[i] = cx_of(i);
    cy[i] = cy_of(i);
    cz[i] = cz_of(i);
    const int k = wclass(i);
    w[i] = k == 0 ? 8.0 / 27 : (k == 1 ? 2.0 / 27 : (k == 2 ? 1.0 / 54 : 1.0 / 216));
  }
}

// ---- per-cell functions ----------------------------------------------------
void orc_equ_f64(double rho, double vx, double vy, double vz, double* f) { equ<double>(rho, vx, vy, vz, f); }
void orc_equ_f32(float rho, float vx, float vy, float vz, float* f) { equ<float>(rho, vx, vy, vz, f); }

void orc_moments_f64(const double* f, double* out4) { moments<double>(f, out4[0], out4[1], out4[2], out4[3]); }
void orc_moments_f32(const float* f, float* out4) { moments<float>(f, out4[0], out4[1], out4[2], out4[3]); }

void orc_collide_f64(double inv_tau, double* f) { collide<double>(inv_tau, f); }
void orc_collide_f32(float inv_tau, float* f) { collide<float>(inv_tau, f); }

// Batched per-cell collide: f is [ncell][27].
void orc_collide_batch_f32(float inv_tau, float* f, int64_t ncell) {
  for (int64_t c = 0; c < ncell; ++c) collide<float>(inv_tau, f + 27 * c);
}
void orc_equ_batch_f32(const float* rho, const float* vx, const float* vy, const float* vz, float* f,
                       int64_t ncell) {
  for (int64_t c = 0; c < ncell; ++c) equ<float>(rho[c], vx[c], vy[c], vz[c], f + 27 * c);
}
void orc_moments_batch_f32(const float* f, float* out, int64_t ncell) {
  for (int64_t c = 0; c < ncell; ++c)
    moments<float>(f + 27 * c, out[4 * c], out[4 * c + 1], out[4 * c + 2], out[4 * c + 3]);
}

// Equilibrium of rho/u fields as SoA populations f[27][n] (initial state).
void orc_feq_soa_f64(const float* rho, const float* vx, const float* vy, const float* vz, double* f,
                     int64_t n) {
  for (int64_t c = 0; c < n; ++c) {
    double g[27];
    equ<double>(rho[c], vx[c], vy[c], vz[c], g);
    for (int i = 0; i < 27; ++i) f[int64_t(i) * n + c] = g[i];
  }
}
void orc_feq_soa_f32(const float* rho, const float* vx, const float* vy, const float* vz, float* f,
                     int64_t n) {
  for (int64_t c = 0; c < n; ++c) {
    float g[27];
    equ<float>(rho[c], vx[c], vy[c], vz[c], g);
    for (int i = 0; i < 27; ++i) f[int64_t(i) * n + c] = g[i];
  }
}

// ---- lattice steps (SoA [27][nz][ny][nx], x fastest) ------------------------
// One plain pull step, double buffered (kernel.cu:2489-2632).  Cells without
// STORE leave dst untouched.
void orc_step_f64(int nx, int ny, int nz, const uint32_t* flag, const float* rho, const float* vx,
                  const float* vy, const float* vz, double inv_tau, const double* src, double* dst,
                  int nthreads) {
  const Dom d{nx, ny, nz, int64_t(nx) * ny * nz};
  step_generic<double>(
      d, flag, rho, vx, vy, vz, inv_tau,
      [&](int i, int64_t s) { return src[int64_t(i) * d.n + s]; },
      [&](int i, int64_t c, double v) { dst[int64_t(i) * d.n + c] = v; }, nthreads);
}

void orc_step_f32(int nx, int ny, int nz, const uint32_t* flag, const float* rho, const float* vx,
                  const float* vy, const float* vz, float inv_tau, const float* src, float* dst,
                  int nthreads) {
  const Dom d{nx, ny, nz, int64_t(nx) * ny * nz};
  step_generic<float>(
      d, flag, rho, vx, vy, vz, inv_tau,
      [&](int i, int64_t s) { return src[int64_t(i) * d.n + s]; },
      [&](int i, int64_t c, float v) { dst[int64_t(i) * d.n + c] = v; }, nthreads);
}

// 16-bit shifted storage: stored h = encode(scale * f); decoded f = decode(h) * (1/scale).
// fmt 1 = IEEE binary16, fmt 2 = bfloat16.  Collision in FP32.
void orc_step_s16(int nx, int ny, int nz, const uint32_t* flag, const float* rho, const float* vx,
                  const float* vy, const float* vz, float inv_tau, int fmt, float scale,
                  const uint16_t* src, uint16_t* dst, int nthreads) {
  const Dom d{nx, ny, nz, int64_t(nx) * ny * nz};
  const float inv_scale = 1.0f / scale;
  step_generic<float>(
      d, flag, rho, vx, vy, vz, inv_tau,
      [&](int i, int64_t s) {
        const uint16_t h = src[int64_t(i) * d.n + s];
        return (fmt == 1 ? f16_to_f32(h) : bf16_to_f32(h)) * inv_scale;
      },
      [&](int i, int64_t c, float v) {
        const float t = v * scale;
        dst[int64_t(i) * d.n + c] = fmt == 1 ? f32_to_f16(t + 0.0f) : f32_to_bf16(t);
      },
      nthreads);
}

// orc_step_s16 with stochastic rounding of the stored values; `state` is the
// number of the state the step produces (the step counter after the step).
void orc_step_s16_sr(int nx, int ny, int nz, const uint32_t* flag, const float* rho, const float* vx,
                     const float* vy, const float* vz, float inv_tau, int fmt, float scale,
                     uint32_t state, const uint16_t* src, uint16_t* dst, int nthreads) {
  const Dom d{nx, ny, nz, int64_t(nx) * ny * nz};
  const float inv_scale = 1.0f / scale;
  step_generic<float>(
      d, flag, rho, vx, vy, vz, inv_tau,
      [&](int i, int64_t s) {
        const uint16_t h = src[int64_t(i) * d.n + s];
        return (fmt == 1 ? f16_to_f32(h) : bf16_to_f32(h)) * inv_scale;
      },
      [&](int i, int64_t c, float v) {
        const float t = v * scale;
        const uint32_t b = sr_byte(sr_key(uint32_t(c), state), i);
        dst[int64_t(i) * d.n + c] = fmt == 1 ? f32_to_f16_sr(t, b) : f32_to_bf16_sr(t, b);
      },
      nthreads);
}

// Vector form for the codec tests: encode already-scaled values with the
// random bytes of (cell k mod cells, direction k / cells) for `state`.
void orc_encode16_sr(int fmt, const float* in, uint16_t* out, int64_t n, int64_t cells, uint32_t state) {
  for (int64_t k = 0; k < n; ++k) {
    const uint32_t b = sr_byte(sr_key(uint32_t(k % cells), state), int(k / cells));
    out[k] = fmt == 1 ? f32_to_f16_sr(in[k], b) : f32_to_bf16_sr(in[k], b);
  }
}

// Dump: rho/u from the pull-gathered populations of src, written only where
// the DUMP bits are set (kernel.cu:2634-2735).
void orc_dump_f32(int nx, int ny, int nz, const uint32_t* flag, const float* src, float* rho,
                  float* vx, float* vy, float* vz, int nthreads) {
  const Dom d{nx, ny, nz, int64_t(nx) * ny * nz};
  dump_generic<float>(
      d, flag, [&](int i, int64_t s) { return src[int64_t(i) * d.n + s]; }, rho, vx, vy, vz,
      nthreads);
}

void orc_dump_s16(int nx, int ny, int nz, const uint32_t* flag, int fmt, float scale,
                  const uint16_t* src, float* rho, float* vx, float* vy, float* vz, int nthreads) {
  const Dom d{nx, ny, nz, int64_t(nx) * ny * nz};
  const float inv_scale = 1.0f / scale;
  dump_generic<float>(
      d, flag,
      [&](int i, int64_t s) {
        const uint16_t h = src[int64_t(i) * d.n + s];
        return (fmt == 1 ? f16_to_f32(h) : bf16_to_f32(h)) * inv_scale;
      },
      rho, vx, vy, vz, nthreads);
}

void orc_dump_f64(int nx, int ny, int nz, const uint32_t* flag, const double* src, double* rho,
                  double* vx, double* vy, double* vz, int nthreads) {
  const Dom d{nx, ny, nz, int64_t(nx) * ny * nz};
  dump_f64_impl(d, flag, src, rho, vx, vy, vz, nthreads);
}

// ---- codecs (vector forms, for tests) --------------------------------------
// Storage encode of already-scaled values; binary16 stores -0 as +0.
void orc_encode16(int fmt, const float* in, uint16_t* out, int64_t n) {
  for (int64_t k = 0; k < n; ++k) out[k] = fmt == 1 ? f32_to_f16(in[k] + 0.0f) : f32_to_bf16(in[k]);
}
void orc_decode16(int fmt, const uint16_t* in, float* out, int64_t n) {
  for (int64_t k = 0; k < n; ++k) out[k] = fmt == 1 ? f16_to_f32(in[k]) : bf16_to_f32(in[k]);
}

// ---- halo-blocking axis math (blocking_alogrithm.hpp:8-41) -----------------
// Restated exactly, including the reference's validity test and the unsigned
// 32-bit arithmetic of its call site (getBlockingNumDim, simulator_expt_platform.cu:523-540).
int orc_blk_valid(int dom_, int blk_, int iter_) {
  const uint32_t dom = uint32_t(dom_), blk = uint32_t(blk_), iter = uint32_t(iter_);
  if (blk < dom && blk + 1u <= iter) return 0;
  if (2u * (blk + 1u - iter) < dom && blk + 2u <= iter) return 0;
  return 1;
}

int orc_blk_num(int dom_, int blk_, int iter_) {
  const uint32_t dom = uint32_t(dom_), blk = uint32_t(blk_), iter = uint32_t(iter_);
  if (dom <= blk) return 1;
  if (dom <= 2u * (blk + 1u - iter)) return 2;
  const uint32_t num = dom - 2u * (blk + 1u - iter);
  const uint32_t den = blk + 2u - 2u * iter;
  return int(2u + (num + den - 1u) / den);
}

int orc_valid_prev(int idx, int blk, int num, int iter, int dom) {
  if (idx == num) return dom;
  if (idx == 0) return 0;
  return (blk + 1 - iter) + (idx - 1) * (blk + 2 - 2 * iter);
}

int orc_max_threads(void) { return omp_get_max_threads(); }

// ---- discrete adjoint (FP64; no reference counterpart) ---------------------
//
// Forward: f^{t+1}(x) = Phi_x(g^t(x)), g^t = P f^t (the clamped pull), Phi_x the
// per-cell LOAD/EQU/BB/COLLIDE/STORE update.  Objective
//   J = sum_{t in [t0, t1]} sum_x w(x) u_x(g^t(x)),
// the weighted x-velocity of the populations each cell pulls (what the dump
// kernel reports for the state after t steps).  Adjoint: mu^t(x) = dJ/dg^t(x),
//   mu^t(x) = (dPhi_x/dg)^T lambda^{t+1}(x) + [t in window] w(x) du_x/dg,
//   lambda^t = P^T mu^t  (each population summed over the cells that pull it).
// Parameters: the EQU inputs (rho, u) of every cell and invTau.

// lam_h = (dC/dh)^T lam_o for the BGK collision at pre-collision h;
// returns sum_i lam_o_i * dC_i/dinvTau = sum_i lam_o_i (feq_i - h_i).
double adj_collide_f64(double inv_tau, const double* h, const double* lam_o, double* lam_h) {
  double rho, ux, uy, uz;
  moments(h, rho, ux, uy, uz);
  double feq[27];
  equ<double>(rho, ux, uy, uz, feq);
  const double u2 = ux * ux + uy * uy + uz * uz;
  double a = 0.0, bx = 0.0, by = 0.0, bz = 0.0, s = 0.0, dtau = 0.0;
  for (int i = 0; i < 27; ++i) {
    const int k = wclass(i);
    const double w = k == 0 ? 8.0 / 27 : (k == 1 ? 2.0 / 27 : (k == 2 ? 1.0 / 54 : 1.0 / 216));
    const double cu = cx_of(i) * ux + cy_of(i) * uy + cz_of(i) * uz;
    const double lw = lam_o[i] * w;
    a += lw * (1.0 - 4.5 * cu * cu + 1.5 * u2);  // d feq_i / d rho at fixed momentum
    const double q = lw * (3.0 + 9.0 * cu);      // d feq_i / d m = w (3c + 9(c.u)c - 3u)
    bx += q * cx_of(i);
    by += q * cy_of(i);
    bz += q * cz_of(i);
    s += lw;
    dtau += lam_o[i] * (feq[i] - h[i]);
  }
  bx -= 3.0 * ux * s;
  by -= 3.0 * uy * s;
  bz -= 3.0 * uz * s;
  for (int j = 0; j < 27; ++j)
    lam_h[j] = (1.0 - inv_tau) * lam_o[j] + inv_tau * (a + bx * cx_of(j) + by * cy_of(j) + bz * cz_of(j));
  return dtau;
}

// grad4 += (d feq / d(rho, ux, uy, uz))^T lam at fixed u (EQU inputs).
void adj_feq_inputs_f64(double rho, double ux, double uy, double uz, const double* lam, double* grad4) {
  const double u2 = ux * ux + uy * uy + uz * uz;
  for (int i = 0; i < 27; ++i) {
    const int k = wclass(i);
    const double w = k == 0 ? 8.0 / 27 : (k == 1 ? 2.0 / 27 : (k == 2 ? 1.0 / 54 : 1.0 / 216));
    const double cu = cx_of(i) * ux + cy_of(i) * uy + cz_of(i) * uz;
    grad4[0] += lam[i] * w * (1.0 + 3.0 * cu + 4.5 * cu * cu - 1.5 * u2);
    const double q = lam[i] * w * rho * (3.0 + 9.0 * cu);
    grad4[1] += q * cx_of(i) - lam[i] * w * rho * 3.0 * ux;
    grad4[2] += q * cy_of(i) - lam[i] * w * rho * 3.0 * uy;
    grad4[3] += q * cz_of(i) - lam[i] * w * rho * 3.0 * uz;
  }
}

// Pulled populations of cell (x,y,z) from state f (dense [27][n]).
inline void pull_cell(const Dom& d, const double* f, int x, int y, int z, double* g) {
  for (int i = 0; i < 27; ++i) g[i] = f[int64_t(i) * d.n + pull_src(d, x, y, z, i)];
}

// Forward run with all states kept, objective J, adjoint sweep; returns J.
// EQU inputs are taken as double (rho/ux/uy/uz) so finite differences can
// perturb them.  grad_* have n entries (accumulated over every step);
// grad_f0 (27n, may be NULL) receives dJ/df^0.
double orc_adjoint_f64(int nx, int ny, int nz, const uint32_t* flag, const double* rho_in,
                       const double* ux_in, const double* uy_in, const double* uz_in,
                       double inv_tau, int nsteps, int t0, int t1, const double* weight,
                       const double* f0, double* grad_rho, double* grad_ux, double* grad_uy,
                       double* grad_uz, double* grad_tau, double* grad_f0) {
  const Dom d{nx, ny, nz, int64_t(nx) * ny * nz};
  const int64_t n = d.n;
  std::vector<double> states(size_t(27 * n) * size_t(nsteps + 1));
  std::memcpy(states.data(), f0, sizeof(double) * 27 * n);
  double J = 0.0;
  auto cell = [&](int x, int y, int z) { return (int64_t(z) * ny + y) * nx + x; };
  auto objective = [&](const double* f) {
    double s = 0.0;
    for (int z = 0; z < nz; ++z)
      for (int y = 0; y < ny; ++y)
        for (int x = 0; x < nx; ++x) {
          const int64_t c = cell(x, y, z);
          if (weight[c] == 0.0) continue;
          double g[27], r, vx, vy, vz;
          pull_cell(d, f, x, y, z, g);
          moments(g, r, vx, vy, vz);
          s += weight[c] * vx;
        }
    return s;
  };
  for (int t = 0; t < nsteps; ++t) {
    const double* src = states.data() + size_t(27 * n) * t;
    double* dst = states.data() + size_t(27 * n) * (t + 1);
    if (t >= t0 && t <= t1) J += objective(src);
    for (int z = 0; z < nz; ++z)
      for (int y = 0; y < ny; ++y)
        for (int x = 0; x < nx; ++x) {
          const int64_t c = cell(x, y, z);
          const uint32_t fl = flag[c];
          double f[27] = {0};
          if (fl & F_LOAD) pull_cell(d, src, x, y, z, f);
          if (fl & F_EQU) equ<double>(rho_in[c], ux_in[c], uy_in[c], uz_in[c], f);
          if (fl & F_BB) bounce_back(f);
          if (fl & F_COLLIDE) collide<double>(inv_tau, f);
          for (int i = 0; i < 27; ++i) dst[int64_t(i) * n + c] = (fl & F_STORE) ? f[i] : src[int64_t(i) * n + c];
        }
  }
  if (nsteps >= t0 && nsteps <= t1) J += objective(states.data() + size_t(27 * n) * nsteps);

  // lam holds lambda^{t+1} = dJ/df^{t+1}; mu holds dJ/dg^t.  A cell without
  // STORE keeps f^{t+1}(x) = f^t(x) (its populations never change), so its
  // lambda carries straight over to the previous state.
  std::vector<double> mu(size_t(27 * n), 0.0), lam(size_t(27 * n), 0.0), nxt(size_t(27 * n), 0.0);
  double gt = 0.0;
  for (int64_t c = 0; c < n; ++c) grad_rho[c] = grad_ux[c] = grad_uy[c] = grad_uz[c] = 0.0;
  auto add_objective = [&](const double* f) {  // mu += w du_x/dg
    for (int z = 0; z < nz; ++z)
      for (int y = 0; y < ny; ++y)
        for (int x = 0; x < nx; ++x) {
          const int64_t c = cell(x, y, z);
          if (weight[c] == 0.0) continue;
          double g[27], r, vx, vy, vz;
          pull_cell(d, f, x, y, z, g);
          moments(g, r, vx, vy, vz);
          for (int i = 0; i < 27; ++i) mu[int64_t(i) * n + c] += weight[c] * (cx_of(i) - vx) / r;
        }
  };
  auto transpose_pull = [&](std::vector<double>& out) {  // out += P^T mu
    for (int z = 0; z < nz; ++z)
      for (int y = 0; y < ny; ++y)
        for (int x = 0; x < nx; ++x)
          for (int i = 0; i < 27; ++i)
            out[int64_t(i) * n + pull_src(d, x, y, z, i)] += mu[int64_t(i) * n + cell(x, y, z)];
  };
  // lambda^N
  if (nsteps >= t0 && nsteps <= t1) add_objective(states.data() + size_t(27 * n) * nsteps);
  transpose_pull(lam);
  for (int t = nsteps - 1; t >= 0; --t) {
    const double* src = states.data() + size_t(27 * n) * t;
    std::fill(mu.begin(), mu.end(), 0.0);
    for (int z = 0; z < nz; ++z)
      for (int y = 0; y < ny; ++y)
        for (int x = 0; x < nx; ++x) {
          const int64_t c = cell(x, y, z);
          const uint32_t fl = flag[c];
          if (!(fl & F_STORE)) continue;
          double g[27] = {0}, h[27], lo[27], lh[27];
          if (fl & F_LOAD) pull_cell(d, src, x, y, z, g);
          for (int i = 0; i < 27; ++i) h[i] = g[i];
          if (fl & F_EQU) equ<double>(rho_in[c], ux_in[c], uy_in[c], uz_in[c], h);
          if (fl & F_BB) bounce_back(h);
          for (int i = 0; i < 27; ++i) lo[i] = lam[int64_t(i) * n + c];
          if (fl & F_COLLIDE) {
            gt += adj_collide_f64(inv_tau, h, lo, lh);
          } else {
            for (int i = 0; i < 27; ++i) lh[i] = lo[i];
          }
          if (fl & F_BB) bounce_back(lh);
          if (fl & F_EQU) {
            double g4[4] = {0, 0, 0, 0};
            adj_feq_inputs_f64(rho_in[c], ux_in[c], uy_in[c], uz_in[c], lh, g4);
            grad_rho[c] += g4[0];
            grad_ux[c] += g4[1];
            grad_uy[c] += g4[2];
            grad_uz[c] += g4[3];
          } else if (fl & F_LOAD) {
            for (int i = 0; i < 27; ++i) mu[int64_t(i) * n + c] = lh[i];
          }
        }
    if (t >= t0 && t <= t1) add_objective(src);
    std::fill(nxt.begin(), nxt.end(), 0.0);
    for (int64_t c = 0; c < n; ++c)
      if (!(flag[c] & F_STORE))
        for (int i = 0; i < 27; ++i) nxt[int64_t(i) * n + c] = lam[int64_t(i) * n + c];
    transpose_pull(nxt);
    std::swap(lam, nxt);  // lambda^t
  }
  if (grad_f0) std::memcpy(grad_f0, lam.data(), sizeof(double) * 27 * n);
  if (grad_tau) *grad_tau = gt;
  return J;
}

}  // extern "C"

// ---- FP64 forward objective and checkpointed, parallel FP64 adjoint --------
//
// The same model as orc_adjoint_f64 for grids where keeping every state does
// not fit: states are kept every `every` steps and each segment is recomputed
// once.  Per-cell work runs in parallel; the transposed pull runs in parallel
// over the 27 directions (each direction writes its own plane) and every sum
// keeps the cell order of orc_adjoint_f64, so both return the same bits.

namespace {

struct AdjIn {
  Dom d;
  const uint32_t* flag;
  const double *rho, *ux, *uy, *uz;
  double inv_tau;
  const double* weight;
};

// Offsets c - src of the unclamped pull, per direction (interior cells).
inline void pull_offsets(const Dom& d, int64_t* off) {
  for (int i = 0; i < 27; ++i)
    off[i] = int64_t(i) * d.n - (cx_of(i) + int64_t(cy_of(i)) * d.nx + int64_t(cz_of(i)) * d.nx * d.ny);
}

// One FP64 pull step with double EQU inputs; cells without STORE keep dst.
void step_f64_in(const AdjIn& in, const double* src, double* dst, int nth) {
  const Dom& d = in.d;
  int64_t off[27];
  pull_offsets(d, off);
#pragma omp parallel for collapse(2) schedule(static) num_threads(nth)
  for (int z = 0; z < d.nz; ++z)
    for (int y = 0; y < d.ny; ++y)
      for (int x = 0; x < d.nx; ++x) {
        const int64_t c = (int64_t(z) * d.ny + y) * d.nx + x;
        const uint32_t fl = in.flag[c];
        if (!(fl & F_STORE)) continue;
        double f[27] = {0};
        const bool inner = x > 0 && x < d.nx - 1 && y > 0 && y < d.ny - 1 && z > 0 && z < d.nz - 1;
        if ((fl & F_LOAD) && inner)
          for (int i = 0; i < 27; ++i) f[i] = src[c + off[i]];
        else if (fl & F_LOAD)
          pull_cell(d, src, x, y, z, f);
        if (fl & F_EQU) equ<double>(in.rho[c], in.ux[c], in.uy[c], in.uz[c], f);
        if (fl & F_BB) bounce_back(f);
        if (fl & F_COLLIDE) collide<double>(in.inv_tau, f);
        for (int i = 0; i < 27; ++i) dst[int64_t(i) * d.n + c] = f[i];
      }
}

// sum_x w(x) u_x(g(x)) in cell order; with mu: mu += w du_x/dg.
double objective_f64(const AdjIn& in, const double* f, double* mu, std::vector<double>& part, int nth) {
  const Dom& d = in.d;
#pragma omp parallel for collapse(2) schedule(static) num_threads(nth)
  for (int z = 0; z < d.nz; ++z)
    for (int y = 0; y < d.ny; ++y)
      for (int x = 0; x < d.nx; ++x) {
        const int64_t c = (int64_t(z) * d.ny + y) * d.nx + x;
        part[size_t(c)] = 0.0;
        if (in.weight[c] == 0.0) continue;
        double g[27], r, vx, vy, vz;
        pull_cell(d, f, x, y, z, g);
        moments(g, r, vx, vy, vz);
        part[size_t(c)] = in.weight[c] * vx;
        if (mu)
          for (int i = 0; i < 27; ++i) mu[int64_t(i) * d.n + c] += in.weight[c] * (cx_of(i) - vx) / r;
      }
  double s = 0.0;
  for (int64_t c = 0; c < d.n; ++c)
    if (in.weight[c] != 0.0) s += part[size_t(c)];
  return s;
}

}  // namespace

extern "C" {

// FP64 forward of nsteps steps from f0 (EQU inputs as double) and
// J = sum_{t in [t0, t1]} sum_x w(x) u_x(g^t(x)).  f_end (27n, may be NULL)
// receives f^nsteps.
double orc_objective_f64(int nx, int ny, int nz, const uint32_t* flag, const double* rho_in,
                         const double* ux_in, const double* uy_in, const double* uz_in,
                         double inv_tau, int nsteps, int t0, int t1, const double* weight,
                         const double* f0, double* f_end, int nthreads) {
  const AdjIn in{{nx, ny, nz, int64_t(nx) * ny * nz}, flag, rho_in, ux_in, uy_in, uz_in, inv_tau, weight};
  const int64_t n = in.d.n;
  std::vector<double> a(f0, f0 + 27 * n), b(a), part(static_cast<size_t>(n));
  double J = 0.0;
  for (int t = 0;; ++t) {
    if (t >= t0 && t <= t1) J += objective_f64(in, a.data(), nullptr, part, nthreads);
    if (t == nsteps) break;
    step_f64_in(in, a.data(), b.data(), nthreads);
    std::swap(a, b);
  }
  if (f_end) std::memcpy(f_end, a.data(), sizeof(double) * 27 * n);
  return J;
}

// orc_adjoint_f64 with states kept every `every` steps (every >= 1) and the
// work spread over `nthreads` threads.  Same outputs, same bits.
double orc_adjoint_ckpt_f64(int nx, int ny, int nz, const uint32_t* flag, const double* rho_in,
                            const double* ux_in, const double* uy_in, const double* uz_in,
                            double inv_tau, int nsteps, int t0, int t1, const double* weight,
                            const double* f0, int every, double* grad_rho, double* grad_ux,
                            double* grad_uy, double* grad_uz, double* grad_tau, double* grad_f0,
                            int nthreads) {
  const AdjIn in{{nx, ny, nz, int64_t(nx) * ny * nz}, flag, rho_in, ux_in, uy_in, uz_in, inv_tau, weight};
  const Dom& d = in.d;
  const int64_t n = d.n;
  const size_t S = size_t(27 * n);
  if (every < 1) every = 1;
  const int nck = (nsteps + every - 1) / every;
  // checkpoints f^{j*every}, j < nck; f^N kept for the objective at N
  std::vector<std::vector<double>> ck(static_cast<size_t>(nck));
  std::vector<double> cur(f0, f0 + S), nxt(cur), part(static_cast<size_t>(n));
  double J = 0.0;
  for (int t = 0;; ++t) {
    if (t % every == 0 && t < nsteps) ck[size_t(t / every)] = cur;
    if (t >= t0 && t <= t1) J += objective_f64(in, cur.data(), nullptr, part, nthreads);
    if (t == nsteps) break;
    step_f64_in(in, cur.data(), nxt.data(), nthreads);
    std::swap(cur, nxt);
  }
  std::vector<double> mu(S, 0.0), lam(S, 0.0), lnext(S, 0.0), dtau(static_cast<size_t>(n), 0.0);
  for (int64_t c = 0; c < n; ++c) grad_rho[c] = grad_ux[c] = grad_uy[c] = grad_uz[c] = 0.0;
  auto transpose_pull = [&](std::vector<double>& out) {  // out += P^T mu, cell order per direction
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
    for (int i = 0; i < 27; ++i)
      for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y)
          for (int x = 0; x < nx; ++x)
            out[size_t(int64_t(i) * n + pull_src(d, x, y, z, i))] +=
                mu[size_t(int64_t(i) * n + (int64_t(z) * ny + y) * nx + x)];
  };
  if (nsteps >= t0 && nsteps <= t1) objective_f64(in, cur.data(), mu.data(), part, nthreads);
  transpose_pull(lam);
  double gt = 0.0;
  std::vector<std::vector<double>> seg(static_cast<size_t>(every));
  for (int j = nck - 1; j >= 0; --j) {
    const int a = j * every, b = std::min(nsteps, (j + 1) * every);
    seg[0] = ck[size_t(j)];
    for (int t = a + 1; t < b; ++t) {
      seg[size_t(t - a)].resize(S);
      std::copy(seg[size_t(t - a - 1)].begin(), seg[size_t(t - a - 1)].end(), seg[size_t(t - a)].begin());
      step_f64_in(in, seg[size_t(t - a - 1)].data(), seg[size_t(t - a)].data(), nthreads);
    }
    for (int t = b - 1; t >= a; --t) {
      const double* src = seg[size_t(t - a)].data();
      std::fill(mu.begin(), mu.end(), 0.0);
#pragma omp parallel for collapse(2) schedule(static) num_threads(nthreads)
      for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y)
          for (int x = 0; x < nx; ++x) {
            const int64_t c = (int64_t(z) * ny + y) * nx + x;
            const uint32_t fl = flag[c];
            dtau[size_t(c)] = 0.0;
            if (!(fl & F_STORE)) continue;
            double g[27] = {0}, h[27], lo[27], lh[27];
            if (fl & F_LOAD) pull_cell(d, src, x, y, z, g);
            for (int i = 0; i < 27; ++i) h[i] = g[i];
            if (fl & F_EQU) equ<double>(rho_in[c], ux_in[c], uy_in[c], uz_in[c], h);
            if (fl & F_BB) bounce_back(h);
            for (int i = 0; i < 27; ++i) lo[i] = lam[size_t(int64_t(i) * n + c)];
            if (fl & F_COLLIDE) {
              dtau[size_t(c)] = adj_collide_f64(inv_tau, h, lo, lh);
            } else {
              for (int i = 0; i < 27; ++i) lh[i] = lo[i];
            }
            if (fl & F_BB) bounce_back(lh);
            if (fl & F_EQU) {
              double g4[4] = {0, 0, 0, 0};
              adj_feq_inputs_f64(rho_in[c], ux_in[c], uy_in[c], uz_in[c], lh, g4);
              grad_rho[c] += g4[0];
              grad_ux[c] += g4[1];
              grad_uy[c] += g4[2];
              grad_uz[c] += g4[3];
            } else if (fl & F_LOAD) {
              for (int i = 0; i < 27; ++i) mu[size_t(int64_t(i) * n + c)] = lh[i];
            }
          }
      for (int64_t c = 0; c < n; ++c)
        if ((flag[c] & F_STORE) && (flag[c] & F_COLLIDE)) gt += dtau[size_t(c)];
      if (t >= t0 && t <= t1) objective_f64(in, src, mu.data(), part, nthreads);
#pragma omp parallel for schedule(static) num_threads(nthreads)
      for (int64_t c = 0; c < n; ++c)
        for (int i = 0; i < 27; ++i)
          lnext[size_t(int64_t(i) * n + c)] = (flag[c] & F_STORE) ? 0.0 : lam[size_t(int64_t(i) * n + c)];
      transpose_pull(lnext);
      std::swap(lam, lnext);
    }
  }
  if (grad_f0) std::memcpy(grad_f0, lam.data(), sizeof(double) * S);
  if (grad_tau) *grad_tau = gt;
  return J;
}

}  // extern "C"
