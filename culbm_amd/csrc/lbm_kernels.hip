// gfx950 kernels of the D3Q27 collide+stream path.
//
// One thread per cell, 256-thread workgroups (four wave64s).  The flat cell
// index runs x fastest, so every wave reads and writes 64 consecutive cells
// of each direction plane: coalesced loads of all 27 directions.
//
// Two population layouts:
//  * pull (double buffer): f_i(x) <- src_i(x - c_i) with the reference's face
//    clamp (kernel.cu:2504-2569), stored at dst_i(x).
//  * in-place "esoteric twist" (single buffer, padded by one cell per axis).
//    With p(c) = max(c, 0) per axis, a step of parity O loads slot i at
//    x + p(-c_i) and stores slot 26-i at x + p(c_i); parity E loads slot 26-i
//    at x + p(-c_i) and stores slot i at x + p(c_i) (the reference applies
//    the same rules inside its tiles, kernel.cu:954-1019, 1195-1348).  Each
//    cell's load set equals its store set, so a step needs no second buffer.
//    The reference's face clamp is carried over exactly: a producer on a
//    domain face also writes each value whose consumer is clamped to it into
//    that consumer's load slot ("ghost" store).  Ghost slots are never in
//    any cell's current load set (their regular writer would lie outside the
//    domain), so ghost stores cannot race with loads of the same step.
#include <hip/hip_cooperative_groups.h>
#include <hip/hip_runtime.h>

#include "lbm_kernels.hpp"

namespace culbm {

constexpr int kBlock = 256;

// Workgroups are dispatched round-robin over the 8 XCDs; remap so that each
// XCD works on one contiguous slab of cells and the cache lines its waves
// share (the +-x neighbours of a wave's edge cells) stay in its own L2.
__device__ __forceinline__ uint32_t xcd_block(uint32_t b, uint32_t nblk) {
  const uint32_t q = nblk >> 3, r = nblk & 7u;
  const uint32_t x = b & 7u, j = b >> 3;
  return x * q + (x < r ? x : r) + j;
}

struct Cell {
  uint32_t t;     // index within the launch range
  int32_t x, y, z;
  uint32_t c;     // dense domain index
  uint32_t l;     // layout index
  bool ok;
};

__device__ __forceinline__ Cell locate(const StepArgs& a, uint32_t blk) {
  Cell k;
  k.t = blk * kBlock + threadIdx.x;
  k.ok = k.t < a.ncells;
  const uint32_t r = a.divx.div(k.t);
  k.x = int32_t(k.t - r * uint32_t(a.nx));
  const uint32_t zz = a.divy.div(r);
  k.y = int32_t(r - zz * uint32_t(a.ny));
  k.z = a.z0 + int32_t(zz);
  k.c = uint32_t(a.z0) * uint32_t(a.nx) * uint32_t(a.ny) + k.t;
  k.l = uint32_t(k.x) + uint32_t(a.px) * (uint32_t(k.y) + uint32_t(a.py) * uint32_t(k.z));
  return k;
}

// ---------------------------------------------------------------------------
// pull layout

template <class S>
struct PullGather {
  const typename S::T* base;
  int64_t ds;
  int32_t xm, xp, ym, yp, zm, zp;  // clamped offsets to x-1, x+1, ...
  uint32_t l;
  float inv_s;

  template <int I>
  __device__ __forceinline__ void run(float* f) const {
    constexpr int cx = cx_of(I), cy = cy_of(I), cz = cz_of(I);
    // source cell x - c_i
    const int32_t ox = cx == 1 ? xm : (cx == -1 ? xp : 0);
    const int32_t oy = cy == 1 ? ym : (cy == -1 ? yp : 0);
    const int32_t oz = cz == 1 ? zm : (cz == -1 ? zp : 0);
    const typename S::T* p = base + I * ds;
    f[I] = S::dec(p[uint32_t(int32_t(l) + ox + oy + oz)], inv_s);
    if constexpr (I < 26) run<I + 1>(f);
  }
};

template <class S>
__device__ __forceinline__ PullGather<S> make_gather(const StepArgs& a, const Cell& k,
                                                     const void* src) {
  PullGather<S> g;
  g.base = static_cast<const typename S::T*>(src);
  g.ds = a.dstride;
  const int32_t plane = a.px * a.py;
  g.xm = k.x == 0 ? 0 : -1;
  g.xp = k.x == a.nx - 1 ? 0 : 1;
  g.ym = k.y == 0 ? 0 : -a.px;
  g.yp = k.y == a.ny - 1 ? 0 : a.px;
  g.zm = k.z == 0 ? 0 : -plane;
  g.zp = k.z == a.nz - 1 ? 0 : plane;
  g.l = k.l;
  g.inv_s = a.inv_scale;
  return g;
}

template <class S, int I>
__device__ __forceinline__ void store_plain(typename S::T* base, int64_t ds, uint32_t l,
                                            const float* f, float s) {
  base[I * ds + l] = S::enc(f[I], s);
  if constexpr (I < 26) store_plain<S, I + 1>(base, ds, l, f, s);
}

// The per-cell update shared by every variant: LOAD -> EQU -> BB -> COLLIDE
// (kernel.cu:2501-2631).  f holds the loaded values on entry.
__device__ __forceinline__ void update_cell(uint32_t fl, const StepArgs& a, uint32_t c, float* f) {
  if (fl & kEqu) equilibrium(a.rho[c], a.vx[c], a.vy[c], a.vz[c], f);
  if (fl & kBounceBack) bounce_back(f);
  if (fl & kCollide) collide(a.inv_tau, a.omt, f);
}

template <class S>
__global__ __launch_bounds__(kBlock) void k_pull(const StepArgs a) {
  const Cell k = locate(a, xcd_block(blockIdx.x, a.nblk));
  if (!k.ok) return;
  const uint32_t fl = a.sflag[k.c];
  float f[27];
#pragma unroll
  for (int i = 0; i < 27; ++i) f[i] = 0.0f;
  if (fl & kLoad) make_gather<S>(a, k, a.src).template run<0>(f);
  update_cell(fl, a, k.c, f);
  if (fl & kStore)
    store_plain<S, 0>(static_cast<typename S::T*>(a.dst), a.dstride, k.l, f, a.scale);
}

// ---------------------------------------------------------------------------
// in-place (twisted) layout

// Loads of a step of parity P (1 = O: slot i, 0 = E: slot 26-i) at x + p(-c_i).
template <class S, int P>
struct TwistLoad {
  const typename S::T* base;
  int64_t ds;
  int32_t sx, sy, sz;  // layout strides of one step along x, y, z
  uint32_t l;
  float inv_s;
  template <int I>
  __device__ __forceinline__ void run(float* f) const {
    constexpr int cx = cx_of(I), cy = cy_of(I), cz = cz_of(I);
    constexpr int slot = P == 1 ? I : 26 - I;
    const int32_t o = (cx == -1 ? sx : 0) + (cy == -1 ? sy : 0) + (cz == -1 ? sz : 0);
    const typename S::T* p = base + slot * ds;
    f[I] = S::dec(p[l + uint32_t(o)], inv_s);
    if constexpr (I < 26) run<I + 1>(f);
  }
};

// Stores of a step of parity P (1 = O: slot 26-i, 0 = E: slot i) at
// x + p(c_i), plus the ghost copies for consumers clamped onto this cell.
template <class S, int P>
struct TwistStore {
  typename S::T* base;
  int64_t ds;
  int32_t sx, sy, sz;
  uint32_t l;
  float s;
  // face[a] bit 0: coordinate is 0 on axis a; bit 1: coordinate is n_a - 1
  uint32_t fx, fy, fz;
  bool face;

  template <int I>
  __device__ __forceinline__ void run(const float* f) const {
    constexpr int cx = cx_of(I), cy = cy_of(I), cz = cz_of(I);
    constexpr int slot = P == 1 ? 26 - I : I;
    const typename S::T v = S::enc(f[I], s);
    typename S::T* p = base + slot * ds;
    const int32_t ox = cx == 1 ? sx : 0, oy = cy == 1 ? sy : 0, oz = cz == 1 ? sz : 0;
    p[l + uint32_t(ox + oy + oz)] = v;
    if (face) {
      // a consumer is clamped on axis a when x_a - c_a leaves the domain:
      // c_a = +1 on the low face, c_a = -1 on the high face.
      const uint32_t gx = cx == 1 ? (fx & 1u) : (cx == -1 ? (fx >> 1) : 0u);
      const uint32_t gy = cy == 1 ? (fy & 1u) : (cy == -1 ? (fy >> 1) : 0u);
      const uint32_t gz = cz == 1 ? (fz & 1u) : (cz == -1 ? (fz >> 1) : 0u);
      const uint32_t gm = gx | (gy << 1) | (gz << 2);
      if (gm) {
        // ghost offset on a clamped axis is p(-c_a)
        const int32_t qx = cx == -1 ? sx : 0, qy = cy == -1 ? sy : 0, qz = cz == -1 ? sz : 0;
#pragma unroll
        for (uint32_t m = 1; m < 8; ++m) {
          if ((m & ~gm) == 0) {
            const int32_t o = ((m & 1u) ? qx : ox) + ((m & 2u) ? qy : oy) + ((m & 4u) ? qz : oz);
            p[l + uint32_t(o)] = v;
          }
        }
      }
    }
    if constexpr (I < 26) run<I + 1>(f);
  }
};

template <class S, int P>
__device__ __forceinline__ TwistStore<S, P> make_twist_store(const StepArgs& a, const Cell& k,
                                                             void* dst) {
  TwistStore<S, P> w;
  w.base = static_cast<typename S::T*>(dst);
  w.ds = a.dstride;
  w.sx = 1;
  w.sy = a.px;
  w.sz = a.px * a.py;
  w.l = k.l;
  w.s = a.scale;
  w.fx = uint32_t(k.x == 0) | (uint32_t(k.x == a.nx - 1) << 1);
  w.fy = uint32_t(k.y == 0) | (uint32_t(k.y == a.ny - 1) << 1);
  w.fz = uint32_t(k.z == 0) | (uint32_t(k.z == a.nz - 1) << 1);
  w.face = (w.fx | w.fy | w.fz) != 0;
  return w;
}

template <class S, int P>
__device__ __forceinline__ TwistLoad<S, P> make_twist_load(const StepArgs& a, const Cell& k,
                                                           const void* src) {
  TwistLoad<S, P> r;
  r.base = static_cast<const typename S::T*>(src);
  r.ds = a.dstride;
  r.sx = 1;
  r.sy = a.px;
  r.sz = a.px * a.py;
  r.l = k.l;
  r.inv_s = a.inv_scale;
  return r;
}

template <class S, int P>
__global__ __launch_bounds__(kBlock) void k_twist(const StepArgs a) {
  const Cell k = locate(a, xcd_block(blockIdx.x, a.nblk));
  if (!k.ok) return;
  const uint32_t fl = a.sflag[k.c];
  float f[27];
#pragma unroll
  for (int i = 0; i < 27; ++i) f[i] = 0.0f;
  if (fl & kLoad) make_twist_load<S, P>(a, k, a.dst).template run<0>(f);
  update_cell(fl, a, k.c, f);
  if (fl & kStore) make_twist_store<S, P>(a, k, a.dst).template run<0>(f);
}

// ---------------------------------------------------------------------------
// cooperative wavefront: all chunk steps of one temporal-blocking pass in a
// single launch, grid-wide barrier between dependent steps.

struct WaveArgs {
  StepArgs a;        // z0/z1/ncells are set per chunk step inside the kernel
  int32_t depth;     // planes per chunk
  int32_t nsteps;    // steps of this pass
  int32_t parity0;   // in-place: parity of the first step; pull: unused
  const void* buf0;  // pull: buffer holding the pass input
  void* buf1;        // pull: the other buffer
};

// One cell of a step: load by layout, update, store by layout.
template <class S, bool INPLACE, int P>
__device__ __forceinline__ void step_cell(const StepArgs& a, const Cell& k) {
  const uint32_t fl = a.sflag[k.c];
  float f[27];
#pragma unroll
  for (int i = 0; i < 27; ++i) f[i] = 0.0f;
  if constexpr (INPLACE) {
    if (fl & kLoad) make_twist_load<S, P>(a, k, a.dst).template run<0>(f);
    update_cell(fl, a, k.c, f);
    if (fl & kStore) make_twist_store<S, P>(a, k, a.dst).template run<0>(f);
  } else {
    if (fl & kLoad) make_gather<S>(a, k, a.src).template run<0>(f);
    update_cell(fl, a, k.c, f);
    if (fl & kStore)
      store_plain<S, 0>(static_cast<typename S::T*>(a.dst), a.dstride, k.l, f, a.scale);
  }
}

template <class S, bool INPLACE>
__global__ __launch_bounds__(kBlock) void k_wave_coop(WaveArgs w) {
  namespace cg = cooperative_groups;
  cg::grid_group grid = cg::this_grid();
  const int32_t nz = w.a.nz;
  const uint32_t plane_cells = uint32_t(w.a.nx) * uint32_t(w.a.ny);
  const int32_t nchunk = (nz + w.nsteps - 1 + w.depth - 1) / w.depth;
  const uint32_t nblk = gridDim.x;
  const uint32_t vblk = xcd_block(blockIdx.x, nblk);
  for (int32_t k = 0; k < nchunk; ++k) {
    for (int32_t t = 1; t <= w.nsteps; ++t) {
      const int32_t lo = max(0, k * w.depth - (t - 1));
      const int32_t hi = min(nz, (k + 1) * w.depth - (t - 1));
      if (lo < hi) {
        StepArgs a = w.a;
        a.z0 = lo;
        a.z1 = hi;
        a.ncells = uint32_t(hi - lo) * plane_cells;
        const int par = (w.parity0 + t - 1) & 1;
        if (!INPLACE) {
          const bool odd = ((t - 1) & 1) != 0;
          a.src = odd ? w.buf1 : w.buf0;
          a.dst = odd ? const_cast<void*>(w.buf0) : w.buf1;
        }
        const uint32_t nround = (a.ncells + nblk * kBlock - 1) / (nblk * kBlock);
        for (uint32_t r = 0; r < nround; ++r) {
          const Cell kc = locate(a, r * nblk + vblk);
          if (kc.ok) {
            if (!INPLACE || par == 0)
              step_cell<S, INPLACE, 0>(a, kc);
            else
              step_cell<S, INPLACE, 1>(a, kc);
          }
        }
      }
      grid.sync();
    }
  }
}

// ---------------------------------------------------------------------------
// dump: rho/u of the populations the next step would load, written only
// where the DUMP bits are set (kernel.cu:2634-2735).  mode 0: pull layout;
// mode 1: in-place layout awaiting an O step; mode 2: awaiting an E step.

struct DumpArgs {
  StepArgs a;
  const uint32_t* flag;
  float* rho;
  float* vx;
  float* vy;
  float* vz;
};

template <class S, int MODE>
__global__ __launch_bounds__(kBlock) void k_dump(const DumpArgs d) {
  const Cell k = locate(d.a, xcd_block(blockIdx.x, d.a.nblk));
  if (!k.ok) return;
  float f[27];
  if constexpr (MODE == 0) make_gather<S>(d.a, k, d.a.src).template run<0>(f);
  if constexpr (MODE == 1) make_twist_load<S, 1>(d.a, k, d.a.src).template run<0>(f);
  if constexpr (MODE == 2) make_twist_load<S, 0>(d.a, k, d.a.src).template run<0>(f);
  const Moments m = moments(f);
  const uint32_t fl = d.flag[k.c];
  if (fl & kDumpRho) d.rho[k.c] = m.rho;
  if (fl & kDumpVx) d.vx[k.c] = m.ux;
  if (fl & kDumpVy) d.vy[k.c] = m.uy;
  if (fl & kDumpVz) d.vz[k.c] = m.uz;
}

// ---------------------------------------------------------------------------
// layout conversion and initialisation.  `plain` is dense FP32 [27][nz][ny][nx]
// holding the post-collision populations of every cell.

// plain FP32 -> storage, pull layout
template <class S>
__global__ __launch_bounds__(kBlock) void k_plain_to_pull(const StepArgs a, const float* plain,
                                                          void* out) {
  const Cell k = locate(a, xcd_block(blockIdx.x, a.nblk));
  if (!k.ok) return;
  const int64_t n = int64_t(a.nx) * a.ny * a.nz;
  float f[27];
#pragma unroll
  for (int i = 0; i < 27; ++i) f[i] = plain[i * n + k.c];
  store_plain<S, 0>(static_cast<typename S::T*>(out), a.dstride, k.l, f, a.scale);
}

template <class S>
__global__ __launch_bounds__(kBlock) void k_pull_to_plain(const StepArgs a, const void* in,
                                                          float* plain) {
  const Cell k = locate(a, xcd_block(blockIdx.x, a.nblk));
  if (!k.ok) return;
  const int64_t n = int64_t(a.nx) * a.ny * a.nz;
  const typename S::T* b = static_cast<const typename S::T*>(in);
#pragma unroll
  for (int i = 0; i < 27; ++i) plain[i * n + k.c] = S::dec(b[i * a.dstride + k.l], a.inv_scale);
}

// plain FP32 -> in-place layout as left by a step of parity P (store rules of P)
template <class S, int P>
__global__ __launch_bounds__(kBlock) void k_plain_to_twist(const StepArgs a, const float* plain,
                                                           void* out) {
  const Cell k = locate(a, xcd_block(blockIdx.x, a.nblk));
  if (!k.ok) return;
  const int64_t n = int64_t(a.nx) * a.ny * a.nz;
  float f[27];
#pragma unroll
  for (int i = 0; i < 27; ++i) f[i] = plain[i * n + k.c];
  make_twist_store<S, P>(a, k, out).template run<0>(f);
}

// in-place layout left by a step of parity P -> plain FP32 (regular slots)
template <class S, int P>
__global__ __launch_bounds__(kBlock) void k_twist_to_plain(const StepArgs a, const void* in,
                                                           float* plain) {
  const Cell k = locate(a, xcd_block(blockIdx.x, a.nblk));
  if (!k.ok) return;
  const int64_t n = int64_t(a.nx) * a.ny * a.nz;
  const typename S::T* b = static_cast<const typename S::T*>(in);
  const int32_t sy = a.px, sz = a.px * a.py;
#pragma unroll
  for (int i = 0; i < 27; ++i) {
    const int slot = P == 1 ? 26 - i : i;
    const int32_t o = (cx_of(i) == 1 ? 1 : 0) + (cy_of(i) == 1 ? sy : 0) + (cz_of(i) == 1 ? sz : 0);
    plain[i * n + k.c] = S::dec(b[slot * a.dstride + (k.l + uint32_t(o))], a.inv_scale);
  }
}

// Equilibrium of the rho/u fields as post-collision state (plain FP32).
__global__ __launch_bounds__(kBlock) void k_feq_plain(const StepArgs a, float* plain) {
  const Cell k = locate(a, xcd_block(blockIdx.x, a.nblk));
  if (!k.ok) return;
  const int64_t n = int64_t(a.nx) * a.ny * a.nz;
  float f[27];
  equilibrium(a.rho[k.c], a.vx[k.c], a.vy[k.c], a.vz[k.c], f);
#pragma unroll
  for (int i = 0; i < 27; ++i) plain[i * n + k.c] = f[i];
}

// Fill bytes with zero (encoded shifted rest equilibrium is all-zero bits in
// every storage format).  A kernel rather than a memset so callers may capture it.
__global__ __launch_bounds__(kBlock) void k_zero(uint32_t* p, uint64_t nwords) {
  const uint64_t stride = uint64_t(gridDim.x) * kBlock;
  for (uint64_t i = uint64_t(blockIdx.x) * kBlock + threadIdx.x; i < nwords; i += stride) p[i] = 0u;
}

// Compact per-cell step flags.
__global__ __launch_bounds__(kBlock) void k_compact_flags(const uint32_t* flag, uint8_t* sflag,
                                                          uint64_t n) {
  const uint64_t i = uint64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (i < n) sflag[i] = uint8_t(flag[i] & kStepBits);
}

// Codec probe: encode then decode, for the storage bit-exactness tests.
template <class S>
__global__ __launch_bounds__(kBlock) void k_codec(const float* in, uint16_t* enc, float* dec,
                                                  uint32_t n, float s, float inv_s) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const typename S::T h = S::enc(in[i], s);
  if constexpr (S::kBytes == 2) {
    uint16_t bits;
    __builtin_memcpy(&bits, &h, 2);
    enc[i] = bits;
  }
  dec[i] = S::dec(h, inv_s);
}

// Per-cell physics probe (one cell per thread, [n][27] arrays): op = 0 collide,
// 1 equilibrium from in4 = (rho, vx, vy, vz), 2 moments -> out4.
__global__ __launch_bounds__(kBlock) void k_cell_probe(int op, float inv_tau, const float* fin,
                                                       const float* in4, float* fout, float* out4,
                                                       uint32_t n) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  float f[27];
  if (op == 1) {
    equilibrium(in4[4 * i], in4[4 * i + 1], in4[4 * i + 2], in4[4 * i + 3], f);
  } else {
#pragma unroll
    for (int k = 0; k < 27; ++k) f[k] = fin[27 * i + k];
  }
  if (op == 0) collide(inv_tau, 1.0f - inv_tau, f);
  if (op == 2) {
    const Moments m = moments(f);
    out4[4 * i] = m.rho;
    out4[4 * i + 1] = m.ux;
    out4[4 * i + 2] = m.uy;
    out4[4 * i + 3] = m.uz;
    return;
  }
#pragma unroll
  for (int k = 0; k < 27; ++k) fout[27 * i + k] = f[k];
}

// ---------------------------------------------------------------------------
// bandwidth probes (the MI355X counterparts of the reference's DRAM/L2
// bandwidth-ratio and partial-store microbenchmarks, analysis/lambda/measure.cu,
// analysis/gamma/measure.cu): 16-byte vector accesses, grid-stride.

__global__ __launch_bounds__(kBlock) void k_bw_copy(const float4* __restrict__ in,
                                                    float4* __restrict__ out, uint64_t n) {
  const uint64_t stride = uint64_t(gridDim.x) * kBlock;
  for (uint64_t i = uint64_t(blockIdx.x) * kBlock + threadIdx.x; i < n; i += stride) out[i] = in[i];
}

__global__ __launch_bounds__(kBlock) void k_bw_read(const float4* __restrict__ in, float* out,
                                                    uint64_t n) {
  const uint64_t stride = uint64_t(gridDim.x) * kBlock;
  float acc = 0.0f;
  for (uint64_t i = uint64_t(blockIdx.x) * kBlock + threadIdx.x; i < n; i += stride) {
    const float4 v = in[i];
    acc += (v.x + v.y) + (v.z + v.w);
  }
  if (acc == 1.2345e-30f) out[0] = acc;  // keeps the loads alive
}

__global__ __launch_bounds__(kBlock) void k_bw_write(float4* __restrict__ out, uint64_t n) {
  const uint64_t stride = uint64_t(gridDim.x) * kBlock;
  for (uint64_t i = uint64_t(blockIdx.x) * kBlock + threadIdx.x; i < n; i += stride)
    out[i] = make_float4(0.f, 0.f, 0.f, 0.f);
}

// Scalar 4-byte copy shifted by `shift` elements: shift 0 is aligned, shift 1
// makes every wave's store straddle one extra cache line (partial-line stores).
__global__ __launch_bounds__(kBlock) void k_bw_copy_shift(const float* __restrict__ in,
                                                          float* __restrict__ out, uint64_t n,
                                                          uint32_t shift) {
  const uint64_t stride = uint64_t(gridDim.x) * kBlock;
  for (uint64_t i = uint64_t(blockIdx.x) * kBlock + threadIdx.x; i < n; i += stride)
    out[i + shift] = in[i];
}

// ---------------------------------------------------------------------------
// explicit instantiation table used by the host dispatcher

template <class S>
struct KernelSet {
  static void* pull() { return reinterpret_cast<void*>(&k_pull<S>); }
  static void* twist(int p) {
    return p ? reinterpret_cast<void*>(&k_twist<S, 1>) : reinterpret_cast<void*>(&k_twist<S, 0>);
  }
  static void* wave(bool inplace) {
    return inplace ? reinterpret_cast<void*>(&k_wave_coop<S, true>)
                   : reinterpret_cast<void*>(&k_wave_coop<S, false>);
  }
  static void* dump(int mode) {
    return mode == 0 ? reinterpret_cast<void*>(&k_dump<S, 0>)
                     : (mode == 1 ? reinterpret_cast<void*>(&k_dump<S, 1>)
                                  : reinterpret_cast<void*>(&k_dump<S, 2>));
  }
  static void* plain_to_pull() { return reinterpret_cast<void*>(&k_plain_to_pull<S>); }
  static void* pull_to_plain() { return reinterpret_cast<void*>(&k_pull_to_plain<S>); }
  static void* plain_to_twist(int p) {
    return p ? reinterpret_cast<void*>(&k_plain_to_twist<S, 1>)
             : reinterpret_cast<void*>(&k_plain_to_twist<S, 0>);
  }
  static void* twist_to_plain(int p) {
    return p ? reinterpret_cast<void*>(&k_twist_to_plain<S, 1>)
             : reinterpret_cast<void*>(&k_twist_to_plain<S, 0>);
  }
  static void* codec() { return reinterpret_cast<void*>(&k_codec<S>); }
};

}  // namespace culbm
