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

// Runs of kXcdGroup consecutive workgroups per XCD, the runs dealt round robin:
// the 8 XCDs still advance through the cells side by side (one streaming front,
// as in dispatch order), while the lines two neighbouring workgroups share (the
// one-element x shift of 9 directions at every 256-cell edge) mostly stay in one
// L2.  Whole super-groups of 8 * kXcdGroup workgroups are permuted; a partial
// last one keeps dispatch order.  At 288x272x280 against dispatch order: FP32
// in-place wavefront +1.4 %, FP16 +4 %, FP32 pull +2.3 % (runs of 2 to 32
// measured; 8 and 16 lead in every form).
constexpr uint32_t kXcdGroup = 8;

template <uint32_t RUN = kXcdGroup>
__device__ __forceinline__ uint32_t xcd_group(uint32_t b, uint32_t nblk) {
  constexpr uint32_t G = 8u * RUN;
  const uint32_t sg = b / G;
  if ((sg + 1) * G > nblk) return b;
  const uint32_t r = b - sg * G;
  return sg * G + (r & 7u) * RUN + (r >> 3);
}

// The workgroup order of a cell launch, chosen by the host per launch (a.order):
// kOrderDispatch, kOrderSlabs (xcd_block) or kOrderRuns (xcd_group).
__device__ __forceinline__ uint32_t cell_block(const StepArgs& a) {
  const uint32_t b = blockIdx.x;
  return a.order == kOrderSlabs ? xcd_block(b, a.nblk) : a.order == kOrderRuns ? xcd_group(b, a.nblk) : b;
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
// memory access: a wave-uniform base address (SGPRs) plus one 32-bit per-lane
// byte offset, the form gfx950 issues as global_load/store with `saddr`, so a
// direction costs no 64-bit address arithmetic.  Direction planes are
// `dsb` bytes apart; every per-lane offset stays below 2^32 (checked on the host).

template <class T>
__device__ __forceinline__ T ld_at(const char* base, uint32_t off) {
  return *reinterpret_cast<const T*>(base + off);
}

template <class T>
__device__ __forceinline__ void st_at(char* base, uint32_t off, T v) {
  *reinterpret_cast<T*>(base + off) = v;
}

// ---------------------------------------------------------------------------
// pull layout

template <class S>
struct PullGather {
  const char* base;
  int64_t dsb;
  uint32_t o[3][3];  // byte offset of the clamped source row (y - cy, z - cz), [cy+1][cz+1]
  int32_t xm, xp;    // byte offset to x-1 / x+1 (0 when clamped on that face)
  float inv_s, nz;

  template <int I>
  __device__ __forceinline__ void run(float* f) const {
    constexpr int cx = cx_of(I), cy = cy_of(I), cz = cz_of(I);
    const int32_t ox = cx == 1 ? xm : (cx == -1 ? xp : 0);  // source x - c_x
    f[I] = S::dec_fast(ld_at<typename S::T>(base + I * dsb, o[cy + 1][cz + 1] + uint32_t(ox)), inv_s, nz);
    if constexpr (I < 26) run<I + 1>(f);
  }
};

template <class S>
__device__ __forceinline__ PullGather<S> make_gather(const StepArgs& a, const Cell& k,
                                                     const void* src) {
  constexpr int32_t B = sizeof(typename S::T);
  PullGather<S> g;
  g.base = static_cast<const char*>(src);
  g.dsb = a.dstride * B;
  const int32_t row = a.px * B, plane = a.px * a.py * B;
  const uint32_t lb = k.l * uint32_t(B);
  // source offsets for c = -1, 0, +1 (the source sits at -c), clamped at faces
  const int32_t oy[3] = {k.y == a.ny - 1 ? 0 : row, 0, k.y == 0 ? 0 : -row};
  const int32_t oz[3] = {k.z == a.nz - 1 ? 0 : plane, 0, k.z == 0 ? 0 : -plane};
#pragma unroll
  for (int jy = 0; jy < 3; ++jy)
#pragma unroll
    for (int jz = 0; jz < 3; ++jz) g.o[jy][jz] = lb + uint32_t(oy[jy] + oz[jz]);
  g.xm = k.x == 0 ? 0 : -B;
  g.xp = k.x == a.nx - 1 ? 0 : B;
  g.inv_s = a.inv_scale;
  g.nz = a.neg_zero;
  return g;
}

// Storage encoding of direction I of one cell: round to nearest, or (SR)
// stochastically with the cell's random bytes.
template <class S, bool SR>
struct Encoder {
  float s;
  SrHash h;  // used only when SR
  template <int I>
  __device__ __forceinline__ typename S::T get(float v) const {
    if constexpr (SR)
      return S::enc_sr(v, s, h.template byte<I>());
    else
      return S::enc(v, s);
  }
};

template <class S, bool SR>
__device__ __forceinline__ Encoder<S, SR> make_encoder(const StepArgs& a, uint32_t cell, uint32_t state) {
  Encoder<S, SR> e;
  e.s = a.scale;
  if constexpr (SR) e.h.make(cell, state);
  return e;
}

template <class S, bool SR, int I>
__device__ __forceinline__ void store_plain(char* base, int64_t dsb, uint32_t lb, const float* f,
                                            const Encoder<S, SR>& e) {
  st_at<typename S::T>(base + I * dsb, lb, e.template get<I>(f[I]));
  if constexpr (I < 26) store_plain<S, SR, I + 1>(base, dsb, lb, f, e);
}

template <class S, bool SR = false>
__device__ __forceinline__ void store_plain_all(const StepArgs& a, const Cell& k, void* dst,
                                                const float* f, const Encoder<S, SR>& e) {
  constexpr uint32_t B = sizeof(typename S::T);
  store_plain<S, SR, 0>(static_cast<char*>(dst), a.dstride * B, k.l * B, f, e);
}

// Cells without LOAD start from zero populations.
__device__ __forceinline__ void clear(float* f) {
#pragma unroll
  for (int i = 0; i < 27; ++i) f[i] = 0.0f;
}

// EQU inputs of a cell on an x face (x = 0 or nx-1), loaded in the same batch
// as its populations.  An EQU cell's inputs are otherwise loaded only once its
// flag has arrived: a second round trip, and, since the counter of vector
// memory operations retires in order, a wait for every load still in flight.
// The reference convention puts its EQU cells (inlet, equilibrium outlet) on
// the x faces; an EQU cell anywhere else still loads its inputs on demand.
struct EquIn {
  float rho, vx, vy, vz;
  bool ok;  // the values are loaded (x-face cell)
};

// Only waves holding an x-face cell issue these loads.  Letting every lane
// load (the others reading cell 0) cost the in-place wavefront 2.7 %: every
// wave of the chip then read the same four lines, each of which one L2
// channel per XCD serves.
__device__ __forceinline__ EquIn equ_prefetch(const StepArgs& a, uint32_t c, int32_t x) {
  EquIn q;
  q.ok = x == 0 || x == a.nx - 1;
  q.rho = q.vx = q.vy = q.vz = 0.0f;
  if (q.ok) {
    q.rho = a.rho[c];
    q.vx = a.vx[c];
    q.vy = a.vy[c];
    q.vz = a.vz[c];
  }
  return q;
}

// The per-cell update shared by every variant: LOAD -> EQU -> BB -> COLLIDE
// (kernel.cu:2501-2631).  f holds the loaded values on entry (for every cell:
// the loads are issued before the flag arrives).  A cell with EQU replaces
// them by feq(rho, u); a cell with neither LOAD nor EQU starts from zero.  Both
// replacements sit in one divergent block that only waves holding such a cell
// enter, so interior waves pay no per-value selects for the zeroing.
__device__ __forceinline__ void update_cell(uint32_t fl, const StepArgs& a, uint32_t c, const EquIn& q,
                                            float* f) {
  if ((fl & kEqu) || !(fl & kLoad)) {
    if (fl & kEqu) {
      float rho = q.rho, vx = q.vx, vy = q.vy, vz = q.vz;
      if (!q.ok) {  // an EQU cell off the x faces
        rho = a.rho[c];
        vx = a.vx[c];
        vy = a.vy[c];
        vz = a.vz[c];
      }
      equilibrium(rho, vx, vy, vz, f);
    } else {
      clear(f);
    }
  }
  if (fl & kBounceBack) bounce_back(f);
  if (fl & kCollide) collide(a.inv_tau, a.omt, f);
}

// Fused dump: the reference's dump pass (StaticD3Q27PullDumpKernel,
// kernel.cu:2634-2735) riding on the first step after a window.  The step has
// just gathered exactly the populations the dump pass would (pull(src), or the
// twisted loads, for every cell whatever its flag), so it writes their rho/u
// where the DUMP bits are set, as the dump pass does.  A dumped field is also
// that cell's EQU input from then on (the two share one buffer, as in the
// reference), so an EQU cell takes the dumped values in this very step.
__device__ __forceinline__ void fused_dump(const StepArgs& a, uint32_t c, const float* f, EquIn& q) {
  const uint32_t fw = a.flag[c];
  if (!(fw & (kDumpRho | kDumpVx | kDumpVy | kDumpVz))) return;
  const Moments m = moments(f);
  if (fw & kDumpRho) {
    a.mrho[c] = m.rho;
    q.rho = m.rho;
  }
  if (fw & kDumpVx) {
    a.mvx[c] = m.ux;
    q.vx = m.ux;
  }
  if (fw & kDumpVy) {
    a.mvy[c] = m.uy;
    q.vy = m.uy;
  }
  if (fw & kDumpVz) {
    a.mvz[c] = m.uz;
    q.vz = m.uz;
  }
}

template <class S, bool SR, bool DUMP = false>
__global__ __launch_bounds__(kBlock) void k_pull(const StepArgs a) {
  const Cell k = locate(a, cell_block(a));
  if (!k.ok) return;
  // The gather is in bounds for every cell (clamped), so it is issued before
  // the flag arrives instead of waiting on it.
  float f[27];
  make_gather<S>(a, k, a.src).template run<0>(f);
  EquIn q = equ_prefetch(a, k.c, k.x);
  if constexpr (DUMP) fused_dump(a, k.c, f, q);
  const uint32_t fl = a.sflag[k.c];
  update_cell(fl, a, k.c, q, f);
  if (fl & kStore) store_plain_all<S, SR>(a, k, a.dst, f, make_encoder<S, SR>(a, k.c, a.state));
}

// ---------------------------------------------------------------------------
// in-place (twisted) layout

// Byte offsets of the four row/plane combinations {x, x+row, x+plane, x+row+plane}.
struct TwistRows {
  uint32_t o[2][2];
  uint32_t lb;
  int32_t row, plane;
};

template <int B>
__device__ __forceinline__ TwistRows twist_rows(const StepArgs& a, const Cell& k) {
  TwistRows r;
  r.row = a.px * B;
  r.plane = a.px * a.py * B;
  r.lb = k.l * uint32_t(B);
  r.o[0][0] = r.lb;
  r.o[1][0] = r.lb + uint32_t(r.row);
  r.o[0][1] = r.lb + uint32_t(r.plane);
  r.o[1][1] = r.lb + uint32_t(r.row + r.plane);
  return r;
}

// Loads of a step of parity P (1 = O: slot i, 0 = E: slot 26-i) at x + p(-c_i).
//
// XR (x faces read in place): a consumer clamped along x (c_x = +1 at x = 0,
// c_x = -1 at x = nx-1) pulls its own value of the step before, which that
// step stored regularly at x + p(c_x), one cell inward: it is read there
// instead of from a ghost copy.  That location belongs to the neighbouring
// cell's load/store set, so it is overwritten in the same step; the read comes
// first because the two cells are adjacent lanes of one wave (x = 0 and 1,
// nx-2 and nx-1: rows start at even lanes when nx is even) and every load of
// a wave returns before any store is issued (the update waits for all 27).
template <class S, int P, bool XR = false>
struct TwistLoad {
  const char* base;
  int64_t dsb;
  TwistRows r;
  float inv_s, nz;
  uint32_t xp, xm;  // XR: byte x offsets of the c_x = +1 / -1 loads (B or 0 at the faces)
  template <int I>
  __device__ __forceinline__ void run(float* f) const {
    constexpr int cx = cx_of(I), cy = cy_of(I), cz = cz_of(I);
    constexpr int slot = P == 1 ? I : 26 - I;
    constexpr uint32_t ox = cx == -1 ? sizeof(typename S::T) : 0;
    uint32_t off = r.o[cy == -1][cz == -1] + ox;
    if constexpr (XR && cx == 1) off = r.o[cy == -1][cz == -1] + xp;
    if constexpr (XR && cx == -1) off = r.o[cy == -1][cz == -1] + xm;
    f[I] = S::dec_fast(ld_at<typename S::T>(base + slot * dsb, off), inv_s, nz);
    if constexpr (I < 26) run<I + 1>(f);
  }
};

// Stores of a step of parity P (1 = O: slot 26-i, 0 = E: slot i) at x + p(c_i).
//
// The c_x = +1 stores of the last cell of a row land at x = nx, the ghost
// column.  With padded rows it is column nx of the row; with rows packed at
// pitch nx (a.gcol != 0) it is a separate (ny+1) x (nz+1) region per direction
// array, index a.gcol + y' + py z', so that no store reaches the next row.  Only
// the layout conversion and the moment sums read it back (its values are the
// last cell's post-collision c_x = +1 populations); no step does.
//
// Ghost copies of a face cell: a consumer is clamped onto this cell on axis a
// when x_a - c_a leaves the domain (c_a = +1 on the low face, -1 on the high
// face); its load slot on that axis is at p(-c_a) instead of p(c_a).  The
// copies for consumers clamped along x alone (the frequent case: two cells of
// every row) are stored next to each regular store; the rest (y or z faces:
// whole rows and planes) by TwistGhostYZ afterwards, from the values just
// stored, so that no population has to stay in a register for them.
template <class S, int P, bool SR, bool XG = true>
struct TwistStore {
  char* base;
  int64_t dsb;
  TwistRows r;
  Encoder<S, SR> e;
  uint32_t fx;  // bit 0: x = 0; bit 1: x = nx-1
  bool gl;      // this lane's c_x = +1 stores go to the packed layout's ghost column
  uint32_t gb;  // byte offset of its ghost-column entry (y, z)
  uint32_t gpy; // byte offset of one ghost-column plane
  template <int I>
  __device__ __forceinline__ void run(const float* f) const {
    using T = typename S::T;
    constexpr int cx = cx_of(I), cy = cy_of(I), cz = cz_of(I);
    constexpr int slot = P == 1 ? 26 - I : I;
    constexpr uint32_t ox = cx == 1 ? sizeof(T) : 0;
    const T v = e.template get<I>(f[I]);
    uint32_t off = r.o[cy == 1][cz == 1] + ox;
    if constexpr (cx == 1 && sizeof(T) == 4)  // only FP32 rows are ever packed
      if (gl) off = gb + (cy == 1 ? uint32_t(sizeof(T)) : 0u) + (cz == 1 ? gpy : 0u);
    st_at<T>(base + slot * dsb, off, v);
    if constexpr (XG && cx != 0) {
      constexpr uint32_t qx = cx == -1 ? sizeof(T) : 0;
      if (cx == 1 ? (fx & 1u) : (fx >> 1)) st_at<T>(base + slot * dsb, r.o[cy == 1][cz == 1] + qx, v);
    }
    if constexpr (I < 26) run<I + 1>(f);
  }
};

// Ghost copies that involve a y or z face (clamped along y and/or z, with or
// without x).  The values are read back from the regular slots this cell has
// just written (its own stores, so the same bits), all loads first.  A c_x = +1
// value of the last cell of a row has no such consumer unless it is clamped
// along x as well (nx = 1): the others would sit at x = nx, outside the domain,
// so they are not written (its read-back value, from x = nx, goes unused).
template <class S, int P, bool XG = true>
struct TwistGhostYZ {
  using T = typename S::T;
  char* base;
  int64_t dsb;
  TwistRows r;
  uint32_t fx, fy, fz;  // bit 0: coordinate 0 on the axis; bit 1: coordinate n-1
  template <int I, int E>
  __device__ __forceinline__ void load_range(T* v) const {
    constexpr int cx = cx_of(I), cy = cy_of(I), cz = cz_of(I);
    constexpr int slot = P == 1 ? 26 - I : I;
    constexpr uint32_t ox = cx == 1 ? sizeof(T) : 0;
    if constexpr (cy != 0 || cz != 0) v[I] = ld_at<T>(base + slot * dsb, r.o[cy == 1][cz == 1] + ox);
    if constexpr (I + 1 < E) load_range<I + 1, E>(v);
  }
  template <int I, int E>
  __device__ __forceinline__ void store_range(const T* v) const {
    constexpr int cx = cx_of(I), cy = cy_of(I), cz = cz_of(I);
    if constexpr (cy != 0 || cz != 0) {
      constexpr int slot = P == 1 ? 26 - I : I;
      constexpr int32_t B = sizeof(T);
      const uint32_t gx = cx == 1 ? (fx & 1u) : (cx == -1 ? (fx >> 1) : 0u);
      const uint32_t gy = cy == 1 ? (fy & 1u) : (cy == -1 ? (fy >> 1) : 0u);
      const uint32_t gz = cz == 1 ? (fz & 1u) : (cz == -1 ? (fz >> 1) : 0u);
      const uint32_t gm = gx | (gy << 1) | (gz << 2);
      char* p = base + slot * dsb;
      const int32_t sx = cx == 1 ? B : 0, sy = cy == 1 ? r.row : 0, sz = cz == 1 ? r.plane : 0;
      const int32_t qx = cx == -1 ? B : 0, qy = cy == -1 ? r.row : 0, qz = cz == -1 ? r.plane : 0;
      const bool out_x = cx == 1 && (fx >> 1);  // consumer x + 1 = nx unless clamped along x
      for (uint32_t m = 2; m < 8; ++m)
        if ((m & ~gm) == 0 && (XG || !(m & 1u)) && !(out_x && !(m & 1u)))
          st_at<T>(p, r.lb + uint32_t(((m & 1u) ? qx : sx) + ((m & 2u) ? qy : sy) + ((m & 4u) ? qz : sz)),
                   v[I]);
    }
    if constexpr (I + 1 < E) store_range<I + 1, E>(v);
  }
};

template <class S, int P, bool XR = false>
__device__ __forceinline__ void twist_load_all(const StepArgs& a, const Cell& k, const void* buf,
                                               float* f) {
  TwistLoad<S, P, XR> t;
  constexpr uint32_t B = sizeof(typename S::T);
  t.xp = k.x == 0 ? B : 0u;
  t.xm = k.x == a.nx - 1 ? 0u : B;
  t.base = static_cast<const char*>(buf);
  t.dsb = a.dstride * int64_t(sizeof(typename S::T));
  t.r = twist_rows<sizeof(typename S::T)>(a, k);
  t.inv_s = a.inv_scale;
  t.nz = a.neg_zero;
  t.template run<0>(f);
}

template <class S, int P, bool SR = false, bool XG = true>
__device__ __forceinline__ void twist_store_all(const StepArgs& a, const Cell& k, void* buf,
                                                const float* f, const Encoder<S, SR>& e) {
  TwistStore<S, P, SR, XG> t;
  t.base = static_cast<char*>(buf);
  t.dsb = a.dstride * int64_t(sizeof(typename S::T));
  t.r = twist_rows<sizeof(typename S::T)>(a, k);
  t.e = e;
  t.fx = uint32_t(k.x == 0) | (uint32_t(k.x == a.nx - 1) << 1);
  constexpr uint32_t B = sizeof(typename S::T);
  t.gl = a.gcol != 0 && k.x == a.nx - 1;
  t.gb = (a.gcol + uint32_t(k.y) + uint32_t(a.py) * uint32_t(k.z)) * B;
  t.gpy = uint32_t(a.py) * B;
  t.template run<0>(f);
  const uint32_t fy = uint32_t(k.y == 0) | (uint32_t(k.y == a.ny - 1) << 1);
  const uint32_t fz = uint32_t(k.z == 0) | (uint32_t(k.z == a.nz - 1) << 1);
  if (fy | fz) {  // waves on a y or z face only
    TwistGhostYZ<S, P, XG> g;
    g.base = t.base;
    g.dsb = t.dsb;
    g.r = t.r;
    g.fx = t.fx;
    g.fy = fy;
    g.fz = fz;
    typename S::T v[27];
    // three batches of nine directions (one c_z each), loads first
    g.template load_range<0, 9>(v);
    g.template store_range<0, 9>(v);
    g.template load_range<9, 18>(v);
    g.template store_range<9, 18>(v);
    g.template load_range<18, 27>(v);
    g.template store_range<18, 27>(v);
  }
}

template <class S, int P, bool SR, bool DUMP = false, bool XR = false>
__global__ __launch_bounds__(kBlock) void k_twist(const StepArgs a) {
  const Cell k = locate(a, cell_block(a));
  if (!k.ok) return;
  // twisted loads stay inside the padded layout for every cell: issue them
  // before the flag arrives
  float f[27];
  twist_load_all<S, P, XR>(a, k, a.dst, f);
  EquIn q = equ_prefetch(a, k.c, k.x);
  if constexpr (DUMP) fused_dump(a, k.c, f, q);
  const uint32_t fl = a.sflag[k.c];
  update_cell(fl, a, k.c, q, f);
  if (fl & kStore) twist_store_all<S, P, SR, !XR>(a, k, a.dst, f, make_encoder<S, SR>(a, k.c, a.state));
}

// ---------------------------------------------------------------------------
// Two steps per launch with the intermediate populations kept in LDS (the
// mixed L1/L2 blocking of the reference, HALO_BLOCKING_L1L2,
// kernel.cu:138-329, redesigned for the 160 KB LDS of a gfx950 CU).
//
// A workgroup owns a TX x TY tile of cells and marches along z over a chunk
// of planes.  For each plane it computes the first step on the tile plus a
// one-cell xy halo ("A" cells, pulled from src in HBM) into a 3-plane LDS ring,
// then the second step of the tile ("B" cells, pulled from the ring) into dst.
// HBM sees the first step's reads (tile + halo) and the second step's writes
// only.  The intermediate is kept in the storage format, exactly as a plain
// step would have stored it, so two fused steps equal two plain steps bit for
// bit.  A cell without STORE keeps its populations in both buffers (see
// culbm_set_state), so its intermediate value is its src value.

// Tile of a workgroup per storage width: TX x TY B cells, (TX+2)(TY+2) A cells
// (the tile and a one-cell xy halo), one thread per A cell.
template <int BYTES>
struct FusedTile;
template <>
struct FusedTile<4> {
  static constexpr int TX = 32, TY = 16, BLOCK = 640;  // 612 A cells, ten waves
};
template <>
struct FusedTile<2> {
  static constexpr int TX = 96, TY = 8, BLOCK = 1024;  // 980 A cells, sixteen waves
};
template <int BYTES>
struct FusedShape : FusedTile<BYTES> {
  using Base = FusedTile<BYTES>;
  static constexpr int AX = Base::TX + 2, AY = Base::TY + 2, A = AX * AY;
  static_assert(A <= Base::BLOCK && Base::BLOCK % 64 == 0 && Base::BLOCK <= 1024, "one A cell per thread");
};

struct FusedArgs {
  StepArgs a;
  int32_t tiles_x, tiles_y;
  int32_t zchunk;  // B planes per workgroup
  int32_t nchunk;
};

// LDS of one workgroup.  A plane's first-step populations are grouped by c_z:
// the c_z = +1 group is read by the B step of the next plane, the c_z = 0 group
// by the same plane, the c_z = -1 group by the previous plane, so the groups
// live 3, 2 and 1 plane(s): 54 values per A cell instead of 81.
template <class T>
struct FusedRing {
  using F = FusedShape<sizeof(T)>;
  T up[3][9][F::A];  // directions 18..26 (c_z = +1), slot = plane % 3
  T mid[2][9][F::A]; // directions 9..17 (c_z = 0), slot = plane % 2
  T dn[9][F::A];     // directions 0..8 (c_z = -1), the latest plane
  uint8_t fl[2][F::A];  // step bits of the A cells, slot = plane % 2 (the B step's flags)
  float eq[2][2][F::TY][4];  // EQU inputs of the B cells on the low / high x face, slot = plane % 2
};

// Pulled (not yet updated) populations of one A cell, gathered a plane ahead
// so the HBM latency overlaps the previous plane's second step.
struct APending {
  float f[27];
  uint8_t fl;  // kept 8-bit: a widened copy across the loop would wait for every load
  uint32_t c;
  bool ok;
};

template <class S>
// Every lane gathers, at a position clamped into the domain (the caller clamps
// the plane), and only q.ok tells whether the cell is used: loads issued on
// every path let the compiler wait for exactly one gather while the next one
// is still in flight, instead of for every outstanding load.
__device__ __forceinline__ void fused_a_gather(const StepArgs& a, int32_t x0, int32_t y0, int ac,
                                               int32_t p, APending& q) {
  using F = FusedShape<sizeof(typename S::T)>;
  const int32_t gx = x0 - 1 + ac % F::AX, gy = y0 - 1 + ac / F::AX;
  q.ok = ac < F::A && gx >= 0 && gx < a.nx && gy >= 0 && gy < a.ny;
  Cell k;
  k.x = min(max(gx, 0), a.nx - 1);
  k.y = min(max(gy, 0), a.ny - 1);
  k.z = p;
  k.c = (uint32_t(p) * uint32_t(a.ny) + uint32_t(k.y)) * uint32_t(a.nx) + uint32_t(k.x);
  k.l = k.c;
  q.c = k.c;
  make_gather<S>(a, k, a.src).template run<0>(q.f);
  q.fl = a.sflag[k.c];
}

template <class S, bool SR, int I>
__device__ __forceinline__ void encode_all(const float* f, typename S::T* v, const Encoder<S, SR>& e) {
  v[I] = e.template get<I>(f[I]);
  if constexpr (I < 26) encode_all<S, SR, I + 1>(f, v, e);
}

// First step of a gathered A cell into the ring: a plain pull update (the state
// numbered a.state), or the cell's own populations when it has no STORE.
template <class S, bool SR>
__device__ __forceinline__ void fused_a_finish(const StepArgs& a, int ac, int32_t p, APending& q,
                                               FusedRing<typename S::T>& r) {
  using T = typename S::T;
  if (!q.ok) return;
  // An x-face EQU cell loads its inputs here, once its flag is known (carrying
  // them in the two pending sets would spill registers in FP32), and passes
  // them to the second step through LDS, where a load would wait for the
  // gathers in flight.
  EquIn qe;
  const uint32_t gx = q.c % uint32_t(a.nx);
  qe.ok = (q.fl & kEqu) && (gx == 0 || gx == uint32_t(a.nx - 1));
  qe.rho = qe.vx = qe.vy = qe.vz = 0.0f;
  if (qe.ok) {
    qe.rho = a.rho[q.c];
    qe.vx = a.vx[q.c];
    qe.vy = a.vy[q.c];
    qe.vz = a.vz[q.c];
  }
  update_cell(q.fl, a, q.c, qe, q.f);
  T v[27];
  if (q.fl & kStore) {
    encode_all<S, SR, 0>(q.f, v, make_encoder<S, SR>(a, q.c, a.state));
  } else {
    constexpr uint32_t B = sizeof(T);
    const char* src = static_cast<const char*>(a.src);
#pragma unroll
    for (int i = 0; i < 27; ++i) v[i] = ld_at<T>(src + i * (a.dstride * B), q.c * B);
  }
  const int su = p % 3, sm = p % 2;
  r.fl[sm][ac] = uint8_t(q.fl);
  // EQU inputs of the x-face cells among the B cells, for the second step
  const int ly = ac / FusedShape<sizeof(T)>::AX - 1;
  if (qe.ok && ly >= 0 && ly < FusedShape<sizeof(T)>::TY) {
    float* d = r.eq[sm][gx == 0 ? 0 : 1][ly];
    d[0] = qe.rho;
    d[1] = qe.vx;
    d[2] = qe.vy;
    d[3] = qe.vz;
  }
#pragma unroll
  for (int j = 0; j < 9; ++j) {
    r.dn[j][ac] = v[j];
    r.mid[sm][j][ac] = v[9 + j];
    r.up[su][j][ac] = v[18 + j];
  }
}

template <class S, bool SR>
// At most 168 VGPRs (three waves per SIMD), room for the ten waves of an FP32
// workgroup; the sixteen waves of a 16-bit workgroup bound it to 128.
__global__ __launch_bounds__(FusedShape<S::kBytes>::BLOCK) __attribute__((amdgpu_waves_per_eu(3))) void k_fused2(
    const FusedArgs fa) {
  using T = typename S::T;
  using F = FusedShape<sizeof(T)>;
  __shared__ FusedRing<T> r;
  const uint32_t ntile = uint32_t(fa.tiles_x) * fa.tiles_y * fa.nchunk;
  // runs of 16 tiles per XCD, tiles in x-major order: x neighbours, whose
  // halos share lines, mostly run on one XCD at the same time.  At 288x272x280,
  // against dispatching the x-face tile columns first: FP32 +2.6 % (runs of 8:
  // +2.3 %); against runs of 8: FP16 +0.4 %.
  const uint32_t w = xcd_group<16>(blockIdx.x, ntile);
  const int32_t tx = int32_t(w % uint32_t(fa.tiles_x));
  const int32_t ty = int32_t((w / uint32_t(fa.tiles_x)) % uint32_t(fa.tiles_y));
  const int32_t zc = int32_t(w / (uint32_t(fa.tiles_x) * fa.tiles_y));
  const int32_t x0 = tx * F::TX, y0 = ty * F::TY;
  const int32_t nx = fa.a.nx, ny = fa.a.ny, nz = fa.a.nz;
  const int32_t zs = zc * fa.zchunk, ze = min(nz, zs + fa.zchunk);
  const int t = threadIdx.x;
  const bool a_ok = t < F::A;

  // A planes gathered ahead of their first step: two for FP32 (the gather
  // latency then spans two planes' B steps), one for 16-bit storage
  constexpr int kAhead = sizeof(T) == 4 ? 2 : 1;
  APending q0, q1;
  q0.ok = false;
  q1.ok = false;
  // prologue: A planes zs-1 (if any) and zs, then the gathers of planes zs+1 (, zs+2)
  for (int32_t p = max(zs - 1, 0); p <= zs; ++p) {
    fused_a_gather<S>(fa.a, x0, y0, t, p, q0);
    if (a_ok) fused_a_finish<S, SR>(fa.a, t, p, q0, r);
    __syncthreads();  // the dn group of plane zs-1 is overwritten by plane zs
  }
  // planes past the chunk (or the domain) are gathered too, clamped, and unused
  fused_a_gather<S>(fa.a, x0, y0, t, min(zs + 1, nz - 1), q0);
  if (kAhead == 2) fused_a_gather<S>(fa.a, x0, y0, t, min(zs + 2, nz - 1), q1);

  const int bx = t % F::TX, by = t / F::TX;
  const int32_t gx = x0 + bx, gy = y0 + by;
  const bool b_ok = t < F::TX * F::TY && gx < nx && gy < ny;
  // local A coordinates of the (clamped) source columns
  const int lxm = gx == 0 ? bx + 1 : bx, lxp = gx == nx - 1 ? bx + 1 : bx + 2;
  const int lym = gy == 0 ? by + 1 : by, lyp = gy == ny - 1 ? by + 1 : by + 2;
  const uint32_t cxy = uint32_t(gy) * uint32_t(nx) + uint32_t(gx), pl = uint32_t(nx) * uint32_t(ny);
  const int bc = (by + 1) * F::AX + bx + 1;  // the B cell among the A cells
  // B plane z; q holds the gathered A plane z+1 and receives A plane z+1+kAhead.
  // The B step issues no global load: its flag byte comes from the A step's
  // copy in LDS, so no wait on a global load can hold back the A gathers
  // still in flight (the counter of vector memory operations is in order).
  auto plane = [&](int32_t z, APending& q) {
    if (a_ok && z + 1 < nz) fused_a_finish<S, SR>(fa.a, t, z + 1, q, r);
    fused_a_gather<S>(fa.a, x0, y0, t, min(z + 1 + kAhead, nz - 1), q);
    __syncthreads();
    if (b_ok) {
      // c_z = +1 pulls from plane z-1 (clamped to z), c_z = -1 from z+1 (the
      // dn group still holds plane z when z is the last plane)
      const int su = (z == 0 ? z : z - 1) % 3, sm = z % 2;
      float f[27];
#pragma unroll
      for (int i = 0; i < 27; ++i) {
        const int cx = cx_of(i), cy = cy_of(i), cz = cz_of(i);
        const int lx = cx == 1 ? lxm : (cx == -1 ? lxp : bx + 1);
        const int ly = cy == 1 ? lym : (cy == -1 ? lyp : by + 1);
        const int a = ly * F::AX + lx;
        const T v = cz == 1 ? r.up[su][i - 18][a] : (cz == 0 ? r.mid[sm][i - 9][a] : r.dn[i][a]);
        f[i] = S::dec(v, fa.a.inv_scale);
      }
      Cell k;
      k.x = gx;
      k.y = gy;
      k.z = z;
      k.c = uint32_t(z) * pl + cxy;
      k.l = k.c;
      const uint32_t fl = r.fl[z % 2][bc];
      EquIn e;
      e.ok = gx == 0 || gx == nx - 1;
      const float* q4 = r.eq[z % 2][gx == 0 ? 0 : 1][by];
      e.rho = q4[0];
      e.vx = q4[1];
      e.vy = q4[2];
      e.vz = q4[3];
      update_cell(fl, fa.a, k.c, e, f);
      if (fl & kStore)  // the second step produces state a.state + 1
        store_plain_all<S, SR>(fa.a, k, fa.a.dst, f, make_encoder<S, SR>(fa.a, k.c, fa.a.state + 1));
    }
    __syncthreads();
  };
  int32_t z = zs;
  if constexpr (kAhead == 2) {
    // both halves on every trip (an odd last plane follows the loop): the
    // compiler then counts the other set's gather as in flight at every wait
    for (; z + 1 < ze; z += 2) {
      plane(z, q0);
      plane(z + 1, q1);
    }
    if (z < ze) plane(z, q0);
  } else {
    for (; z < ze; ++z) plane(z, q0);
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
  const Cell k = locate(d.a, cell_block(d.a));
  if (!k.ok) return;
  float f[27];
  if constexpr (MODE == 0) make_gather<S>(d.a, k, d.a.src).template run<0>(f);
  // between steps the regular locations hold every value (ghost copies may be absent)
  if constexpr (MODE == 1) twist_load_all<S, 1, true>(d.a, k, d.a.src, f);
  if constexpr (MODE == 2) twist_load_all<S, 0, true>(d.a, k, d.a.src, f);
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
  const Cell k = locate(a, cell_block(a));
  if (!k.ok) return;
  const int64_t n = int64_t(a.nx) * a.ny * a.nz;
  float f[27];
#pragma unroll
  for (int i = 0; i < 27; ++i) f[i] = plain[i * n + k.c];
  store_plain_all<S>(a, k, out, f, make_encoder<S, false>(a, 0u, 0u));
}

template <class S>
__global__ __launch_bounds__(kBlock) void k_pull_to_plain(const StepArgs a, const void* in,
                                                          float* plain) {
  const Cell k = locate(a, cell_block(a));
  if (!k.ok) return;
  const int64_t n = int64_t(a.nx) * a.ny * a.nz;
  const typename S::T* b = static_cast<const typename S::T*>(in);
#pragma unroll
  for (int i = 0; i < 27; ++i) plain[i * n + k.c] = S::dec(b[i * a.dstride + k.l], a.inv_scale);
}

// plain FP32 -> in-place layout as left by a step of parity P (store rules of P);
// XG: with the x ghost copies (odd nx; for even nx the steps read x faces in place)
template <class S, int P, bool XG>
__global__ __launch_bounds__(kBlock) void k_plain_to_twist(const StepArgs a, const float* plain,
                                                           void* out) {
  const Cell k = locate(a, cell_block(a));
  if (!k.ok) return;
  const int64_t n = int64_t(a.nx) * a.ny * a.nz;
  float f[27];
#pragma unroll
  for (int i = 0; i < 27; ++i) f[i] = plain[i * n + k.c];
  twist_store_all<S, P, false, XG>(a, k, out, f, make_encoder<S, false>(a, 0u, 0u));
}

// Element index of population i of cell k in the in-place layout left by a step
// of parity P (its regular store slot; the ghost column for a c_x = +1 value of
// the last cell of a row when the rows are packed).
template <int P>
__device__ __forceinline__ int64_t twist_stored_index(const StepArgs& a, const Cell& k, int i) {
  const int slot = P == 1 ? 26 - i : i;
  const uint32_t oy = cy_of(i) == 1 ? 1u : 0u, oz = cz_of(i) == 1 ? 1u : 0u;
  if (cx_of(i) == 1 && a.gcol != 0 && k.x == a.nx - 1)
    return slot * a.dstride + int64_t(a.gcol + (uint32_t(k.y) + oy) + uint32_t(a.py) * (uint32_t(k.z) + oz));
  const uint32_t o = (cx_of(i) == 1 ? 1u : 0u) + oy * uint32_t(a.px) + oz * uint32_t(a.px) * uint32_t(a.py);
  return slot * a.dstride + int64_t(k.l + o);
}

// in-place layout left by a step of parity P -> plain FP32 (regular slots)
template <class S, int P>
__global__ __launch_bounds__(kBlock) void k_twist_to_plain(const StepArgs a, const void* in,
                                                           float* plain) {
  const Cell k = locate(a, cell_block(a));
  if (!k.ok) return;
  const int64_t n = int64_t(a.nx) * a.ny * a.nz;
  const typename S::T* b = static_cast<const typename S::T*>(in);
#pragma unroll
  for (int i = 0; i < 27; ++i) plain[i * n + k.c] = S::dec(b[twist_stored_index<P>(a, k, i)], a.inv_scale);
}

// Stored populations of cell k (the post-collision view of culbm_get_ddf):
// MODE 0 pull layout, MODE 1 / 2 in-place layout left by a step of parity 1 / 0.
template <class S, int MODE>
__device__ __forceinline__ void stored_cell(const StepArgs& a, const Cell& k, const void* buf, float* f) {
  const typename S::T* b = static_cast<const typename S::T*>(buf);
#pragma unroll
  for (int i = 0; i < 27; ++i) {
    if constexpr (MODE == 0) {
      f[i] = S::dec(b[i * a.dstride + k.l], a.inv_scale);
    } else {
      constexpr int P = MODE == 1 ? 1 : 0;
      f[i] = S::dec(b[twist_stored_index<P>(a, k, i)], a.inv_scale);
    }
  }
}

// FP64 sums of rho = 1 + sum f and of the momentum sum c f over the cells whose
// flag shares a bit with `mask` (every cell for mask 0).  A bounded grid walks
// the cell blocks (workgroup g takes blocks g, g + G, ...), so the partial sums
// number gridDim.x whatever the domain; k_sum_parts then adds them in a fixed
// order.  One FP64 partial of each sum per workgroup.
template <class S, int MODE>
__global__ __launch_bounds__(kBlock) void k_moment_sums(const StepArgs a, const void* buf,
                                                        const uint32_t* flag, uint32_t mask,
                                                        double* part) {
  double m[4] = {0.0, 0.0, 0.0, 0.0};
  for (uint32_t blk = blockIdx.x; blk < a.nblk; blk += gridDim.x) {
    const Cell k = locate(a, blk);
    if (k.ok && (mask == 0u || (flag[k.c] & mask))) {
      float f[27];
      stored_cell<S, MODE>(a, k, buf, f);
      double rho = 1.0, jx = 0.0, jy = 0.0, jz = 0.0;
#pragma unroll
      for (int i = 0; i < 27; ++i) {
        const double v = f[i];
        rho += v;
        jx += cx_of(i) * v;
        jy += cy_of(i) * v;
        jz += cz_of(i) * v;
      }
      m[0] += rho;
      m[1] += jx;
      m[2] += jy;
      m[3] += jz;
    }
  }
  __shared__ double red[4][kBlock / 64];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    double v = m[q];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if ((threadIdx.x & 63) == 0) red[q][threadIdx.x >> 6] = v;
  }
  __syncthreads();
  if (threadIdx.x < 4) {
    double s = 0.0;
#pragma unroll
    for (int w = 0; w < kBlock / 64; ++w) s += red[threadIdx.x][w];
    part[4 * blockIdx.x + threadIdx.x] = s;
  }
}

// The four totals of `nparts` partial sums, added in index order.
__global__ __launch_bounds__(64) void k_sum_parts(const double* part, uint32_t nparts, double* out) {
  if (threadIdx.x < 4) {
    double s = 0.0;
    for (uint32_t b = 0; b < nparts; ++b) s += part[4 * b + threadIdx.x];
    out[threadIdx.x] = s;
  }
}

// Equilibrium of the rho/u fields as post-collision state (plain FP32).
__global__ __launch_bounds__(kBlock) void k_feq_plain(const StepArgs a, float* plain) {
  const Cell k = locate(a, cell_block(a));
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

// Cells without STORE never change; keep their two pull buffers identical
// (copy the current buffer into the other) so that any order of buffer reads
// sees the same value.
template <class S>
__global__ __launch_bounds__(kBlock) void k_sync_nostore(const StepArgs a, const void* cur,
                                                         void* other) {
  const Cell k = locate(a, cell_block(a));
  if (!k.ok || (a.sflag[k.c] & kStore)) return;
  constexpr uint32_t B = sizeof(typename S::T);
  const char* s = static_cast<const char*>(cur);
  char* d = static_cast<char*>(other);
  const int64_t dsb = a.dstride * B;
#pragma unroll
  for (int i = 0; i < 27; ++i)
    st_at<typename S::T>(d + i * dsb, k.l * B, ld_at<typename S::T>(s + i * dsb, k.l * B));
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

// One element per thread: a grid-stride loop keeps one load per wave in flight
// and measures the latency of the loop, not the memory (≈ 4.5 instead of
// 6.1 TB/s for a 2 GiB copy on MI355X).
__global__ __launch_bounds__(kBlock) void k_bw_copy(const float4* __restrict__ in,
                                                    float4* __restrict__ out, uint64_t n) {
  const uint64_t i = uint64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (i < n) out[i] = in[i];
}

// Read probe: each workgroup writes the XOR of the words it read (one word per
// 4 KiB read, so the loads stay live and the check can recompute it).
__global__ __launch_bounds__(kBlock) void k_bw_read(const uint4* __restrict__ in, uint32_t* out,
                                                    uint64_t n) {
  const uint64_t i = uint64_t(blockIdx.x) * kBlock + threadIdx.x;
  uint32_t acc = 0u;
  if (i < n) {
    const uint4 v = in[i];
    acc = (v.x ^ v.y) ^ (v.z ^ v.w);
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) acc ^= __shfl_xor(acc, off, 64);
  __shared__ uint32_t red[kBlock / 64];
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) out[blockIdx.x] = (red[0] ^ red[1]) ^ (red[2] ^ red[3]);
}

__global__ __launch_bounds__(kBlock) void k_bw_write(float4* __restrict__ out, uint64_t n) {
  const uint64_t i = uint64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (i < n) out[i] = make_float4(0.f, 0.f, 0.f, 0.f);
}

// Scalar 4-byte copy shifted by `shift` elements: shift 0 is aligned, shift 1
// makes every wave's store straddle one extra cache line (partial-line stores).
__global__ __launch_bounds__(kBlock) void k_bw_copy_shift(const float* __restrict__ in,
                                                          float* __restrict__ out, uint64_t n,
                                                          uint32_t shift) {
  const uint64_t i = uint64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (i < n) out[i + shift] = in[i];
}

// The population access pattern without the physics: 27 arrays of `cells`
// floats, one cell per thread, every value read once and written once; with
// `shift` the 18 directions with c_x != 0 read the neighbour at x -+ 1 (the
// pull gather's misaligned loads).
__global__ __launch_bounds__(kBlock) void k_bw_pop27(const float* __restrict__ in,
                                                     float* __restrict__ out, uint64_t cells,
                                                     uint32_t shift) {
  const uint64_t c = uint64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (c >= cells) return;
  float f[27];
#pragma unroll
  for (int i = 0; i < 27; ++i) {
    const int cx = cx_of(i);
    uint64_t cc = c;
    if (shift && cx < 0) cc = c + 1 < cells ? c + 1 : c;
    if (shift && cx > 0) cc = c > 0 ? c - 1 : c;
    f[i] = in[uint64_t(i) * cells + cc];
  }
#pragma unroll
  for (int i = 0; i < 27; ++i) out[uint64_t(i) * cells + c] = f[i];
}

// Data checks of the probes: the buffers start from known words, and after the
// timed launches every word a probe claims to have moved must hold what it
// should.  Word pattern of the source: a hash of the word index.
__device__ __forceinline__ uint32_t probe_word(uint64_t i, uint32_t seed) {
  return mix32(uint32_t(i) ^ mix32(uint32_t(i >> 32) + seed)) | 1u;  // never zero
}

__global__ __launch_bounds__(kBlock) void k_bw_fill(uint32_t* p, uint64_t nwords, uint32_t seed) {
  const uint64_t i = uint64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (i < nwords) p[i] = probe_word(i, seed);
}

// Counts the words of `dst` that differ from what probe `mode` leaves there
// (`src` filled with seed `seed`); grid-stride, one atomic per mismatching wave.
// `wrong` != 0 checks against a deliberately wrong expectation (the copy one
// word off), so a test can show the check catches a probe that moved the
// wrong data.
__global__ __launch_bounds__(kBlock) void k_bw_verify(int mode, const uint32_t* src, const uint32_t* dst,
                                                      uint64_t nwords, uint64_t cells, uint64_t nread_blk,
                                                      uint32_t seed, uint32_t wrong, uint32_t* bad) {
  const uint64_t stride = uint64_t(gridDim.x) * kBlock;
  const uint64_t total = mode == 1 ? nread_blk : (mode >= 5 ? 27 * cells : nwords);
  uint32_t miss = 0;
  for (uint64_t i = uint64_t(blockIdx.x) * kBlock + threadIdx.x; i < total; i += stride) {
    uint32_t want;
    uint64_t at = i;
    if (mode == 1) {  // XOR of the 1024 words block i read
      want = 0u;
      for (uint64_t w = 0; w < uint64_t(4 * kBlock); ++w) {
        const uint64_t k = i * 4 * kBlock + w;
        if (k < nwords) want ^= probe_word(k, seed);
      }
    } else if (mode == 2) {
      want = 0u;
    } else if (mode >= 5) {
      const uint64_t d = i / cells, c = i - d * cells;
      const int cx = int(d % 3) - 1;
      uint64_t cc = c;
      if (mode == 6 && cx < 0) cc = c + 1 < cells ? c + 1 : c;
      if (mode == 6 && cx > 0) cc = c > 0 ? c - 1 : c;
      want = probe_word(d * cells + cc, seed);
    } else {
      if (mode == 4) at = i + 1;  // shifted by one element
      want = probe_word(i, seed);
    }
    if (wrong) at += 1;
    miss += dst[at] != want ? 1u : 0u;
  }
  const uint64_t ballot = __ballot(miss != 0u);
  uint32_t sum = miss;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) sum += __shfl_xor(sum, off, 64);
  if (ballot && (threadIdx.x & 63) == 0) atomicAdd(bad, sum);
}

}  // namespace culbm
