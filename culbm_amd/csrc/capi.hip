// Host side of the C ABI declared in include/culbm.h: context, buffers,
// policy dispatch and the temporal-blocking wavefront schedule.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/culbm.h"
#include "blocking.hpp"
#include "lbm_kernels.hip"
#include "adjoint.hip"

using namespace culbm;

struct culbm_ctx {
  culbm_config cfg{};
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  std::string err;

  int64_t n = 0;           // cells
  int32_t px = 0, py = 0, pz = 0;
  int64_t gcol = 0;        // in-place, rows packed: first ghost-column element per direction
  int64_t dstride = 0;     // population elements per direction
  int32_t bpp = 4;         // bytes per population
  bool inplace = false;
  bool sr = false;         // stochastic rounding of the 16-bit stores
  void* buf[2] = {nullptr, nullptr};
  int cur = 0;             // pull: index of the buffer holding the current state
  int parity = 1;          // in-place: parity of the next step (1 = O, 0 = E)
  uint32_t* flag = nullptr;
  uint8_t* sflag = nullptr;
  float* rho = nullptr;
  float* vx = nullptr;
  float* vy = nullptr;
  float* vz = nullptr;
  int64_t step = 0;
  float scale = 1.0f, inv_scale = 1.0f;
  double state_bound = 0.0;  // feq_bound over the cells of the last culbm_set_state
  int32_t wave_depth = 0;
  // launch shape of k_fused2: tile columns and rows, z planes per workgroup and
  // z chunks (the mixed policy; the adjoint's forward sweeps on every context)
  int32_t tiles_x = 0, tiles_y = 0, fzchunk = 0, nchunk = 0;
  // static policy: one instantiated graph per starting state of a full pass
  // (pull: index of the current buffer; in-place: parity of the next step)
  hipGraphExec_t pass_graph[2] = {nullptr, nullptr};
  int64_t device_bytes = 0;
  // a dump armed by culbm_dump_deferred: the next step writes it (fused), or
  // the next call that reads or replaces the fields or populations runs k_dump
  bool dump_pending = false;
  // per-launch timing (culbm_profile_steps): an event pair around every launch
  bool prof = false;
  size_t prof_used = 0;
  std::vector<hipEvent_t> prof_ev;
};

namespace {

thread_local std::string g_create_err = "no error";

int fail(culbm_ctx* ctx, int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  if (ctx)
    ctx->err = buf;
  else
    g_create_err = buf;
  return code;
}

#define HIPCHK(ctx, expr)                                                                   \
  do {                                                                                      \
    hipError_t e_ = (expr);                                                                 \
    if (e_ != hipSuccess)                                                                   \
      return fail((ctx), CULBM_ERR_HIP, "%s failed: %s", #expr, hipGetErrorString(e_));     \
  } while (0)

int64_t round_up(int64_t v, int64_t m) { return (v + m - 1) / m * m; }

bool is_pow2_scale(float s) {
  if (!(s > 0.0f) || !std::isfinite(s)) return false;
  int e;
  const float m = std::frexp(s, &e);
  return m == 0.5f;
}

bool blocked(const culbm_config& c) { return c.opt_policy != CULBM_OPT_NONE; }

// Scale of the 16-bit formats, derived per run (storage_scale 0).  FP16 keeps
// (f_i - w_i) * s in binary16, finite up to 65504 and normal from 2^-14: the
// scale is the largest power of two that keeps kScaleHeadroom times the run's
// bound on |f_i - w_i| finite, so that the small deviations stay as far from
// the subnormals as the range allows.  A bound below kBoundFloor (a state at
// rest) counts as kBoundFloor, which leaves room for a flow of |u| ~ 0.1 that
// a later culbm_set_state may bring without re-initialising the populations.
// BF16 has FP32's exponent range and takes 1.
constexpr double kScaleHeadroom = 16.0;
constexpr double kBoundFloor = 1.0 / 32.0;
constexpr int kMaxScaleExp = 20;

float derived_scale(int32_t storage, double bound) {
  if (storage != CULBM_STORAGE_FP16) return 1.0f;
  const double q = 65504.0 / (kScaleHeadroom * std::max(bound, kBoundFloor));
  int e;
  (void)std::frexp(q, &e);  // q = m 2^e with m in [0.5, 1): the largest power of two <= q is 2^(e-1)
  return std::ldexp(1.0f, std::min(std::max(e - 1, 0), kMaxScaleExp));
}

// Bound on |feq_i(rho, u) - w_i| over the 27 directions of a cell:
// w_i (|rho - 1| + rho (3 |c.u| + 4.5 (c.u)^2 + 1.5 u^2)) with w_i <= 8/27 and
// |c.u| <= |ux| + |uy| + |uz|.
double feq_bound(float rho, float ux, float uy, float uz) {
  const double r = rho, s = std::fabs(double(ux)) + std::fabs(double(uy)) + std::fabs(double(uz));
  const double u2 = double(ux) * ux + double(uy) * uy + double(uz) * uz;
  return (8.0 / 27.0) * (std::fabs(r - 1.0) + std::fabs(r) * (3.0 * s + 4.5 * s * s + 1.5 * u2));
}

// FP16 with storage_scale 0: the scale follows the run (BF16's is 1 either way)
bool auto_scale(const culbm_ctx* ctx) {
  return ctx->cfg.storage == CULBM_STORAGE_FP16 && ctx->cfg.storage_scale == 0.0f;
}

void set_scale(culbm_ctx* ctx, float s) {
  if (s == ctx->scale) return;
  ctx->scale = s;
  ctx->inv_scale = 1.0f / s;
  // captured passes carry the old scale in their kernel arguments
  (void)hipStreamSynchronize(ctx->stream);
  for (hipGraphExec_t& g : ctx->pass_graph)
    if (g) {
      (void)hipGraphExecDestroy(g);
      g = nullptr;
    }
}

// Population layout of a configuration: pitches and direction stride.  The one
// place that defines it, so the 32-bit offset check of culbm_validate tests
// exactly what culbm_create allocates.
int32_t bytes_per_pop(int32_t storage) {
  return (storage == CULBM_STORAGE_FP16 || storage == CULBM_STORAGE_BF16) ? 2 : 4;
}

struct Layout {
  int32_t px, py, pz;
  int64_t gcol;     // packed rows: first element of the ghost column region (0: padded rows)
  int64_t dstride;  // elements per direction array
};

Layout layout_of(const culbm_config& c) {
  Layout L{};
  if (c.stream_policy == CULBM_STREAM_INPLACE) {
    // Row pitch nx + 1 rounded up to 16 elements (64 B in FP32, 32 B in 16-bit),
    // not to a 128-byte line: with whole-line rows the ghost column of a row
    // whose nx + 1 just passes a line boundary (nx = 288: 289 of 320 elements)
    // sits alone in a line written 4 bytes at a time.  Measured against the
    // 128-byte pitch: FP32 in place + wavefront +2.7 % at 288x272x280, +6.6 %
    // at 289x272x280, +8.7 % at 384^3 (-3 % at 512^3); FP16 +1 to +7 %; FP32
    // in place without blocking +7 to +16 %.  (An unpadded pitch, nx + 1,
    // helped the unblocked step most but cost 10 % at 289x272x280 blocked.)
    // Without blocking, 8 elements measured faster still: FP32 +3 to +13 %
    // (21.1k vs 18.8k MLUPS at 288x272x280), 16-bit +1.5 to +3 %.
    const int64_t align = c.opt_policy == CULBM_OPT_NONE ? 8 : 16;
    L.px = int32_t(round_up(int64_t(c.nx) + 1, align));
    L.py = c.ny + 1;
    L.pz = c.nz + 1;
    // FP32 with even nx: rows packed at pitch nx, the ghost column (the c_x = +1
    // stores of x = nx-1) in a (ny+1) x (nz+1) region after the planes.  Even nx
    // reads x faces in place, so no load ever reaches x = nx.  Measured against
    // the padded pitch, FP32 in place + wavefront, two runs per setting in
    // alternating order: +2-4 % at 288x272x280, +14-19 % at 128^3, +5 % at
    // 192^3, +4 % at 384^3, +1 % at 512^3, +-0.5 % at 256^3; unblocked +6.5 % at
    // 288x272x280.  16-bit storage, whose 576-byte rows then share lines, lost
    // 4-6 % and keeps padded rows.
    if (bytes_per_pop(c.storage) == 4 && c.nx % 2 == 0) {
      L.px = c.nx;
      L.gcol = int64_t(L.px) * L.py * L.pz;
    }
  } else {
    L.px = c.nx;
    L.py = c.ny;
    L.pz = c.nz;
  }
  L.dstride = round_up(int64_t(L.px) * L.py * L.pz + (L.gcol ? int64_t(L.py) * L.pz : 0), 256);
  return L;
}

}  // namespace

extern "C" {

void culbm_config_default(culbm_config* c) {
  if (!c) return;
  std::memset(c, 0, sizeof(*c));
  c->nx = c->ny = c->nz = 256;
  c->inv_tau = 0.5f;
  c->stream_policy = CULBM_STREAM_PULL;
  c->opt_policy = CULBM_OPT_NONE;
  c->inner_loop = 5;
  c->block_dim[0] = 32;
  c->block_dim[1] = 16;
  c->block_dim[2] = 2;
  c->grid_dim[0] = 2;
  c->grid_dim[1] = 3;
  c->grid_dim[2] = 19;
  c->use_ring = 0;
  c->device = 0;
  c->storage = CULBM_STORAGE_FP32;
  c->storage_scale = 0.0f;
  c->wave_depth = 0;
}

int culbm_validate(const culbm_config* c, char* msg, size_t msg_len) {
  auto bad = [&](int code, const char* fmt, ...) {
    if (msg && msg_len) {
      va_list ap;
      va_start(ap, fmt);
      vsnprintf(msg, msg_len, fmt, ap);
      va_end(ap);
    }
    return code;
  };
  if (!c) return bad(CULBM_ERR_INVALID, "config is NULL");
  if (c->nx < 1 || c->ny < 1 || c->nz < 1)
    return bad(CULBM_ERR_INVALID, "domain dimensions must be >= 1 (got %d x %d x %d)", c->nx, c->ny,
               c->nz);
  const int64_t n = int64_t(c->nx) * c->ny * c->nz;
  if (n > (int64_t(1) << 31) - 1)
    return bad(CULBM_ERR_INVALID, "domain of %lld cells exceeds the 32-bit cell index", (long long)n);
  {
    // the kernels address one direction array with 32-bit per-lane byte offsets
    // (cell index x + px * (y + py * z) of the layout culbm_create allocates)
    const Layout L = layout_of(*c);
    const int64_t cells = int64_t(L.px) * L.py * L.pz + (L.gcol ? int64_t(L.py) * L.pz : 0);
    const int64_t per_dir = cells * bytes_per_pop(c->storage);
    if (cells > (int64_t(1) << 32) - 1 || per_dir > (int64_t(1) << 32))  // the last element starts at per_dir - bpp
      return bad(CULBM_ERR_INVALID,
                 "domain too large: one direction array takes %lld bytes, above the 4 GiB the "
                 "kernels' 32-bit offsets reach (use 16-bit storage or a smaller domain)",
                 (long long)per_dir);
  }
  if (!std::isfinite(c->inv_tau)) return bad(CULBM_ERR_INVALID, "invTau must be finite");
  if (c->stream_policy != CULBM_STREAM_PULL && c->stream_policy != CULBM_STREAM_INPLACE)
    return bad(CULBM_ERR_INVALID, "streamPolicy must be 0 (pull) or 1 (in-place), got %d",
               c->stream_policy);
  if (c->opt_policy < CULBM_OPT_NONE || c->opt_policy > CULBM_OPT_DYNAMIC)
    return bad(CULBM_ERR_INVALID, "optPolicy must be in 0..3, got %d", c->opt_policy);
  if (c->opt_policy == CULBM_OPT_MIXED && c->stream_policy != CULBM_STREAM_PULL)
    return bad(CULBM_ERR_UNSUPPORTED,
               "optPolicy 2 (HALO_BLOCKING_L1L2) runs with --streamPolicy 0 only");
  // same rule as the reference (simulator_expt_platform.cu:365-368)
  if (c->use_ring &&
      (c->stream_policy != CULBM_STREAM_INPLACE || c->opt_policy != CULBM_OPT_DYNAMIC))
    return bad(CULBM_ERR_INVALID,
               "--useRingDDFBuf is unimplemented unless --streamPolicy 1 --optPolicy 3.");
  if (blocked(*c) && c->inner_loop < 1)
    return bad(CULBM_ERR_INVALID, "innerLoop must be >= 1, got %d", c->inner_loop);
  if (blocked(*c)) {
    // the reference's tile check on blockDim*gridDim (simulator_expt_platform.cu:523-534),
    // kept so that a command line it rejects is rejected here too
    const int32_t dom[3] = {c->nx, c->ny, c->nz};
    for (int k = 0; k < 3; ++k) {
      if (c->block_dim[k] < 1 || c->grid_dim[k] < 1)
        return bad(CULBM_ERR_INVALID, "blockDim and gridDim entries must be >= 1");
      const uint64_t tile = uint64_t(c->block_dim[k]) * uint64_t(c->grid_dim[k]);
      if (tile > 0xffffffffull ||
          !blk_axis_valid_ref(uint32_t(dom[k]), uint32_t(tile), uint32_t(c->inner_loop)))
        return bad(CULBM_ERR_INVALID, "<Domain Dimension, Blocking Dimension and InnerLoop> don't match!");
    }
  }
  if (c->storage < CULBM_STORAGE_FP32 || c->storage > CULBM_STORAGE_BF16)
    return bad(CULBM_ERR_INVALID, "storage must be 0 (fp32), 1 (fp16) or 2 (bf16), got %d", c->storage);
  if (c->storage == CULBM_STORAGE_FP32 && c->storage_scale != 0.0f && c->storage_scale != 1.0f)
    return bad(CULBM_ERR_INVALID, "FP32 storage takes no scale");
  if (c->storage != CULBM_STORAGE_FP32 && c->storage_scale != 0.0f && !is_pow2_scale(c->storage_scale))
    return bad(CULBM_ERR_INVALID, "storage scale must be a positive power of two, got %g",
               double(c->storage_scale));
  if (c->wave_depth < 0) return bad(CULBM_ERR_INVALID, "wave depth must be >= 0");
  if (c->rounding != CULBM_ROUND_NEAREST && c->rounding != CULBM_ROUND_STOCHASTIC)
    return bad(CULBM_ERR_INVALID, "rounding must be 0 (nearest) or 1 (stochastic), got %d", c->rounding);
  if (c->rounding == CULBM_ROUND_STOCHASTIC && c->storage == CULBM_STORAGE_FP32)
    return bad(CULBM_ERR_INVALID, "stochastic rounding applies to FP16 or BF16 storage");
  if (c->device < 0) return bad(CULBM_ERR_INVALID, "devId must be >= 0");
  if (msg && msg_len) msg[0] = 0;
  return CULBM_OK;
}

const char* culbm_error_string(int code) {
  switch (code) {
    case CULBM_OK: return "CULBM_OK";
    case CULBM_ERR_INVALID: return "CULBM_ERR_INVALID";
    case CULBM_ERR_HIP: return "CULBM_ERR_HIP";
    case CULBM_ERR_STATE: return "CULBM_ERR_STATE";
    case CULBM_ERR_UNSUPPORTED: return "CULBM_ERR_UNSUPPORTED";
    default: return "CULBM_ERR_UNKNOWN";
  }
}

const char* culbm_last_error(const culbm_ctx* ctx) {
  return ctx ? ctx->err.c_str() : g_create_err.c_str();
}

}  // extern "C"

namespace {

StepArgs base_args(const culbm_ctx* ctx) {
  StepArgs a{};
  a.nx = ctx->cfg.nx;
  a.ny = ctx->cfg.ny;
  a.nz = ctx->cfg.nz;
  a.px = ctx->px;
  a.py = ctx->py;
  a.gcol = uint32_t(ctx->gcol);
  a.z0 = 0;
  a.z1 = ctx->cfg.nz;
  a.ncells = uint32_t(ctx->n);
  a.dstride = ctx->dstride;
  a.divx = FastDiv::make(uint32_t(ctx->cfg.nx));
  a.divy = FastDiv::make(uint32_t(ctx->cfg.ny));
  a.sflag = ctx->sflag;
  a.rho = ctx->rho;
  a.vx = ctx->vx;
  a.vy = ctx->vy;
  a.vz = ctx->vz;
  a.inv_tau = ctx->cfg.inv_tau;
  a.omt = 1.0f - ctx->cfg.inv_tau;
  a.scale = ctx->scale;
  a.inv_scale = ctx->inv_scale;
  a.neg_zero = -0.0f;
  a.flag = ctx->flag;
  a.mrho = ctx->rho;
  a.mvx = ctx->vx;
  a.mvy = ctx->vy;
  a.mvz = ctx->vz;
  return a;
}

// Workgroup order of a launch over planes [z0, z1) (see cell_block), measured
// per form from 128^3 to 384^3.  Runs of 8 workgroups per XCD, dealt round
// robin, keep the 8 XCDs on one streaming front; they beat XCD-contiguous slabs
// for every wavefront chunk (288x272x280 at the chosen depths: FP32 in place
// 26.1k vs 21.3k MLUPS) and whole-domain step, and plain dispatch order in
// place (+0.2 to +8 %) and for FP32 pull (+0.5 to +5.6 %).  16-bit pull stays in
// dispatch order: there runs gain 1-2 % on most sizes but lose 13 % on the
// 128^3 wavefront and 4 % at 256^3.  Slabs stay on whole-domain gather launches
// (pull, adjoint) below 2^22 cells, where each XCD's L2 serving its own slab's
// neighbour planes wins (FP32 pull at 128^3: +5 % over dispatch order).
constexpr uint32_t kRemapGatherCells = 1u << 22;

void set_range(StepArgs& a, int32_t z0, int32_t z1, bool gather, bool narrow) {
  a.z0 = z0;
  a.z1 = z1;
  a.ncells = uint32_t(z1 - z0) * uint32_t(a.nx) * uint32_t(a.ny);
  a.nblk = (a.ncells + kBlock - 1) / kBlock;
  const bool whole = z0 == 0 && z1 == a.nz;
  a.order = (whole && gather && a.ncells < kRemapGatherCells) ? kOrderSlabs
          : (gather && narrow)                              ? kOrderDispatch
                                                            : kOrderRuns;
}

// Event pair around one launch when profiling is on.
int prof_begin(culbm_ctx* ctx) {
  if (!ctx->prof) return CULBM_OK;
  while (ctx->prof_ev.size() < 2 * (ctx->prof_used + 1)) {
    hipEvent_t e;
    HIPCHK(ctx, hipEventCreate(&e));
    ctx->prof_ev.push_back(e);
  }
  HIPCHK(ctx, hipEventRecord(ctx->prof_ev[2 * ctx->prof_used], ctx->stream));
  return CULBM_OK;
}

int prof_end(culbm_ctx* ctx) {
  if (!ctx->prof) return CULBM_OK;
  HIPCHK(ctx, hipEventRecord(ctx->prof_ev[2 * ctx->prof_used + 1], ctx->stream));
  ++ctx->prof_used;
  return CULBM_OK;
}

// One step over planes [z0, z1); a.state is the number of the state it produces.
// DUMP: the launch also writes the pending dump of the planes it covers.
template <class S, bool SR, bool DUMP>
void launch_step_kernel(const culbm_ctx* ctx, const StepArgs& a, int parity) {
  if (ctx->inplace) {
    // even nx: x-face consumers read in place (no x ghost copies); rows then
    // start at even lanes, which keeps x = 0 / 1 and nx-2 / nx-1 in one wave
    if (ctx->cfg.nx % 2 == 0) {
      if (parity)
        hipLaunchKernelGGL((k_twist<S, 1, SR, DUMP, true>), dim3(a.nblk), dim3(kBlock), 0, ctx->stream, a);
      else
        hipLaunchKernelGGL((k_twist<S, 0, SR, DUMP, true>), dim3(a.nblk), dim3(kBlock), 0, ctx->stream, a);
      return;
    }
    if (parity)
      hipLaunchKernelGGL((k_twist<S, 1, SR, DUMP>), dim3(a.nblk), dim3(kBlock), 0, ctx->stream, a);
    else
      hipLaunchKernelGGL((k_twist<S, 0, SR, DUMP>), dim3(a.nblk), dim3(kBlock), 0, ctx->stream, a);
  } else {
    hipLaunchKernelGGL((k_pull<S, SR, DUMP>), dim3(a.nblk), dim3(kBlock), 0, ctx->stream, a);
  }
}

template <class S, bool SR>
void launch_step_kernel(const culbm_ctx* ctx, const StepArgs& a, int parity, bool dump) {
  if (dump)
    launch_step_kernel<S, SR, true>(ctx, a, parity);
  else
    launch_step_kernel<S, SR, false>(ctx, a, parity);
}

template <class S>
int launch_step(culbm_ctx* ctx, StepArgs a, int32_t z0, int32_t z1, int parity, bool dump = false) {
  set_range(a, z0, z1, !ctx->inplace, ctx->bpp == 2);
  if (a.ncells == 0) return CULBM_OK;
  if (int rc = prof_begin(ctx)) return rc;
  if (ctx->inplace) {
    a.src = ctx->buf[0];
    a.dst = ctx->buf[0];
  }
  if constexpr (S::kBytes == 2) {
    if (ctx->sr)
      launch_step_kernel<S, true>(ctx, a, parity, dump);
    else
      launch_step_kernel<S, false>(ctx, a, parity, dump);
  } else {
    launch_step_kernel<S, false>(ctx, a, parity, dump);
  }
  HIPCHK(ctx, hipGetLastError());
  return prof_end(ctx);
}

// The launches of one temporal-blocking pass of T steps as a skewed
// z-wavefront, in issue order: f(lo, hi, t) for step t (1-based) over planes
// [lo, hi).  Chunk k at step t covers planes [kD - (t-1), (k+1)D - (t-1)):
// every plane it reads at level t-1 was produced either by an earlier chunk or
// by the same chunk one launch earlier, and no (plane, step) is computed twice.
// A chunk's working set (D + T planes) is sized for the Infinity Cache.  The
// schedule is exported (culbm_wave_schedule) for the host-side hazard check.
template <class F>
int for_each_wave_launch(int32_t nz, int32_t D, int32_t T, F&& f) {
  const int32_t K = (nz + T - 1 + D - 1) / D;
  for (int32_t k = 0; k < K; ++k) {
    for (int32_t t = 1; t <= T; ++t) {
      const int32_t lo = std::max(0, k * D - (t - 1));
      const int32_t hi = std::min(nz, (k + 1) * D - (t - 1));
      if (lo >= hi) continue;
      if (int rc = f(lo, hi, t)) return rc;
    }
  }
  return CULBM_OK;
}

template <class S>
int wave_pass(culbm_ctx* ctx, int32_t T) {
  StepArgs a = base_args(ctx);
  const uint32_t base = uint32_t(ctx->step);
  // a pending dump rides on the first step: its launches cover every plane once
  const bool dump = ctx->dump_pending;
  const int rc = for_each_wave_launch(ctx->cfg.nz, ctx->wave_depth, T, [&](int32_t lo, int32_t hi, int32_t t) {
    if (!ctx->inplace) {
      a.src = ctx->buf[(ctx->cur + t - 1) & 1];
      a.dst = ctx->buf[(ctx->cur + t) & 1];
    }
    a.state = base + uint32_t(t);
    return launch_step<S>(ctx, a, lo, hi, (ctx->parity + t - 1) & 1, dump && t == 1);
  });
  if (rc) return rc;
  ctx->dump_pending = false;
  ctx->cur = (ctx->cur + T) & 1;
  ctx->parity = (ctx->parity + T) & 1;
  ctx->step += T;
  return CULBM_OK;
}

// Static blocking: the wavefront schedule of a full pass is fixed by the
// configuration, so it is captured once into a HIP graph and replayed; the
// device's command processor then issues the chunk steps back to back with
// no host round trip.  A shorter final pass uses plain launches, and so does
// every pass with stochastic rounding (each launch carries its step number).
template <class S>
int static_pass(culbm_ctx* ctx, int32_t T) {
  if (T != ctx->cfg.inner_loop || ctx->sr || ctx->dump_pending) return wave_pass<S>(ctx, T);
  const int key = ctx->inplace ? ctx->parity : ctx->cur;
  if (!ctx->pass_graph[key]) {
    const int cur = ctx->cur, parity = ctx->parity;
    const int64_t step = ctx->step;
    const bool prof = ctx->prof;
    ctx->prof = false;
    HIPCHK(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    const int rc = wave_pass<S>(ctx, T);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(ctx->stream, &g);
    ctx->prof = prof;
    ctx->cur = cur;
    ctx->parity = parity;
    ctx->step = step;
    if (rc) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    if (e != hipSuccess) return fail(ctx, CULBM_ERR_HIP, "stream capture failed: %s", hipGetErrorString(e));
    const hipError_t ei = hipGraphInstantiate(&ctx->pass_graph[key], g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ei != hipSuccess) {
      ctx->pass_graph[key] = nullptr;
      return fail(ctx, CULBM_ERR_HIP, "graph instantiation failed: %s", hipGetErrorString(ei));
    }
  }
  if (int rc = prof_begin(ctx)) return rc;
  HIPCHK(ctx, hipGraphLaunch(ctx->pass_graph[key], ctx->stream));
  if (int rc = prof_end(ctx)) return rc;
  ctx->cur = (ctx->cur + T) & 1;
  ctx->parity = (ctx->parity + T) & 1;
  ctx->step += T;
  return CULBM_OK;
}

// Mixed blocking: two steps per launch with the intermediate in LDS, and one
// plain step for an odd remainder.
template <class S>
int fused_pass(culbm_ctx* ctx, int32_t T) {
  FusedArgs fa{};
  fa.a = base_args(ctx);
  fa.tiles_x = ctx->tiles_x;
  fa.tiles_y = ctx->tiles_y;
  fa.zchunk = ctx->fzchunk;
  fa.nchunk = ctx->nchunk;
  const uint32_t nblk = uint32_t(ctx->tiles_x) * ctx->tiles_y * ctx->nchunk;
  for (int32_t s = 0; s + 1 < T; s += 2) {
    fa.a.src = ctx->buf[ctx->cur];
    fa.a.dst = ctx->buf[ctx->cur ^ 1];
    fa.a.state = uint32_t(ctx->step) + 1u;  // the launch produces states state and state + 1
    if (int rc = prof_begin(ctx)) return rc;
    bool sr = false;
    if constexpr (S::kBytes == 2) sr = ctx->sr;
    if (sr)
      hipLaunchKernelGGL((k_fused2<S, true>), dim3(nblk), dim3(FusedShape<S::kBytes>::BLOCK), 0, ctx->stream, fa);
    else
      hipLaunchKernelGGL((k_fused2<S, false>), dim3(nblk), dim3(FusedShape<S::kBytes>::BLOCK), 0, ctx->stream, fa);
    HIPCHK(ctx, hipGetLastError());
    if (int rc = prof_end(ctx)) return rc;
    ctx->cur ^= 1;
    ctx->step += 2;
  }
  if (T & 1) {
    StepArgs a = base_args(ctx);
    a.src = ctx->buf[ctx->cur];
    a.dst = ctx->buf[ctx->cur ^ 1];
    a.state = uint32_t(ctx->step) + 1u;
    if (int rc = launch_step<S>(ctx, a, 0, ctx->cfg.nz, 0)) return rc;
    ctx->cur ^= 1;
    ctx->step += 1;
  }
  return CULBM_OK;
}

template <class S>
int dump_impl(culbm_ctx* ctx);

template <class S>
int do_steps(culbm_ctx* ctx, int32_t nsteps) {
  const culbm_config& c = ctx->cfg;
  if (nsteps > 0 && ctx->dump_pending && c.opt_policy == CULBM_OPT_MIXED) {
    // the two-step kernel gathers tiles with halos: the dump runs as its own pass
    if (int rc = dump_impl<S>(ctx)) return rc;
    ctx->dump_pending = false;
  }
  if (!blocked(c)) {
    StepArgs a = base_args(ctx);
    for (int32_t s = 0; s < nsteps; ++s) {
      if (!ctx->inplace) {
        a.src = ctx->buf[ctx->cur];
        a.dst = ctx->buf[ctx->cur ^ 1];
      }
      a.state = uint32_t(ctx->step) + 1u;
      const int rc = launch_step<S>(ctx, a, 0, c.nz, ctx->parity, ctx->dump_pending);
      if (rc) return rc;
      ctx->dump_pending = false;
      ctx->cur ^= 1;
      ctx->parity ^= 1;
      ctx->step += 1;
    }
  } else if (c.opt_policy == CULBM_OPT_MIXED) {
    // two steps per launch across the passes of innerLoop steps (the bits do
    // not depend on the grouping), one plain step for an odd remainder
    return fused_pass<S>(ctx, nsteps);
  } else {
    int32_t left = nsteps;
    while (left > 0) {
      const int32_t T = std::min(left, c.inner_loop);
      const int rc = c.opt_policy == CULBM_OPT_STATIC ? static_pass<S>(ctx, T) : wave_pass<S>(ctx, T);
      if (rc) return rc;
      left -= T;
    }
  }
  return CULBM_OK;
}

int steps_dispatch(culbm_ctx* ctx, int32_t nsteps) {
  switch (ctx->cfg.storage) {
    case CULBM_STORAGE_FP16: return do_steps<StoreF16>(ctx, nsteps);
    case CULBM_STORAGE_BF16: return do_steps<StoreBF16>(ctx, nsteps);
    default: return do_steps<StoreF32>(ctx, nsteps);
  }
}

StepArgs full_args(const culbm_ctx* ctx) {
  StepArgs a = base_args(ctx);
  set_range(a, 0, ctx->cfg.nz, false, false);
  return a;
}

// plain FP32 device array -> populations in the context's layout
template <class S>
int upload_plain(culbm_ctx* ctx, const float* dplain) {
  StepArgs a = full_args(ctx);
  if (ctx->inplace) {
    // written with the E-step store rules, so the next step is an O step; the x
    // ghost copies only where the steps read them (odd nx)
    if (ctx->cfg.nx % 2 == 0)
      hipLaunchKernelGGL((k_plain_to_twist<S, 0, false>), dim3(a.nblk), dim3(kBlock), 0, ctx->stream, a,
                         dplain, ctx->buf[0]);
    else
      hipLaunchKernelGGL((k_plain_to_twist<S, 0, true>), dim3(a.nblk), dim3(kBlock), 0, ctx->stream, a,
                         dplain, ctx->buf[0]);
    ctx->parity = 1;
  } else {
    for (int b = 0; b < 2; ++b)
      hipLaunchKernelGGL((k_plain_to_pull<S>), dim3(a.nblk), dim3(kBlock), 0, ctx->stream, a,
                         dplain, ctx->buf[b]);
    ctx->cur = 0;
  }
  HIPCHK(ctx, hipGetLastError());
  return CULBM_OK;
}

template <class S>
int download_plain(culbm_ctx* ctx, float* dplain) {
  StepArgs a = full_args(ctx);
  if (ctx->inplace) {
    if (ctx->parity == 1)  // last stores used the E rules
      hipLaunchKernelGGL((k_twist_to_plain<S, 0>), dim3(a.nblk), dim3(kBlock), 0, ctx->stream, a,
                         ctx->buf[0], dplain);
    else
      hipLaunchKernelGGL((k_twist_to_plain<S, 1>), dim3(a.nblk), dim3(kBlock), 0, ctx->stream, a,
                         ctx->buf[0], dplain);
  } else {
    hipLaunchKernelGGL((k_pull_to_plain<S>), dim3(a.nblk), dim3(kBlock), 0, ctx->stream, a,
                       ctx->buf[ctx->cur], dplain);
  }
  HIPCHK(ctx, hipGetLastError());
  return CULBM_OK;
}

template <class S>
int dump_impl(culbm_ctx* ctx) {
  DumpArgs d{};
  d.a = full_args(ctx);
  d.flag = ctx->flag;
  d.rho = ctx->rho;
  d.vx = ctx->vx;
  d.vy = ctx->vy;
  d.vz = ctx->vz;
  if (ctx->inplace) {
    d.a.src = ctx->buf[0];
    if (ctx->parity == 1)
      hipLaunchKernelGGL((k_dump<S, 1>), dim3(d.a.nblk), dim3(kBlock), 0, ctx->stream, d);
    else
      hipLaunchKernelGGL((k_dump<S, 2>), dim3(d.a.nblk), dim3(kBlock), 0, ctx->stream, d);
  } else {
    d.a.src = ctx->buf[ctx->cur];
    hipLaunchKernelGGL((k_dump<S, 0>), dim3(d.a.nblk), dim3(kBlock), 0, ctx->stream, d);
  }
  HIPCHK(ctx, hipGetLastError());
  return CULBM_OK;
}

template <class F>
int by_storage(int32_t storage, F&& f) {
  switch (storage) {
    case CULBM_STORAGE_FP16: return f(StoreF16{});
    case CULBM_STORAGE_BF16: return f(StoreBF16{});
    default: return f(StoreF32{});
  }
}

// Run a dump armed by culbm_dump_deferred that no step has taken yet (as a
// separate pass), before a call reads or replaces the fields or populations.
int flush_dump(culbm_ctx* ctx) {
  if (!ctx->dump_pending) return CULBM_OK;
  const int rc = by_storage(ctx->cfg.storage, [&](auto s) { return dump_impl<decltype(s)>(ctx); });
  if (rc) return rc;
  ctx->dump_pending = false;
  return CULBM_OK;
}

int zero_device(culbm_ctx* ctx, void* p, size_t bytes) {
  const uint64_t words = bytes / 4;
  const uint32_t nblk = uint32_t(std::min<uint64_t>((words + kBlock - 1) / kBlock, 65536));
  if (words)
    hipLaunchKernelGGL(k_zero, dim3(nblk), dim3(kBlock), 0, ctx->stream,
                       static_cast<uint32_t*>(p), words);
  HIPCHK(ctx, hipGetLastError());
  return CULBM_OK;
}

size_t ddf_bytes(const culbm_ctx* ctx) { return size_t(27) * size_t(ctx->dstride) * size_t(ctx->bpp); }

int alloc(culbm_ctx* ctx, void** p, size_t bytes) {
  HIPCHK(ctx, hipMalloc(p, bytes));
  ctx->device_bytes += int64_t(bytes);
  return CULBM_OK;
}

// Planes per wavefront chunk.  FP32 in place and pull: a launch-size target
// (1.4M cells in place, 1.25M pull) capped by the cache-fitting depth (D + T + 1
// planes within ~160 MB of the 256 MB Infinity Cache) or half the domain,
// whichever is larger.  16-bit in place: the deepest chunk whose working set,
// D + T + 1 planes of the padded in-place layout, stays within 190 MB, but at
// least 1.6M cells per launch.  Measured with the chunks in runs of 8 workgroups
// per XCD, innerLoop 5: FP32 in place at 288x272x280 runs 26.1k MLUPS at 18
// planes, 24.0k at 22, 22.7k at 32 (the rate falls off once the working set
// nears the cache; at 256^3 and 384^3 the optimum moved by a plane or two from
// run to run, within +-4 %); FP16 pull 46.7k at 16 planes, 44.5k at 32.  FP16
// in place peaked at 172-190 MB in every run: 256^3 46.2k at 33 planes, 42.2k
// at 40; 384^3 47.7k at 12, 44.7k at 18; 288x272x280 44.9-48.2k at 32-34.  At
// 128^3 the cache-fitting 83 planes beat one whole-domain chunk by 16 % (FP32).
int32_t auto_depth(const culbm_ctx* ctx) {
  const double cells = double(ctx->cfg.nx) * ctx->cfg.ny;
  const int32_t T = std::max(1, ctx->cfg.inner_loop);
  const int32_t nz = ctx->cfg.nz;
  if (ctx->inplace && ctx->bpp == 2) {
    const double plane = 27.0 * double(ctx->px) * double(ctx->py) * ctx->bpp + cells;
    const int32_t fit = int32_t(190e6 / plane) - T - 1;
    const int32_t d = std::max(fit, int32_t(std::ceil(1.6e6 / cells)));
    return std::max(2, std::min(d, nz));
  }
  const double plane = cells * (27.0 * ctx->bpp * (ctx->inplace ? 1 : 2) + 1.0);
  const double budget = 160.0 * 1024 * 1024;
  const int32_t fit = std::max(int32_t(budget / plane) - T - 1, 2);
  const double target = ctx->inplace ? 1.4e6 : 1.25e6;
  const int32_t floor_ = int32_t(std::ceil(target / cells));
  // 16-bit pull: when even a floor-sized chunk's two-buffer working set passes
  // 190 MB the cache cannot serve the later steps, and small chunks only add
  // launches; chunks of half the domain run at the plain-pull rate (FP16 at
  // 512^3: 38.3k MLUPS at 5 planes, 43.2k at 256, plain pull 43.8k; 384^3: 42.4k
  // at 9, 44.3k at 192).  FP32 pull measured within 2 % either way and keeps
  // the launch-size rule.
  if (!ctx->inplace && ctx->bpp == 2 && (floor_ + T + 1) * plane > 190e6) return std::max(2, nz / 2);
  const int32_t d = std::min(floor_, std::max(fit, nz / 2));
  return std::max(2, std::min(d, nz));
}

// Moment sums: at most kSumParts workgroups, each writing one partial of the
// four sums, then the four totals after them: the work array holds
// 4 * kSumParts + 4 doubles whatever the domain.
constexpr uint32_t kSumParts = 1024;

constexpr size_t moment_sums_work_bytes() { return (4 * size_t(kSumParts) + 4) * sizeof(double); }

// Launch the partial sums and their fixed-order total into `work` (device
// memory of moment_sums_work_bytes()).  Every partial the total reads is
// written by the same call, so `work` needs no clearing between calls.
int moment_sums_launch(culbm_ctx* ctx, uint32_t flag_mask, double* work) {
  StepArgs a = full_args(ctx);
  const uint32_t nparts = std::min(a.nblk, kSumParts);
  if (nparts == 0 || nparts > kSumParts) return fail(ctx, CULBM_ERR_INVALID, "bad partial-sum count %u", nparts);
  const void* buf = ctx->inplace ? ctx->buf[0] : ctx->buf[ctx->cur];
  const int rc = by_storage(ctx->cfg.storage, [&](auto s) {
    using S = decltype(s);
    if (!ctx->inplace)
      hipLaunchKernelGGL((k_moment_sums<S, 0>), dim3(nparts), dim3(kBlock), 0, ctx->stream, a, buf, ctx->flag,
                         flag_mask, work);
    else if (ctx->parity == 1)  // last stores used the E rules
      hipLaunchKernelGGL((k_moment_sums<S, 2>), dim3(nparts), dim3(kBlock), 0, ctx->stream, a, buf, ctx->flag,
                         flag_mask, work);
    else
      hipLaunchKernelGGL((k_moment_sums<S, 1>), dim3(nparts), dim3(kBlock), 0, ctx->stream, a, buf, ctx->flag,
                         flag_mask, work);
    return hipGetLastError() == hipSuccess ? CULBM_OK : CULBM_ERR_HIP;
  });
  if (rc) return fail(ctx, rc, "moment-sum kernel launch failed");
  hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(64), 0, ctx->stream, static_cast<const double*>(work), nparts,
                     work + 4 * kSumParts);
  HIPCHK(ctx, hipGetLastError());
  return CULBM_OK;
}

}  // namespace

// Work array of culbm_moment_sums_noalloc: the partial sums and the totals,
// 4 * kSumParts + 4 doubles (32.8 KB), one copy on each device the library
// runs on, held until the process ends.  In host code this name stands for the
// symbol only; kernels and copies get the device address from scratch_on().
static __device__ __align__(256) double scratch[4 * kSumParts + 4];
static_assert(sizeof(scratch) == moment_sums_work_bytes(), "scratch holds the moment-sum work array");

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch[kMaxDevices] = {};

// Device address of the current device's scratch copy, looked up once per
// device (the caller has made ctx's device current).
int scratch_on(culbm_ctx* ctx, double** out) {
  const int dev = ctx->cfg.device;
  if (dev < 0 || dev >= kMaxDevices) return fail(ctx, CULBM_ERR_UNSUPPORTED, "device %d beyond %d", dev, kMaxDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(ctx, CULBM_ERR_HIP, "hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
    g_scratch[dev] = static_cast<double*>(p);
  }
  *out = g_scratch[dev];
  return CULBM_OK;
}

}  // namespace

extern "C" {

int culbm_create(const culbm_config* cfg, culbm_ctx** out) {
  if (!out) return fail(nullptr, CULBM_ERR_INVALID, "out is NULL");
  *out = nullptr;
  char msg[512];
  const int v = culbm_validate(cfg, msg, sizeof(msg));
  if (v) return fail(nullptr, v, "%s", msg);

  int ndev = 0;
  hipError_t e = hipGetDeviceCount(&ndev);
  if (e != hipSuccess || ndev == 0)
    return fail(nullptr, CULBM_ERR_HIP, "no HIP device available (%s)", hipGetErrorString(e));
  if (cfg->device >= ndev)
    return fail(nullptr, CULBM_ERR_INVALID, "devId %d out of range (%d devices)", cfg->device, ndev);

  culbm_ctx* ctx = new culbm_ctx();
  ctx->cfg = *cfg;
  auto bail = [&](int rc) {
    g_create_err = ctx->err;
    culbm_destroy(ctx);
    return rc;
  };
#define CK(expr)                   \
  do {                             \
    const int rc_ = (expr);        \
    if (rc_) return bail(rc_);     \
  } while (0)
#define HK(expr)                                                                    \
  do {                                                                              \
    hipError_t e_ = (expr);                                                         \
    if (e_ != hipSuccess) {                                                         \
      fail(ctx, CULBM_ERR_HIP, "%s failed: %s", #expr, hipGetErrorString(e_));      \
      return bail(CULBM_ERR_HIP);                                                   \
    }                                                                               \
  } while (0)

  HK(hipSetDevice(cfg->device));
  HK(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
  HK(hipEventCreate(&ctx->ev0));
  HK(hipEventCreate(&ctx->ev1));

  const culbm_config& c = ctx->cfg;
  ctx->n = int64_t(c.nx) * c.ny * c.nz;
  ctx->inplace = c.stream_policy == CULBM_STREAM_INPLACE;
  ctx->sr = c.rounding == CULBM_ROUND_STOCHASTIC;
  ctx->bpp = bytes_per_pop(c.storage);
  ctx->scale = c.storage == CULBM_STORAGE_FP32 ? 1.0f
               : (c.storage_scale != 0.0f ? c.storage_scale : derived_scale(c.storage, 0.0));
  ctx->inv_scale = 1.0f / ctx->scale;
  {
    const Layout L = layout_of(c);
    ctx->px = L.px;
    ctx->py = L.py;
    ctx->pz = L.pz;
    ctx->gcol = L.gcol;
    ctx->dstride = L.dstride;
  }

  const size_t nb = size_t(ctx->n);
  CK(alloc(ctx, reinterpret_cast<void**>(&ctx->flag), nb * 4));
  CK(alloc(ctx, reinterpret_cast<void**>(&ctx->sflag), round_up(int64_t(nb), 4)));
  CK(alloc(ctx, reinterpret_cast<void**>(&ctx->rho), nb * 4));
  CK(alloc(ctx, reinterpret_cast<void**>(&ctx->vx), nb * 4));
  CK(alloc(ctx, reinterpret_cast<void**>(&ctx->vy), nb * 4));
  CK(alloc(ctx, reinterpret_cast<void**>(&ctx->vz), nb * 4));
  CK(alloc(ctx, &ctx->buf[0], ddf_bytes(ctx)));
  if (!ctx->inplace) CK(alloc(ctx, &ctx->buf[1], ddf_bytes(ctx)));

  if (blocked(c)) {
    ctx->wave_depth = c.wave_depth > 0 ? std::min(c.wave_depth, c.nz) : auto_depth(ctx);
  }
  {
    // k_fused2's launch shape, for the mixed policy and for the adjoint's
    // forward sweeps: z-chunks per tile column balance the workgroup count over
    // the resident workgroup slots against the two extra A planes a chunk
    // recomputes
    int ncu = 256;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, c.device) == hipSuccess) ncu = prop.multiProcessorCount;
    int per_cu = 1;
    by_storage(c.storage, [&](auto s) {
      using S = decltype(s);
      if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(&k_fused2<S, false>),
                                                       FusedShape<S::kBytes>::BLOCK, 0) != hipSuccess)
        per_cu = 1;
      return 0;
    });
    const int slots = ncu * std::max(per_cu, 1);
    const int ftx = ctx->bpp == 4 ? FusedShape<4>::TX : FusedShape<2>::TX;
    const int fty = ctx->bpp == 4 ? FusedShape<4>::TY : FusedShape<2>::TY;
    ctx->tiles_x = (c.nx + ftx - 1) / ftx;
    ctx->tiles_y = (c.ny + fty - 1) / fty;
    const double tiles = double(ctx->tiles_x) * ctx->tiles_y;
    if (c.opt_policy == CULBM_OPT_MIXED && c.wave_depth > 0) {
      ctx->fzchunk = std::min(c.wave_depth, c.nz);
    } else {
      double best = 1e30;
      for (int n = 1; n <= std::min(64, c.nz); ++n) {
        const int len = (c.nz + n - 1) / n;
        const double wgs = tiles * n;
        const double rounds = std::ceil(wgs / slots);
        const double cost = rounds / (wgs / slots) * (1.0 + 2.0 / len);
        if (cost < best - 1e-9) {
          best = cost;
          ctx->fzchunk = len;
        }
      }
    }
    ctx->nchunk = (c.nz + ctx->fzchunk - 1) / ctx->fzchunk;
    if (c.opt_policy == CULBM_OPT_MIXED) ctx->wave_depth = ctx->fzchunk;  // reported as the depth
  }

  // reference defaults: all FLUID + DUMP, rho 1, u 0, rest populations
  {
    std::vector<uint32_t> fl(nb, CULBM_FLAG_LOAD | CULBM_FLAG_COLLIDE | CULBM_FLAG_STORE |
                                     CULBM_FLAG_DUMP_RHO | CULBM_FLAG_DUMP_VX |
                                     CULBM_FLAG_DUMP_VY | CULBM_FLAG_DUMP_VZ);
    CK(culbm_set_state(ctx, fl.data(), nullptr, nullptr, nullptr, nullptr));
  }
  CK(culbm_init_ddf(ctx, CULBM_INIT_REST));
#undef CK
#undef HK
  *out = ctx;
  return CULBM_OK;
}

int culbm_destroy(culbm_ctx* ctx) {
  if (!ctx) return CULBM_OK;
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  void* ptrs[] = {ctx->buf[0], ctx->buf[1], ctx->flag, ctx->sflag, ctx->rho, ctx->vx, ctx->vy, ctx->vz};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  for (hipGraphExec_t g : ctx->pass_graph)
    if (g) (void)hipGraphExecDestroy(g);
  for (hipEvent_t e : ctx->prof_ev) (void)hipEventDestroy(e);
  if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
  if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return CULBM_OK;
}

int culbm_set_state(culbm_ctx* ctx, const uint32_t* flag, const float* rho, const float* vx,
                    const float* vy, const float* vz) {
  if (!ctx) return fail(nullptr, CULBM_ERR_INVALID, "ctx is NULL");
  HIPCHK(ctx, hipSetDevice(ctx->cfg.device));
  if (int rc = flush_dump(ctx)) return rc;
  const size_t nb = size_t(ctx->n);
  if (flag) {
    HIPCHK(ctx, hipMemcpyAsync(ctx->flag, flag, nb * 4, hipMemcpyHostToDevice, ctx->stream));
    const uint32_t nblk = uint32_t((nb + kBlock - 1) / kBlock);
    hipLaunchKernelGGL(k_compact_flags, dim3(nblk), dim3(kBlock), 0, ctx->stream, ctx->flag,
                       ctx->sflag, uint64_t(nb));
    HIPCHK(ctx, hipGetLastError());
    if (!ctx->inplace && ctx->buf[1]) {
      StepArgs a = full_args(ctx);
      const int rc = by_storage(ctx->cfg.storage, [&](auto s) {
        using S = decltype(s);
        hipLaunchKernelGGL((k_sync_nostore<S>), dim3(a.nblk), dim3(kBlock), 0, ctx->stream, a,
                           ctx->buf[ctx->cur], ctx->buf[ctx->cur ^ 1]);
        return hipGetLastError() == hipSuccess ? CULBM_OK : CULBM_ERR_HIP;
      });
      if (rc) return fail(ctx, rc, "buffer sync kernel launch failed");
    }
  }
  const float* src[4] = {rho, vx, vy, vz};
  float* dst[4] = {ctx->rho, ctx->vx, ctx->vy, ctx->vz};
  const float dflt[4] = {1.0f, 0.0f, 0.0f, 0.0f};
  if (auto_scale(ctx)) {
    double b = 0.0;
    for (size_t i = 0; i < nb; ++i)
      b = std::max(b, feq_bound(rho ? rho[i] : 1.0f, vx ? vx[i] : 0.0f, vy ? vy[i] : 0.0f, vz ? vz[i] : 0.0f));
    ctx->state_bound = b;
  }
  std::vector<float> fill;
  for (int k = 0; k < 4; ++k) {
    if (src[k]) {
      HIPCHK(ctx, hipMemcpyAsync(dst[k], src[k], nb * 4, hipMemcpyHostToDevice, ctx->stream));
    } else {
      if (fill.empty() || fill[0] != dflt[k]) fill.assign(nb, dflt[k]);
      HIPCHK(ctx, hipMemcpyAsync(dst[k], fill.data(), nb * 4, hipMemcpyHostToDevice, ctx->stream));
      HIPCHK(ctx, hipStreamSynchronize(ctx->stream));  // `fill` is reused
    }
  }
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return CULBM_OK;
}

int culbm_init_ddf(culbm_ctx* ctx, int32_t mode) {
  if (!ctx) return fail(nullptr, CULBM_ERR_INVALID, "ctx is NULL");
  if (mode != CULBM_INIT_REST && mode != CULBM_INIT_FEQ)
    return fail(ctx, CULBM_ERR_INVALID, "init mode must be 0 (rest) or 1 (feq), got %d", mode);
  HIPCHK(ctx, hipSetDevice(ctx->cfg.device));
  if (int rc = flush_dump(ctx)) return rc;
  // every population is encoded anew: the point where a derived scale is set
  if (auto_scale(ctx)) set_scale(ctx, derived_scale(ctx->cfg.storage, ctx->state_bound));
  for (int b = 0; b < 2; ++b)
    if (ctx->buf[b]) {
      const int rc = zero_device(ctx, ctx->buf[b], ddf_bytes(ctx));
      if (rc) return rc;
    }
  ctx->cur = 0;
  ctx->parity = 1;
  if (mode == CULBM_INIT_FEQ) {
    float* plain = nullptr;
    HIPCHK(ctx, hipMalloc(&plain, size_t(27) * size_t(ctx->n) * 4));
    StepArgs a = full_args(ctx);
    hipLaunchKernelGGL(k_feq_plain, dim3(a.nblk), dim3(kBlock), 0, ctx->stream, a, plain);
    int rc = CULBM_OK;
    if (hipGetLastError() != hipSuccess) rc = fail(ctx, CULBM_ERR_HIP, "feq kernel launch failed");
    if (!rc) rc = by_storage(ctx->cfg.storage, [&](auto s) { return upload_plain<decltype(s)>(ctx, plain); });
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipFree(plain);
    if (rc) return rc;
  }
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  ctx->step = 0;
  return CULBM_OK;
}

int culbm_step(culbm_ctx* ctx, int32_t nsteps) {
  if (!ctx) return fail(nullptr, CULBM_ERR_INVALID, "ctx is NULL");
  if (nsteps < 0) return fail(ctx, CULBM_ERR_INVALID, "nsteps must be >= 0");
  HIPCHK(ctx, hipSetDevice(ctx->cfg.device));
  const int rc = steps_dispatch(ctx, nsteps);
  if (rc) return rc;
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return CULBM_OK;
}

int culbm_time_steps(culbm_ctx* ctx, int32_t nsteps, float* ms) {
  if (!ctx) return fail(nullptr, CULBM_ERR_INVALID, "ctx is NULL");
  if (nsteps < 0) return fail(ctx, CULBM_ERR_INVALID, "nsteps must be >= 0");
  HIPCHK(ctx, hipSetDevice(ctx->cfg.device));
  HIPCHK(ctx, hipEventRecord(ctx->ev0, ctx->stream));
  const int rc = steps_dispatch(ctx, nsteps);
  if (rc) return rc;
  HIPCHK(ctx, hipEventRecord(ctx->ev1, ctx->stream));
  HIPCHK(ctx, hipEventSynchronize(ctx->ev1));
  float t = 0.0f;
  HIPCHK(ctx, hipEventElapsedTime(&t, ctx->ev0, ctx->ev1));
  if (ms) *ms = t;
  return CULBM_OK;
}

int culbm_profile_steps(culbm_ctx* ctx, int32_t nsteps, float* kernel_ms, int32_t* launches) {
  if (!ctx) return fail(nullptr, CULBM_ERR_INVALID, "ctx is NULL");
  if (nsteps < 0) return fail(ctx, CULBM_ERR_INVALID, "nsteps must be >= 0");
  HIPCHK(ctx, hipSetDevice(ctx->cfg.device));
  ctx->prof = true;
  ctx->prof_used = 0;
  const int rc = steps_dispatch(ctx, nsteps);
  ctx->prof = false;
  if (rc) return rc;
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  double total = 0.0;
  for (size_t k = 0; k < ctx->prof_used; ++k) {
    float t = 0.0f;
    HIPCHK(ctx, hipEventElapsedTime(&t, ctx->prof_ev[2 * k], ctx->prof_ev[2 * k + 1]));
    total += t;
  }
  if (kernel_ms) *kernel_ms = float(total);
  if (launches) *launches = int32_t(ctx->prof_used);
  return CULBM_OK;
}

int culbm_run_window(culbm_ctx* ctx, int32_t dstep, int32_t* steps_done, float* ms) {
  if (!ctx) return fail(nullptr, CULBM_ERR_INVALID, "ctx is NULL");
  if (dstep < 0) return fail(ctx, CULBM_ERR_INVALID, "dstep must be >= 0");
  int32_t steps = dstep;
  if (blocked(ctx->cfg)) {
    const int32_t I = ctx->cfg.inner_loop;
    steps = (dstep + I - 1) / I * I;
  }
  const int rc = culbm_time_steps(ctx, steps, ms);
  if (rc) return rc;
  if (steps_done) *steps_done = steps;
  return CULBM_OK;
}

int culbm_dump(culbm_ctx* ctx) {
  if (!ctx) return fail(nullptr, CULBM_ERR_INVALID, "ctx is NULL");
  HIPCHK(ctx, hipSetDevice(ctx->cfg.device));
  ctx->dump_pending = false;  // this pass delivers it
  const int rc = by_storage(ctx->cfg.storage, [&](auto s) { return dump_impl<decltype(s)>(ctx); });
  if (rc) return rc;
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return CULBM_OK;
}

int culbm_dump_deferred(culbm_ctx* ctx) {
  if (!ctx) return fail(nullptr, CULBM_ERR_INVALID, "ctx is NULL");
  ctx->dump_pending = true;
  return CULBM_OK;
}

int culbm_get_macros(culbm_ctx* ctx, float* rho, float* vx, float* vy, float* vz) {
  if (!ctx) return fail(nullptr, CULBM_ERR_INVALID, "ctx is NULL");
  HIPCHK(ctx, hipSetDevice(ctx->cfg.device));
  if (int rc = flush_dump(ctx)) return rc;
  float* dst[4] = {rho, vx, vy, vz};
  const float* src[4] = {ctx->rho, ctx->vx, ctx->vy, ctx->vz};
  for (int k = 0; k < 4; ++k)
    if (dst[k])
      HIPCHK(ctx, hipMemcpyAsync(dst[k], src[k], size_t(ctx->n) * 4, hipMemcpyDeviceToHost,
                                 ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return CULBM_OK;
}

int culbm_get_ddf(culbm_ctx* ctx, float* ddf) {
  if (!ctx) return fail(nullptr, CULBM_ERR_INVALID, "ctx is NULL");
  if (!ddf) return fail(ctx, CULBM_ERR_INVALID, "ddf is NULL");
  HIPCHK(ctx, hipSetDevice(ctx->cfg.device));
  const size_t bytes = size_t(27) * size_t(ctx->n) * 4;
  float* plain = nullptr;
  HIPCHK(ctx, hipMalloc(&plain, bytes));
  int rc = by_storage(ctx->cfg.storage, [&](auto s) { return download_plain<decltype(s)>(ctx, plain); });
  if (!rc && hipMemcpyAsync(ddf, plain, bytes, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess)
    rc = fail(ctx, CULBM_ERR_HIP, "device-to-host copy failed");
  (void)hipStreamSynchronize(ctx->stream);
  (void)hipFree(plain);
  return rc;
}

int culbm_set_ddf(culbm_ctx* ctx, const float* ddf) {
  if (!ctx) return fail(nullptr, CULBM_ERR_INVALID, "ctx is NULL");
  if (!ddf) return fail(ctx, CULBM_ERR_INVALID, "ddf is NULL");
  HIPCHK(ctx, hipSetDevice(ctx->cfg.device));
  if (int rc = flush_dump(ctx)) return rc;
  const size_t bytes = size_t(27) * size_t(ctx->n) * 4;
  if (auto_scale(ctx)) {
    // the populations given (shifted, f_i - w_i) and the state's equilibria both have to fit
    float m = 0.0f;
    const size_t total = size_t(27) * size_t(ctx->n);
    for (size_t c = 0; c < total; ++c) m = std::max(m, std::fabs(ddf[c]));
    set_scale(ctx, derived_scale(ctx->cfg.storage, std::max(ctx->state_bound, double(m))));
  }
  float* plain = nullptr;
  HIPCHK(ctx, hipMalloc(&plain, bytes));
  int rc = CULBM_OK;
  if (hipMemcpyAsync(plain, ddf, bytes, hipMemcpyHostToDevice, ctx->stream) != hipSuccess)
    rc = fail(ctx, CULBM_ERR_HIP, "host-to-device copy failed");
  if (!rc) {
    for (int b = 0; b < 2 && !rc; ++b)
      if (ctx->buf[b]) rc = zero_device(ctx, ctx->buf[b], ddf_bytes(ctx));
  }
  if (!rc) rc = by_storage(ctx->cfg.storage, [&](auto s) { return upload_plain<decltype(s)>(ctx, plain); });
  (void)hipStreamSynchronize(ctx->stream);
  (void)hipFree(plain);
  return rc;
}

int culbm_moment_sums(culbm_ctx* ctx, uint32_t flag_mask, double* sums) {
  if (!ctx) return fail(nullptr, CULBM_ERR_INVALID, "ctx is NULL");
  if (!sums) return fail(ctx, CULBM_ERR_INVALID, "sums is NULL");
  HIPCHK(ctx, hipSetDevice(ctx->cfg.device));
  double* work = nullptr;
  HIPCHK(ctx, hipMalloc(&work, moment_sums_work_bytes()));
  int rc = moment_sums_launch(ctx, flag_mask, work);
  if (!rc && hipMemcpyAsync(sums, work + 4 * kSumParts, 4 * sizeof(double), hipMemcpyDeviceToHost,
                            ctx->stream) != hipSuccess)
    rc = fail(ctx, CULBM_ERR_HIP, "device-to-host copy failed");
  if (hipStreamSynchronize(ctx->stream) != hipSuccess && !rc) rc = fail(ctx, CULBM_ERR_HIP, "moment-sum kernels failed");
  (void)hipFree(work);
  return rc;
}

int culbm_moment_sums_noalloc(culbm_ctx* ctx, uint32_t flag_mask, double* sums) {
  if (!ctx) return fail(nullptr, CULBM_ERR_INVALID, "ctx is NULL");
  if (!sums) return fail(ctx, CULBM_ERR_INVALID, "sums is NULL");
  HIPCHK(ctx, hipSetDevice(ctx->cfg.device));
  double* work = nullptr;
  if (int rc = scratch_on(ctx, &work)) return rc;
  // the partial count is capped at kSumParts for every domain, so the work
  // always fits the scratch array (moment_sums_launch rejects anything else)
  if (int rc = moment_sums_launch(ctx, flag_mask, work)) return rc;
  HIPCHK(ctx, hipMemcpyAsync(sums, work + 4 * kSumParts, 4 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return CULBM_OK;
}

int culbm_get_info(culbm_ctx* ctx, culbm_info* info) {
  if (!ctx) return fail(nullptr, CULBM_ERR_INVALID, "ctx is NULL");
  if (!info) return fail(ctx, CULBM_ERR_INVALID, "info is NULL");
  std::memset(info, 0, sizeof(*info));
  info->step = ctx->step;
  info->device_bytes = ctx->device_bytes;
  info->wave_depth = ctx->wave_depth;
  info->bytes_per_pop = ctx->bpp;
  info->storage_scale = ctx->scale;
  const culbm_config& c = ctx->cfg;
  if (!blocked(c)) {
    info->launches_per_pass = 1;
  } else if (c.opt_policy == CULBM_OPT_MIXED) {
    info->launches_per_pass = (c.inner_loop + 1) / 2;  // two steps per launch (pairs run across passes)
  } else {
    int32_t cnt = 0;
    for_each_wave_launch(c.nz, ctx->wave_depth, c.inner_loop, [&](int32_t, int32_t, int32_t) {
      ++cnt;
      return CULBM_OK;
    });
    info->launches_per_pass = cnt;
  }
  return CULBM_OK;
}

int culbm_wave_schedule(int32_t nz, int32_t depth, int32_t steps, int32_t* launches, int32_t max_launches,
                        int32_t* count) {
  if (!count) return fail(nullptr, CULBM_ERR_INVALID, "count is NULL");
  if (nz < 1 || depth < 1 || steps < 1)
    return fail(nullptr, CULBM_ERR_INVALID, "nz, depth and steps must be >= 1");
  int32_t n = 0;
  const int rc = for_each_wave_launch(nz, depth, steps, [&](int32_t lo, int32_t hi, int32_t t) {
    if (launches) {
      if (n >= max_launches) return fail(nullptr, CULBM_ERR_INVALID, "more than %d launches", max_launches);
      launches[3 * n] = lo;
      launches[3 * n + 1] = hi;
      launches[3 * n + 2] = t;
    }
    ++n;
    return int(CULBM_OK);
  });
  if (rc) return rc;
  *count = n;
  return CULBM_OK;
}

// ---- probes -----------------------------------------------------------------

int culbm_probe_cells(int32_t device, int32_t op, float inv_tau, float* f, const float* in4,
                      float* out4, int64_t n) {
  if (n <= 0 || n > (int64_t(1) << 26)) return fail(nullptr, CULBM_ERR_INVALID, "bad n");
  if (op < 0 || op > 2) return fail(nullptr, CULBM_ERR_INVALID, "bad op");
  if (hipSetDevice(device) != hipSuccess) return fail(nullptr, CULBM_ERR_HIP, "hipSetDevice failed");
  float *df = nullptr, *din = nullptr, *dout = nullptr;
  int rc = CULBM_OK;
  const size_t fb = size_t(n) * 27 * 4, qb = size_t(n) * 4 * 4;
  if (hipMalloc(&df, fb) != hipSuccess || hipMalloc(&din, qb) != hipSuccess ||
      hipMalloc(&dout, qb) != hipSuccess)
    rc = fail(nullptr, CULBM_ERR_HIP, "allocation failed");
  if (!rc && op != 1 && hipMemcpy(df, f, fb, hipMemcpyHostToDevice) != hipSuccess)
    rc = fail(nullptr, CULBM_ERR_HIP, "copy failed");
  if (!rc && op == 1 && hipMemcpy(din, in4, qb, hipMemcpyHostToDevice) != hipSuccess)
    rc = fail(nullptr, CULBM_ERR_HIP, "copy failed");
  if (!rc) {
    hipLaunchKernelGGL(k_cell_probe, dim3(uint32_t((n + kBlock - 1) / kBlock)), dim3(kBlock), 0, 0,
                       op, inv_tau, df, din, df, dout, uint32_t(n));
    if (hipDeviceSynchronize() != hipSuccess) rc = fail(nullptr, CULBM_ERR_HIP, "probe kernel failed");
  }
  if (!rc && op != 2 && hipMemcpy(f, df, fb, hipMemcpyDeviceToHost) != hipSuccess)
    rc = fail(nullptr, CULBM_ERR_HIP, "copy failed");
  if (!rc && op == 2 && hipMemcpy(out4, dout, qb, hipMemcpyDeviceToHost) != hipSuccess)
    rc = fail(nullptr, CULBM_ERR_HIP, "copy failed");
  if (df) (void)hipFree(df);
  if (din) (void)hipFree(din);
  if (dout) (void)hipFree(dout);
  return rc;
}

int culbm_probe_codec(int32_t device, int32_t storage, float scale, const float* in,
                      uint16_t* enc, float* dec, int64_t n) {
  if (n <= 0 || n > (int64_t(1) << 28)) return fail(nullptr, CULBM_ERR_INVALID, "bad n");
  if (storage < 0 || storage > 2) return fail(nullptr, CULBM_ERR_INVALID, "bad storage");
  if (!(scale > 0.0f)) return fail(nullptr, CULBM_ERR_INVALID, "bad scale");
  if (hipSetDevice(device) != hipSuccess) return fail(nullptr, CULBM_ERR_HIP, "hipSetDevice failed");
  float *din = nullptr, *ddec = nullptr;
  uint16_t* denc = nullptr;
  int rc = CULBM_OK;
  if (hipMalloc(&din, size_t(n) * 4) != hipSuccess || hipMalloc(&ddec, size_t(n) * 4) != hipSuccess ||
      hipMalloc(&denc, size_t(n) * 2) != hipSuccess)
    rc = fail(nullptr, CULBM_ERR_HIP, "allocation failed");
  if (!rc && hipMemcpy(din, in, size_t(n) * 4, hipMemcpyHostToDevice) != hipSuccess)
    rc = fail(nullptr, CULBM_ERR_HIP, "copy failed");
  if (!rc) {
    const dim3 g(uint32_t((n + kBlock - 1) / kBlock));
    const float inv = 1.0f / scale;
    if (storage == 1)
      hipLaunchKernelGGL(k_codec<StoreF16>, g, dim3(kBlock), 0, 0, din, denc, ddec, uint32_t(n), scale, inv);
    else if (storage == 2)
      hipLaunchKernelGGL(k_codec<StoreBF16>, g, dim3(kBlock), 0, 0, din, denc, ddec, uint32_t(n), scale, inv);
    else
      hipLaunchKernelGGL(k_codec<StoreF32>, g, dim3(kBlock), 0, 0, din, denc, ddec, uint32_t(n), scale, inv);
    if (hipDeviceSynchronize() != hipSuccess) rc = fail(nullptr, CULBM_ERR_HIP, "codec kernel failed");
  }
  if (!rc && enc && storage != 0 && hipMemcpy(enc, denc, size_t(n) * 2, hipMemcpyDeviceToHost) != hipSuccess)
    rc = fail(nullptr, CULBM_ERR_HIP, "copy failed");
  if (!rc && dec && hipMemcpy(dec, ddec, size_t(n) * 4, hipMemcpyDeviceToHost) != hipSuccess)
    rc = fail(nullptr, CULBM_ERR_HIP, "copy failed");
  if (din) (void)hipFree(din);
  if (ddec) (void)hipFree(ddec);
  if (denc) (void)hipFree(denc);
  return rc;
}

int culbm_blocking_axis(int32_t dom, int32_t blk, int32_t iter, int32_t* num, int32_t* starts,
                        int32_t max_starts) {
  if (!num) return fail(nullptr, CULBM_ERR_INVALID, "num is NULL");
  if (!blk_axis_valid(dom, blk, iter))
    return fail(nullptr, CULBM_ERR_INVALID,
                "axis of %d cells cannot be cut into tiles of %d advancing %d steps", dom, blk, iter);
  const int64_t n = blk_num(dom, blk, iter);
  *num = int32_t(n);
  if (starts) {
    if (max_starts < n + 1) return fail(nullptr, CULBM_ERR_INVALID, "starts needs %lld entries", (long long)(n + 1));
    for (int64_t k = 0; k <= n; ++k) starts[k] = int32_t(blk_valid_start(k, blk, n, iter, dom));
  }
  return CULBM_OK;
}

int culbm_probe_bandwidth(int32_t device, int32_t mode, int64_t bytes, int32_t iters, float* gbs) {
  const bool wrong = (mode & CULBM_PROBE_CHECK_WRONG) != 0;
  mode &= ~CULBM_PROBE_CHECK_WRONG;
  // modes 3 and 4 launch one work-item per 4 bytes: 8 GiB keeps every mode's
  // grid below 2^31 work-items
  if (bytes < 4096 || bytes > (int64_t(8) << 30) || iters < 1 || mode < 0 || mode > 6)
    return fail(nullptr, CULBM_ERR_INVALID, "bad bandwidth probe arguments");
  if (mode >= 5 && uint64_t(bytes) / 4 / 27 < 256)  // at least one 256-cell block per direction array
    return fail(nullptr, CULBM_ERR_INVALID, "population-pattern probe needs at least %d bytes", 27 * 256 * 4);
  if (hipSetDevice(device) != hipSuccess) return fail(nullptr, CULBM_ERR_HIP, "hipSetDevice failed");
  const uint64_t n4 = uint64_t(bytes) / 16;
  const uint64_t nwords = 4 * n4;
  void *a = nullptr, *b = nullptr;
  uint32_t* bad = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = CULBM_OK;
  auto hip_ok = [&](hipError_t e, const char* what) {
    if (e != hipSuccess && !rc) rc = fail(nullptr, CULBM_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
    return rc == CULBM_OK;
  };
  if (hipMalloc(&a, n4 * 16 + 256) != hipSuccess || hipMalloc(&b, n4 * 16 + 256) != hipSuccess ||
      hipMalloc(&bad, 256) != hipSuccess)
    rc = fail(nullptr, CULBM_ERR_HIP, "allocation failed");
  if (!rc && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess))
    rc = fail(nullptr, CULBM_ERR_HIP, "event creation failed");
  // one element (or one cell) per thread
  const uint32_t nb4 = uint32_t((n4 + kBlock - 1) / kBlock), nb1 = uint32_t((nwords + kBlock - 1) / kBlock);
  // modes 5, 6: 27 arrays per buffer, each starting on a 1 KiB boundary like
  // the population layout's direction stride
  const uint64_t cells = nwords / 27 / 256 * 256;
  const uint32_t nbc = uint32_t((cells + kBlock - 1) / kBlock);
  const uint32_t seed_a = 0x1234567u, seed_b = 0x89abcdefu;
  if (!rc) {
    // known words in both buffers (64 words past the end included), so the
    // check below sees exactly what the probe wrote
    const uint64_t fw = nwords + 64;
    const uint32_t nbf = uint32_t((fw + kBlock - 1) / kBlock);
    hipLaunchKernelGGL(k_bw_fill, dim3(nbf), dim3(kBlock), 0, 0, static_cast<uint32_t*>(a), fw, seed_a);
    hipLaunchKernelGGL(k_bw_fill, dim3(nbf), dim3(kBlock), 0, 0, static_cast<uint32_t*>(b), fw, seed_b);
    hipLaunchKernelGGL(k_zero, dim3(1), dim3(kBlock), 0, 0, bad, uint64_t(64));
    hip_ok(hipGetLastError(), "probe set-up launch");
  }
  if (!rc) {
    auto launch = [&]() {
      switch (mode) {
        case 0: hipLaunchKernelGGL(k_bw_copy, dim3(nb4), dim3(kBlock), 0, 0, static_cast<const float4*>(a), static_cast<float4*>(b), n4); break;
        case 1: hipLaunchKernelGGL(k_bw_read, dim3(nb4), dim3(kBlock), 0, 0, static_cast<const uint4*>(a), static_cast<uint32_t*>(b), n4); break;
        case 2: hipLaunchKernelGGL(k_bw_write, dim3(nb4), dim3(kBlock), 0, 0, static_cast<float4*>(b), n4); break;
        case 3: hipLaunchKernelGGL(k_bw_copy_shift, dim3(nb1), dim3(kBlock), 0, 0, static_cast<const float*>(a), static_cast<float*>(b), nwords, 0u); break;
        case 4: hipLaunchKernelGGL(k_bw_copy_shift, dim3(nb1), dim3(kBlock), 0, 0, static_cast<const float*>(a), static_cast<float*>(b), nwords, 1u); break;
        case 5: hipLaunchKernelGGL(k_bw_pop27, dim3(nbc), dim3(kBlock), 0, 0, static_cast<const float*>(a), static_cast<float*>(b), cells, 0u); break;
        default: hipLaunchKernelGGL(k_bw_pop27, dim3(nbc), dim3(kBlock), 0, 0, static_cast<const float*>(a), static_cast<float*>(b), cells, 1u); break;
      }
    };
    launch();  // warm-up
    if (hip_ok(hipGetLastError(), "probe launch") && hip_ok(hipEventRecord(e0, 0), "hipEventRecord")) {
      for (int32_t k = 0; k < iters; ++k) launch();
      hip_ok(hipGetLastError(), "probe launch");
    }
    if (!rc && hip_ok(hipEventRecord(e1, 0), "hipEventRecord") &&
        hip_ok(hipEventSynchronize(e1), "probe kernels")) {
      float ms = 0.0f;
      if (hip_ok(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime") && !(ms > 0.0f))
        rc = fail(nullptr, CULBM_ERR_HIP, "probe timing returned %g ms", double(ms));
      const double moved = (mode >= 5 ? double(cells) * 27.0 * 4.0 : double(n4) * 16.0) *
                           (mode == 1 || mode == 2 ? 1.0 : 2.0) * iters;
      if (!rc && gbs) *gbs = float(moved / (double(ms) * 1e-3) / 1e9);
    }
  }
  if (!rc) {
    // the data check: every word the probe claims to have moved holds what it should
    hipLaunchKernelGGL(k_bw_verify, dim3(4096), dim3(kBlock), 0, 0, int(mode), static_cast<const uint32_t*>(a),
                       static_cast<const uint32_t*>(b), nwords, cells, uint64_t(nb4), seed_a, uint32_t(wrong), bad);
    uint32_t nbad = 0;
    if (hip_ok(hipGetLastError(), "probe check launch") &&
        hip_ok(hipMemcpy(&nbad, bad, sizeof(nbad), hipMemcpyDeviceToHost), "probe check copy") && nbad)
      rc = fail(nullptr, CULBM_ERR_STATE, "bandwidth probe mode %d moved wrong data: %u words differ", int(mode),
                nbad);
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (a) (void)hipFree(a);
  if (b) (void)hipFree(b);
  if (bad) (void)hipFree(bad);
  return rc;
}

}  // extern "C"

#include "adjoint_capi.hip"
