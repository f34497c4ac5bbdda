/*
 * culbm.h - C ABI of the MI355X D3Q27 lattice-Boltzmann engine.
 *
 * The reference exposes no library API: its boundary is the process
 * (culbm::simulator::single_dev_expt::Simulator(argc, argv), run(), ~Simulator(),
 * src/simulator/single_dev_expt/simulator_expt_platform.hpp:7-16, driven by
 * src/single_dev_expt_main.cpp:5-17).  These entry points split that object
 * into the calls its constructor and run() make, so a host CLI, ctypes or any
 * FFI can drive the same path:
 *
 *   culbm_create        Simulator::Data::Data, buffer allocation
 *                       (simulator_expt_platform.cu:308-445)
 *   culbm_set_state     initFlag / initRhoU (simulator_expt_platform.cu:95-118, 200-228)
 *   culbm_init_ddf      initDoubleDDFBuf / initRingDDFBuf (simulator_expt_platform.cu:230-277)
 *   culbm_step          the step launches of the policy runs (simulator_expt_platform.cu:749-991)
 *   culbm_run_window    one timed window of run() (event record, dstep steps,
 *                       MLUPS), simulator_expt_platform.cu:963-985
 *   culbm_dump          the *DumpKernel launches (kernel.cu:2634-2735, 1829-1936)
 *   culbm_get_macros    dumpRes' device-to-host copies (simulator_expt_platform.cu:572-606)
 *   culbm_destroy       Simulator::Data::~Data (simulator_expt_platform.cu:447-502)
 *
 * Conventions: every call returns CULBM_OK (0) or a CULBM_ERR_* code and never
 * throws; culbm_last_error() gives the message.  The library owns all device
 * memory; callers own host arrays.  Host arrays are dense x-fastest
 * [nz][ny][nx] (populations: [27][nz][ny][nx]), little-endian, exactly the
 * layout of the reference's .dat files.  One context owns one HIP stream and is
 * not thread safe.
 */
#ifndef CULBM_H
#define CULBM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CULBM_ABI_VERSION 1

/* return codes */
#define CULBM_OK 0
#define CULBM_ERR_INVALID 1     /* bad argument or configuration              */
#define CULBM_ERR_HIP 2         /* HIP runtime error                          */
#define CULBM_ERR_STATE 3       /* call not valid in the current state        */
#define CULBM_ERR_UNSUPPORTED 4 /* combination not provided                   */

/* flag bits (src/simulator/single_dev_expt/local_config.hpp:12-22) */
#define CULBM_FLAG_LOAD (1u << 0)
#define CULBM_FLAG_EQU (1u << 1)
#define CULBM_FLAG_COLLIDE (1u << 2)
#define CULBM_FLAG_BOUNCE_BACK (1u << 3)
#define CULBM_FLAG_STORE (1u << 4)
#define CULBM_FLAG_DUMP_RHO (1u << 27)
#define CULBM_FLAG_DUMP_VX (1u << 28)
#define CULBM_FLAG_DUMP_VY (1u << 29)
#define CULBM_FLAG_DUMP_VZ (1u << 30)

/* --streamPolicy */
#define CULBM_STREAM_PULL 0
#define CULBM_STREAM_INPLACE 1
/* --optPolicy */
#define CULBM_OPT_NONE 0
#define CULBM_OPT_STATIC 1  /* reference: HALO_BLOCKING_STATIC_L2  */
#define CULBM_OPT_MIXED 2   /* reference: HALO_BLOCKING_L1L2       */
#define CULBM_OPT_DYNAMIC 3 /* reference: HALO_BLOCKING_DYNAMIC_L2 */
/* population storage (no reference counterpart: the reference stores FP32) */
#define CULBM_STORAGE_FP32 0
#define CULBM_STORAGE_FP16 1
#define CULBM_STORAGE_BF16 2
/* rounding of the 16-bit stores (no reference counterpart) */
#define CULBM_ROUND_NEAREST 0    /* round to nearest even                       */
#define CULBM_ROUND_STOCHASTIC 1 /* unbiased, keyed by (cell, direction, step)  */
/* culbm_init_ddf modes */
#define CULBM_INIT_REST 0 /* feq(1,0) = 0 everywhere, as the reference   */
#define CULBM_INIT_FEQ 1  /* feq(rho, u) of the rho/u fields (opt-in)     */

typedef struct culbm_ctx culbm_ctx;

/* Knobs of the reference CLI (simulator_expt_platform.cu:312-354) plus the
 * storage format and the wavefront depth of the blocked policies. */
typedef struct culbm_config {
  int32_t nx, ny, nz;       /* --domainDim                                   */
  float inv_tau;            /* --invTau                                      */
  int32_t stream_policy;    /* --streamPolicy                                */
  int32_t opt_policy;       /* --optPolicy                                   */
  int32_t inner_loop;       /* --innerLoop: steps per temporal-blocking pass */
  int32_t block_dim[3];     /* --blockDim (checked; kernels pick their own)  */
  int32_t grid_dim[3];      /* --gridDim                                     */
  int32_t use_ring;         /* --useRingDDFBuf                               */
  int32_t device;           /* --devId                                       */
  int32_t storage;          /* CULBM_STORAGE_*                               */
  float storage_scale;      /* 16-bit scale (power of two); 0 = per run      */
  int32_t wave_depth;       /* planes per wavefront chunk; 0 = automatic     */
  int32_t rounding;         /* CULBM_ROUND_* of the 16-bit stores            */
  int32_t reserved[6];
} culbm_config;

typedef struct culbm_info {
  int64_t step;               /* cumulative steps taken                      */
  int64_t device_bytes;       /* bytes of device memory held                 */
  int32_t wave_depth;         /* chunk depth in use (blocked policies)       */
  int32_t launches_per_pass;  /* kernel launches per pass of inner_loop steps */
  int32_t bytes_per_pop;      /* 4 (FP32) or 2 (FP16/BF16)                   */
  float storage_scale;        /* scale in use                                */
  int32_t reserved[8];
} culbm_info;

/* Fill cfg with the reference CLI defaults (simulator_expt_platform.cu:312-354). */
void culbm_config_default(culbm_config* cfg);

/* Validate a configuration without touching the device.  Returns CULBM_OK or
 * an error code; msg (may be NULL) receives the reason. */
int culbm_validate(const culbm_config* cfg, char* msg, size_t msg_len);

/* Validate, select the device, allocate buffers.  Flags default to the
 * reference's no-init-folder state (all FLUID + DUMP), rho = 1, u = 0, and
 * the populations to the rest equilibrium. */
int culbm_create(const culbm_config* cfg, culbm_ctx** out);

/* Upload the flag field and the rho/vx/vy/vz fields (host arrays of nx*ny*nz
 * values).  flag == NULL keeps the current flags; a NULL field is reset to its
 * default (rho 1, u 0), as initRhoU does for a missing file. */
int culbm_set_state(culbm_ctx* ctx, const uint32_t* flag, const float* rho, const float* vx,
                    const float* vy, const float* vz);

/* (Re)initialise the populations: CULBM_INIT_REST or CULBM_INIT_FEQ.  Resets
 * the step counter. */
int culbm_init_ddf(culbm_ctx* ctx, int32_t mode);

/* Advance exactly nsteps steps (any policy; blocked policies run passes of
 * inner_loop steps and a shorter final pass). */
int culbm_step(culbm_ctx* ctx, int32_t nsteps);

/* One window of the reference's run(): time the step launches of `dstep`
 * steps with device events.  Blocked policies advance in whole passes, so
 * they take ceil(dstep / inner_loop) * inner_loop steps, as the reference's
 * window loop does (simulator_expt_platform.cu:785).  Outputs may be NULL. */
int culbm_run_window(culbm_ctx* ctx, int32_t dstep, int32_t* steps_done, float* ms);

/* Device-timed run of nsteps steps (event pair on the context stream around
 * the launches only); *ms receives the elapsed milliseconds. */
int culbm_time_steps(culbm_ctx* ctx, int32_t nsteps, float* ms);

/* Device-timed run of nsteps steps with an event pair around every kernel
 * launch: *kernel_ms is the summed kernel time (launch gaps excluded),
 * *launches the number of launches.  Measurement aid; slower than
 * culbm_time_steps. */
int culbm_profile_steps(culbm_ctx* ctx, int32_t nsteps, float* kernel_ms, int32_t* launches);

/* Dump kernel: rho/u of the current populations written into the device
 * rho/vx/vy/vz buffers where the DUMP bits are set. */
int culbm_dump(culbm_ctx* ctx);

/* The same dump, fused into the next step instead of a pass of its own (the
 * reference's separate dump pass, kernel.cu:2634-2735 with dumpRes,
 * simulator_expt_platform.cu:572-606, made to ride on the step that follows a
 * window): the next step launch(es) gather exactly the populations the dump
 * reads and write the same rho/u values.  The dump is delivered by whichever
 * comes first: the next culbm_step / culbm_run_window / culbm_time_steps (its
 * first step; the mixed policy runs it as a separate pass there), or a call
 * that reads or replaces the fields or populations (culbm_get_macros,
 * culbm_dump, culbm_set_state, culbm_init_ddf, culbm_set_ddf, culbm_adjoint),
 * which runs it as a separate pass first.  A dumped field is that cell's EQU
 * input afterwards, as after culbm_dump. */
int culbm_dump_deferred(culbm_ctx* ctx);

/* Copy the device rho/vx/vy/vz buffers to host arrays (NULL skips a field). */
int culbm_get_macros(culbm_ctx* ctx, float* rho, float* vx, float* vy, float* vz);

/* Populations as dense FP32 [27][nz][ny][nx], post-collision view of the last
 * step (decoded from 16-bit storage). */
int culbm_get_ddf(culbm_ctx* ctx, float* ddf);

/* Replace the populations (same layout as culbm_get_ddf; encoded to the
 * storage format).  The step counter is kept. */
int culbm_set_ddf(culbm_ctx* ctx, const float* ddf);

/* FP64 reductions of the stored populations (the culbm_get_ddf view):
 * sums[0] = sum of rho = 1 + sum_i f_i, sums[1..3] = sum of the momentum
 * sum_i c_i f_i, over the cells whose flag shares a bit with flag_mask (every
 * cell for 0).  The mass-conservation check; no reference counterpart. */
int culbm_moment_sums(culbm_ctx* ctx, uint32_t flag_mask, double* sums);

/* culbm_moment_sums without any allocation.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The same
 * kernels, and so the same bits, as culbm_moment_sums; their partial sums live
 * in a fixed 32.8 KB array of the library (4 * 1024 + 4 doubles, one copy per
 * device, held until the process ends), whatever the domain size.  Calls that
 * overlap on one device share that array and race: issue them one at a time
 * per device. */
int culbm_moment_sums_noalloc(culbm_ctx* ctx, uint32_t flag_mask, double* sums);

int culbm_get_info(culbm_ctx* ctx, culbm_info* info);

/* Release device memory and the stream.  NULL is accepted. */
int culbm_destroy(culbm_ctx* ctx);

/* Message of the last failing call on ctx (or of the last failing
 * culbm_create when ctx is NULL).  Never NULL. */
const char* culbm_last_error(const culbm_ctx* ctx);

/* Name of a return code. */
const char* culbm_error_string(int code);

/* ---- discrete adjoint (no reference counterpart) ----------------------------
 * Objective over the states f^t reached after t steps from the current state:
 *   J = sum_{t = t_begin..t_end} sum_x w(x) u_x(x, t),
 * u_x(x, t) being what the dump kernel reports for cell x after t steps.
 * Gradients: dJ/d(rho, ux, uy, uz) of the EQU inputs (the inlet profile),
 * dJ/dinvTau and, on request, dJ/df^0.  The forward trajectory is the plain
 * pull step in the context's storage format (every policy produces the same
 * bits); the adjoint runs in FP32 with FP64 accumulators.  Forward states are
 * kept as checkpoints and recomputed under a device-memory cap. */
typedef struct culbm_adjoint_spec {
  int32_t nsteps;            /* N >= 1 forward steps                                   */
  int32_t t_begin, t_end;    /* objective window, 0 <= t_begin <= t_end <= N          */
  int32_t checkpoint_every;  /* k > 0: keep f^0, f^k, f^2k, ...; 0: binomial schedule */
  int32_t want_f0;           /* also return dJ/df^0                                    */
  int64_t mem_cap;           /* device bytes the call may allocate; 0: 90% of free     */
  int32_t timing;            /* fill the ms_* fields (device events per phase)         */
  int32_t max_slots;         /* binomial: at most this many states held (0: no limit)  */
  int32_t reserved[4];
} culbm_adjoint_spec;

typedef struct culbm_adjoint_stats {
  double objective;          /* J                                                      */
  double grad_inv_tau;       /* dJ/dinvTau                                             */
  int64_t forward_steps;     /* forward steps run, the first sweep included            */
  int64_t device_bytes;      /* workspace allocated by the call                        */
  int32_t slots;             /* states held as checkpoints                             */
  int32_t launches;          /* kernel launches                                        */
  float ms_sweep;            /* first forward sweep (with the objective)               */
  float ms_recompute;        /* forward steps recomputed during the reverse sweep      */
  float ms_reverse;          /* adjoint steps                                          */
  float ms_total;            /* whole call, set-up and copies included                 */
  int32_t reserved[6];
} culbm_adjoint_stats;

/* Run N steps from the current state (the context ends N steps further on,
 * as after culbm_step) and the reverse sweep.  weight: nx*ny*nz doubles (the
 * objective's w).  grad_rho/ux/uy/uz: nx*ny*nz doubles each or NULL, zero
 * outside EQU cells.  grad_f0: 27*nx*ny*nz floats ([27][nz][ny][nx]) or NULL;
 * required when spec->want_f0.  Fails with CULBM_ERR_INVALID when the
 * checkpoints do not fit mem_cap. */
int culbm_adjoint(culbm_ctx* ctx, const culbm_adjoint_spec* spec, const double* weight,
                  culbm_adjoint_stats* stats, double* grad_rho, double* grad_ux,
                  double* grad_uy, double* grad_uz, float* grad_f0);

/* Checkpoint schedule without running it: forward steps and state slots the
 * reverse sweep of nsteps steps takes with checkpoint_every (> 0) or, for 0,
 * the binomial schedule with `slots` states.  Host only. */
int culbm_adjoint_plan(int32_t nsteps, int32_t checkpoint_every, int32_t slots,
                       int64_t* forward_steps, int32_t* slots_used);

/* ---- test probes (single cells, codecs); device arrays are managed here ---- */

/* op 0: collide f[n][27] in place; op 1: equilibrium of in4[n][4] = (rho,u)
 * into f; op 2: moments of f into out4[n][4]. */
int culbm_probe_cells(int32_t device, int32_t op, float inv_tau, float* f, const float* in4,
                      float* out4, int64_t n);

/* Encode+decode n floats through a storage format on the device. */
int culbm_probe_codec(int32_t device, int32_t storage, float scale, const float* in,
                      uint16_t* enc, float* dec, int64_t n);

/* Halo-blocking decomposition of one axis (src/blocking_core/blocking_alogrithm.hpp:8-41):
 * *num tiles of length blk advancing iter steps cover [0, dom); starts[k], k = 0..num,
 * are the valid-range boundaries (starts[0] = 0, starts[num] = dom); tile k loads
 * [max(starts[k]-(iter-1), 0), min(starts[k+1]+(iter-1), dom)).  starts may be NULL. */
int culbm_blocking_axis(int32_t dom, int32_t blk, int32_t iter, int32_t* num, int32_t* starts,
                        int32_t max_starts);

/* Launch schedule of one temporal-blocking pass of the dynamic and static
 * policies (the z-wavefront that replaces the reference's per-tile launch loop,
 * simulator_expt_platform.cu:749-872): for a domain of nz planes, chunks of
 * `depth` planes and `steps` steps, launches[3k..3k+2] = (z0, z1, t) of the
 * k-th launch in issue order, step t (1-based) over planes [z0, z1).
 * launches may be NULL to count only.  Host only (the hazard checker's input). */
int culbm_wave_schedule(int32_t nz, int32_t depth, int32_t steps, int32_t* launches, int32_t max_launches,
                        int32_t* count);

/* Bandwidth calibration (counterpart of the reference's analysis/lambda and
 * analysis/gamma probes, analysis/lambda/measure.cu:14-63,
 * analysis/gamma/measure.cu:14-69), one element per thread: mode 0 copy,
 * 1 read (each workgroup writes one XOR word of what it read), 2 write
 * (16-byte accesses), 3 scalar copy, 4 scalar copy with every store shifted
 * by one element, 5 the population access pattern (27 arrays, one cell per
 * thread, each value read and written once), 6 the same with the pull
 * gather's +-1 x offsets on 18 of the 27 loads.  `bytes` per buffer (4 KiB to
 * 8 GiB); *gbs = bytes moved / time.  After the timed launches the device
 * checks every word the probe claims to have moved (the source holds known
 * words); a mismatch returns CULBM_ERR_STATE.  mode | CULBM_PROBE_CHECK_WRONG
 * runs the same probe but checks it against a deliberately wrong expectation
 * (one word off), which must fail: the test of the check itself. */
#define CULBM_PROBE_CHECK_WRONG 0x100
int culbm_probe_bandwidth(int32_t device, int32_t mode, int64_t bytes, int32_t iters, float* gbs);

#ifdef __cplusplus
}
#endif

#endif /* CULBM_H */
