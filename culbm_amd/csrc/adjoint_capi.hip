// Host side of the discrete adjoint: checkpoint schedules and the reverse sweep
// (culbm_adjoint, culbm_adjoint_plan in include/culbm.h).  Part of the capi.hip
// translation unit, after the context helpers it uses.
//
// Schedules.  A first forward sweep runs f^0 .. f^N (and J); states are kept in
// "slots" of the dense pull layout.  The reverse sweep needs f^t for t = N-1 .. 0
// and recomputes what no slot holds:
//  * every-k: the sweep keeps f^0, f^k, f^2k, ...; each segment is recomputed
//    once from its checkpoint into k-1 segment slots and reversed, except the
//    last, whose states the sweep itself leaves in the segment slots;
//  * binomial: with s slots, reverse(a, b) stores a state at the split m that
//    minimises the forward steps of reverse(m, b) with one slot fewer plus
//    reverse(a, m) (Griewank's binomial checkpointing, "revolve"); the
//    cost of reversing l steps from a held state with s slots is
//    t(l, s) = r l - beta(s + 1, r - 1), beta(s, r) = C(s + r, s), r minimal
//    with beta(s, r) >= l.  The splits of the ranges that end at N lie on the
//    first sweep's path, so the sweep stores them instead of recomputing them.

namespace {

double adj_beta(int64_t s, int64_t r) {
  if (r < 0) return 0.0;
  double v = 1.0;
  for (int64_t k = 1; k <= r; ++k) {
    v = v * double(s + k) / double(k);
    if (v > 1e18) return 1e18;
  }
  return v;
}

// Forward steps to reverse l steps from a held state with `extra` more slots.
int64_t adj_cost(int64_t l, int64_t extra) {
  if (l <= 1) return 0;
  if (extra <= 0) return l * (l - 1) / 2;
  if (extra >= l - 1) return l - 1;
  const int64_t s = extra + 1;
  int64_t r = 0;
  while (adj_beta(s, r) < double(l)) ++r;
  return r * l - int64_t(adj_beta(s + 1, r - 1));
}

int32_t adj_split(int32_t l, int32_t extra) {
  if (extra >= l - 1) return 1;
  int32_t best = 1;
  int64_t bc = INT64_MAX;
  for (int32_t m = 1; m < l; ++m) {
    const int64_t c = m + adj_cost(l - m, extra - 1) + adj_cost(m, extra);
    if (c < bc) {
      bc = c;
      best = m;
    }
  }
  return best;
}

struct AdjOp {
  int32_t kind;   // 0 advance, 1 reverse
  int32_t t;      // advance: first state; reverse: step reversed (needs f^t)
  int32_t steps;  // advance: steps to run
  int32_t src;    // slot holding f^t, or -1: the state the last advance left in a work buffer
  int32_t dst;    // advance: slot receiving the result, or -1: a work buffer
};

struct AdjPlan {
  int32_t nsteps = 0;
  int32_t slots = 0;                                // slot 0 holds f^0
  std::vector<std::pair<int32_t, int32_t>> sweep;   // (t, slot) stored by the first sweep
  std::vector<AdjOp> ops;
  int64_t forward = 0;                              // forward steps, the sweep included
};

struct AdjPlanner {
  AdjPlan& p;
  std::vector<int32_t> free;
  bool chain;

  // f^a is in slot sa; emits the reverse steps b-1 .. a.
  void rev(int32_t a, int32_t b, int32_t sa) {
    while (b - a > 1) {
      if (free.empty()) {
        for (int32_t t = b - 1; t > a; --t) {
          p.ops.push_back({0, a, t - a, sa, -1});
          p.forward += t - a;
          p.ops.push_back({1, t, 0, -1, -1});
        }
        break;
      }
      const int32_t m = a + adj_split(b - a, int32_t(free.size()));
      const int32_t sm = free.back();
      free.pop_back();
      if (chain && b == p.nsteps) {
        p.sweep.push_back({m, sm});
      } else {
        p.ops.push_back({0, a, m - a, sa, sm});
        p.forward += m - a;
      }
      rev(m, b, sm);
      free.push_back(sm);
      b = m;
    }
    p.ops.push_back({1, a, 0, sa, -1});
  }
};

AdjPlan adj_make_plan(int32_t nsteps, int32_t every, int32_t slots) {
  AdjPlan p;
  p.nsteps = nsteps;
  p.forward = nsteps;
  if (every > 0) {
    const int32_t C = (nsteps + every - 1) / every;
    const int32_t seg = std::min(every, nsteps) - 1;
    p.slots = C + seg;
    AdjPlanner pl{p, {}, true};
    for (int32_t s = C + seg - 1; s >= C; --s) pl.free.push_back(s);
    for (int32_t j = 1; j < C; ++j) p.sweep.push_back({j * every, j});
    for (int32_t j = C - 1; j >= 0; --j) pl.rev(j * every, std::min(nsteps, (j + 1) * every), j);
  } else {
    const int32_t S = std::max(1, std::min(slots, nsteps));
    p.slots = S;
    AdjPlanner pl{p, {}, true};
    for (int32_t s = S - 1; s >= 1; --s) pl.free.push_back(s);
    pl.rev(0, nsteps, 0);
  }
  std::sort(p.sweep.begin(), p.sweep.end());
  return p;
}

// Symbolic run of a plan: every op reads the state it expects, the reverse
// steps come exactly in the order N-1 .. 0.
bool adj_check_plan(const AdjPlan& p) {
  std::vector<int32_t> held(size_t(p.slots), -1);
  held[0] = 0;
  for (const auto& s : p.sweep) {
    if (s.second <= 0 || s.second >= p.slots || s.first <= 0 || s.first >= p.nsteps) return false;
    held[size_t(s.second)] = s.first;
  }
  int32_t work = -1, next = p.nsteps - 1;
  for (const AdjOp& o : p.ops) {
    const int32_t have = o.src >= 0 ? held[size_t(o.src)] : work;
    if (have != o.t) return false;
    if (o.kind == 0) {
      if (o.steps < 1) return false;
      if (o.dst >= 0) held[size_t(o.dst)] = o.t + o.steps;
      else work = o.t + o.steps;
    } else {
      if (o.t != next) return false;
      --next;
    }
  }
  return next == -1;
}

// Device workspace of one culbm_adjoint call, released on every exit path.
struct AdjBuffers {
  culbm_ctx* ctx;
  std::vector<void*> ptrs;
  int64_t bytes = 0;
  explicit AdjBuffers(culbm_ctx* c) : ctx(c) {}
  ~AdjBuffers() {
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    for (void* q : ptrs) (void)hipFree(q);
  }
  template <class T>
  int get(T** out, size_t nbytes) {
    void* q = nullptr;
    const size_t b = std::max<size_t>(nbytes, 256);
    if (hipMalloc(&q, b) != hipSuccess) {
      (void)hipGetLastError();
      return fail(ctx, CULBM_ERR_HIP, "adjoint workspace: hipMalloc of %zu bytes failed", b);
    }
    ptrs.push_back(q);
    bytes += int64_t(b);
    *out = static_cast<T*>(q);
    return CULBM_OK;
  }
};

// Event pairs per phase (0 sweep, 1 recompute, 2 reverse), summed at the end:
// one pair per run of consecutive operations of the same phase (a pair around
// every launch would add its own gaps between the launches it measures).
struct AdjTimer {
  culbm_ctx* ctx;
  bool on;
  std::vector<std::pair<hipEvent_t, int>> ev;
  int open = -1;  // phase of the pair begun last and not yet ended
  ~AdjTimer() {
    for (auto& e : ev) (void)hipEventDestroy(e.first);
  }
  int record(int phase) {
    hipEvent_t e;
    HIPCHK(ctx, hipEventCreate(&e));
    ev.push_back({e, phase});
    HIPCHK(ctx, hipEventRecord(e, ctx->stream));
    return CULBM_OK;
  }
  int enter(int phase) {  // the work enqueued next belongs to `phase`
    if (!on || open == phase) return CULBM_OK;
    if (open >= 0)
      if (const int rc = record(open)) return rc;
    open = phase;
    return record(phase);
  }
  int close() {  // ends the running pair
    if (!on || open < 0) return CULBM_OK;
    const int p = open;
    open = -1;
    return record(p);
  }
  int sum(float* ms) {  // pairs (begin, end) in order of recording
    for (int k = 0; k < 3; ++k) ms[k] = 0.0f;
    for (size_t i = 0; i + 1 < ev.size(); i += 2) {
      float t = 0.0f;
      HIPCHK(ctx, hipEventElapsedTime(&t, ev[i].first, ev[i + 1].first));
      ms[ev[i].second] += t;
    }
    return CULBM_OK;
  }
};

template <class S>
int adjoint_impl(culbm_ctx* ctx, const culbm_adjoint_spec& sp, const double* weight,
                 culbm_adjoint_stats* st, double* const grad[4], float* grad_f0) {
  const culbm_config& c = ctx->cfg;
  const int64_t n = ctx->n;
  const int32_t N = sp.nsteps;
  hipEvent_t t_begin = nullptr, t_end = nullptr;
  HIPCHK(ctx, hipEventCreate(&t_begin));
  HIPCHK(ctx, hipEventCreate(&t_end));
  struct EvGuard {
    hipEvent_t a, b;
    ~EvGuard() {
      (void)hipEventDestroy(a);
      (void)hipEventDestroy(b);
    }
  } evg{t_begin, t_end};
  HIPCHK(ctx, hipEventRecord(t_begin, ctx->stream));

  // dense pull layout over the whole domain
  StepArgs a = base_args(ctx);
  a.px = c.nx;
  a.py = c.ny;
  a.gcol = 0;
  a.dstride = n;
  set_range(a, 0, c.nz, true, ctx->bpp == 2);

  // host-side index lists
  std::vector<uint8_t> fl(static_cast<size_t>(n));
  HIPCHK(ctx, hipMemcpyAsync(fl.data(), ctx->sflag, size_t(n), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  std::vector<uint32_t> aux(size_t(n), 0u), equ_cells, obj_idx;
  std::vector<double> obj_w;
  std::vector<float> wobj(size_t(n), 0.0f);
  int64_t n_keep = 0;
  for (int64_t i = 0; i < n; ++i) {
    const uint8_t f = fl[size_t(i)];
    if ((f & kStore) && (f & kEqu)) {
      aux[size_t(i)] = uint32_t(equ_cells.size());
      equ_cells.push_back(uint32_t(i));
    } else if (!(f & kStore) && sp.want_f0) {
      aux[size_t(i)] = uint32_t(n_keep++);
    }
    const double w = weight[i];
    if (!std::isfinite(w)) return fail(ctx, CULBM_ERR_INVALID, "weight[%lld] is not finite", (long long)i);
    if (w != 0.0) {
      obj_idx.push_back(uint32_t(i));
      obj_w.push_back(w);
      wobj[size_t(i)] = float(w);
    }
  }
  const int64_t n_equ = int64_t(equ_cells.size());

  // memory plan
  const size_t state = size_t(27) * size_t(n) * sizeof(typename S::T);
  const size_t mu_b = size_t(27) * size_t(n) * 4;
  const uint32_t obj_blk = uint32_t(std::max<size_t>((obj_idx.size() + kBlock - 1) / kBlock, 1));
  const size_t fixed = 2 * mu_b + size_t(n) * 4 + size_t(4 * n_equ) * 8 +
                       size_t(27 * n_keep) * 4 + size_t(a.nblk) * 8 + size_t(obj_blk) * 8 +
                       obj_idx.size() * 12 + 2 * state + size_t(n) * 4 + 20 * 256;
  int64_t cap = sp.mem_cap;
  if (cap == 0) {
    size_t fr = 0, tot = 0;
    HIPCHK(ctx, hipMemGetInfo(&fr, &tot));
    cap = int64_t(double(fr) * 0.9);
  }
  int32_t slots;
  if (sp.checkpoint_every > 0) {
    const int32_t C = (N + sp.checkpoint_every - 1) / sp.checkpoint_every;
    slots = C + std::min(sp.checkpoint_every, N) - 1;
    const int64_t need = int64_t(fixed) + int64_t(slots) * int64_t(state);
    if (need > cap)
      return fail(ctx, CULBM_ERR_INVALID,
                  "checkpoints every %d steps need %lld bytes, above the memory cap of %lld",
                  sp.checkpoint_every, (long long)need, (long long)cap);
  } else {
    const int64_t room = cap - int64_t(fixed);
    slots = room > 0 ? int32_t(std::min<int64_t>(room / int64_t(state), N)) : 0;
    if (sp.max_slots > 0) slots = std::min(slots, sp.max_slots);
    if (slots < 1)
      return fail(ctx, CULBM_ERR_INVALID,
                  "memory cap of %lld bytes is below the %lld the adjoint needs with one checkpoint",
                  (long long)cap, (long long)(int64_t(fixed) + int64_t(state)));
  }
  const AdjPlan plan = adj_make_plan(N, sp.checkpoint_every, slots);
  if (!adj_check_plan(plan)) return fail(ctx, CULBM_ERR_STATE, "inconsistent checkpoint schedule");

  AdjBuffers mem(ctx);
  std::vector<void*> slot(size_t(plan.slots), nullptr);
  void* work[2] = {nullptr, nullptr};
  float* mu[2] = {nullptr, nullptr};
  uint32_t *d_aux = nullptr, *d_idx = nullptr;
  double *d_grad4 = nullptr, *d_tau = nullptr, *d_j = nullptr, *d_w = nullptr;
  float *d_carry = nullptr, *d_wobj = nullptr;
  int rc = CULBM_OK;
  for (auto& s : slot)
    if ((rc = mem.get(&s, state))) return rc;
  for (auto& w : work)
    if ((rc = mem.get(&w, state))) return rc;
  for (auto& m : mu)
    if ((rc = mem.get(&m, mu_b))) return rc;
  if ((rc = mem.get(&d_aux, size_t(n) * 4)) || (rc = mem.get(&d_grad4, size_t(4 * n_equ) * 8)) ||
      (rc = mem.get(&d_tau, size_t(a.nblk) * 8)) || (rc = mem.get(&d_j, size_t(obj_blk) * 8)) ||
      (rc = mem.get(&d_idx, obj_idx.size() * 4)) || (rc = mem.get(&d_w, obj_w.size() * 8)))
    return rc;
  if (sp.want_f0 && (rc = mem.get(&d_carry, size_t(27 * n_keep) * 4))) return rc;
  if ((rc = mem.get(&d_wobj, size_t(n) * 4))) return rc;
  HIPCHK(ctx, hipMemcpyAsync(d_wobj, wobj.data(), size_t(n) * 4, hipMemcpyHostToDevice, ctx->stream));

  HIPCHK(ctx, hipMemcpyAsync(d_aux, aux.data(), size_t(n) * 4, hipMemcpyHostToDevice, ctx->stream));
  if (!obj_idx.empty()) {
    HIPCHK(ctx, hipMemcpyAsync(d_idx, obj_idx.data(), obj_idx.size() * 4, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(ctx, hipMemcpyAsync(d_w, obj_w.data(), obj_w.size() * 8, hipMemcpyHostToDevice, ctx->stream));
  }
  if ((rc = zero_device(ctx, d_grad4, std::max<size_t>(size_t(4 * n_equ) * 8, 4))) ||
      (rc = zero_device(ctx, d_tau, size_t(a.nblk) * 8)) ||
      (rc = zero_device(ctx, d_j, size_t(obj_blk) * 8)))
    return rc;
  if (d_carry && (rc = zero_device(ctx, d_carry, std::max<size_t>(size_t(27 * n_keep) * 4, 4)))) return rc;

  // f^0 into slot 0; the cells without STORE never change, so every other
  // state buffer takes their values once and keeps them.
  if ((rc = download_plain<S>(ctx, mu[0]))) return rc;
  hipLaunchKernelGGL((k_plain_to_pull<S>), dim3(a.nblk), dim3(kBlock), 0, ctx->stream, a, mu[0], slot[0]);
  std::vector<void*> others(slot.begin() + 1, slot.end());
  others.push_back(work[0]);
  others.push_back(work[1]);
  for (void* b : others)
    hipLaunchKernelGGL((k_sync_nostore<S>), dim3(a.nblk), dim3(kBlock), 0, ctx->stream, a, slot[0], b);
  HIPCHK(ctx, hipGetLastError());

  int32_t launches = 0;
  // f^{t_out} from f^{t_out - 1}: the state numbers continue the context's step
  // counter, so stochastic rounding reproduces the context's own trajectory
  const uint32_t s0 = uint32_t(ctx->step);
  auto step = [&](const void* src, void* dst, int32_t t_out) {
    StepArgs s = a;
    s.src = src;
    s.dst = dst;
    s.state = s0 + uint32_t(t_out);
    bool sr = false;
    if constexpr (S::kBytes == 2) sr = ctx->sr;
    if (sr)
      hipLaunchKernelGGL((k_pull<S, true>), dim3(s.nblk), dim3(kBlock), 0, ctx->stream, s);
    else
      hipLaunchKernelGGL((k_pull<S, false>), dim3(s.nblk), dim3(kBlock), 0, ctx->stream, s);
    ++launches;
  };
  // f^{t_out} from f^{t_out - 2} in one launch (two steps through LDS, the same
  // bits as two plain steps), where the state in between is not needed
  auto step2 = [&](const void* src, void* dst, int32_t t_out) {
    FusedArgs fa{};
    fa.a = a;
    fa.a.src = src;
    fa.a.dst = dst;
    fa.a.state = s0 + uint32_t(t_out) - 1u;  // the launch produces states state and state + 1
    fa.tiles_x = ctx->tiles_x;
    fa.tiles_y = ctx->tiles_y;
    fa.zchunk = ctx->fzchunk;
    fa.nchunk = ctx->nchunk;
    const uint32_t nblk = uint32_t(ctx->tiles_x) * ctx->tiles_y * ctx->nchunk;
    bool sr = false;
    if constexpr (S::kBytes == 2) sr = ctx->sr;
    if (sr)
      hipLaunchKernelGGL((k_fused2<S, true>), dim3(nblk), dim3(FusedShape<S::kBytes>::BLOCK), 0, ctx->stream, fa);
    else
      hipLaunchKernelGGL((k_fused2<S, false>), dim3(nblk), dim3(FusedShape<S::kBytes>::BLOCK), 0, ctx->stream, fa);
    ++launches;
  };
  auto in_window = [&](int32_t t) { return t >= sp.t_begin && t <= sp.t_end; };
  // the first sweep accumulates J; the reverse sweep adds the objective's gradient to mu
  auto objective = [&](const void* f, float* m) {
    if (obj_idx.empty()) return;
    ObjArgs o{};
    o.a = a;
    o.a.src = f;
    o.idx = d_idx;
    o.w = d_w;
    o.count = uint32_t(obj_idx.size());
    o.mu = m;
    o.j_acc = m ? nullptr : d_j;
    hipLaunchKernelGGL((k_objective<S>), dim3(obj_blk), dim3(kBlock), 0, ctx->stream, o);
    ++launches;
  };
  // a work buffer that is not `busy`
  auto free_work = [&](const void* busy) { return work[0] != busy ? work[0] : work[1]; };

  AdjTimer timer{ctx, sp.timing != 0, {}};
  // ---- first sweep: J, the sweep checkpoints, f^N
  if ((rc = timer.enter(0))) return rc;
  const void* cur = slot[0];
  size_t next_store = 0;
  auto stored_at = [&](int32_t t) { return next_store < plan.sweep.size() && plan.sweep[next_store].first == t; };
  for (int32_t t = 0;;) {
    if (in_window(t)) objective(cur, nullptr);
    if (t == N) break;
    // two steps in one launch when f^{t+1} is neither evaluated nor kept
    const int32_t n = (t + 2 <= N && !in_window(t + 1) && !stored_at(t + 1)) ? 2 : 1;
    void* dst = stored_at(t + n) ? slot[size_t(plan.sweep[next_store++].second)] : free_work(cur);
    if (n == 2)
      step2(cur, dst, t + 2);
    else
      step(cur, dst, t + 1);
    cur = dst;
    t += n;
  }
  if ((rc = timer.close())) return rc;
  HIPCHK(ctx, hipGetLastError());
  // the context moves on to f^N
  hipLaunchKernelGGL((k_pull_to_plain<S>), dim3(a.nblk), dim3(kBlock), 0, ctx->stream, a, cur, mu[1]);
  for (void* b : ctx->buf)
    if (b && (rc = zero_device(ctx, b, ddf_bytes(ctx)))) return rc;
  if ((rc = upload_plain<S>(ctx, mu[1]))) return rc;
  ctx->step += N;
  // mu^N: the objective's own term
  if ((rc = zero_device(ctx, mu[0], mu_b))) return rc;
  if (in_window(N)) objective(cur, mu[0]);
  int mc = 0;

  // ---- reverse sweep
  AdjArgs j{};
  j.a = a;
  j.aux = d_aux;
  j.grad4 = d_grad4;
  j.n_equ = n_equ;
  j.carry = d_carry;
  j.n_keep = n_keep;
  j.tau_acc = d_tau;
  const void* last_work = nullptr;
  for (const AdjOp& o : plan.ops) {
    const void* src = o.src >= 0 ? slot[size_t(o.src)] : last_work;
    if (o.kind == 0) {
      if ((rc = timer.enter(1))) return rc;
      const void* p = src;
      for (int32_t s = 0; s < o.steps;) {
        const int32_t n = s + 2 <= o.steps ? 2 : 1;  // pairs of steps in one launch
        void* dst = (s + n == o.steps && o.dst >= 0) ? slot[size_t(o.dst)] : free_work(p);
        if (n == 2)
          step2(p, dst, o.t + s + 2);
        else
          step(p, dst, o.t + s + 1);
        p = dst;
        s += n;
      }
      if (o.dst < 0) last_work = p;
    } else {
      if ((rc = timer.enter(2))) return rc;
      j.a.src = src;
      j.mu_in = mu[mc];
      j.mu_out = mu[mc ^ 1];
      j.wobj = in_window(o.t) ? d_wobj : nullptr;
      hipLaunchKernelGGL((k_adjoint<S>), dim3(a.nblk), dim3(kBlock), 0, ctx->stream, j);
      ++launches;
      mc ^= 1;
    }
  }
  if ((rc = timer.close())) return rc;
  HIPCHK(ctx, hipGetLastError());

  // ---- gradients back to the host
  if (sp.want_f0) {
    j.mu_in = mu[mc];
    hipLaunchKernelGGL(k_adjoint_final, dim3(a.nblk), dim3(kBlock), 0, ctx->stream, j, mu[mc ^ 1]);
    ++launches;
    HIPCHK(ctx, hipGetLastError());
    HIPCHK(ctx, hipMemcpyAsync(grad_f0, mu[mc ^ 1], mu_b, hipMemcpyDeviceToHost, ctx->stream));
  }
  std::vector<double> g4(size_t(4 * n_equ)), tau(a.nblk), jj(obj_blk);
  if (n_equ)
    HIPCHK(ctx, hipMemcpyAsync(g4.data(), d_grad4, g4.size() * 8, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipMemcpyAsync(tau.data(), d_tau, tau.size() * 8, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipMemcpyAsync(jj.data(), d_j, jj.size() * 8, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipEventRecord(t_end, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  for (int q = 0; q < 4; ++q) {
    if (!grad[q]) continue;
    std::fill(grad[q], grad[q] + n, 0.0);
    for (int64_t e = 0; e < n_equ; ++e) grad[q][equ_cells[size_t(e)]] = g4[size_t(q * n_equ + e)];
  }
  std::memset(st, 0, sizeof(*st));
  for (double v : jj) st->objective += v;
  for (double v : tau) st->grad_inv_tau += v;
  st->forward_steps = plan.forward;
  st->device_bytes = mem.bytes;
  st->slots = plan.slots;
  st->launches = launches;
  if (sp.timing) {
    float ms[3];
    if ((rc = timer.sum(ms))) return rc;
    st->ms_sweep = ms[0];
    st->ms_recompute = ms[1];
    st->ms_reverse = ms[2];
  }
  HIPCHK(ctx, hipEventElapsedTime(&st->ms_total, t_begin, t_end));
  return CULBM_OK;
}

}  // namespace

extern "C" {

int culbm_adjoint(culbm_ctx* ctx, const culbm_adjoint_spec* spec, const double* weight,
                  culbm_adjoint_stats* stats, double* grad_rho, double* grad_ux,
                  double* grad_uy, double* grad_uz, float* grad_f0) {
  if (!ctx) return fail(nullptr, CULBM_ERR_INVALID, "ctx is NULL");
  if (!spec || !weight || !stats) return fail(ctx, CULBM_ERR_INVALID, "spec, weight and stats are required");
  const culbm_adjoint_spec& s = *spec;
  if (s.nsteps < 1) return fail(ctx, CULBM_ERR_INVALID, "adjoint nsteps must be >= 1, got %d", s.nsteps);
  if (s.t_begin < 0 || s.t_begin > s.t_end || s.t_end > s.nsteps)
    return fail(ctx, CULBM_ERR_INVALID, "objective window [%d, %d] must satisfy 0 <= begin <= end <= %d",
                s.t_begin, s.t_end, s.nsteps);
  if (s.checkpoint_every < 0) return fail(ctx, CULBM_ERR_INVALID, "checkpoint_every must be >= 0");
  if (s.mem_cap < 0 || s.max_slots < 0)
    return fail(ctx, CULBM_ERR_INVALID, "mem_cap and max_slots must be >= 0");
  if (s.want_f0 && !grad_f0) return fail(ctx, CULBM_ERR_INVALID, "want_f0 needs a grad_f0 array");
  if (size_t(27) * size_t(ctx->n) * 4 >= (size_t(1) << 32))
    return fail(ctx, CULBM_ERR_UNSUPPORTED, "adjoint buffers need 27*4*cells < 2^32 bytes");
  HIPCHK(ctx, hipSetDevice(ctx->cfg.device));
  if (int rc = flush_dump(ctx)) return rc;
  double* g[4] = {grad_rho, grad_ux, grad_uy, grad_uz};
  return by_storage(ctx->cfg.storage, [&](auto st) {
    return adjoint_impl<decltype(st)>(ctx, s, weight, stats, g, s.want_f0 ? grad_f0 : nullptr);
  });
}

int culbm_adjoint_plan(int32_t nsteps, int32_t checkpoint_every, int32_t slots,
                       int64_t* forward_steps, int32_t* slots_used) {
  if (nsteps < 1 || checkpoint_every < 0 || (checkpoint_every == 0 && slots < 1))
    return fail(nullptr, CULBM_ERR_INVALID, "plan needs nsteps >= 1 and every > 0 or slots >= 1");
  const AdjPlan p = adj_make_plan(nsteps, checkpoint_every, slots);
  if (!adj_check_plan(p)) return fail(nullptr, CULBM_ERR_STATE, "inconsistent checkpoint schedule");
  if (forward_steps) *forward_steps = p.forward;
  if (slots_used) *slots_used = p.slots;
  return CULBM_OK;
}

}  // extern "C"
