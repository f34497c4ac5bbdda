// Discrete adjoint of the pull step (no reference counterpart).
//
// Forward: f^{t+1}(x) = Phi_x(g^t(x)) with g^t = P f^t the clamped pull and
// Phi_x the LOAD/EQU/BB/COLLIDE/STORE update of one cell; cells without STORE
// keep their populations.  Objective
//   J = sum_{t in [t_begin, t_end]} sum_x w(x) u_x(g^t(x)),
// the weighted x-velocity each cell reports after t steps.  With
// mu^t = dJ/dg^t and lambda^t = dJ/df^t:
//   lambda^{t+1}(x) = sum of mu^{t+1} over the cells that pull x's populations
//                     (the transposed pull, face clamp included)
//                     + lambda^{t+2}(x) for cells without STORE,
//   mu^t(x)         = (dPhi_x/dg)^T lambda^{t+1}(x) + [t in window] w(x) du_x/dg.
// A cell without STORE never changes, so its lambda only reaches dJ/df^0; the
// running sum ("carry") is kept only when that gradient is asked for.
// Parameter gradients: the EQU inputs (rho, u) of every EQU cell and invTau.
//
// Layout: every buffer here uses the dense pull layout (px = nx, py = ny,
// direction stride = cells); populations in the storage format S, mu in FP32.
// Gradients accumulate in FP64: one accumulator per EQU cell and input, one per
// workgroup for invTau and J, each summed in a fixed order, so a run is
// bit-reproducible.
// Included by capi.hip after lbm_kernels.hip.

namespace culbm {

struct AdjArgs {
  StepArgs a;            // geometry, flags, EQU inputs; a.src = f^t
  const float* mu_in;    // mu^{t+1}, [27][dstride]
  float* mu_out;         // mu^t
  const uint32_t* aux;   // per cell: rank among EQU+STORE cells, or among cells without STORE
  double* grad4;         // [4][n_equ]: dJ/d(rho, ux, uy, uz) of the EQU inputs
  int64_t n_equ;
  float* carry;          // [27][n_keep] running lambda of the cells without STORE, or null
  int64_t n_keep;
  double* tau_acc;       // one accumulator per workgroup
  const float* wobj;     // objective weights (FP32, per cell) when t is in the window, else null
};

// Sum of mu_i over the consumers of population i of cell k.  Along each axis a
// the consumer coordinate is k_a + c_a when that cell exists, and k_a itself
// when k_a - c_a lies outside the domain (a face cell clamps its pull to itself).
// This gathers the all-shifted consumer for every cell (its address clamped
// and its value dropped where it would leave the domain); on the x faces the
// consumer clamped along x only (the element next to the shifted one, in the
// same cache line); on the y and z faces the consumers clamped along y or z.
// Every load is independent of the others, so none waits on another.
struct ConsumerOffsets {
  int32_t ox[3], oy[3], oz[3];
  bool vx[3], vy[3], vz[3];  // consumer exists along the axis, for c = -1, 0, +1
};

__device__ __forceinline__ ConsumerOffsets consumer_offsets(const StepArgs& a, const Cell& k) {
  ConsumerOffsets o;
  const int32_t row = a.px * 4, plane = a.px * a.py * 4;
  const bool xm = k.x > 0, xp = k.x < a.nx - 1, ym = k.y > 0, yp = k.y < a.ny - 1,
             zm = k.z > 0, zp = k.z < a.nz - 1;
  o.ox[0] = xm ? -4 : 0;     o.ox[1] = 0; o.ox[2] = xp ? 4 : 0;
  o.oy[0] = ym ? -row : 0;   o.oy[1] = 0; o.oy[2] = yp ? row : 0;
  o.oz[0] = zm ? -plane : 0; o.oz[1] = 0; o.oz[2] = zp ? plane : 0;
  o.vx[0] = xm; o.vx[1] = true; o.vx[2] = xp;
  o.vy[0] = ym; o.vy[1] = true; o.vy[2] = yp;
  o.vz[0] = zm; o.vz[1] = true; o.vz[2] = zp;
  return o;
}

__device__ __forceinline__ void adj_gather_consumers(const AdjArgs& j, const Cell& k,
                                                     const float* mu, float* lo) {
  const StepArgs& a = j.a;
  const char* base = reinterpret_cast<const char*>(mu);
  const int64_t dsb = a.dstride * 4;
  const uint32_t lb = k.l * 4u;
  const ConsumerOffsets o = consumer_offsets(a, k);
#pragma unroll
  for (int i = 0; i < 27; ++i) {
    const int cx = cx_of(i), cy = cy_of(i), cz = cz_of(i);
    const float v = ld_at<float>(base + i * dsb, lb + uint32_t(o.ox[cx + 1] + o.oy[cy + 1] + o.oz[cz + 1]));
    lo[i] = (o.vx[cx + 1] && o.vy[cy + 1] && o.vz[cz + 1]) ? v : 0.0f;
  }
  if (!o.vx[0] || !o.vx[2]) {  // x face: the cell itself consumes along x
    // every lane of the wave loads (all addresses are in the domain) and keeps
    // the value where it applies: one wait for the batch instead of one
    // round trip per conditional load
    float e[27];
#pragma unroll
    for (int i = 0; i < 27; ++i) {
      const int cy = cy_of(i), cz = cz_of(i);
      if (cx_of(i) != 0) e[i] = ld_at<float>(base + i * dsb, lb + uint32_t(o.oy[cy + 1] + o.oz[cz + 1]));
    }
#pragma unroll
    for (int i = 0; i < 27; ++i) {
      const int cx = cx_of(i), cy = cy_of(i), cz = cz_of(i);
      if (cx == 0) continue;
      if (!o.vx[1 - cx] && o.vy[cy + 1] && o.vz[cz + 1]) lo[i] += e[i];  // x - c_x leaves the domain
    }
  }
  if (!(o.vy[0] && o.vy[2] && o.vz[0] && o.vz[2])) {
    // y or z face (whole rows and planes, so interior waves skip this): the
    // consumers clamped on a subset of the axes that contains y or z
    // in three batches of nine directions (one c_z each), loads first
#pragma unroll
    for (int g = 0; g < 3; ++g) {
      float e[9][6];
#pragma unroll
      for (int i = 9 * g; i < 9 * g + 9; ++i) {
        const int cx = cx_of(i), cy = cy_of(i), cz = cz_of(i);
#pragma unroll
        for (int m = 2; m < 8; ++m) {  // bit a of m: the consumer is the cell itself along axis a
          if (((m & 1) && cx == 0) || ((m & 2) && cy == 0) || ((m & 4) && cz == 0)) continue;
          const int32_t off = ((m & 1) ? 0 : o.ox[cx + 1]) + ((m & 2) ? 0 : o.oy[cy + 1]) +
                              ((m & 4) ? 0 : o.oz[cz + 1]);
          e[i - 9 * g][m - 2] = ld_at<float>(base + i * dsb, lb + uint32_t(off));
        }
      }
#pragma unroll
      for (int i = 9 * g; i < 9 * g + 9; ++i) {
        const int cx = cx_of(i), cy = cy_of(i), cz = cz_of(i);
        const bool sx = cx != 0 && !o.vx[1 - cx], sy = cy != 0 && !o.vy[1 - cy], sz = cz != 0 && !o.vz[1 - cz];
#pragma unroll
        for (int m = 2; m < 8; ++m) {
          if (((m & 1) && cx == 0) || ((m & 2) && cy == 0) || ((m & 4) && cz == 0)) continue;
          const bool ok = ((m & 1) ? sx : o.vx[cx + 1]) && ((m & 2) ? sy : o.vy[cy + 1]) &&
                          ((m & 4) ? sz : o.vz[cz + 1]);
          if (ok) lo[i] += e[i - 9 * g][m - 2];
        }
      }
    }
  }
}

// Weighted moments of lam: M0 = sum w_i lam_i, M1_a = sum w_i lam_i c_ia,
// M2_ab = sum w_i lam_i c_ia c_ib.  Every derivative of the equilibrium below is
// quadratic in c_i, so these ten sums carry all the lam-dependence.
struct LamMoments {
  float m0, x, y, z, xx, yy, zz, xy, xz, yz;
};

__device__ __forceinline__ LamMoments lam_moments(const float* lam) {
  LamMoments m{};
#pragma unroll
  for (int i = 0; i < 27; ++i) {
    constexpr float kw[4] = {kW0, kWS, kWE, kWC};
    const int cx = cx_of(i), cy = cy_of(i), cz = cz_of(i);
    const float lw = lam[i] * kw[(cx != 0) + (cy != 0) + (cz != 0)];
    m.m0 += lw;
    if (cx != 0) { m.x += cx > 0 ? lw : -lw; m.xx += lw; }
    if (cy != 0) { m.y += cy > 0 ? lw : -lw; m.yy += lw; }
    if (cz != 0) { m.z += cz > 0 ? lw : -lw; m.zz += lw; }
    if (cx != 0 && cy != 0) m.xy += cx * cy > 0 ? lw : -lw;
    if (cx != 0 && cz != 0) m.xz += cx * cz > 0 ? lw : -lw;
    if (cy != 0 && cz != 0) m.yz += cy * cz > 0 ? lw : -lw;
  }
  return m;
}

// u^T M2 u
__device__ __forceinline__ float umu(const LamMoments& m, float ux, float uy, float uz) {
  return m.xx * ux * ux + m.yy * uy * uy + m.zz * uz * uz +
         2.0f * (m.xy * ux * uy + m.xz * ux * uz + m.yz * uy * uz);
}

// In place: lam <- (dC/dh)^T lam for the BGK collision C at pre-collision h
// (shifted populations, rho = 1 + sum h).  Returns sum_i lam_i (feq_i - h_i),
// the derivative of lam . C(h) by invTau.  With m = sum_i c_i h_i (= rho u) the
// equilibrium is a function of (rho, m):
//   d feq_i / d rho |m = w_i (1 - 4.5 (c_i.u)^2 + 1.5 |u|^2)
//   d feq_i / d m      = w_i (3 c_i + 9 (c_i.u) c_i - 3 u)
// and d rho / d h_j = 1, d m / d h_j = c_j, so
//   (dC/dh)^T lam = (1 - invTau) lam + invTau (A + B . c_j),
//   A = M0 (1 + 1.5 |u|^2) - 4.5 u^T M2 u,  B = 3 M1 + 9 M2 u - 3 M0 u.
// h is consumed by the first pass (moments and lam . h).
// With bb the collision acts on the bounced state h'_i = h_{26-i}: the same rho,
// the opposite momentum, and lam . h' = sum lam_i h_{26-i}, so h is never permuted.
__device__ __forceinline__ float adj_collide(float inv_tau, const float* h, float* lam, bool bb) {
  float rho = 1.0f, mx = 0.0f, my = 0.0f, mz = 0.0f, l0 = 0.0f, l1 = 0.0f;
#pragma unroll
  for (int i = 0; i < 27; ++i) {
    rho += h[i];
    if (cx_of(i) != 0) mx += cx_of(i) > 0 ? h[i] : -h[i];
    if (cy_of(i) != 0) my += cy_of(i) > 0 ? h[i] : -h[i];
    if (cz_of(i) != 0) mz += cz_of(i) > 0 ? h[i] : -h[i];
    l0 += lam[i] * h[i];
    l1 += lam[i] * h[26 - i];
  }
  const float lh = bb ? l1 : l0;
  const float inv = bb ? -1.0f / rho : 1.0f / rho;
  const float ux = mx * inv, uy = my * inv, uz = mz * inv;
  const float u2 = ux * ux + uy * uy + uz * uz;
  const LamMoments m = lam_moments(lam);
  const float q = umu(m, ux, uy, uz);
  const float mu1 = m.x * ux + m.y * uy + m.z * uz;
  const float A = m.m0 * (1.0f + 1.5f * u2) - 4.5f * q;
  const float bx = 3.0f * m.x + 9.0f * (m.xx * ux + m.xy * uy + m.xz * uz) - 3.0f * m.m0 * ux;
  const float by = 3.0f * m.y + 9.0f * (m.xy * ux + m.yy * uy + m.yz * uz) - 3.0f * m.m0 * uy;
  const float bz = 3.0f * m.z + 9.0f * (m.xz * ux + m.yz * uy + m.zz * uz) - 3.0f * m.m0 * uz;
  // sum_i lam_i (feq_i - w_i) = rho (M0 (1 - 1.5 |u|^2) + 3 M1.u + 4.5 u^T M2 u) - M0
  const float lf = rho * (m.m0 * (1.0f - 1.5f * u2) + 3.0f * mu1 + 4.5f * q) - m.m0;
  const float omt = 1.0f - inv_tau;
  const float ta = inv_tau * A, tx = inv_tau * bx, ty = inv_tau * by, tz = inv_tau * bz;
#pragma unroll
  for (int j = 0; j < 27; ++j) {
    float v = ta;
    if (cx_of(j) != 0) v += cx_of(j) > 0 ? tx : -tx;
    if (cy_of(j) != 0) v += cy_of(j) > 0 ? ty : -ty;
    if (cz_of(j) != 0) v += cz_of(j) > 0 ? tz : -tz;
    lam[j] = omt * lam[j] + v;
  }
  return lf - lh;
}

// (d feq / d(rho, ux, uy, uz))^T lam at the EQU inputs:
//   d/d rho = M0 (1 - 1.5 |u|^2) + 3 M1.u + 4.5 u^T M2 u,
//   d/d u   = rho (3 M1 + 9 M2 u - 3 M0 u).
__device__ __forceinline__ void adj_feq_inputs(float rho, float ux, float uy, float uz,
                                               const float* lam, float* g4) {
  const LamMoments m = lam_moments(lam);
  const float u2 = ux * ux + uy * uy + uz * uz;
  g4[0] = m.m0 * (1.0f - 1.5f * u2) + 3.0f * (m.x * ux + m.y * uy + m.z * uz) + 4.5f * umu(m, ux, uy, uz);
  g4[1] = rho * (3.0f * m.x + 9.0f * (m.xx * ux + m.xy * uy + m.xz * uz) - 3.0f * m.m0 * ux);
  g4[2] = rho * (3.0f * m.y + 9.0f * (m.xy * ux + m.yy * uy + m.yz * uz) - 3.0f * m.m0 * uy);
  g4[3] = rho * (3.0f * m.z + 9.0f * (m.xz * ux + m.yz * uy + m.zz * uz) - 3.0f * m.m0 * uz);
}

// Workgroup sum of one double per thread in a fixed order: a butterfly within
// each wave64, then the four wave sums in order.
__device__ __forceinline__ double block_sum(double v) {
  __shared__ double part[kBlock / 64];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = 0.0;
#pragma unroll
  for (int k = 0; k < kBlock / 64; ++k) s += part[k];
  return s;
}

// One reverse step: mu^t from mu^{t+1} and f^t.  Both gathers are in bounds for every cell, so they are issued
// before the flag arrives.  With wobj the objective's term is added to mu^t.
template <class S>
// Four waves per SIMD: the batched face loads would otherwise take the register
// count past 128.
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(4))) void k_adjoint(
    const AdjArgs j) {
  const StepArgs& a = j.a;
  const uint32_t blk = cell_block(a);
  const Cell k = locate(a, blk);
  float dtau = 0.0f;
  if (k.ok) {
    float h[27], lam[27];
    make_gather<S>(a, k, a.src).template run<0>(h);
    adj_gather_consumers(j, k, j.mu_in, lam);
    // x-face cells (the reference convention's EQU cells) fetch their EQU
    // inputs and gradient slot with the rest of the batch
    EquIn q = equ_prefetch(a, k.c, k.x);
    const uint32_t aux_q = j.aux[q.ok ? k.c : 0u];
    const uint32_t fl = a.sflag[k.c];
    // objective term w (c_x - u_x) / rho of the pulled populations, added to mu^t last
    float wq = 0.0f, wux = 0.0f;
    if (j.wobj) {
      const float w = j.wobj[k.c];
      if (w != 0.0f) {
        const Moments m = moments(h);
        wq = w / m.rho;
        wux = m.ux;
      }
    }
    if (fl & kStore) {
      const bool bb = fl & kBounceBack;
      if ((fl & kEqu) && !q.ok) {  // an EQU cell off the x faces
        q.rho = a.rho[k.c];
        q.vx = a.vx[k.c];
        q.vy = a.vy[k.c];
        q.vz = a.vz[k.c];
      }
      if (fl & kEqu)
        equilibrium(q.rho, q.vx, q.vy, q.vz, h);
      else if (!(fl & kLoad))
        clear(h);
      if (fl & kCollide) dtau = adj_collide(a.inv_tau, h, lam, bb);
      if (bb) bounce_back(lam);
      if (fl & kEqu) {
        float g4[4];
        adj_feq_inputs(q.rho, q.vx, q.vy, q.vz, lam, g4);
        const uint32_t r = q.ok ? aux_q : j.aux[k.c];
#pragma unroll
        for (int q = 0; q < 4; ++q) j.grad4[q * j.n_equ + r] += double(g4[q]);
      }
      if ((fl & kEqu) || !(fl & kLoad)) clear(lam);  // the pulled values are not used
    } else {
      if (j.carry) {
        const uint32_t r = j.aux[k.c];
#pragma unroll
        for (int i = 0; i < 27; ++i) j.carry[i * j.n_keep + r] += lam[i];
      }
      clear(lam);
    }
    if (wq != 0.0f) {
#pragma unroll
      for (int i = 0; i < 27; ++i) lam[i] += wq * (float(cx_of(i)) - wux);
    }
    char* out = reinterpret_cast<char*>(j.mu_out);
#pragma unroll
    for (int i = 0; i < 27; ++i) st_at<float>(out + i * (a.dstride * 4), k.l * 4u, lam[i]);
  }
  // one accumulator per block of cells (not per hardware workgroup), so the
  // fixed-order host sum does not depend on the workgroup order
  const double s = block_sum(double(dtau));
  if (threadIdx.x == 0) j.tau_acc[blk] += s;
}

// lambda^0 = dJ/df^0 (plain FP32 [27][n]): consumers' mu^0, plus the carry of
// the cells without STORE.
__global__ __launch_bounds__(kBlock) void k_adjoint_final(const AdjArgs j, float* out) {
  const StepArgs& a = j.a;
  const Cell k = locate(a, cell_block(a));
  if (!k.ok) return;
  float lam[27];
  adj_gather_consumers(j, k, j.mu_in, lam);
  if (!(a.sflag[k.c] & kStore) && j.carry) {
    const uint32_t r = j.aux[k.c];
#pragma unroll
    for (int i = 0; i < 27; ++i) lam[i] += j.carry[i * j.n_keep + r];
  }
  const int64_t n = int64_t(a.nx) * a.ny * a.nz;
#pragma unroll
  for (int i = 0; i < 27; ++i) out[i * n + k.c] = lam[i];
}

struct ObjArgs {
  StepArgs a;            // a.src = f^t
  const uint32_t* idx;   // cells with a non-zero weight, each once
  const double* w;
  uint32_t count;
  float* mu;             // mu^t to add the objective's gradient to, or null
  double* j_acc;         // one accumulator per workgroup, or null (J not accumulated)
};

// With j_acc: J += sum w u_x(g^t) over the listed cells; with mu: mu^t += w (c_x - u_x) / rho.
template <class S>
__global__ __launch_bounds__(kBlock) void k_objective(const ObjArgs o) {
  const StepArgs& a = o.a;
  const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
  double jv = 0.0;
  if (t < o.count) {
    const uint32_t c = o.idx[t];
    Cell k;
    k.t = t;
    k.c = c;
    k.l = c;
    k.ok = true;
    const uint32_t r = a.divx.div(c);
    k.x = int32_t(c - r * uint32_t(a.nx));
    const uint32_t zz = a.divy.div(r);
    k.y = int32_t(r - zz * uint32_t(a.ny));
    k.z = int32_t(zz);
    float g[27];
    make_gather<S>(a, k, a.src).template run<0>(g);
    const Moments m = moments(g);
    jv = o.w[t] * double(m.ux);
    if (o.mu) {
      char* base = reinterpret_cast<char*>(o.mu);
      const float q = float(o.w[t]) / m.rho;
#pragma unroll
      for (int i = 0; i < 27; ++i) {
        float* p = reinterpret_cast<float*>(base + i * (a.dstride * 4) + uint64_t(c) * 4u);
        *p += q * (float(cx_of(i)) - m.ux);
      }
    }
  }
  if (o.j_acc) {
    const double s = block_sum(jv);
    if (threadIdx.x == 0) o.j_acc[blockIdx.x] += s;
  }
}

}  // namespace culbm
