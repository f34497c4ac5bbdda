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
};

// Sum of mu_i over the consumers of population i of cell k.  Along each axis a
// the consumer coordinate is k_a + c_a when that cell exists, and k_a itself
// when k_a - c_a lies outside the domain (a face cell clamps its pull to itself).
__device__ __forceinline__ void adj_gather_consumers(const StepArgs& a, const Cell& k,
                                                     const float* mu, float* lo) {
  const char* base = reinterpret_cast<const char*>(mu);
  const int64_t dsb = a.dstride * 4;
  const int32_t sx = 4, sy = a.px * 4, sz = a.px * a.py * 4;
  const uint32_t lb = k.l * 4u;
  const bool face = k.x == 0 || k.x == a.nx - 1 || k.y == 0 || k.y == a.ny - 1 || k.z == 0 ||
                    k.z == a.nz - 1;
  if (!face) {
#pragma unroll
    for (int i = 0; i < 27; ++i)
      lo[i] = ld_at<float>(base + i * dsb,
                           lb + uint32_t(cx_of(i) * sx + cy_of(i) * sy + cz_of(i) * sz));
    return;
  }
  const int32_t co[3] = {k.x, k.y, k.z};
  const int32_t nn[3] = {a.nx, a.ny, a.nz};
  const int32_t st[3] = {sx, sy, sz};
#pragma unroll
  for (int i = 0; i < 27; ++i) {
    const int32_t c[3] = {cx_of(i), cy_of(i), cz_of(i)};
    float sum = 0.0f;
    for (int m = 0; m < 8; ++m) {  // bit a of m: consumer is the cell itself along axis a
      bool ok = true;
      int32_t off = 0;
#pragma unroll
      for (int ax = 0; ax < 3; ++ax) {
        if ((m >> ax) & 1) {
          ok = ok && c[ax] != 0 && (co[ax] - c[ax] < 0 || co[ax] - c[ax] >= nn[ax]);
        } else {
          const int32_t q = co[ax] + c[ax];
          ok = ok && q >= 0 && q < nn[ax];
          off += c[ax] * st[ax];
        }
      }
      if (ok) sum += ld_at<float>(base + i * dsb, lb + uint32_t(off));
    }
    lo[i] = sum;
  }
}

// In place: lam <- (dC/dh)^T lam for the BGK collision C at pre-collision h.
// Returns sum_i lam_i (feq_i - h_i), the derivative of lam . C(h) by invTau.
// With m = sum_i c_i h_i (= rho u) the equilibrium is a function of (rho, m):
//   d feq_i / d rho |m = w_i (1 - 4.5 (c_i.u)^2 + 1.5 |u|^2)
//   d feq_i / d m      = w_i (3 c_i + 9 (c_i.u) c_i - 3 u)
// and d rho / d h_j = 1, d m / d h_j = c_j.
__device__ __forceinline__ float adj_collide(float inv_tau, const float* h, float* lam) {
  const Moments m = moments(h);
  float fe[27];
  equilibrium(m.rho, m.ux, m.uy, m.uz, fe);
  const float u2 = m.ux * m.ux + m.uy * m.uy + m.uz * m.uz;
  float a = 0.0f, bx = 0.0f, by = 0.0f, bz = 0.0f, s = 0.0f, dtau = 0.0f;
#pragma unroll
  for (int i = 0; i < 27; ++i) {
    const int cls = (cx_of(i) != 0) + (cy_of(i) != 0) + (cz_of(i) != 0);
    const float w = cls == 0 ? kW0 : (cls == 1 ? kWS : (cls == 2 ? kWE : kWC));
    const float cu = float(cx_of(i)) * m.ux + float(cy_of(i)) * m.uy + float(cz_of(i)) * m.uz;
    const float lw = lam[i] * w;
    a += lw * (1.0f - 4.5f * cu * cu + 1.5f * u2);
    const float q = lw * (3.0f + 9.0f * cu);
    bx += q * float(cx_of(i));
    by += q * float(cy_of(i));
    bz += q * float(cz_of(i));
    s += lw;
    dtau += lam[i] * (fe[i] - h[i]);
  }
  bx -= 3.0f * m.ux * s;
  by -= 3.0f * m.uy * s;
  bz -= 3.0f * m.uz * s;
  const float omt = 1.0f - inv_tau;
#pragma unroll
  for (int j = 0; j < 27; ++j)
    lam[j] = omt * lam[j] +
             inv_tau * (a + bx * float(cx_of(j)) + by * float(cy_of(j)) + bz * float(cz_of(j)));
  return dtau;
}

// (d feq / d(rho, ux, uy, uz))^T lam at the EQU inputs.
__device__ __forceinline__ void adj_feq_inputs(float rho, float ux, float uy, float uz,
                                               const float* lam, float* g4) {
  const float u2 = ux * ux + uy * uy + uz * uz;
  g4[0] = g4[1] = g4[2] = g4[3] = 0.0f;
#pragma unroll
  for (int i = 0; i < 27; ++i) {
    const int cls = (cx_of(i) != 0) + (cy_of(i) != 0) + (cz_of(i) != 0);
    const float w = cls == 0 ? kW0 : (cls == 1 ? kWS : (cls == 2 ? kWE : kWC));
    const float cu = float(cx_of(i)) * ux + float(cy_of(i)) * uy + float(cz_of(i)) * uz;
    const float lw = lam[i] * w;
    g4[0] += lw * (1.0f + 3.0f * cu + 4.5f * cu * cu - 1.5f * u2);
    const float q = lw * rho * (3.0f + 9.0f * cu);
    const float r3 = lw * rho * 3.0f;
    g4[1] += q * float(cx_of(i)) - r3 * ux;
    g4[2] += q * float(cy_of(i)) - r3 * uy;
    g4[3] += q * float(cz_of(i)) - r3 * uz;
  }
}

// Workgroup sum of one double per thread (fixed tree order).
__device__ __forceinline__ double block_sum(double v) {
  __shared__ double red[kBlock];
  red[threadIdx.x] = v;
  __syncthreads();
#pragma unroll
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if (int(threadIdx.x) < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  return red[0];
}

// One reverse step: mu^t from mu^{t+1} and f^t.
template <class S>
__global__ __launch_bounds__(kBlock) void k_adjoint(const AdjArgs j) {
  const StepArgs& a = j.a;
  const Cell k = locate(a, xcd_block(blockIdx.x, a.nblk));
  double dtau = 0.0;
  if (k.ok) {
    const uint32_t fl = a.sflag[k.c];
    float lam[27];
    if (fl & kStore) {
      float h[27];
      make_gather<S>(a, k, a.src).template run<0>(h);
      adj_gather_consumers(a, k, j.mu_in, lam);
      if (!(fl & kLoad)) clear(h);
      if (fl & kEqu) equilibrium(a.rho[k.c], a.vx[k.c], a.vy[k.c], a.vz[k.c], h);
      if (fl & kBounceBack) bounce_back(h);
      if (fl & kCollide) dtau = double(adj_collide(a.inv_tau, h, lam));
      if (fl & kBounceBack) bounce_back(lam);
      if (fl & kEqu) {
        float g4[4];
        adj_feq_inputs(a.rho[k.c], a.vx[k.c], a.vy[k.c], a.vz[k.c], lam, g4);
        const uint32_t r = j.aux[k.c];
#pragma unroll
        for (int q = 0; q < 4; ++q) j.grad4[q * j.n_equ + r] += double(g4[q]);
      }
      if ((fl & kEqu) || !(fl & kLoad)) clear(lam);  // the pulled values are not used
    } else {
      if (j.carry) {
        adj_gather_consumers(a, k, j.mu_in, lam);
        const uint32_t r = j.aux[k.c];
#pragma unroll
        for (int i = 0; i < 27; ++i) j.carry[i * j.n_keep + r] += lam[i];
      }
      clear(lam);
    }
    char* out = reinterpret_cast<char*>(j.mu_out);
#pragma unroll
    for (int i = 0; i < 27; ++i) st_at<float>(out + i * (a.dstride * 4), k.l * 4u, lam[i]);
  }
  const double s = block_sum(dtau);
  if (threadIdx.x == 0) j.tau_acc[blockIdx.x] += s;
}

// lambda^0 = dJ/df^0 (plain FP32 [27][n]): consumers' mu^0, plus the carry of
// the cells without STORE.
__global__ __launch_bounds__(kBlock) void k_adjoint_final(const AdjArgs j, float* out) {
  const StepArgs& a = j.a;
  const Cell k = locate(a, xcd_block(blockIdx.x, a.nblk));
  if (!k.ok) return;
  float lam[27];
  adj_gather_consumers(a, k, j.mu_in, lam);
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
