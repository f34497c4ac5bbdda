// Host binding of the reference's own physics and blocking headers.
//
// TEST INFRASTRUCTURE ONLY.  Built by oracle/Makefile from the headers where
// they lie in the read-only reference checkout (never copied) into
// oracle/_ref/libculbm_ref.so, which the CPU tests and the golden-vector
// script use to pin the oracle.  The reference's kernels (kernel.cu) need
// nvcc and a CUDA GPU and are not built; only these host-compilable headers:
//   src/lbm_core/device_function.hpp   calcEqu / calcRhoU / collision / collision2
//   src/basic/velocity_set.hpp         VelSet3D<27>
//   src/blocking_core/blocking_alogrithm.hpp
#include <cstdint>

#include "blocking_alogrithm.hpp"
#include "device_function.hpp"
#include "velocity_set.hpp"

extern "C" {

void ref_velset(int32_t* cx, int32_t* cy, int32_t* cz, float* w) {
  using V = culbm::basic::detail::VelSet3D<27>;
  const auto dx = V::getDxArr();
  const auto dy = V::getDyArr();
  const auto dz = V::getDzArr();
  // the weight table is only reachable through the per-direction accessor
  float ws[27] = {
      V::getOmega<0>(),  V::getOmega<1>(),  V::getOmega<2>(),  V::getOmega<3>(),
      V::getOmega<4>(),  V::getOmega<5>(),  V::getOmega<6>(),  V::getOmega<7>(),
      V::getOmega<8>(),  V::getOmega<9>(),  V::getOmega<10>(), V::getOmega<11>(),
      V::getOmega<12>(), V::getOmega<13>(), V::getOmega<14>(), V::getOmega<15>(),
      V::getOmega<16>(), V::getOmega<17>(), V::getOmega<18>(), V::getOmega<19>(),
      V::getOmega<20>(), V::getOmega<21>(), V::getOmega<22>(), V::getOmega<23>(),
      V::getOmega<24>(), V::getOmega<25>(), V::getOmega<26>()};
  for (int i = 0; i < 27; ++i) {
    cx[i] = dx[i];
    cy[i] = dy[i];
    cz[i] = dz[i];
    w[i] = ws[i];
  }
}

void ref_equ(float rho, float vx, float vy, float vz, float* f) {
  culbm::lbm_core::bgk::calcEqu<27>(rho, vx, vy, vz, f);
}

void ref_moments(const float* f, float* out4) {
  culbm::lbm_core::bgk::calcRhoU<27>(out4[0], out4[1], out4[2], out4[3], f);
}

void ref_collide(float inv_tau, float* f) { culbm::lbm_core::bgk::collision2<27>(inv_tau, f); }

// collision<27> also reports the pre-collision rho/u it used
void ref_collide_moments(float inv_tau, float* f, float* out4) {
  culbm::lbm_core::bgk::collision<27>(inv_tau, out4[0], out4[1], out4[2], out4[3], f);
}

void ref_collide_batch(float inv_tau, float* f, int64_t ncell) {
  for (int64_t c = 0; c < ncell; ++c) culbm::lbm_core::bgk::collision2<27>(inv_tau, f + 27 * c);
}

void ref_equ_batch(const float* rho, const float* vx, const float* vy, const float* vz, float* f,
                   int64_t ncell) {
  for (int64_t c = 0; c < ncell; ++c)
    culbm::lbm_core::bgk::calcEqu<27>(rho[c], vx[c], vy[c], vz[c], f + 27 * c);
}

void ref_moments_batch(const float* f, float* out, int64_t ncell) {
  for (int64_t c = 0; c < ncell; ++c)
    culbm::lbm_core::bgk::calcRhoU<27>(out[4 * c], out[4 * c + 1], out[4 * c + 2], out[4 * c + 3],
                                        f + 27 * c);
}

int ref_blk_valid(uint32_t dom, uint32_t blk, uint32_t iter) {
  return culbm::blocking_core::validBlkAxisConfig<uint32_t>(dom, blk, iter) ? 1 : 0;
}

uint32_t ref_blk_num(uint32_t dom, uint32_t blk, uint32_t iter) {
  return culbm::blocking_core::calcBlkNum<uint32_t>(dom, blk, iter);
}

int32_t ref_valid_prev(int32_t idx, int32_t blk, int32_t num, int32_t iter, int32_t dom) {
  return culbm::blocking_core::calcValidPrev<int32_t>(idx, blk, num, iter, dom);
}

}  // extern "C"
