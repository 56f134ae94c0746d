// Boundary fills shared by the acoustic, stage-term and scalar kernels.
#pragma once
#include "cm1_common.h"

namespace gm1 {

// w(k=1) on physical column (j, i) from the terrain slope (bc.F:1151-1178):
// ((u(i)+u(i+1))*dzdx + (v(j)+v(j+1))*dzdy) * 0.5 * gz in that order.
GM1_DEV float bcwsfc_value(const float* __restrict__ u, const float* __restrict__ v,
                           const float* __restrict__ dzdx, const float* __restrict__ dzdy,
                           const float* __restrict__ gz, int j, int i, int ni, int nj) {
  const int64_t h = (int64_t)j * ni + i;
  const float* u1 = u + ix3(1, j + NG, i + NG, sy(nj), ux(ni));
  const float* v1 = v + ix3(1, j + NG, i + NG, vy(nj), sx(ni));
  const float ut = (u1[0] + u1[1]) * dzdx[h];
  const float vt = (v1[0] + v1[sx(ni)]) * dzdy[h];
  return (0.5f * (ut + vt)) * gz[h];
}

// Host launchers (enqueue on stream s; return hipError_t).
// kind: 0 scalar, 1 u, 2 v, 3 w.  sides: CM1 (wbc, ebc, sbc, nbc) word.
int fill_family(float* a, int kind, unsigned sides, int ni, int nj, int nk, hipStream_t s);
int fill_pressure(float* p, unsigned sides, int ni, int nj, int nk, hipStream_t s);
int w_rigid(float* w, int ni, int nj, int nk, hipStream_t s);
int bcwsfc(const float* u, const float* v, const float* dzdx, const float* dzdy, const float* gz, float* w, int ni,
           int nj, hipStream_t s);
int bc2d_open(float* w, int ni, int nj, hipStream_t s);

}  // namespace gm1
