// Lateral and vertical boundary fills of the padded CM1 carriers (gfx950).
//
// Replaces the fill kernels of cuda_boundary_families.py / cuda_sound3_open_boundary.py:
// family fills of scalar/u/v/w carriers for periodic (1), open (2) and rigid
// (3/4) sides, the one-cell pressure-family fill, the rigid floor/lid of w,
// bcwsfc on physical bottom faces and the open bc2d fill of w(k=1).
#include "boundary.h"

using namespace gm1;

namespace {

// One ghost plane per (k, row, layer, side): every source is an interior
// plane, so all planes of one direction are written in one race-free launch.
__global__ __launch_bounds__(256) void k_family_we(float* __restrict__ a, int ni, int nj, int nk, int kind,
                                                   unsigned wbc, unsigned ebc) {
  const int nx = (kind == 1) ? ux(ni) : sx(ni);
  const int ny = (kind == 2) ? vy(nj) : sy(nj);
  const int kfirst = (kind == 3) ? 2 : 1, kcount = (kind == 3) ? nk - 1 : nk;
  const int jhi = (kind == 2) ? nj + 2 : nj + 1;
  const int64_t count = (int64_t)kcount * ny * NG * 2;
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < count; q += (int64_t)gridDim.x * blockDim.x) {
    int64_t r = q;
    const int layer = (int)(r % NG); r /= NG;
    const int side = (int)(r % 2); r /= 2;
    const int jp = (int)(r % ny);
    const int k = (int)(r / ny) + kfirst;
    const unsigned mode = side == 0 ? wbc : ebc;
    if (mode == 0u) continue;
    const int jc = jp - (NG - 1);
    if (mode != 1u && (jc < 0 || jc > jhi)) continue;
    int dst, src, wall = -1;
    bool neg = false;
    if (side == 0) {
      if (mode == 1u) { dst = NG - 1 - layer; src = ni + NG - 1 - layer; }
      else if (mode == 2u) { dst = layer; src = NG; }
      else if (kind == 1) { dst = NG - 1 - layer; src = NG + 1 + layer; neg = true; wall = NG; }
      else { dst = NG - 1 - layer; src = NG + layer; }
    } else {
      if (mode == 1u) {
        if (kind == 1) { dst = ni + NG + 1 + layer; src = NG + 1 + layer; }
        else { dst = ni + NG + layer; src = NG + layer; }
      } else if (mode == 2u) {
        if (kind == 1) { dst = ni + NG + 1 + layer; src = ni + NG; }
        else { dst = ni + NG + layer; src = ni + NG - 1; }
      } else if (kind == 1) { dst = ni + NG + 1 + layer; src = ni + NG - 1 - layer; neg = true; wall = ni + NG; }
      else { dst = ni + NG + layer; src = ni + NG - 1 - layer; }
    }
    const float v = a[ix3(k, jp, src, ny, nx)];
    a[ix3(k, jp, dst, ny, nx)] = neg ? -v : v;
    if (wall >= 0 && layer == 0) a[ix3(k, jp, wall, ny, nx)] = 0.0f;
  }
}

__global__ __launch_bounds__(256) void k_family_sn(float* __restrict__ a, int ni, int nj, int nk, int kind,
                                                   unsigned sbc, unsigned nbc) {
  const int nx = (kind == 1) ? ux(ni) : sx(ni);
  const int ny = (kind == 2) ? vy(nj) : sy(nj);
  const int kfirst = (kind == 3) ? 2 : 1, kcount = (kind == 3) ? nk - 1 : nk;
  const int ihi = (kind == 1) ? ni + 2 : ni + 1;
  const int64_t count = (int64_t)kcount * nx * NG * 2;
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < count; q += (int64_t)gridDim.x * blockDim.x) {
    int64_t r = q;
    const int layer = (int)(r % NG); r /= NG;
    const int side = (int)(r % 2); r /= 2;
    const int ip = (int)(r % nx);
    const int k = (int)(r / nx) + kfirst;
    const unsigned mode = side == 0 ? sbc : nbc;
    if (mode == 0u) continue;
    const int ic = ip - (NG - 1);
    if (mode != 1u && (ic < 0 || ic > ihi)) continue;
    int dst, src, wall = -1;
    bool neg = false;
    if (side == 0) {
      if (mode == 1u) { dst = NG - 1 - layer; src = nj + NG - 1 - layer; }
      else if (mode == 2u) { dst = layer; src = NG; }
      else if (kind == 2) { dst = NG - 1 - layer; src = NG + 1 + layer; neg = true; wall = NG; }
      else { dst = NG - 1 - layer; src = NG + layer; }
    } else {
      if (mode == 1u) {
        if (kind == 2) { dst = nj + NG + 1 + layer; src = NG + 1 + layer; }
        else { dst = nj + NG + layer; src = NG + layer; }
      } else if (mode == 2u) {
        if (kind == 2) { dst = nj + NG + 1 + layer; src = nj + NG; }
        else { dst = nj + NG + layer; src = nj + NG - 1; }
      } else if (kind == 2) { dst = nj + NG + 1 + layer; src = nj + NG - 1 - layer; neg = true; wall = nj + NG; }
      else { dst = nj + NG + layer; src = nj + NG - 1 - layer; }
    }
    const float v = a[ix3(k, src, ip, ny, nx)];
    a[ix3(k, dst, ip, ny, nx)] = neg ? -v : v;
    if (wall >= 0 && layer == 0) a[ix3(k, wall, ip, ny, nx)] = 0.0f;
  }
}

// Pressure family (nk+2, nj+2, ni+2): west/east columns of every row.
// bcp on the one-cell halo ring of levels 1..nk in one pass: bc.F fills the
// west/east columns then the south/north rows (corners from the filled
// columns), so ring cell (x, y) takes cell (xs(x), ys(y)), where xs maps
// column 0 to ni (periodic west) or 1, column ni+1 to 1 (periodic east) or
// ni, and leaves the others (a side with code 0 keeps its ring); ys likewise.
// No source is a ring cell this pass writes.
__global__ __launch_bounds__(256) void k_pressure_ring(float* __restrict__ p, int ni, int nj, unsigned w, unsigned e,
                                                       unsigned so, unsigned n) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  const int k = blockIdx.y + 1;
  if (q >= 2 * (ni + 2) + 2 * nj) return;
  int x, y;
  if (q < ni + 2) { y = 0; x = q; }
  else if (q < 2 * (ni + 2)) { y = nj + 1; x = q - (ni + 2); }
  else { const int r = q - 2 * (ni + 2); y = 1 + r / 2; x = (r & 1) ? ni + 1 : 0; }
  int xs = x, ys = y;
  if (x == 0 && w) xs = (w == 1u) ? ni : 1;
  else if (x == ni + 1 && e) xs = (e == 1u) ? 1 : ni;
  if (y == 0 && so) ys = (so == 1u) ? nj : 1;
  else if (y == nj + 1 && n) ys = (n == 1u) ? 1 : nj;
  if (xs == x && ys == y) return;
  float* plane = p + (int64_t)k * (nj + 2) * (ni + 2);
  plane[(int64_t)y * (ni + 2) + x] = plane[(int64_t)ys * (ni + 2) + xs];
}

__global__ __launch_bounds__(256) void k_w_rigid(float* __restrict__ w, int ni, int nj, int nk) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;  // 0..ni+1 -> padded x+NG-1
  const int y = blockIdx.y;                             // 0..nj+1
  if (x >= ni + 2) return;
  const int nxs = sx(ni), nys = sy(nj);
  w[ix3(1, y + NG - 1, x + NG - 1, nys, nxs)] = 0.0f;
  w[ix3(nk + 1, y + NG - 1, x + NG - 1, nys, nxs)] = 0.0f;
}

__global__ __launch_bounds__(256) void k_bcwsfc(const float* __restrict__ u, const float* __restrict__ v,
                                                const float* __restrict__ dzdx, const float* __restrict__ dzdy,
                                                const float* __restrict__ gz, float* __restrict__ w, int ni, int nj) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  if (i >= ni) return;
  w[ix3(1, j + NG, i + NG, sy(nj), sx(ni))] = bcwsfc_value(u, v, dzdx, dzdy, gz, j, i, ni, nj);
}

// bc2d: w(k=1) zero-gradient halo in W, E, S, N order (one launch per direction).
__global__ __launch_bounds__(256) void k_bc2d(float* __restrict__ w, int ni, int nj, int direction) {
  const int nx = sx(ni), ny = sy(nj);
  const int n = (direction < 2) ? ny * NG : nx * NG;
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  int x, y, sxx, syy;
  if (direction < 2) {
    y = q / NG;
    const int h = q % NG;
    syy = y;
    if (direction == 0) { x = h; sxx = NG; } else { x = NG + ni + h; sxx = NG + ni - 1; }
  } else {
    x = q % nx;
    const int h = q / nx;
    sxx = x;
    if (direction == 2) { y = h; syy = NG; } else { y = NG + nj + h; syy = NG + nj - 1; }
  }
  w[ix3(1, y, x, ny, nx)] = w[ix3(1, syy, sxx, ny, nx)];
}

}  // namespace

namespace gm1 {

int fill_family(float* a, int kind, unsigned sides, int ni, int nj, int nk, hipStream_t s) {
  const unsigned w = side_of(sides, 0), e = side_of(sides, 1), so = side_of(sides, 2), n = side_of(sides, 3);
  const int nx = (kind == 1) ? ni + 1 + 2 * NG : ni + 2 * NG;
  const int ny = (kind == 2) ? nj + 1 + 2 * NG : nj + 2 * NG;
  const int kcount = (kind == 3) ? nk - 1 : nk;
  const int bwe = gm1_blocks((int64_t)kcount * ny * NG * 2, 256);
  const int bsn = gm1_blocks((int64_t)kcount * nx * NG * 2, 256);
  if (kind == 2) {
    hipLaunchKernelGGL(k_family_sn, dim3(bsn), dim3(256), 0, s, a, ni, nj, nk, kind, so, n);
    hipLaunchKernelGGL(k_family_we, dim3(bwe), dim3(256), 0, s, a, ni, nj, nk, kind, w, e);
  } else {
    hipLaunchKernelGGL(k_family_we, dim3(bwe), dim3(256), 0, s, a, ni, nj, nk, kind, w, e);
    hipLaunchKernelGGL(k_family_sn, dim3(bsn), dim3(256), 0, s, a, ni, nj, nk, kind, so, n);
  }
  GM1_LAUNCH_CHECK();
  return 0;
}

int fill_pressure(float* p, unsigned sides, int ni, int nj, int nk, hipStream_t s) {
  hipLaunchKernelGGL(k_pressure_ring, dim3((2 * (ni + 2) + 2 * nj + 255) / 256, nk), dim3(256), 0, s, p, ni, nj,
                     side_of(sides, 0), side_of(sides, 1), side_of(sides, 2), side_of(sides, 3));
  GM1_LAUNCH_CHECK();
  return 0;
}

int w_rigid(float* w, int ni, int nj, int nk, hipStream_t s) {
  hipLaunchKernelGGL(k_w_rigid, dim3((ni + 2 + 255) / 256, nj + 2), dim3(256), 0, s, w, ni, nj, nk);
  GM1_LAUNCH_CHECK();
  return 0;
}

int bcwsfc(const float* u, const float* v, const float* dzdx, const float* dzdy, const float* gz, float* w, int ni,
           int nj, hipStream_t s) {
  hipLaunchKernelGGL(k_bcwsfc, dim3((ni + 63) / 64, nj), dim3(64), 0, s, u, v, dzdx, dzdy, gz, w, ni, nj);
  GM1_LAUNCH_CHECK();
  return 0;
}

int bc2d_open(float* w, int ni, int nj, hipStream_t s) {
  for (int d = 0; d < 4; ++d) {
    const int n = (d < 2) ? (nj + 2 * NG) * NG : (ni + 2 * NG) * NG;
    hipLaunchKernelGGL(k_bc2d, dim3((n + 255) / 256), dim3(256), 0, s, w, ni, nj, d);
  }
  GM1_LAUNCH_CHECK();
  return 0;
}

}  // namespace gm1

extern "C" {

int gm1hip_fill_family(float* a, int kind, unsigned sides, int ni, int nj, int nk, void* stream) {
  return gm1::fill_family(a, kind, sides, ni, nj, nk, (hipStream_t)stream);
}

int gm1hip_fill_pressure(float* p, unsigned sides, int ni, int nj, int nk, void* stream) {
  return gm1::fill_pressure(p, sides, ni, nj, nk, (hipStream_t)stream);
}

int gm1hip_w_rigid(float* w, int ni, int nj, int nk, void* stream) {
  return gm1::w_rigid(w, ni, nj, nk, (hipStream_t)stream);
}

int gm1hip_bcwsfc(const float* u3d, const float* v3d, const float* dzdx, const float* dzdy, const float* gz,
                  float* w3d, int ni, int nj, void* stream) {
  return gm1::bcwsfc(u3d, v3d, dzdx, dzdy, gz, w3d, ni, nj, (hipStream_t)stream);
}

int gm1hip_bc2d_open(float* w3d, int ni, int nj, void* stream) {
  return gm1::bc2d_open(w3d, ni, nj, (hipStream_t)stream);
}

}  // extern "C"
