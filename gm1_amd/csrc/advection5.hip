// Scalar fifth-order / WENO5 advection and the terrain mass flux for gfx950.
//
// Replaces the kernel family of cuda_advection5.py (terrain mass flux,
// divergence, scalar face fluxes and tendency).  Design:
//   * one fused tendency kernel per scalar: a 64x4 (i,j) tile marches up its
//     column chunk; x- and y-face fluxes of each level are formed once into
//     LDS and shared by the two cells on either side, the z-face flux is
//     carried in a register from one level to the next.  Face fluxes never
//     round-trip HBM (they are written only when the caller asks for them).
//   * i is the fastest index of every access, so a wavefront's 64 lanes read
//     one contiguous 256-byte row segment.
#include "cm1_common.h"
#include "timing.h"

using namespace gm1;

namespace {

constexpr int TX = 64;
constexpr int TY = 4;
constexpr int NT = TX * TY;

// ---------------------------------------------------------------------------
// terrain mass flux (cuda_advection5.py:270-437): rru/rrv on interior faces,
// rrw on interior columns, divx on interior cells.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_tmf_horizontal(
    const float* __restrict__ u, const float* __restrict__ v,
    const float* __restrict__ rho0, const float* __restrict__ rgzu,
    const float* __restrict__ rgzv, float* __restrict__ rru,
    float* __restrict__ rrv, float* __restrict__ rrw, int ni, int nj, int nk) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  const int k0 = blockIdx.z;
  if (i >= ni) return;
  const int k = k0 + 1;
  const int nxu = ux(ni), nyu = sy(nj), nxv = sx(ni), nyv = vy(nj);
  const float half = 0.5f;
  if (i >= 1) {
    float d = rho0[((int64_t)k0 * nj + j) * ni + i - 1] + rho0[((int64_t)k0 * nj + j) * ni + i];
    int64_t p = ix3(k, j + NG, i + NG, nyu, nxu);
    rru[p] = (half * d) * (u[p] * rgzu[(int64_t)j * (ni + 1) + i]);
  }
  if (j >= 1) {
    float d = rho0[((int64_t)k0 * nj + j - 1) * ni + i] + rho0[((int64_t)k0 * nj + j) * ni + i];
    int64_t p = ix3(k, j + NG, i + NG, nyv, nxv);
    rrv[p] = (half * d) * (v[p] * rgzv[(int64_t)j * ni + i]);
  }
  if (k == 1) {
    rrw[ix3(1, j + NG, i + NG, sy(nj), sx(ni))] = 0.0f;
    rrw[ix3(nk + 1, j + NG, i + NG, sy(nj), sx(ni))] = 0.0f;
  }
}

__global__ __launch_bounds__(256) void k_tmf_vertical(
    const float* __restrict__ w, const float* __restrict__ rf0,
    const float* __restrict__ dzdx, const float* __restrict__ dzdy,
    const float* __restrict__ gz, const float* __restrict__ rru,
    const float* __restrict__ rrv, float* __restrict__ rrw, int ni, int nj,
    int nk, float dz) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  const int k = blockIdx.z + 2;
  if (i < 1 || i + 1 >= ni || j < 1 || j + 1 >= nj) return;
  const int nxu = ux(ni), nyu = sy(nj), nxv = sx(ni), nyv = vy(nj);
  const float half = 0.5f;
  const float zt = (float)nk * dz;
  const float rzt = 1.0f / zt;
  const float sigmaf = (float)(k - 1) * dz;
  const float sigma = (float)((double)k - 1.5) * dz;
  float r2 = (sigmaf - sigma) * (1.0f / dz);
  const float r1 = half * (1.0f - r2);
  r2 = half * r2;
  const int x = i + NG, y = j + NG;
  float uxv = r2 * (rru[ix3(k, y, x, nyu, nxu)] + rru[ix3(k, y, x + 1, nyu, nxu)]) +
              r1 * (rru[ix3(k - 1, y, x, nyu, nxu)] + rru[ix3(k - 1, y, x + 1, nyu, nxu)]);
  float vyv = r2 * (rrv[ix3(k, y, x, nyv, nxv)] + rrv[ix3(k, y + 1, x, nyv, nxv)]) +
              r1 * (rrv[ix3(k - 1, y, x, nyv, nxv)] + rrv[ix3(k - 1, y + 1, x, nyv, nxv)]);
  const int64_t h = (int64_t)j * ni + i;
  float t = uxv * dzdx[h] + vyv * dzdy[h];
  t = t * (sigmaf - zt);
  t = t * gz[h];
  t = t * rzt;
  const int64_t p = ix3(k, y, x, sy(nj), sx(ni));
  rrw[p] = rf0[((int64_t)(k - 1) * nj + j) * ni + i] * w[p] + t;
}

__global__ __launch_bounds__(256) void k_tmf_divergence(
    const float* __restrict__ rru, const float* __restrict__ rrv,
    const float* __restrict__ rrw, float* __restrict__ divx, int ni, int nj,
    int nk, float rdx, float rdy, float rdsf) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  const int k0 = blockIdx.z;
  if (i < 1 || i + 1 >= ni || j < 1 || j + 1 >= nj) return;
  const int k = k0 + 1, x = i + NG, y = j + NG;
  const int nxu = ux(ni), nyu = sy(nj), nxv = sx(ni), nyv = vy(nj), nxw = sx(ni), nyw = sy(nj);
  float d = (rru[ix3(k, y, x + 1, nyu, nxu)] - rru[ix3(k, y, x, nyu, nxu)]) * rdx +
            (rrv[ix3(k, y + 1, x, nyv, nxv)] - rrv[ix3(k, y, x, nyv, nxv)]) * rdy;
  d = d + (rrw[ix3(k + 1, y, x, nyw, nxw)] - rrw[ix3(k, y, x, nyw, nxw)]) * rdsf;
  divx[((int64_t)k0 * nj + j) * ni + i] = (fabs((double)d) < 1.0e-20) ? 0.0f : d;
}

// ---------------------------------------------------------------------------
// scalar face fluxes and tendency (cuda_advection5.py:440-579, :661-669)
// ---------------------------------------------------------------------------
template <bool WENO>
__device__ __forceinline__ float xflux(const float* __restrict__ s, const float* __restrict__ rru,
                                       int k, int y, int x, int nys, int nxs, int nxu, double weps) {
  float m = rru[ix3(k, y, x, nys, nxu)];
  const float* row = s + ix3(k, y, 0, nys, nxs);
  float r = (m >= 0.0f) ? recon5(row[x - 3], row[x - 2], row[x - 1], row[x], row[x + 1], WENO, weps)
                        : recon5(row[x + 2], row[x + 1], row[x], row[x - 1], row[x - 2], WENO, weps);
  return m * r;
}

template <bool WENO>
__device__ __forceinline__ float yflux(const float* __restrict__ s, const float* __restrict__ rrv,
                                       int k, int y, int x, int nys, int nxs, int nyv, double weps) {
  float m = rrv[ix3(k, y, x, nyv, nxs)];
  const float* col = s + ix3(k, 0, x, nys, nxs);
  const int64_t st = nxs;
  float r = (m >= 0.0f)
                ? recon5(col[(y - 3) * st], col[(y - 2) * st], col[(y - 1) * st], col[y * st], col[(y + 1) * st], WENO, weps)
                : recon5(col[(y + 2) * st], col[(y + 1) * st], col[y * st], col[(y - 1) * st], col[(y - 2) * st], WENO, weps);
  return m * r;
}

// z-face flux at face k (2..nk) of one column with the lower-order closures.
template <bool WENO>
__device__ __forceinline__ float zflux(const float* __restrict__ col, int64_t st,
                                       const float* __restrict__ rrw, int64_t wp,
                                       float cc1, float cc2, int k, int nk, double weps) {
  const float m = rrw[wp];
  auto S = [&](int level) { return col[(int64_t)level * st]; };
  const float wf = (float)weps;
  float r;
  if (k == 2) {
    if (m >= 0.0f) r = cc1 * S(k - 1) + cc2 * S(k);
    else r = WENO ? upstrpd(S(k + 1), S(k), S(k - 1), wf) : flx3(S(k + 1), S(k), S(k - 1));
  } else if (k == nk) {
    bool up = WENO ? (m > 0.0f) : (m >= 0.0f);
    if (up) r = WENO ? upstrpd(S(k - 2), S(k - 1), S(k), wf) : flx3(S(k - 2), S(k - 1), S(k));
    else r = cc1 * S(k - 1) + cc2 * S(k);
  } else if (k == 3) {
    if (m >= 0.0f) r = WENO ? upstrpd(S(k - 2), S(k - 1), S(k), wf) : flx3(S(k - 2), S(k - 1), S(k));
    else r = recon5(S(k + 2), S(k + 1), S(k), S(k - 1), S(k - 2), WENO, weps);
  } else if (k == nk - 1) {
    if (m >= 0.0f) r = recon5(S(k - 3), S(k - 2), S(k - 1), S(k), S(k + 1), WENO, weps);
    else r = WENO ? upstrpd(S(k + 1), S(k), S(k - 1), wf) : flx3(S(k + 1), S(k), S(k - 1));
  } else {
    if (m >= 0.0f) r = recon5(S(k - 3), S(k - 2), S(k - 1), S(k), S(k + 1), WENO, weps);
    else r = recon5(S(k + 2), S(k + 1), S(k), S(k - 1), S(k - 2), WENO, weps);
  }
  return m * r;
}

struct ScalarAdvArgs {
  const float* s;
  const float* rru;
  const float* rrv;
  const float* rrw;
  const float* c1;
  const float* c2;
  const float* divx;
  const float* rr0;
  const float* gz;
  float* tend;
  float* fx;  // optional flux outputs (nullptr: not written)
  float* fy;
  float* fz;
  int ni, nj, nk, kchunk;
  float rdx, rdy, rdsf;
  double weps;
  unsigned obc;
};

template <bool WENO>
__global__ __launch_bounds__(NT) void k_scalar_adv5(ScalarAdvArgs a) {
  __shared__ float lfx[TY][TX + 1];
  __shared__ float lfy[TY + 1][TX];
  const int ni = a.ni, nj = a.nj, nk = a.nk;
  const int tx = threadIdx.x % TX, ty = threadIdx.x / TX;
  const int i0 = blockIdx.x * TX, j0 = blockIdx.y * TY;
  const int i = i0 + tx, j = j0 + ty;
  const int kb = 1 + blockIdx.z * a.kchunk;
  const int ke = min(nk, kb + a.kchunk - 1);
  const int nxs = sx(ni), nys = sy(nj), nxu = ux(ni), nyv = vy(nj);
  const bool live = (i < ni) && (j < nj);
  const unsigned obc = a.obc;
  const bool owned = live && (obc != 0u || (i >= 3 && i + 3 < ni && j >= 3 && j + 3 < nj));
  const int x = i + NG, y = j + NG;
  const int64_t st = (int64_t)nys * nxs;  // level stride of scalar and w arrays
  const float* col = a.s + (int64_t)y * nxs + x;
  const int64_t wcol = (int64_t)y * nxs + x;
  const int64_t mcol = (int64_t)j * ni + i;
  const int64_t mst = (int64_t)nj * ni;

  // z-face flux at the bottom face of the first level of the chunk.
  float fzb = 0.0f;
  if (live && kb >= 2)
    fzb = zflux<WENO>(col, st, a.rrw, kb * st + wcol, a.c1[(kb - 1) * mst + mcol],
                      a.c2[(kb - 1) * mst + mcol], kb, nk, a.weps);
  for (int k = kb; k <= ke; ++k) {
    // x-face fluxes: faces i0..i0+TX of the TY rows.
    for (int slot = threadIdx.x; slot < TY * (TX + 1); slot += NT) {
      const int r = slot / (TX + 1), t = slot % (TX + 1);
      const int f = i0 + t, jj = j0 + r;
      if (f <= ni && jj < nj) {
        float v = xflux<WENO>(a.s, a.rru, k, jj + NG, f + NG, nys, nxs, nxu, a.weps);
        lfx[r][t] = v;
        if (a.fx && (t < TX || f == ni)) a.fx[ix3(k, jj + NG, f + NG, nys, nxu)] = v;
      }
    }
    // y-face fluxes: rows j0..j0+TY of the TX columns.
    for (int slot = threadIdx.x; slot < (TY + 1) * TX; slot += NT) {
      const int r = slot / TX, t = slot % TX;
      const int ii = i0 + t, g = j0 + r;
      if (ii < ni && g <= nj) {
        float v = yflux<WENO>(a.s, a.rrv, k, g + NG, ii + NG, nys, nxs, nyv, a.weps);
        lfy[r][t] = v;
        if (a.fy && (r < TY || g == nj)) a.fy[ix3(k, g + NG, ii + NG, nyv, nxs)] = v;
      }
    }
    float fzt = 0.0f;
    if (live && k + 1 <= nk)
      fzt = zflux<WENO>(col, st, a.rrw, (int64_t)(k + 1) * st + wcol, a.c1[(int64_t)k * mst + mcol],
                        a.c2[(int64_t)k * mst + mcol], k + 1, nk, a.weps);
    __syncthreads();
    if (live && a.fz) {
      if (k == 1) a.fz[1 * st + wcol] = 0.0f;
      a.fz[(int64_t)(k + 1) * st + wcol] = (k + 1 <= nk) ? fzt : 0.0f;
      if (k == kb && kb >= 2) a.fz[(int64_t)kb * st + wcol] = fzb;
    }
    if (owned) {
      float fw = lfx[ty][tx], fe = lfx[ty][tx + 1];
      float fs = lfy[ty][tx], fn = lfy[ty + 1][tx];
      const float sv = col[(int64_t)k * st];
      float ax, ay, az;
      if (obc != 0u) {
        const float uw = a.rru[ix3(k, y, x, nys, nxu)], ue = a.rru[ix3(k, y, x + 1, nys, nxu)];
        const float vs = a.rrv[ix3(k, y, x, nyv, nxs)], vn = a.rrv[ix3(k, y + 1, x, nyv, nxs)];
        const bool iw = (i == 0) && side_of(obc, 0) == 2u && uw >= 0.0f;
        const bool ie = (i + 1 == ni) && side_of(obc, 1) == 2u && ue <= 0.0f;
        const bool js = (j == 0) && side_of(obc, 2) == 2u && vs >= 0.0f;
        const bool jn = (j + 1 == nj) && side_of(obc, 3) == 2u && vn <= 0.0f;
        if (iw) fw = fe;
        if (ie) fe = fw;
        if (js) fs = fn;
        if (jn) fn = fs;
        ax = (fw - fe) * a.rdx;
        ay = (fs - fn) * a.rdy;
        if (iw) ax = ax - (sv * (ue - uw)) * a.rdx;
        if (ie) ax = ax - (sv * (ue - uw)) * a.rdx;
        if (js) ay = ay - (sv * (vn - vs)) * a.rdy;
        if (jn) ay = ay - (sv * (vn - vs)) * a.rdy;
      } else {
        ax = (fw - fe) * a.rdx;
        ay = (fs - fn) * a.rdy;
      }
      if (k == 1) az = (-fzt) * a.rdsf;
      else if (k == nk) az = fzb * a.rdsf;
      else az = (fzb - fzt) * a.rdsf;
      float total = (ax + ay) + az;
      const int64_t q = (int64_t)(k - 1) * mst + mcol;
      total = total + sv * a.divx[q];
      a.tend[q] = (total * a.rr0[q]) * a.gz[mcol];
    }
    fzb = fzt;
    __syncthreads();
  }
}

}  // namespace

extern "C" {

int gm1hip_terrain_mass_flux5(const float* u3d, const float* v3d, const float* w3d,
                              float* rru, float* rrv, float* rrw, const float* rho0_3d,
                              const float* rf0, const float* rgzu, const float* rgzv,
                              const float* dzdx, const float* dzdy, const float* gz,
                              float* divx, int ni, int nj, int nk, float dx, float dy,
                              float dz, void* stream) {
  hipStream_t s = (hipStream_t)stream;
  dim3 blk(64);
  dim3 g1((ni + 63) / 64, nj, nk);
  hipLaunchKernelGGL(k_tmf_horizontal, g1, blk, 0, s, u3d, v3d, rho0_3d, rgzu, rgzv, rru, rrv, rrw, ni, nj, nk);
  GM1_LAUNCH_CHECK();
  dim3 g2((ni + 63) / 64, nj, nk - 1);
  hipLaunchKernelGGL(k_tmf_vertical, g2, blk, 0, s, w3d, rf0, dzdx, dzdy, gz, rru, rrv, rrw, ni, nj, nk, dz);
  GM1_LAUNCH_CHECK();
  const float rdx = 1.0f / dx, rdy = 1.0f / dy, rdsf = 1.0f / dz;
  hipLaunchKernelGGL(k_tmf_divergence, g1, blk, 0, s, rru, rrv, rrw, divx, ni, nj, nk, rdx, rdy, rdsf);
  GM1_LAUNCH_CHECK();
  return 0;
}

int gm1hip_scalar_advection5(const float* scalar, const float* rru, const float* rrv,
                             const float* rrw, const float* c1, const float* c2,
                             const float* rr0, const float* gz, const float* divx,
                             float* tendency, float* flux_x, float* flux_y, float* flux_z,
                             int ni, int nj, int nk, int use_weno, double weps, float dx,
                             float dy, float dz, unsigned obc, void* stream) {
  ScalarAdvArgs a;
  a.s = scalar; a.rru = rru; a.rrv = rrv; a.rrw = rrw; a.c1 = c1; a.c2 = c2;
  a.divx = divx; a.rr0 = rr0; a.gz = gz; a.tend = tendency;
  a.fx = flux_x; a.fy = flux_y; a.fz = flux_z;
  a.ni = ni; a.nj = nj; a.nk = nk;
  a.rdx = 1.0f / dx; a.rdy = 1.0f / dy; a.rdsf = 1.0f / dz;
  a.weps = weps; a.obc = obc;
  const int tiles = ((ni + TX - 1) / TX) * ((nj + TY - 1) / TY);
  // enough column chunks to put >= 2048 tiles in flight on 256 CUs
  int chunks = (2048 + tiles - 1) / tiles;
  if (chunks > nk) chunks = nk;
  if (chunks < 1) chunks = 1;
  a.kchunk = (nk + chunks - 1) / chunks;
  chunks = (nk + a.kchunk - 1) / a.kchunk;
  dim3 grid((ni + TX - 1) / TX, (nj + TY - 1) / TY, chunks);
  hipStream_t s = (hipStream_t)stream;
  if (use_weno) GM1_TIMED(TK_SCALAR_ADV5, s, hipLaunchKernelGGL(k_scalar_adv5<true>, grid, dim3(NT), 0, s, a));
  else GM1_TIMED(TK_SCALAR_ADV5, s, hipLaunchKernelGGL(k_scalar_adv5<false>, grid, dim3(NT), 0, s, a));
  GM1_LAUNCH_CHECK();
  return 0;
}

}  // extern "C"
