// Fifth-order flux-form momentum advection for gfx950.
//
// Replaces the kernel family of cuda_momentum_advection5.py: terrain mass
// fluxes, divx, and the advu/advv/advw increments added to uten/vten/wten,
// for any CM1 lateral family (periodic sides read across the seam at the
// wrapped index; open sides take the inflow edits).
//   * k_mom_mass_h   : rru on faces 0..ni, rrv on faces 0..nj.
//   * k_mom_mass_v   : one thread per column marches k, forms rrw at every
//                      w level and divx from the rrw pair it holds in
//                      registers (one launch instead of two).
//   * k_mom_adv{u,v,w}: one thread per output word, i fastest.
#include "cm1_common.h"
#include "timing.h"

using namespace gm1;

namespace {

struct MomArgs {
  const float* u;
  const float* v;
  const float* w;
  const float* rru;
  const float* rrv;
  const float* rrw;
  const float* divx;
  const float* rho0;
  const float* c2;
  const float* gzu;
  const float* gzv;
  const float* rrf0;
  const float* gz;
  const float* uh;
  const float* vh;
  float* uten;
  float* vten;
  float* wten;
  int ni, nj, nk;
  float rdx, rdy, rdz;
  unsigned obc;
  // Flat stretched column (terrain_flag=.false.): mh (nk+2) then mf (nk+3),
  // CM1-indexed; null on the terrain-coordinate family.  rr0 is the base
  // state's reciprocal density, read at the first column.
  const float* flat;
  const float* rr0;
};

GM1_DEV bool per(unsigned obc, unsigned s) { return side_of(obc, s) == 1u; }
GM1_DEV bool opn(unsigned obc, unsigned s) { return side_of(obc, s) == 2u; }
// periodic wrap of a cell index (0..n-1) and of a face index (0..n)
GM1_DEV int wc(int i, int n, bool lo, bool hi) { return (lo && i < 0) ? i + n : ((hi && i >= n) ? i - n : i); }
GM1_DEV int wf(int i, int n, bool lo, bool hi) { return (lo && i < 0) ? i + n : ((hi && i > n) ? i - n : i); }
GM1_DEV int clampi(int a, int n) { return a < 0 ? 0 : (a >= n ? n - 1 : a); }

// ------------------------------------------------------------- mass fluxes
// Levels per thread of the mass-flux kernels (loads of MKB levels in flight).
constexpr int MKB = 8;

__global__ __launch_bounds__(256) void k_mom_mass_h(
    const float* __restrict__ u, const float* __restrict__ v, const float* __restrict__ rho0,
    const float* __restrict__ rgzu, const float* __restrict__ rgzv, float* __restrict__ rru,
    float* __restrict__ rrv, int ni, int nj, int nk) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;  // 0..ni (face or cell)
  const int j = blockIdx.y;                             // 0..nj (row or face)
  const float half = 0.5f;
  // MKB levels per thread, the loads of all of them in flight together (past
  // the top the level is clamped and its words are stored again)
#pragma unroll
  for (int uu = 0; uu < MKB; ++uu) {
    const int k0 = min((int)blockIdx.z * MKB + uu, nk - 1), k = k0 + 1;  // clamped: the top level again
    const int64_t plane = (int64_t)k0 * nj * ni;
    if (i <= ni && j < nj) {
      const int l = (i == 0) ? 0 : i - 1, r = (i == ni) ? ni - 1 : i;
      float d = rho0[plane + (int64_t)j * ni + l] + rho0[plane + (int64_t)j * ni + r];
      const int64_t p = ix3(k, j + NG, i + NG, sy(nj), ux(ni));
      rru[p] = (half * d) * (u[p] * rgzu[(int64_t)j * (ni + 1) + i]);
    }
    if (i < ni && j <= nj) {
      const int b = (j == 0) ? 0 : j - 1, t = (j == nj) ? nj - 1 : j;
      float d = rho0[plane + (int64_t)b * ni + i] + rho0[plane + (int64_t)t * ni + i];
      const int64_t p = ix3(k, j + NG, i + NG, vy(nj), sx(ni));
      rrv[p] = (half * d) * (v[p] * rgzv[(int64_t)j * ni + i]);
    }
  }
}

__global__ __launch_bounds__(256) void k_mom_mass_v(
    const float* __restrict__ w, const float* __restrict__ rf0, const float* __restrict__ dzdx,
    const float* __restrict__ dzdy, const float* __restrict__ gz, const float* __restrict__ rru,
    const float* __restrict__ rrv, float* __restrict__ rrw, float* __restrict__ divx, int ni,
    int nj, int nk, float dz, float rdx, float rdy) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  if (i >= ni) return;
  const int nxu = ux(ni), nyu = sy(nj), nxs = sx(ni), nyv = vy(nj);
  const int x = i + NG, y = j + NG;
  const float half = 0.5f;
  const float zt = (float)nk * dz, rzt = 1.0f / zt, rdz = 1.0f / dz;
  const int64_t h = (int64_t)j * ni + i;
  const float ddx = dzdx[h], ddy = dzdy[h], g = gz[h];
  const int64_t wst = (int64_t)nyu * nxs;
  const int64_t wcol = (int64_t)y * nxs + x;
  // rrw at w level kf (0 at kf = 1 and nk + 1)
  auto rrw_at = [&](int kf) {
    if (kf <= 1 || kf > nk) return 0.0f;
    const float sf = (float)(kf - 1) * dz, sm = (float)((double)kf - 1.5) * dz;
    float r2 = (sf - sm) * rdz;
    const float r1 = half * (1.0f - r2);
    r2 = half * r2;
    float uxv = r2 * (rru[ix3(kf, y, x, nyu, nxu)] + rru[ix3(kf, y, x + 1, nyu, nxu)]) +
                r1 * (rru[ix3(kf - 1, y, x, nyu, nxu)] + rru[ix3(kf - 1, y, x + 1, nyu, nxu)]);
    float vyv = r2 * (rrv[ix3(kf, y, x, nyv, nxs)] + rrv[ix3(kf, y + 1, x, nyv, nxs)]) +
                r1 * (rrv[ix3(kf - 1, y, x, nyv, nxs)] + rrv[ix3(kf - 1, y + 1, x, nyv, nxs)]);
    float t = uxv * ddx + vyv * ddy;
    t = t * (sf - zt);
    t = t * g;
    t = t * rzt;
    return rf0[(int64_t)(kf - 1) * nj * ni + h] * w[(int64_t)kf * wst + wcol] + t;
  };
  // this thread's chunk: w levels kf0 .. kf0+MKB-1 of 2..nk+1 (rrw(1) = 0 by the first chunk)
  const int kf0 = 2 + (int)blockIdx.z * MKB;
  if (kf0 == 2) rrw[wst + wcol] = 0.0f;
  float below = rrw_at(kf0 - 1);
#pragma unroll
  for (int q = 0; q < MKB; ++q) {
    // clamped past w level nk+1: the top pair is formed and stored again
    const int kf = min(kf0 + q, nk + 1);
    const float val = rrw_at(kf);
    rrw[(int64_t)kf * wst + wcol] = val;
    // divx at scalar level k = kf-1 from the rrw pair (k, k+1)
    const int k = kf - 1;
    float d = (rru[ix3(k, y, x + 1, nyu, nxu)] - rru[ix3(k, y, x, nyu, nxu)]) * rdx +
              (rrv[ix3(k, y + 1, x, nyv, nxs)] - rrv[ix3(k, y, x, nyv, nxs)]) * rdy;
    d = d + (val - below) * rdz;
    divx[(int64_t)(k - 1) * nj * ni + h] = (fabsf(d) < 1.0e-20f) ? 0.0f : d;
    below = (kf <= nk) ? val : below;  // the top pair keeps its lower word for the clamped repeats
  }
}

// Flat stretched column (solve2.F:493-520, 605-617 with terrain_flag=.false.):
// rru = rho0(1,1,k)*u, rrv = rho0(1,1,k)*v, rrw = rf0(1,1,k)*w (zero at both
// lids), divx's vertical term ((d*rdz)*mh(k)) and smeps 1e-30.
__global__ __launch_bounds__(256) void k_mom_mass_h_flat(
    const float* __restrict__ u, const float* __restrict__ v, const float* __restrict__ rho0,
    float* __restrict__ rru, float* __restrict__ rrv, int ni, int nj, int nk) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
#pragma unroll
  for (int uu = 0; uu < MKB; ++uu) {
    const int k0 = min((int)blockIdx.z * MKB + uu, nk - 1), k = k0 + 1;  // clamped: the top level again
    const float rc = rho0[(int64_t)k0 * nj * ni];
    if (i <= ni && j < nj) {
      const int64_t p = ix3(k, j + NG, i + NG, sy(nj), ux(ni));
      rru[p] = rc * u[p];
    }
    if (i < ni && j <= nj) {
      const int64_t p = ix3(k, j + NG, i + NG, vy(nj), sx(ni));
      rrv[p] = rc * v[p];
    }
  }
}

__global__ __launch_bounds__(256) void k_mom_mass_v_flat(
    const float* __restrict__ w, const float* __restrict__ rf0, const float* __restrict__ rru,
    const float* __restrict__ rrv, const float* __restrict__ mh, float* __restrict__ rrw,
    float* __restrict__ divx, int ni, int nj, int nk, float rdx, float rdy, float rdz) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  if (i >= ni) return;
  const int nxu = ux(ni), nyu = sy(nj), nxs = sx(ni), nyv = vy(nj);
  const int x = i + NG, y = j + NG;
  const int64_t h = (int64_t)j * ni + i;
  const int64_t wst = (int64_t)nyu * nxs;
  const int64_t wcol = (int64_t)y * nxs + x;
  auto rrw_at = [&](int kf) {
    return (kf <= 1 || kf > nk) ? 0.0f : rf0[(int64_t)(kf - 1) * nj * ni] * w[(int64_t)kf * wst + wcol];
  };
  // this thread's chunk: w levels kf0 .. kf0+MKB-1 of 2..nk+1 (rrw(1) = 0 by the first chunk)
  const int kf0 = 2 + (int)blockIdx.z * MKB;
  if (kf0 == 2) rrw[wst + wcol] = 0.0f;
  float below = rrw_at(kf0 - 1);
#pragma unroll
  for (int q = 0; q < MKB; ++q) {
    const int kf = min(kf0 + q, nk + 1);  // clamped past w level nk+1: the top pair again
    const float val = rrw_at(kf);
    rrw[(int64_t)kf * wst + wcol] = val;
    const int k = kf - 1;
    float d = (rru[ix3(k, y, x + 1, nyu, nxu)] - rru[ix3(k, y, x, nyu, nxu)]) * rdx +
              (rrv[ix3(k, y + 1, x, nyv, nxs)] - rrv[ix3(k, y, x, nyv, nxs)]) * rdy;
    d = d + ((val - below) * rdz) * mh[k];
    divx[(int64_t)(k - 1) * nj * ni + h] = (fabsf(d) < 1.0e-30f) ? 0.0f : d;
    below = (kf <= nk) ? val : below;  // the top pair keeps its lower word for the clamped repeats
  }
}

// ------------------------------------------------------------- advu
GM1_DEV float ufx(const MomArgs& a, int k, int j, int f) {
  const bool pw = per(a.obc, 0), pe = per(a.obc, 1);
  const int nxu = ux(a.ni), nyu = sy(a.nj);
  const int y = j + NG, x = f + NG;
  const float* R = a.rru + ix3(k, y, 0, nyu, nxu);
  const float m = 0.5f * (R[wf(f, a.ni, pw, pe) + NG] + R[wf(f + 1, a.ni, pw, pe) + NG]);
  const float* U = a.u + ix3(k, y, 0, nyu, nxu);
  const float r = (m >= 0.0f) ? flx5(U[x - 2], U[x - 1], U[x], U[x + 1], U[x + 2])
                              : flx5(U[x + 3], U[x + 2], U[x + 1], U[x], U[x - 1]);
  return m * r;
}

GM1_DEV float ufy_base(const MomArgs& a, int k, int p, int f) {
  const bool pw = per(a.obc, 0), pe = per(a.obc, 1);
  const int nxu = ux(a.ni), nyu = sy(a.nj), nxv = sx(a.ni), nyv = vy(a.nj);
  const int y = p + NG, x = f + NG;
  const float* R = a.rrv + ix3(k, y, 0, nyv, nxv);
  const float m = 0.5f * (R[wc(f, a.ni, pw, pe) + NG] + R[wc(f - 1, a.ni, pw, pe) + NG]);
  const float* U = a.u + ix3(k, 0, x, nyu, nxu);
  auto q = [&](int yy) { return U[(int64_t)yy * nxu]; };
  const float r = (m >= 0.0f) ? flx5(q(y - 3), q(y - 2), q(y - 1), q(y), q(y + 1))
                              : flx5(q(y + 2), q(y + 1), q(y), q(y - 1), q(y - 2));
  return m * r;
}

GM1_DEV float ufy(const MomArgs& a, int k, int p, int f) {
  const bool pw = per(a.obc, 0), pe = per(a.obc, 1);
  const int nxv = sx(a.ni), nyv = vy(a.nj);
  const float* R = a.rrv + ix3(k, p + NG, 0, nyv, nxv);
  const float mt = R[wc(f, a.ni, pw, pe) + NG] + R[wc(f - 1, a.ni, pw, pe) + NG];
  if (p == 0 && opn(a.obc, 2) && mt >= 0.0f) return ufy_base(a, k, 1, f);
  if (p == a.nj && opn(a.obc, 3) && mt <= 0.0f) return ufy_base(a, k, a.nj - 1, f);
  return ufy_base(a, k, p, f);
}

// vertical flux of u (stag 0) or v (stag 1) at w level kf of column (j, i)
template <int STAG>
GM1_DEV float uvz(const MomArgs& a, int kf, int j, int i) {
  const int ni = a.ni, nj = a.nj, nk = a.nk;
  const int nxs = sx(ni), nys = sy(nj);
  float m, c2h;
  const float* c2k = a.c2 + (int64_t)(kf - 1) * nj * ni;
  if (STAG == 0) {
    const bool pw = per(a.obc, 0), pe = per(a.obc, 1);
    const float* R = a.rrw + ix3(kf, j + NG, 0, nys, nxs);
    m = 0.5f * (R[wc(i, ni, pw, pe) + NG] + R[wc(i - 1, ni, pw, pe) + NG]);
    const int jc = clampi(j, nj);
    c2h = 0.5f * (c2k[(int64_t)jc * ni + clampi(i - 1, ni)] + c2k[(int64_t)jc * ni + clampi(i, ni)]);
  } else {
    const bool ps = per(a.obc, 2), pn = per(a.obc, 3);
    m = 0.5f * (a.rrw[ix3(kf, wc(j, nj, ps, pn) + NG, i + NG, nys, nxs)] +
                a.rrw[ix3(kf, wc(j - 1, nj, ps, pn) + NG, i + NG, nys, nxs)]);
    const int ic = clampi(i, ni);
    c2h = 0.5f * (c2k[(int64_t)clampi(j - 1, nj) * ni + ic] + c2k[(int64_t)clampi(j, nj) * ni + ic]);
  }
  const float c1h = 1.0f - c2h;
  const float* q = (STAG == 0) ? a.u : a.v;
  const int nxq = (STAG == 0) ? ux(ni) : sx(ni);
  const int nyq = (STAG == 0) ? sy(nj) : vy(nj);
  const int64_t st = (int64_t)nyq * nxq;
  const float* col = q + (int64_t)(j + NG) * nxq + (i + NG);
  auto Q = [&](int lev) { return col[(int64_t)lev * st]; };
  float r;
  if (kf == 2) {
    r = (m >= 0.0f) ? c1h * Q(1) + c2h * Q(2) : flx3(Q(3), Q(2), Q(1));
  } else if (kf == nk) {
    r = (m >= 0.0f) ? flx3(Q(nk - 2), Q(nk - 1), Q(nk)) : c1h * Q(nk - 1) + c2h * Q(nk);
  } else if (kf == 3) {
    r = (m >= 0.0f) ? flx3(Q(1), Q(2), Q(3)) : flx5(Q(5), Q(4), Q(3), Q(2), Q(1));
  } else if (kf == nk - 1) {
    r = (m >= 0.0f) ? flx5(Q(kf - 3), Q(kf - 2), Q(kf - 1), Q(kf), Q(kf + 1))
                    : flx3(Q(kf + 1), Q(kf), Q(kf - 1));
  } else {
    r = (m >= 0.0f) ? flx5(Q(kf - 3), Q(kf - 2), Q(kf - 1), Q(kf), Q(kf + 1))
                    : flx5(Q(kf + 2), Q(kf + 1), Q(kf), Q(kf - 1), Q(kf - 2));
  }
  return m * r;
}

template <bool FLAT>
__global__ __launch_bounds__(256) void k_mom_advu(MomArgs a, int f0, int nf) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y, k0 = blockIdx.z, k = k0 + 1;
  if (t >= nf) return;
  const int f = f0 + t;
  const int ni = a.ni, nj = a.nj, nk = a.nk;
  const bool pw = per(a.obc, 0), pe = per(a.obc, 1);
  const int west = wc(f - 1, ni, pw, pe), east = wc(f, ni, pw, pe);
  const int nxu = ux(ni), nyu = sy(nj), nxv = sx(ni), nyv = vy(nj);
  const int y = j + NG, x = f + NG;
  float ax = (-(ufx(a, k, j, f) - ufx(a, k, j, f - 1))) * a.rdx;
  float ay = (-(ufy(a, k, j + 1, f) - ufy(a, k, j, f))) * a.rdy;
  const float uc = a.u[ix3(k, y, x, nyu, nxu)];
  if (j == 0 && opn(a.obc, 2)) {
    const float* R0 = a.rrv + ix3(k, NG, 0, nyv, nxv);
    const float* R1 = a.rrv + ix3(k, NG + 1, 0, nyv, nxv);
    const float m = R0[east + NG] + R0[west + NG];
    if (m >= 0.0f) {
      const float c = (R1[west + NG] - R0[west + NG]) + (R1[east + NG] - R0[east + NG]);
      ay = FLAT ? ay - (((uc * 0.5f) * c) * a.rdy) * a.vh[y] : ay - (uc * 0.5f) * (c * (a.rdy * a.vh[y]));
    }
  } else if (j == nj - 1 && opn(a.obc, 3)) {
    const float* RN = a.rrv + ix3(k, NG + nj, 0, nyv, nxv);
    const float* R0 = a.rrv + ix3(k, y, 0, nyv, nxv);
    const float* R1 = a.rrv + ix3(k, y + 1, 0, nyv, nxv);
    const float m = RN[east + NG] + RN[west + NG];
    if (m <= 0.0f) {
      const float c = (R1[west + NG] - R0[west + NG]) + (R1[east + NG] - R0[east + NG]);
      ay = FLAT ? ay - (((uc * 0.5f) * c) * a.rdy) * a.vh[y] : ay - (uc * 0.5f) * (c * (a.rdy * a.vh[y]));
    }
  }
  float az;
  if (k == 1) az = (-uvz<0>(a, 2, j, f)) * a.rdz;
  else if (k == nk) az = uvz<0>(a, nk, j, f) * a.rdz;
  else az = (-(uvz<0>(a, k + 1, j, f) - uvz<0>(a, k, j, f))) * a.rdz;
  if constexpr (FLAT) az = az * a.flat[k];
  float tot = (ax + ay) + az;
  const int64_t row = ((int64_t)k0 * nj + j) * ni;
  tot = tot + uc * (0.5f * (a.divx[row + east] + a.divx[row + west]));
  const float sc = FLAT ? a.rr0[(int64_t)k0 * nj * ni]
                          : a.gzu[(int64_t)j * (ni + 1) + f] / (0.5f * (a.rho0[row + west] + a.rho0[row + east]));
  const int64_t p = ix3(k, y, x, nyu, nxu);
  a.uten[p] = a.uten[p] + tot * sc;
}

// ------------------------------------------------------------- advv
GM1_DEV float vfx_base(const MomArgs& a, int k, int g, int p) {
  const bool ps = per(a.obc, 2), pn = per(a.obc, 3);
  const int nxu = ux(a.ni), nyu = sy(a.nj), nxv = sx(a.ni), nyv = vy(a.nj);
  const int y = g + NG, x = p + NG;
  const float m = 0.5f * (a.rru[ix3(k, wc(g, a.nj, ps, pn) + NG, x, nyu, nxu)] +
                          a.rru[ix3(k, wc(g - 1, a.nj, ps, pn) + NG, x, nyu, nxu)]);
  const float* V = a.v + ix3(k, y, 0, nyv, nxv);
  const float r = (m >= 0.0f) ? flx5(V[x - 3], V[x - 2], V[x - 1], V[x], V[x + 1])
                              : flx5(V[x + 2], V[x + 1], V[x], V[x - 1], V[x - 2]);
  return m * r;
}

GM1_DEV float vfx(const MomArgs& a, int k, int g, int p) {
  const bool ps = per(a.obc, 2), pn = per(a.obc, 3);
  const int nxu = ux(a.ni), nyu = sy(a.nj);
  const int x = p + NG;
  const float mt = a.rru[ix3(k, wc(g, a.nj, ps, pn) + NG, x, nyu, nxu)] +
                   a.rru[ix3(k, wc(g - 1, a.nj, ps, pn) + NG, x, nyu, nxu)];
  if (p == 0 && opn(a.obc, 0) && mt >= 0.0f) return vfx_base(a, k, g, 1);
  if (p == a.ni && opn(a.obc, 1) && mt <= 0.0f) return vfx_base(a, k, g, a.ni - 1);
  return vfx_base(a, k, g, p);
}

GM1_DEV float vfy(const MomArgs& a, int k, int p, int i) {
  const bool ps = per(a.obc, 2), pn = per(a.obc, 3);
  const int nxv = sx(a.ni), nyv = vy(a.nj);
  const int y = p + NG, x = i + NG;
  const float m = 0.5f * (a.rrv[ix3(k, wf(p, a.nj, ps, pn) + NG, x, nyv, nxv)] +
                          a.rrv[ix3(k, wf(p + 1, a.nj, ps, pn) + NG, x, nyv, nxv)]);
  const float* V = a.v + ix3(k, 0, x, nyv, nxv);
  auto q = [&](int yy) { return V[(int64_t)yy * nxv]; };
  const float r = (m >= 0.0f) ? flx5(q(y - 2), q(y - 1), q(y), q(y + 1), q(y + 2))
                              : flx5(q(y + 3), q(y + 2), q(y + 1), q(y), q(y - 1));
  return m * r;
}

template <bool FLAT>
__global__ __launch_bounds__(256) void k_mom_advv(MomArgs a, int g0, int nr) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int r = blockIdx.y, k0 = blockIdx.z, k = k0 + 1;
  const int ni = a.ni, nj = a.nj, nk = a.nk;
  if (i >= ni || r >= nr) return;
  const int g = g0 + r;
  const bool ps = per(a.obc, 2), pn = per(a.obc, 3);
  const int south = wc(g - 1, nj, ps, pn), north = wc(g, nj, ps, pn);
  const int nxu = ux(ni), nyu = sy(nj), nxv = sx(ni), nyv = vy(nj);
  const int y = g + NG, x = i + NG;
  float ax = (-(vfx(a, k, g, i + 1) - vfx(a, k, g, i))) * a.rdx;
  float ay = (-(vfy(a, k, g, i) - vfy(a, k, g - 1, i))) * a.rdy;
  const float vc = a.v[ix3(k, y, x, nyv, nxv)];
  const float* Rs = a.rru + ix3(k, south + NG, 0, nyu, nxu);
  const float* Rn = a.rru + ix3(k, north + NG, 0, nyu, nxu);
  if (i == 0 && opn(a.obc, 0)) {
    const float m = Rn[NG] + Rs[NG];
    if (m >= 0.0f) {
      const float c = (Rs[NG + 1] - Rs[NG]) + (Rn[NG + 1] - Rn[NG]);
      ax = FLAT ? ax - (((vc * 0.5f) * c) * a.rdx) * a.uh[NG] : ax - (vc * 0.5f) * (c * (a.rdx * a.uh[NG]));
    }
  } else if (i == ni - 1 && opn(a.obc, 1)) {
    const float m = Rn[NG + ni] + Rs[NG + ni];
    if (m <= 0.0f) {
      const float c = (Rs[x + 1] - Rs[x]) + (Rn[x + 1] - Rn[x]);
      ax = FLAT ? ax - (((vc * 0.5f) * c) * a.rdx) * a.uh[x] : ax - (vc * 0.5f) * (c * (a.rdx * a.uh[x]));
    }
  }
  float az;
  if (k == 1) az = (-uvz<1>(a, 2, g, i)) * a.rdz;
  else if (k == nk) az = uvz<1>(a, nk, g, i) * a.rdz;
  else az = (-(uvz<1>(a, k + 1, g, i) - uvz<1>(a, k, g, i))) * a.rdz;
  if constexpr (FLAT) az = az * a.flat[k];
  float tot = (ax + ay) + az;
  const int64_t plane = (int64_t)k0 * nj * ni;
  tot = tot + vc * (0.5f * (a.divx[plane + (int64_t)north * ni + i] + a.divx[plane + (int64_t)south * ni + i]));
  const float sc = FLAT ? a.rr0[plane]
                          : a.gzv[(int64_t)g * ni + i] /
                                (0.5f * (a.rho0[plane + (int64_t)south * ni + i] + a.rho0[plane + (int64_t)north * ni + i]));
  const int64_t p = ix3(k, y, x, nyv, nxv);
  a.vten[p] = a.vten[p] + tot * sc;
}

// ------------------------------------------------------------- advw
GM1_DEV float wfx_base(const MomArgs& a, int kf, int j, int p) {
  const int ni = a.ni, nj = a.nj;
  const float* c2k = a.c2 + (int64_t)(kf - 1) * nj * ni + (int64_t)j * ni;
  const float c2h = 0.5f * (c2k[clampi(p - 1, ni)] + c2k[clampi(p, ni)]);
  const float c1h = 1.0f - c2h;
  const int nxu = ux(ni), nyu = sy(nj), nxs = sx(ni), nys = sy(nj);
  const int y = j + NG, x = p + NG;
  const float m = c2h * a.rru[ix3(kf, y, x, nyu, nxu)] + c1h * a.rru[ix3(kf - 1, y, x, nyu, nxu)];
  const float* W = a.w + ix3(kf, y, 0, nys, nxs);
  const float r = (m >= 0.0f) ? flx5(W[x - 3], W[x - 2], W[x - 1], W[x], W[x + 1])
                              : flx5(W[x + 2], W[x + 1], W[x], W[x - 1], W[x - 2]);
  return m * r;
}

GM1_DEV float wfx(const MomArgs& a, int kf, int j, int p) {
  const int ni = a.ni, nj = a.nj;
  const float* c2k = a.c2 + (int64_t)(kf - 1) * nj * ni + (int64_t)j * ni;
  const float c2h = 0.5f * (c2k[clampi(p - 1, ni)] + c2k[clampi(p, ni)]);
  const float c1h = 1.0f - c2h;
  const int nxu = ux(ni), nyu = sy(nj);
  const int y = j + NG, x = p + NG;
  const float m = c2h * a.rru[ix3(kf, y, x, nyu, nxu)] + c1h * a.rru[ix3(kf - 1, y, x, nyu, nxu)];
  if (p == 0 && opn(a.obc, 0) && m >= 0.0f) return wfx_base(a, kf, j, 1);
  if (p == ni && opn(a.obc, 1) && m <= 0.0f) return wfx_base(a, kf, j, ni - 1);
  return wfx_base(a, kf, j, p);
}

GM1_DEV float wfy_base(const MomArgs& a, int kf, int p, int i) {
  const int ni = a.ni, nj = a.nj;
  const float* c2k = a.c2 + (int64_t)(kf - 1) * nj * ni;
  const float c2h = 0.5f * (c2k[(int64_t)clampi(p - 1, nj) * ni + i] + c2k[(int64_t)clampi(p, nj) * ni + i]);
  const float c1h = 1.0f - c2h;
  const int nxs = sx(ni), nyv = vy(nj), nys = sy(nj);
  const int y = p + NG, x = i + NG;
  const float m = c2h * a.rrv[ix3(kf, y, x, nyv, nxs)] + c1h * a.rrv[ix3(kf - 1, y, x, nyv, nxs)];
  const float* W = a.w + ix3(kf, 0, x, nys, nxs);
  auto q = [&](int yy) { return W[(int64_t)yy * nxs]; };
  const float r = (m >= 0.0f) ? flx5(q(y - 3), q(y - 2), q(y - 1), q(y), q(y + 1))
                              : flx5(q(y + 2), q(y + 1), q(y), q(y - 1), q(y - 2));
  return m * r;
}

GM1_DEV float wfy(const MomArgs& a, int kf, int p, int i) {
  const int ni = a.ni, nj = a.nj;
  const float* c2k = a.c2 + (int64_t)(kf - 1) * nj * ni;
  const float c2h = 0.5f * (c2k[(int64_t)clampi(p - 1, nj) * ni + i] + c2k[(int64_t)clampi(p, nj) * ni + i]);
  const float c1h = 1.0f - c2h;
  const int nxs = sx(ni), nyv = vy(nj);
  const int y = p + NG, x = i + NG;
  const float m = c2h * a.rrv[ix3(kf, y, x, nyv, nxs)] + c1h * a.rrv[ix3(kf - 1, y, x, nyv, nxs)];
  if (p == 0 && opn(a.obc, 2) && m >= 0.0f) return wfy_base(a, kf, 1, i);
  if (p == nj && opn(a.obc, 3) && m <= 0.0f) return wfy_base(a, kf, nj - 1, i);
  return wfy_base(a, kf, p, i);
}

GM1_DEV float wz(const MomArgs& a, int k, int j, int i) {
  const int nxs = sx(a.ni), nys = sy(a.nj), nk = a.nk;
  const int64_t st = (int64_t)nys * nxs;
  const int64_t c = (int64_t)(j + NG) * nxs + (i + NG);
  const float m = 0.5f * (a.rrw[(int64_t)k * st + c] + a.rrw[(int64_t)(k + 1) * st + c]);
  auto Q = [&](int lev) { return a.w[(int64_t)lev * st + c]; };
  float r;
  if (k == 1) r = (m >= 0.0f) ? 0.5f * (Q(1) + Q(2)) : flx3(Q(3), Q(2), Q(1));
  else if (k == 2) r = (m >= 0.0f) ? flx3(Q(1), Q(2), Q(3)) : flx5(Q(5), Q(4), Q(3), Q(2), Q(1));
  else if (k == nk - 1) r = (m >= 0.0f) ? flx5(Q(k - 2), Q(k - 1), Q(k), Q(k + 1), Q(k + 2)) : flx3(Q(k + 2), Q(k + 1), Q(k));
  else if (k == nk) r = (m >= 0.0f) ? flx3(Q(k - 1), Q(k), Q(k + 1)) : 0.5f * (Q(k) + Q(k + 1));
  else r = (m >= 0.0f) ? flx5(Q(k - 2), Q(k - 1), Q(k), Q(k + 1), Q(k + 2)) : flx5(Q(k + 3), Q(k + 2), Q(k + 1), Q(k), Q(k - 1));
  return m * r;
}

template <bool FLAT>
__global__ __launch_bounds__(256) void k_mom_advw(MomArgs a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y, kf = blockIdx.z + 2;
  const int ni = a.ni, nj = a.nj;
  if (i >= ni) return;
  const int nxu = ux(ni), nyu = sy(nj), nxs = sx(ni), nys = sy(nj), nyv = vy(nj);
  const int y = j + NG, x = i + NG;
  constexpr bool flat = FLAT;
  float ax, ay;
  if (flat) {
    ax = ((-(wfx(a, kf, j, i + 1) - wfx(a, kf, j, i))) * a.rdx) * a.uh[x];
    ay = ((-(wfy(a, kf, j + 1, i) - wfy(a, kf, j, i))) * a.rdy) * a.vh[y];
  } else {
    ax = (-(wfx(a, kf, j, i + 1) - wfx(a, kf, j, i))) * (a.rdx * a.uh[x]);
    ay = (-(wfy(a, kf, j + 1, i) - wfy(a, kf, j, i))) * (a.rdy * a.vh[y]);
  }
  const int64_t h = (int64_t)j * ni + i;
  const float c2 = a.c2[(int64_t)(kf - 1) * nj * ni + h];
  const float c1 = 1.0f - c2;
  const float wc0 = a.w[ix3(kf, y, x, nys, nxs)];
  if (i == 0 && opn(a.obc, 0)) {
    const float* Rt = a.rru + ix3(kf, y, 0, nyu, nxu);
    const float* Rb = a.rru + ix3(kf - 1, y, 0, nyu, nxu);
    const float m = c2 * Rt[NG] + c1 * Rb[NG];
    if (m >= 0.0f) {
      const float c = c1 * (Rb[NG + 1] - Rb[NG]) + c2 * (Rt[NG + 1] - Rt[NG]);
      ax = flat ? ax - ((wc0 * c) * a.rdx) * a.uh[x] : ax - wc0 * (c * (a.rdx * a.uh[x]));
    }
  } else if (i == ni - 1 && opn(a.obc, 1)) {
    const float* Rt = a.rru + ix3(kf, y, 0, nyu, nxu);
    const float* Rb = a.rru + ix3(kf - 1, y, 0, nyu, nxu);
    const float m = c2 * Rt[NG + ni] + c1 * Rb[NG + ni];
    if (m <= 0.0f) {
      const float c = c1 * (Rb[x + 1] - Rb[x]) + c2 * (Rt[x + 1] - Rt[x]);
      ax = flat ? ax - ((wc0 * c) * a.rdx) * a.uh[x] : ax - wc0 * (c * (a.rdx * a.uh[x]));
    }
  }
  if (j == 0 && opn(a.obc, 2)) {
    const float m = c2 * a.rrv[ix3(kf, NG, x, nyv, nxs)] + c1 * a.rrv[ix3(kf - 1, NG, x, nyv, nxs)];
    if (m >= 0.0f) {
      const float c = c1 * (a.rrv[ix3(kf - 1, NG + 1, x, nyv, nxs)] - a.rrv[ix3(kf - 1, NG, x, nyv, nxs)]) +
                      c2 * (a.rrv[ix3(kf, NG + 1, x, nyv, nxs)] - a.rrv[ix3(kf, NG, x, nyv, nxs)]);
      ay = flat ? ay - ((wc0 * c) * a.rdy) * a.vh[y] : ay - wc0 * (c * (a.rdy * a.vh[y]));
    }
  } else if (j == nj - 1 && opn(a.obc, 3)) {
    const float m = c2 * a.rrv[ix3(kf, NG + nj, x, nyv, nxs)] + c1 * a.rrv[ix3(kf - 1, NG + nj, x, nyv, nxs)];
    if (m <= 0.0f) {
      const float c = c1 * (a.rrv[ix3(kf - 1, y + 1, x, nyv, nxs)] - a.rrv[ix3(kf - 1, y, x, nyv, nxs)]) +
                      c2 * (a.rrv[ix3(kf, y + 1, x, nyv, nxs)] - a.rrv[ix3(kf, y, x, nyv, nxs)]);
      ay = flat ? ay - ((wc0 * c) * a.rdy) * a.vh[y] : ay - wc0 * (c * (a.rdy * a.vh[y]));
    }
  }
  float az = (-(wz(a, kf, j, i) - wz(a, kf - 1, j, i))) * a.rdz;
  if (flat) az = az * a.flat[(a.nk + 2) + kf];
  float tot = (ax + ay) + az;
  tot = tot + wc0 * (c2 * a.divx[(int64_t)(kf - 1) * nj * ni + h] + c1 * a.divx[(int64_t)(kf - 2) * nj * ni + h]);
  const float sc = flat ? a.rrf0[(int64_t)(kf - 1) * nj * ni] : a.rrf0[(int64_t)(kf - 1) * nj * ni + h] * a.gz[h];
  const int64_t p = ix3(kf, y, x, nys, nxs);
  a.wten[p] = a.wten[p] + tot * sc;
}

}  // namespace

extern "C" {

int gm1hip_momentum_advection5(
    const float* u3d, const float* v3d, const float* w3d, float* uten, float* vten, float* wten,
    float* rru, float* rrv, float* rrw, const float* uh, const float* vh, const float* rho0_3d,
    const float* rf0, const float* rrf0, const float* c2, const float* gz, const float* gzu,
    const float* gzv, const float* rgzu, const float* rgzv, const float* dzdx, const float* dzdy,
    float* divx, int ni, int nj, int nk, float dx, float dy, float dz, unsigned obc,
    int stop_after_divergence, void* stream) {
  hipStream_t s = (hipStream_t)stream;
  const float rdx = 1.0f / dx, rdy = 1.0f / dy, rdz = 1.0f / dz;
  dim3 blk(64);
  hipLaunchKernelGGL(k_mom_mass_h, dim3((ni + 1 + 63) / 64, nj + 1, (nk + MKB - 1) / MKB), blk, 0, s, u3d, v3d,
                     rho0_3d, rgzu, rgzv, rru, rrv, ni, nj, nk);
  GM1_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_mom_mass_v, dim3((ni + 63) / 64, nj, (nk + MKB - 1) / MKB), blk, 0, s, w3d, rf0, dzdx, dzdy,
                     gz, rru, rrv, rrw, divx, ni, nj, nk, dz, rdx, rdy);
  GM1_LAUNCH_CHECK();
  if (stop_after_divergence) return 0;
  MomArgs a;
  a.u = u3d; a.v = v3d; a.w = w3d; a.rru = rru; a.rrv = rrv; a.rrw = rrw; a.divx = divx;
  a.rho0 = rho0_3d; a.c2 = c2; a.gzu = gzu; a.gzv = gzv; a.rrf0 = rrf0; a.gz = gz; a.uh = uh; a.vh = vh;
  a.uten = uten; a.vten = vten; a.wten = wten; a.ni = ni; a.nj = nj; a.nk = nk;
  a.rdx = rdx; a.rdy = rdy; a.rdz = rdz; a.obc = obc;
  a.flat = nullptr; a.rr0 = nullptr;
  const bool pw = side_of(obc, 0) == 1u, pe = side_of(obc, 1) == 1u;
  const bool ps = side_of(obc, 2) == 1u, pn = side_of(obc, 3) == 1u;
  const int f0 = pw ? 0 : 1, nf = ni - 1 + (pw ? 1 : 0) + (pe ? 1 : 0);
  const int g0 = ps ? 0 : 1, nr = nj - 1 + (ps ? 1 : 0) + (pn ? 1 : 0);
  GM1_TIMED(TK_MOM_ADVU, s, hipLaunchKernelGGL(k_mom_advu<false>, dim3((nf + 63) / 64, nj, nk), blk, 0, s, a, f0, nf));
  GM1_LAUNCH_CHECK();
  GM1_TIMED(TK_MOM_ADVV, s, hipLaunchKernelGGL(k_mom_advv<false>, dim3((ni + 63) / 64, nr, nk), blk, 0, s, a, g0, nr));
  GM1_LAUNCH_CHECK();
  GM1_TIMED(TK_MOM_ADVW, s, hipLaunchKernelGGL(k_mom_advw<false>, dim3((ni + 63) / 64, nj, nk - 1), blk, 0, s, a));
  GM1_LAUNCH_CHECK();
  return 0;
}

int gm1hip_momentum_transport5(
    const float* u3d, const float* v3d, const float* w3d, float* uten, float* vten, float* wten,
    const float* rru, const float* rrv, const float* rrw, const float* uh, const float* vh,
    const float* rho0_3d, const float* rrf0, const float* c2, const float* gz, const float* gzu,
    const float* gzv, const float* divx, int ni, int nj, int nk, float dx, float dy, float dz,
    unsigned obc, void* stream) {
  hipStream_t s = (hipStream_t)stream;
  MomArgs a;
  a.u = u3d; a.v = v3d; a.w = w3d; a.rru = rru; a.rrv = rrv; a.rrw = rrw; a.divx = divx;
  a.rho0 = rho0_3d; a.c2 = c2; a.gzu = gzu; a.gzv = gzv; a.rrf0 = rrf0; a.gz = gz; a.uh = uh; a.vh = vh;
  a.uten = uten; a.vten = vten; a.wten = wten; a.ni = ni; a.nj = nj; a.nk = nk;
  a.rdx = 1.0f / dx; a.rdy = 1.0f / dy; a.rdz = 1.0f / dz; a.obc = obc;
  a.flat = nullptr; a.rr0 = nullptr;
  dim3 blk(64);
  const bool pw = side_of(obc, 0) == 1u, pe = side_of(obc, 1) == 1u;
  const bool ps = side_of(obc, 2) == 1u, pn = side_of(obc, 3) == 1u;
  const int f0 = pw ? 0 : 1, nf = ni - 1 + (pw ? 1 : 0) + (pe ? 1 : 0);
  const int g0 = ps ? 0 : 1, nr = nj - 1 + (ps ? 1 : 0) + (pn ? 1 : 0);
  GM1_TIMED(TK_MOM_ADVU, s, hipLaunchKernelGGL(k_mom_advu<false>, dim3((nf + 63) / 64, nj, nk), blk, 0, s, a, f0, nf));
  GM1_LAUNCH_CHECK();
  GM1_TIMED(TK_MOM_ADVV, s, hipLaunchKernelGGL(k_mom_advv<false>, dim3((ni + 63) / 64, nr, nk), blk, 0, s, a, g0, nr));
  GM1_LAUNCH_CHECK();
  GM1_TIMED(TK_MOM_ADVW, s, hipLaunchKernelGGL(k_mom_advw<false>, dim3((ni + 63) / 64, nj, nk - 1), blk, 0, s, a));
  GM1_LAUNCH_CHECK();
  return 0;
}

// The flat stretched column's momentum advection (launch_cuda_gm1_child_stretched_momentum5
// with a flat column, cuda_gm1_child_stretched_momentum.py:655-757).
int gm1hip_momentum_advection5_flat(
    const float* u3d, const float* v3d, const float* w3d, float* uten, float* vten, float* wten,
    float* rru, float* rrv, float* rrw, const float* uh, const float* vh, const float* rho0_3d,
    const float* rf0, const float* rr0, const float* rrf0, const float* c2, const float* flat_column,
    float* divx, int ni, int nj, int nk, float dx, float dy, float rdz, unsigned obc, void* stream) {
  if (!flat_column) return (int)hipErrorInvalidValue;
  hipStream_t s = (hipStream_t)stream;
  const float rdx = 1.0f / dx, rdy = 1.0f / dy;
  dim3 blk(64);
  hipLaunchKernelGGL(k_mom_mass_h_flat, dim3((ni + 1 + 63) / 64, nj + 1, (nk + MKB - 1) / MKB), blk, 0, s, u3d, v3d,
                     rho0_3d, rru, rrv, ni, nj, nk);
  GM1_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_mom_mass_v_flat, dim3((ni + 63) / 64, nj, (nk + MKB - 1) / MKB), blk, 0, s, w3d, rf0, rru, rrv,
                     flat_column, rrw, divx, ni, nj, nk, rdx, rdy, rdz);
  GM1_LAUNCH_CHECK();
  MomArgs a;
  a.u = u3d; a.v = v3d; a.w = w3d; a.rru = rru; a.rrv = rrv; a.rrw = rrw; a.divx = divx;
  a.rho0 = rho0_3d; a.c2 = c2; a.gzu = nullptr; a.gzv = nullptr; a.rrf0 = rrf0; a.gz = nullptr; a.uh = uh;
  a.vh = vh; a.uten = uten; a.vten = vten; a.wten = wten; a.ni = ni; a.nj = nj; a.nk = nk;
  a.rdx = rdx; a.rdy = rdy; a.rdz = rdz; a.obc = obc; a.flat = flat_column; a.rr0 = rr0;
  const bool pw = side_of(obc, 0) == 1u, pe = side_of(obc, 1) == 1u;
  const bool ps = side_of(obc, 2) == 1u, pn = side_of(obc, 3) == 1u;
  const int f0 = pw ? 0 : 1, nf = ni - 1 + (pw ? 1 : 0) + (pe ? 1 : 0);
  const int g0 = ps ? 0 : 1, nr = nj - 1 + (ps ? 1 : 0) + (pn ? 1 : 0);
  GM1_TIMED(TK_MOM_ADVU, s, hipLaunchKernelGGL(k_mom_advu<true>, dim3((nf + 63) / 64, nj, nk), blk, 0, s, a, f0, nf));
  GM1_LAUNCH_CHECK();
  GM1_TIMED(TK_MOM_ADVV, s, hipLaunchKernelGGL(k_mom_advv<true>, dim3((ni + 63) / 64, nr, nk), blk, 0, s, a, g0, nr));
  GM1_LAUNCH_CHECK();
  GM1_TIMED(TK_MOM_ADVW, s, hipLaunchKernelGGL(k_mom_advw<true>, dim3((ni + 63) / 64, nj, nk - 1), blk, 0, s, a));
  GM1_LAUNCH_CHECK();
  return 0;
}

}  // extern "C"
