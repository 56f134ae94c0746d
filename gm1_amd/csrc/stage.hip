// Per-RK-stage bridge kernels of the CM1 large step (gfx950):
// eqtset=2 stage terms, buoyancy, post-acoustic getdiv/getdivx, scalar
// packing and the A-anchored RK compose, held-tendency seeding, the held
// compose of solve1 and the solve3 pressure/density finish.
//
// Replaces cuda_moist_stage_bridge.py, the compose/pack kernels of
// cuda_scalar_moist_rk.py, cuda_momentum_rk.py's seed and the forecast
// kernels of cuda_gm1_forecast.py.  All are single-pass, i-fastest
// elementwise or column kernels; getdivx's vertical mass flux and divx share
// one column pass.
#include "boundary.h"

using namespace gm1;

namespace {

constexpr float RD = 287.04f, RV = 461.5f, CP = 1005.7f, CPV = 1870.0f;
constexpr float CPL = 4190.0f, CPI = 2106.0f, GRAV = 9.81f;

// Levels per thread of the per-cell kernels: each thread keeps the loads of
// KB levels in flight (one load per wave would cap HBM throughput at about a
// tenth of its peak).
constexpr int KB = 8;

// ----------------------------------------------------------- stage terms
// q3d: (numq, nk, nj, ni) species in the order qv, qc, qr, qi, qs, qg.
__global__ __launch_bounds__(256) void k_stage_terms(
    const float* __restrict__ th0, const float* __restrict__ qv0, const float* __restrict__ pi0,
    const float* __restrict__ qc0, const float* __restrict__ qi0, const float* __restrict__ th3d,
    const float* __restrict__ q3d, const float* __restrict__ pp3d, float* __restrict__ thv,
    float* __restrict__ ppterm, float* __restrict__ piadv, float* __restrict__ dum8, float* __restrict__ thterm,
    int ni, int nj, int nk, int numq) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y, k = blockIdx.z;
  if (i >= ni) return;
  const float cv = CP - RD, cvv = CPV - RV, rovcp = RD / CP, reps = RV / RD, repsm1 = reps - 1.0f;
  const int64_t n = (int64_t)ni * nj * nk;
  const int64_t z = ((int64_t)k * nj + j) * ni + i;
  // species beyond numq are exact zeros (a dry state, or warm rain without ice)
  const float q0 = numq >= 1 ? q3d[z] : 0.0f;
  const float q1 = numq >= 2 ? q3d[n + z] : 0.0f, q2 = numq >= 3 ? q3d[2 * n + z] : 0.0f;
  const float q3 = numq >= 4 ? q3d[3 * n + z] : 0.0f, q4 = numq >= 5 ? q3d[4 * n + z] : 0.0f;
  const float q5 = numq >= 6 ? q3d[5 * n + z] : 0.0f;
  const float qv = fmaxf(q0, 0.0f);
  const float liq = q1 + q2;
  const float ice = (q3 + q4) + q5;
  const float qli = fmaxf(liq + ice, 0.0f);
  const float cpli = CPL * fmaxf(liq, 0.0f) + CPI * fmaxf(ice, 0.0f);
  const float bt = th0[k], bq = qv0[k], bp = pi0[k];
  const float rth = 1.0f / bt, theta = bt + th3d[z];
  const float bc = qc0 ? qc0[k] : 0.0f, bi = qi0 ? qi0[k] : 0.0f;
  float b = th3d[z] * rth + repsm1 * (qv - bq);
  b = b - ((qli - bc) - bi);
  dum8[ix3(k + 1, j + NG, i + NG, sy(nj), sx(ni))] = GRAV * b;
  const int64_t p = pix(k + 1, j + 1, i + 1, ni, nj);
  thv[p] = (theta * (1.0f + reps * qv)) / ((1.0f + qv) + qli);
  const float cpm = (CP + CPV * qv) + cpli;
  const float cvm = 1.0f / ((cv + cvv * qv) + cpli);
  thterm[z] = (theta * ((RD + RV * qv) - rovcp * cpm)) * cvm;
  const float pi = bp + pp3d[p];
  piadv[p] = pi;
  ppterm[p] = ((pi * rovcp) * cpm) * cvm;
}

// piadv = pi0 + pp3d on the halo ring (i = 0, ni+1 or j = 0, nj+1) of levels
// 1..nk; the stage terms write the interior.
__global__ __launch_bounds__(256) void k_piadv_ring(const float* __restrict__ pi0, const float* __restrict__ pp3d,
                                                    float* __restrict__ piadv, int ni, int nj) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  const int k = blockIdx.y + 1;
  if (q >= 2 * (ni + 2) + 2 * nj) return;
  int i, j;
  if (q < ni + 2) { j = 0; i = q; }
  else if (q < 2 * (ni + 2)) { j = nj + 1; i = q - (ni + 2); }
  else { const int r = q - 2 * (ni + 2); j = 1 + r / 2; i = (r & 1) ? ni + 1 : 0; }
  const int64_t p = pix(k, j, i, ni, nj);
  piadv[p] = pi0[k - 1] + pp3d[p];
}

// ----------------------------------------------------------- buoyancy
struct BuoyArgs {
  const float* dum8; const float* c1; const float* c2; const float* gxu; const float* gyv;
  float* uten; float* vten; float* wten; int ni, nj, nk;
};


// wten(k) += b(k) at w levels 2..nk, and the terrain-horizontal parts on the
// interior u and v faces.  Each thread marches one column through a chunk of
// BUOY_KC levels, carrying dum8 and the w-level buoyancy of its own cell and
// of its west and south neighbours in registers: b at w level k+1 is formed
// once and reused as level k's on the next iteration, so a cell's b is formed
// once per chunk instead of six times.
constexpr int BUOY_KC = 8;

__global__ __launch_bounds__(64) void k_buoyancy_col(BuoyArgs a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  const int ni = a.ni, nj = a.nj, nk = a.nk;
  const int k0 = (int)blockIdx.z * BUOY_KC + 1;
  if (i >= ni || k0 > nk) return;
  const int k1 = min(k0 + BUOY_KC - 1, nk);
  const bool fu = i >= 1, fv = j >= 1;
  const float* __restrict__ c1 = a.c1;
  const float* __restrict__ c2 = a.c2;
  const int64_t plane = (int64_t)nj * ni;
  const int64_t cp = (int64_t)j * ni + i;       // c1/c2 column
  const int64_t wp = fu ? cp - 1 : cp;          // west neighbour (unused when !fu)
  const int64_t sp = fv ? cp - ni : cp;         // south neighbour (unused when !fv)
  const int64_t sst = (int64_t)sy(nj) * sx(ni);
  const float* __restrict__ dm = a.dum8;
  const int64_t dc = ix3(0, j + NG, i + NG, sy(nj), sx(ni));
  const int64_t dw = fu ? dc - 1 : dc;
  const int64_t ds = fv ? dc - sx(ni) : dc;
  // w level k0: zero at the surface (k0 == 1), else b(k0) from dum8(k0-1), dum8(k0).
  float cur_c = dm[dc + k0 * sst], cur_w = dm[dw + k0 * sst], cur_s = dm[ds + k0 * sst];
  float lo_c = 0.0f, lo_w = 0.0f, lo_s = 0.0f;
  if (k0 > 1) {
    const int64_t q = (int64_t)(k0 - 1) * plane;
    const int64_t m = (int64_t)(k0 - 1) * sst;
    lo_c = c1[q + cp] * dm[dc + m] + c2[q + cp] * cur_c;
    lo_w = c1[q + wp] * dm[dw + m] + c2[q + wp] * cur_w;
    lo_s = c1[q + sp] * dm[ds + m] + c2[q + sp] * cur_s;
  }
#pragma unroll
  for (int kk = 0; kk < BUOY_KC; ++kk) {
    const int k = k0 + kk;
    if (k > k1) break;
    float hi_c = 0.0f, hi_w = 0.0f, hi_s = 0.0f, nxt_c = 0.0f, nxt_w = 0.0f, nxt_s = 0.0f;
    if (k < nk) {  // w level k+1 is interior; level nk+1 (the lid) carries no buoyancy
      const int64_t q = (int64_t)k * plane;
      const int64_t m = (int64_t)(k + 1) * sst;
      nxt_c = dm[dc + m];
      nxt_w = dm[dw + m];
      nxt_s = dm[ds + m];
      hi_c = c1[q + cp] * cur_c + c2[q + cp] * nxt_c;
      hi_w = c1[q + wp] * cur_w + c2[q + wp] * nxt_w;
      hi_s = c1[q + sp] * cur_s + c2[q + sp] * nxt_s;
    }
    if (k >= 2) {
      const int64_t d = ix3(k, j + NG, i + NG, sy(nj), sx(ni));
      a.wten[d] = a.wten[d] + lo_c;
    }
    if (fu) {
      float t = (hi_c + hi_w) + (lo_c + lo_w);
      t = 0.125f * t;
      t = t - 0.25f * (cur_c + cur_w);
      t = t * (a.gxu[((int64_t)(k - 1) * nj + j) * (ni + 1) + i] + a.gxu[((int64_t)k * nj + j) * (ni + 1) + i]);
      const int64_t d = ix3(k, j + NG, i + NG, sy(nj), ux(ni));
      a.uten[d] = a.uten[d] + t;
    }
    if (fv) {
      float t = (hi_c + hi_s) + (lo_c + lo_s);
      t = 0.125f * t;
      t = t - 0.25f * (cur_c + cur_s);
      t = t * (a.gyv[((int64_t)(k - 1) * (nj + 1) + j) * ni + i] + a.gyv[((int64_t)k * (nj + 1) + j) * ni + i]);
      const int64_t d = ix3(k, j + NG, i + NG, vy(nj), sx(ni));
      a.vten[d] = a.vten[d] + t;
    }
    lo_c = hi_c;
    lo_w = hi_w;
    lo_s = hi_s;
    cur_c = nxt_c;
    cur_w = nxt_w;
    cur_s = nxt_s;
  }
}

// ----------------------------------------------------------- getdiv
struct DivArgs {
  float* rru; float* rrv; float* rrw;
  const float* rho0; const float* rf0; const float* rgzu; const float* rgzv; const float* dzdx; const float* dzdy;
  const float* gz; const float* thterm; float* div; float* theta_div; float* divx;
  int ni, nj, nk; float rdx, rdy, dz;
};

// one column: w-level terrain terms from the averaged winds, then div per level
__global__ __launch_bounds__(64) void k_getdiv(DivArgs a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  const int ni = a.ni, nj = a.nj, nk = a.nk;
  if (i >= ni) return;
  const int nxu = ux(ni), nyu = sy(nj), nxs = sx(ni), nyv = vy(nj);
  const int64_t h = (int64_t)j * ni + i;
  const float rgw = a.rgzu[(int64_t)j * (ni + 1) + i], rge = a.rgzu[(int64_t)j * (ni + 1) + i + 1];
  const float rgs = a.rgzv[h], rgn = a.rgzv[h + ni];
  const float dz = a.dz, rds = 1.0f / dz, rdsf = 1.0f / dz, zt = (float)nk * dz, rzt = 1.0f / zt;
  const float ddx = a.dzdx[h], ddy = a.dzdy[h], g = a.gz[h];
  auto xw = [&](int k) { return a.rru[ix3(k, j + NG, i + NG, nyu, nxu)] * rgw; };
  auto xe = [&](int k) { return a.rru[ix3(k, j + NG, i + NG + 1, nyu, nxu)] * rge; };
  auto ys = [&](int k) { return a.rrv[ix3(k, j + NG, i + NG, nyv, nxs)] * rgs; };
  auto yn = [&](int k) { return a.rrv[ix3(k, j + NG + 1, i + NG, nyv, nxs)] * rgn; };
  auto wt = [&](int kf) {
    if (kf == 1 || kf == nk + 1) return 0.0f;
    const float sf = (float)(kf - 1) * dz, sm = ((float)(kf - 1) - 0.5f) * dz;
    float r2 = (sf - sm) * rds;
    const float r1 = 0.5f * (1.0f - r2);
    r2 = 0.5f * r2;
    const float ux_ = r2 * (xw(kf) + xe(kf)) + r1 * (xw(kf - 1) + xe(kf - 1));
    const float vy_ = r2 * (ys(kf) + yn(kf)) + r1 * (ys(kf - 1) + yn(kf - 1));
    float t = ux_ * ddx + vy_ * ddy;
    t = ((t * (sf - zt)) * g) * rzt;
    return a.rrw[ix3(kf, j + NG, i + NG, nyu, nxs)] + t;
  };
  // this thread's chunk of KB levels; wt at its bottom w level formed again
  const int kb = 1 + (int)blockIdx.z * KB;
  float wlo = wt(kb);
#pragma unroll
  for (int q = 0; q < KB; ++q) {
    const int k = min(kb + q, nk);  // clamped past the top: the top level is formed and stored again
    const float whi = wt(k + 1);
    float d = (xe(k) - xw(k)) * a.rdx + (yn(k) - ys(k)) * a.rdy;
    d = d + (whi - wlo) * rdsf;
    if (fabsf(d) < 1.0e-20f) d = 0.0f;
    const int64_t z = (int64_t)(k - 1) * nj * ni + h;
    a.div[z] = d;
    a.theta_div[z] = d * a.thterm[z];
    wlo = (k < nk) ? whi : wlo;  // the top level keeps its lower word for the clamped repeats
  }
}

__global__ __launch_bounds__(256) void k_getdivx_h(DivArgs a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;  // 0..ni
  const int j = blockIdx.y;                             // 0..nj
  const int ni = a.ni, nj = a.nj, nk = a.nk;
  if (i > ni) return;
  float* __restrict__ rru = a.rru;
  float* __restrict__ rrv = a.rrv;
  const float* __restrict__ rho0 = a.rho0;
  const float* __restrict__ rgzu = a.rgzu;
  const float* __restrict__ rgzv = a.rgzv;
  // KB levels per thread: the levels' words are all read first (clamped past
  // the top), then each face is rewritten once by this thread alone
  float ru[KB], rv[KB];
#pragma unroll
  for (int u = 0; u < KB; ++u) {
    const int k = min((int)blockIdx.z * KB + u, nk - 1) + 1;
    ru[u] = (j < nj) ? rru[ix3(k, j + NG, i + NG, sy(nj), ux(ni))] : 0.0f;
    rv[u] = (i < ni) ? rrv[ix3(k, j + NG, i + NG, vy(nj), sx(ni))] : 0.0f;
  }
#pragma unroll
  for (int u = 0; u < KB; ++u) {
    if ((int)blockIdx.z * KB + u >= nk) break;
    const int k = (int)blockIdx.z * KB + u + 1;
    const int64_t plane = (int64_t)(k - 1) * nj * ni;
    if (j < nj) {
      const int l = (i == 0) ? 0 : i - 1, r = (i == ni) ? ni - 1 : i;
      const float den = 0.5f * (rho0[plane + (int64_t)j * ni + l] + rho0[plane + (int64_t)j * ni + r]);
      rru[ix3(k, j + NG, i + NG, sy(nj), ux(ni))] = (den * ru[u]) * rgzu[(int64_t)j * (ni + 1) + i];
    }
    if (i < ni) {
      const int s = (j == 0) ? 0 : j - 1, n = (j == nj) ? nj - 1 : j;
      const float den = 0.5f * (rho0[plane + (int64_t)s * ni + i] + rho0[plane + (int64_t)n * ni + i]);
      rrv[ix3(k, j + NG, i + NG, vy(nj), sx(ni))] = (den * rv[u]) * rgzv[(int64_t)j * ni + i];
    }
  }
}

__global__ __launch_bounds__(64) void k_getdivx_v(DivArgs a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  const int ni = a.ni, nj = a.nj, nk = a.nk;
  if (i >= ni) return;
  const int nxu = ux(ni), nyu = sy(nj), nxs = sx(ni), nyv = vy(nj);
  const int x = i + NG, y = j + NG;
  const int64_t h = (int64_t)j * ni + i;
  const float dz = a.dz, rds = 1.0f / dz, rdsf = 1.0f / dz, zt = (float)nk * dz, rzt = 1.0f / zt;
  const float ddx = a.dzdx[h], ddy = a.dzdy[h], g = a.gz[h];
  const int64_t wst = (int64_t)nyu * nxs, wc = (int64_t)y * nxs + x;
  a.rrw[wst + wc] = 0.0f;
  float below = 0.0f;
  for (int kf = 2; kf <= nk + 1; ++kf) {
    float val = 0.0f;
    if (kf <= nk) {
      const float sf = (float)(kf - 1) * dz, sm = ((float)(kf - 1) - 0.5f) * dz;
      float r2 = (sf - sm) * rds;
      const float r1 = 0.5f * (1.0f - r2);
      r2 = 0.5f * r2;
      const float uxv = r2 * (a.rru[ix3(kf, y, x, nyu, nxu)] + a.rru[ix3(kf, y, x + 1, nyu, nxu)]) +
                        r1 * (a.rru[ix3(kf - 1, y, x, nyu, nxu)] + a.rru[ix3(kf - 1, y, x + 1, nyu, nxu)]);
      const float vyv = r2 * (a.rrv[ix3(kf, y, x, nyv, nxs)] + a.rrv[ix3(kf, y + 1, x, nyv, nxs)]) +
                        r1 * (a.rrv[ix3(kf - 1, y, x, nyv, nxs)] + a.rrv[ix3(kf - 1, y + 1, x, nyv, nxs)]);
      float t = uxv * ddx + vyv * ddy;
      t = ((t * (sf - zt)) * g) * rzt;
      val = a.rf0[(int64_t)(kf - 1) * nj * ni + h] * a.rrw[(int64_t)kf * wst + wc] + t;
    }
    a.rrw[(int64_t)kf * wst + wc] = val;
    const int k = kf - 1;
    float d = (a.rru[ix3(k, y, x + 1, nyu, nxu)] - a.rru[ix3(k, y, x, nyu, nxu)]) * a.rdx +
              (a.rrv[ix3(k, y + 1, x, nyv, nxs)] - a.rrv[ix3(k, y, x, nyv, nxs)]) * a.rdy;
    d = d + (val - below) * rdsf;
    a.divx[(int64_t)(k - 1) * nj * ni + h] = (fabsf(d) < 1.0e-20f) ? 0.0f : d;
    below = val;
  }
}

// k_getdivx_v with the column's levels split over GX_SLOTS level slots of one
// workgroup: every slot first reads the raw averaged w of its levels and of
// the level below them (which the slot below rewrites), then -- after the
// barrier -- rewrites its own levels in place.  Columns up to
// GX_SLOTS * GX_MAXL levels.
constexpr int GX_SLOTS = 8;
constexpr int GX_MAXL = 16;

// Level slots pay only while one thread per column leaves the chip short of
// waves (fewer than 2^18 columns: 256 CUs x 16 waves x 64 lanes).
bool slotted_columns(int ni, int nj, int nk) {
  return (int64_t)ni * nj < (int64_t)1 << 18 && nk <= GX_SLOTS * GX_MAXL;
}

__global__ __launch_bounds__(64 * GX_SLOTS) void k_getdivx_v_slots(DivArgs a) {
  const int c = threadIdx.x % 64, slot = threadIdx.x / 64;
  const int i = blockIdx.x * 64 + c;
  const int j = blockIdx.y;
  const int ni = a.ni, nj = a.nj, nk = a.nk;
  const bool live = i < ni;
  const int L = (nk + GX_SLOTS - 1) / GX_SLOTS;
  const int kf0 = 2 + slot * L, kf1 = min(kf0 + L - 1, nk + 1);  // this slot's w levels (2..nk+1)
  const int nxu = ux(ni), nyu = sy(nj), nxs = sx(ni), nyv = vy(nj);
  const int x = i + NG, y = j + NG;
  const int64_t h = (int64_t)j * ni + i;
  const int64_t wst = (int64_t)nyu * nxs, wc = (int64_t)y * nxs + x;
  float raw[GX_MAXL + 1];  // raw[q]: averaged w at level kf0 - 1 + q
#pragma unroll
  for (int q = 0; q <= GX_MAXL; ++q) {
    const int kf = kf0 - 1 + q;
    raw[q] = (live && kf <= kf1 && kf >= 2 && kf <= nk) ? a.rrw[(int64_t)kf * wst + wc] : 0.0f;
  }
  __syncthreads();
  if (!live || kf0 > kf1) return;
  const float dz = a.dz, rds = 1.0f / dz, rdsf = 1.0f / dz, zt = (float)nk * dz, rzt = 1.0f / zt;
  const float ddx = a.dzdx[h], ddy = a.dzdy[h], g = a.gz[h];
  auto val_at = [&](int kf, float w) {
    if (kf <= 1 || kf > nk) return 0.0f;
    const float sf = (float)(kf - 1) * dz, sm = ((float)(kf - 1) - 0.5f) * dz;
    float r2 = (sf - sm) * rds;
    const float r1 = 0.5f * (1.0f - r2);
    r2 = 0.5f * r2;
    const float uxv = r2 * (a.rru[ix3(kf, y, x, nyu, nxu)] + a.rru[ix3(kf, y, x + 1, nyu, nxu)]) +
                      r1 * (a.rru[ix3(kf - 1, y, x, nyu, nxu)] + a.rru[ix3(kf - 1, y, x + 1, nyu, nxu)]);
    const float vyv = r2 * (a.rrv[ix3(kf, y, x, nyv, nxs)] + a.rrv[ix3(kf, y + 1, x, nyv, nxs)]) +
                      r1 * (a.rrv[ix3(kf - 1, y, x, nyv, nxs)] + a.rrv[ix3(kf - 1, y + 1, x, nyv, nxs)]);
    float t = uxv * ddx + vyv * ddy;
    t = ((t * (sf - zt)) * g) * rzt;
    return a.rf0[(int64_t)(kf - 1) * nj * ni + h] * w + t;
  };
  if (kf0 == 2) a.rrw[wst + wc] = 0.0f;
  float below = val_at(kf0 - 1, raw[0]);
#pragma unroll
  for (int q = 1; q <= GX_MAXL; ++q) {
    const int kf = kf0 - 1 + q;
    if (kf > kf1) break;
    const float val = val_at(kf, raw[q]);
    a.rrw[(int64_t)kf * wst + wc] = val;
    const int k = kf - 1;
    float d = (a.rru[ix3(k, y, x + 1, nyu, nxu)] - a.rru[ix3(k, y, x, nyu, nxu)]) * a.rdx +
              (a.rrv[ix3(k, y + 1, x, nyv, nxs)] - a.rrv[ix3(k, y, x, nyv, nxs)]) * a.rdy;
    d = d + (val - below) * rdsf;
    a.divx[(int64_t)(k - 1) * nj * ni + h] = (fabsf(d) < 1.0e-20f) ? 0.0f : d;
    below = val;
  }
}

// ----------------------------------------------------------- flat stretched column
// misclibs.F's flat getdiv/getdivx and solve2.F:897-906's flat buoyancy
// (terrain_flag=.false.): ((d*rdz)*mh(k)) vertically, the base state at
// CM1's (1,1,k), smeps 1e-30, and no horizontal buoyancy term.
__global__ __launch_bounds__(256) void k_buoyancy_flat(const float* __restrict__ dum8, const float* __restrict__ c1,
                                                       const float* __restrict__ c2, float* __restrict__ wten, int ni,
                                                       int nj, int nk) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  if (i >= ni) return;
  const int nxs = sx(ni), nys = sy(nj);
  // KB w levels (2..nk) per thread, all read before the first (in-place) write
  float wt[KB], ad[KB];
#pragma unroll
  for (int u = 0; u < KB; ++u) {
    const int kf = min(2 + (int)blockIdx.z * KB + u, nk);
    const int64_t p = ((int64_t)(kf - 1) * nj + j) * ni + i;
    ad[u] = c1[p] * dum8[ix3(kf - 1, j + NG, i + NG, nys, nxs)] + c2[p] * dum8[ix3(kf, j + NG, i + NG, nys, nxs)];
    wt[u] = wten[ix3(kf, j + NG, i + NG, nys, nxs)];
  }
#pragma unroll
  for (int u = 0; u < KB; ++u) {
    const int kf = 2 + (int)blockIdx.z * KB + u;
    if (kf > nk) break;
    wten[ix3(kf, j + NG, i + NG, nys, nxs)] = wt[u] + ad[u];
  }
}

struct FlatDivArgs {
  float* rru; float* rrv; float* rrw;
  const float* rho0; const float* rf0; const float* mh; const float* thterm;
  float* div; float* theta_div; float* divx;
  int ni, nj, nk; float rdx, rdy, rdz;
};

__global__ __launch_bounds__(64) void k_getdiv_flat(FlatDivArgs a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  const int ni = a.ni, nj = a.nj, nk = a.nk;
  if (i >= ni) return;
  const int nxu = ux(ni), nyu = sy(nj), nxs = sx(ni), nyv = vy(nj);
  const int x = i + NG, y = j + NG;
  const int64_t h = (int64_t)j * ni + i;
  // this thread's chunk of KB levels (clamped past the top: the top level again)
  const int kb = 1 + (int)blockIdx.z * KB;
  float wlo = a.rrw[ix3(kb, y, x, nyu, nxs)];
#pragma unroll
  for (int q = 0; q < KB; ++q) {
    const int k = min(kb + q, nk);
    const float whi = a.rrw[ix3(k + 1, y, x, nyu, nxs)];
    float d = (a.rru[ix3(k, y, x + 1, nyu, nxu)] - a.rru[ix3(k, y, x, nyu, nxu)]) * a.rdx +
              (a.rrv[ix3(k, y + 1, x, nyv, nxs)] - a.rrv[ix3(k, y, x, nyv, nxs)]) * a.rdy;
    d = d + ((whi - wlo) * a.rdz) * a.mh[k];
    if (fabsf(d) < 1.0e-30f) d = 0.0f;
    const int64_t z = (int64_t)(k - 1) * nj * ni + h;
    a.div[z] = d;
    a.theta_div[z] = d * a.thterm[z];
    wlo = (k < nk) ? whi : wlo;  // the top level keeps its lower word for the clamped repeats
  }
}

__global__ __launch_bounds__(256) void k_getdivx_h_flat(FlatDivArgs a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;  // 0..ni
  const int j = blockIdx.y;                             // 0..nj
  const int ni = a.ni, nj = a.nj, nk = a.nk;
  if (i > ni) return;
  float* __restrict__ rru = a.rru;
  float* __restrict__ rrv = a.rrv;
  // KB levels per thread: all faces read first (clamped past the top), then rewritten once
  float ru[KB], rv[KB];
#pragma unroll
  for (int u = 0; u < KB; ++u) {
    const int k = min((int)blockIdx.z * KB + u, nk - 1) + 1;
    ru[u] = (j < nj) ? rru[ix3(k, j + NG, i + NG, sy(nj), ux(ni))] : 0.0f;
    rv[u] = (i < ni) ? rrv[ix3(k, j + NG, i + NG, vy(nj), sx(ni))] : 0.0f;
  }
#pragma unroll
  for (int u = 0; u < KB; ++u) {
    if ((int)blockIdx.z * KB + u >= nk) break;
    const int k = (int)blockIdx.z * KB + u + 1;
    const float rc = a.rho0[(int64_t)(k - 1) * nj * ni];
    if (j < nj) rru[ix3(k, j + NG, i + NG, sy(nj), ux(ni))] = ru[u] * rc;
    if (i < ni) rrv[ix3(k, j + NG, i + NG, vy(nj), sx(ni))] = rv[u] * rc;
  }
}

__global__ __launch_bounds__(64) void k_getdivx_v_flat(FlatDivArgs a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  const int ni = a.ni, nj = a.nj, nk = a.nk;
  if (i >= ni) return;
  const int nxu = ux(ni), nyu = sy(nj), nxs = sx(ni), nyv = vy(nj);
  const int x = i + NG, y = j + NG;
  const int64_t h = (int64_t)j * ni + i;
  const int64_t wst = (int64_t)nyu * nxs, wc = (int64_t)y * nxs + x;
  a.rrw[wst + wc] = 0.0f;
  float below = 0.0f;
  for (int kf = 2; kf <= nk + 1; ++kf) {
    const float val = (kf <= nk) ? a.rrw[(int64_t)kf * wst + wc] * a.rf0[(int64_t)(kf - 1) * nj * ni] : 0.0f;
    a.rrw[(int64_t)kf * wst + wc] = val;
    const int k = kf - 1;
    float d = (a.rru[ix3(k, y, x + 1, nyu, nxu)] - a.rru[ix3(k, y, x, nyu, nxu)]) * a.rdx +
              (a.rrv[ix3(k, y + 1, x, nyv, nxs)] - a.rrv[ix3(k, y, x, nyv, nxs)]) * a.rdy;
    d = d + ((val - below) * a.rdz) * a.mh[k];
    a.divx[(int64_t)(k - 1) * nj * ni + h] = (fabsf(d) < 1.0e-30f) ? 0.0f : d;
    below = val;
  }
}

// getdiv and getdivx's vertical pass of the flat column in one pass over
// GX_SLOTS level slots of a workgroup: each slot reads the raw averaged w of
// its levels (and the level above them) before the barrier, then forms div and
// theta_div from the raw averages, rewrites rrw = rf0*w in place and forms
// divx from the rescaled faces -- its own rho0(k)*rru / rho0(k)*rrv products,
// the words k_getdivx_h_flat writes after it.
__global__ __launch_bounds__(64 * GX_SLOTS) void k_post_acoustic_flat(FlatDivArgs a) {
  const int c = threadIdx.x % 64, slot = threadIdx.x / 64;
  const int i = blockIdx.x * 64 + c;
  const int j = blockIdx.y;
  const int ni = a.ni, nj = a.nj, nk = a.nk;
  const bool live = i < ni;
  const int L = (nk + GX_SLOTS - 1) / GX_SLOTS;
  const int k0 = 1 + slot * L, k1 = min(k0 + L - 1, nk);  // this slot's levels
  const int nxu = ux(ni), nyu = sy(nj), nxs = sx(ni), nyv = vy(nj);
  const int x = i + NG, y = j + NG;
  const int64_t h = (int64_t)j * ni + i, plane = (int64_t)nj * ni;
  const int64_t wst = (int64_t)nyu * nxs, wc = (int64_t)y * nxs + x;
  float raw[GX_MAXL + 1];  // raw[q]: averaged w at level k0 + q
#pragma unroll
  for (int q = 0; q <= GX_MAXL; ++q) {
    const int kf = k0 + q;
    raw[q] = (live && kf <= k1 + 1) ? a.rrw[(int64_t)kf * wst + wc] : 0.0f;
  }
  __syncthreads();
  if (!live || k0 > k1) return;
  auto val_at = [&](int kf, float w) { return (kf >= 2 && kf <= nk) ? w * a.rf0[(int64_t)(kf - 1) * plane] : 0.0f; };
  if (k0 == 1) a.rrw[wst + wc] = 0.0f;
  float below = val_at(k0, raw[0]);
#pragma unroll
  for (int q = 0; q < GX_MAXL; ++q) {
    const int k = k0 + q;
    if (k > k1) break;
    const float ru0 = a.rru[ix3(k, y, x, nyu, nxu)], ru1 = a.rru[ix3(k, y, x + 1, nyu, nxu)];
    const float rv0 = a.rrv[ix3(k, y, x, nyv, nxs)], rv1 = a.rrv[ix3(k, y + 1, x, nyv, nxs)];
    const int64_t z = (int64_t)(k - 1) * plane + h;
    float d = (ru1 - ru0) * a.rdx + (rv1 - rv0) * a.rdy;
    d = d + ((raw[q + 1] - raw[q]) * a.rdz) * a.mh[k];
    if (fabsf(d) < 1.0e-30f) d = 0.0f;
    a.div[z] = d;
    a.theta_div[z] = d * a.thterm[z];
    const float rc = a.rho0[(int64_t)(k - 1) * plane];
    const float val = val_at(k + 1, raw[q + 1]);
    a.rrw[(int64_t)(k + 1) * wst + wc] = val;
    float dx = (ru1 * rc - ru0 * rc) * a.rdx + (rv1 * rc - rv0 * rc) * a.rdy;
    dx = dx + ((val - below) * a.rdz) * a.mh[k];
    a.divx[z] = (fabsf(dx) < 1.0e-30f) ? 0.0f : dx;
    below = val;
  }
}

// ----------------------------------------------------------- scalar RK
__global__ __launch_bounds__(256) void k_pack(const float* __restrict__ th0, const float* __restrict__ src,
                                              float* __restrict__ padded, int ni, int nj, int nk, int theta) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y, k = blockIdx.z;
  if (i >= ni) return;
  const int64_t z = ((int64_t)k * nj + j) * ni + i;
  const float v = theta ? (th0[k] - 300.0f) + src[z] : src[z];
  padded[ix3(k + 1, j + NG, i + NG, sy(nj), sx(ni))] = v;
}

__global__ __launch_bounds__(256) void k_compose(const float* __restrict__ a0, const float* __restrict__ held,
                                                 const float* __restrict__ divterm, const float* __restrict__ adv,
                                                 float* __restrict__ out, int64_t n, float dttmp) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  const float seed = divterm ? held[q] - divterm[q] : held[q];
  const float v = a0[q] + dttmp * (seed + adv[q]);
  out[q] = (fabsf(v) < 1.0e-30f) ? 0.0f : v;
}

// ----------------------------------------------------------- held seed
__global__ __launch_bounds__(256) void k_seed_held(const float* __restrict__ u1, const float* __restrict__ v1,
                                                   const float* __restrict__ w1, float* __restrict__ ut,
                                                   float* __restrict__ vt, float* __restrict__ wt, int ni, int nj,
                                                   int nk) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;  // 0..ni
  const int j = blockIdx.y;                             // 0..nj
  const int k = blockIdx.z + 1;
  if (i > ni) return;
  const int x = i + NG, y = j + NG;
  const int64_t pu = ix3(k, y, x, sy(nj), ux(ni)), pv = ix3(k, y, x, vy(nj), sx(ni)), pw = ix3(k, y, x, sy(nj), sx(ni));
  ut[pu] = u1[pu];
  vt[pv] = v1[pv];
  wt[pw] = w1[pw];
}

// ----------------------------------------------------------- solve1 held compose
__global__ __launch_bounds__(256) void k_held_compose(
    const float* __restrict__ th0, const float* __restrict__ pi0, const float* __restrict__ rho,
    const float* __restrict__ ppi, const float* __restrict__ tha, const float* __restrict__ qa,
    float* __restrict__ thheld, float* __restrict__ qheld, float* __restrict__ ppten,
    const float* __restrict__ qpten, const float* __restrict__ qtten, const float* __restrict__ qvten,
    const float* __restrict__ qcten, int ni, int nj, int nk, float dt, int zero_held) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y, k = blockIdx.z;
  if (i >= ni) return;
  const float rp00 = 1.0f / 100000.0f, rddcv = RD / (CP - RD), tem = 0.0001f * 0.0001f, qsmall = 1.0e-14f;
  const int64_t n = (int64_t)ni * nj * nk, z = ((int64_t)k * nj + j) * ni + i;
  const int64_t p = pix(k + 1, j + 1, i + 1, ni, nj);
  // zero_held: the held fields are +0 (no solve1 contribution) and are not read
  const float ht = zero_held ? 0.0f : thheld[z], qvt = zero_held ? 0.0f : qheld[z];
  if (fabsf(dt * ht) > tem || fabsf(dt * qvt) > qsmall) {
    const float thnew = tha[z] + dt * ht, qvnew = qa[z] + dt * qvt;
    float base = (rho[z] * (th0[k] + thnew)) * (RD + RV * qvnew);
    base = base * rp00;
    const float pinew = powf(base, rddcv) - pi0[k];
    ppten[p] = ((pinew - ppi[p]) / dt) + qpten[z];
  } else {
    ppten[p] = qpten[z];
  }
  thheld[z] = ht + qtten[z];
  qheld[z] = qvt + qvten[z];
  qheld[n + z] = (zero_held ? 0.0f : qheld[n + z]) + qcten[z];
}

// ----------------------------------------------------------- solve3
// With `mass` (the microphysics' (nk,nj,ni) pp3d) the new pp3d comes from it
// and is written back into the pressure family.
__global__ __launch_bounds__(256) void k_solve3_pressure(float* __restrict__ ppx, float* __restrict__ pp3d,
                                                         float* __restrict__ ppi, const float* __restrict__ mass,
                                                         int ni, int nj) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y, k = blockIdx.z + 1;
  if (i >= ni) return;
  const int64_t p = pix(k, j + 1, i + 1, ni, nj);
  float v;
  if (mass) {
    v = mass[((int64_t)(k - 1) * nj + j) * ni + i];
    pp3d[p] = v;
  } else {
    v = pp3d[p];
  }
  ppx[p] = v + ppx[p];
  ppi[p] = v;
}

__global__ __launch_bounds__(256) void k_solve3_density(const float* __restrict__ rho, const float* __restrict__ c1,
                                                        const float* __restrict__ c2, float* __restrict__ rr,
                                                        float* __restrict__ rf, int ni, int nj, int nk) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y, f = blockIdx.z;  // 0..nk
  if (i >= ni) return;
  const int64_t plane = (int64_t)nj * ni, h = (int64_t)j * ni + i;
  if (f < nk) rr[f * plane + h] = 1.0f / rho[f * plane + h];
  float v;
  if (f == 0) v = (1.875f * rho[h] + -1.25f * rho[plane + h]) + 0.375f * rho[2 * plane + h];
  else if (f == nk) v = (1.875f * rho[(nk - 1) * plane + h] + -1.25f * rho[(nk - 2) * plane + h]) + 0.375f * rho[(nk - 3) * plane + h];
  else v = c1[f * plane + h] * rho[(f - 1) * plane + h] + c2[f * plane + h] * rho[f * plane + h];
  rf[f * plane + h] = v;
}

}  // namespace

extern "C" {

int gm1hip_stage_terms(const float* th0, const float* qv0, const float* pi0, const float* qc0, const float* qi0,
                       const float* th3d, const float* q3d, int numq, const float* pp3d, float* thv, float* ppterm,
                       float* piadv, float* dum8, float* thterm, int ni, int nj, int nk, unsigned sides,
                       void* stream) {
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_stage_terms, dim3((ni + 63) / 64, nj, nk), dim3(64), 0, s, th0, qv0, pi0,
                     qc0, qi0, th3d, q3d,
                     pp3d, thv, ppterm, piadv, dum8, thterm, ni, nj, nk, numq);
  hipLaunchKernelGGL(k_piadv_ring, dim3((2 * (ni + 2) + 2 * nj + 255) / 256, nk), dim3(256), 0, s, pi0, pp3d,
                     piadv, ni, nj);
  GM1_LAUNCH_CHECK();
  int rc = fill_pressure(thv, sides, ni, nj, nk, s);
  if (!rc) rc = fill_family(dum8, 0, sides, ni, nj, nk, s);
  return rc;
}

int gm1hip_buoyancy(const float* dum8, const float* c1, const float* c2, const float* gxu, const float* gyv,
                    float* uten, float* vten, float* wten, int ni, int nj, int nk, void* stream) {
  BuoyArgs a{dum8, c1, c2, gxu, gyv, uten, vten, wten, ni, nj, nk};
  hipLaunchKernelGGL(k_buoyancy_col, dim3((ni + 63) / 64, nj, (nk + BUOY_KC - 1) / BUOY_KC), dim3(64), 0,
                     (hipStream_t)stream, a);
  GM1_LAUNCH_CHECK();
  return 0;
}

int gm1hip_post_acoustic(float* rru, float* rrv, float* rrw, const float* rho0_3d, const float* rf0,
                         const float* rgzu, const float* rgzv, const float* dzdx, const float* dzdy, const float* gz,
                         const float* thterm, float* divergence, float* theta_div, float* divx, int ni, int nj,
                         int nk, float dx, float dy, float dz, void* stream) {
  hipStream_t s = (hipStream_t)stream;
  DivArgs a{rru, rrv, rrw, rho0_3d, rf0, rgzu, rgzv, dzdx, dzdy, gz, thterm, divergence, theta_div, divx,
            ni, nj, nk, 1.0f / dx, 1.0f / dy, dz};
  hipLaunchKernelGGL(k_getdiv, dim3((ni + 63) / 64, nj, (nk + KB - 1) / KB), dim3(64), 0, s, a);
  hipLaunchKernelGGL(k_getdivx_h, dim3((ni + 1 + 63) / 64, nj + 1, (nk + KB - 1) / KB), dim3(64), 0, s, a);
  if (slotted_columns(ni, nj, nk))
    hipLaunchKernelGGL(k_getdivx_v_slots, dim3((ni + 63) / 64, nj), dim3(64 * GX_SLOTS), 0, s, a);
  else
    hipLaunchKernelGGL(k_getdivx_v, dim3((ni + 63) / 64, nj), dim3(64), 0, s, a);
  GM1_LAUNCH_CHECK();
  return 0;
}

int gm1hip_buoyancy_flat(const float* dum8, const float* c1, const float* c2, float* wten, int ni, int nj, int nk,
                         void* stream) {
  if (nk < 2) return (int)hipErrorInvalidValue;
  hipLaunchKernelGGL(k_buoyancy_flat, dim3((ni + 63) / 64, nj, (nk - 1 + KB - 1) / KB), dim3(64), 0,
                     (hipStream_t)stream, dum8, c1, c2, wten, ni, nj, nk);
  GM1_LAUNCH_CHECK();
  return 0;
}

int gm1hip_post_acoustic_flat(float* rru, float* rrv, float* rrw, const float* rho0_3d, const float* rf0,
                              const float* flat_column, const float* thterm, float* divergence, float* theta_div,
                              float* divx, int ni, int nj, int nk, float dx, float dy, float rdz, void* stream) {
  if (!flat_column) return (int)hipErrorInvalidValue;
  hipStream_t s = (hipStream_t)stream;
  FlatDivArgs a{rru, rrv, rrw, rho0_3d, rf0, flat_column, thterm, divergence, theta_div, divx,
                ni, nj, nk, 1.0f / dx, 1.0f / dy, rdz};
  if (nk <= GX_SLOTS * GX_MAXL) {  // getdiv and getdivx's vertical pass together, then the faces rescaled
    hipLaunchKernelGGL(k_post_acoustic_flat, dim3((ni + 63) / 64, nj), dim3(64 * GX_SLOTS), 0, s, a);
    hipLaunchKernelGGL(k_getdivx_h_flat, dim3((ni + 1 + 63) / 64, nj + 1, (nk + KB - 1) / KB), dim3(64), 0, s, a);
  } else {
    hipLaunchKernelGGL(k_getdiv_flat, dim3((ni + 63) / 64, nj, (nk + KB - 1) / KB), dim3(64), 0, s, a);
    hipLaunchKernelGGL(k_getdivx_h_flat, dim3((ni + 1 + 63) / 64, nj + 1, (nk + KB - 1) / KB), dim3(64), 0, s, a);
    hipLaunchKernelGGL(k_getdivx_v_flat, dim3((ni + 63) / 64, nj), dim3(64), 0, s, a);
  }
  GM1_LAUNCH_CHECK();
  return 0;
}

int gm1hip_pack_scalar(const float* th0, const float* src, float* padded, int ni, int nj, int nk, int theta,
                       unsigned sides, void* stream) {
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_pack, dim3((ni + 63) / 64, nj, nk), dim3(64), 0, s, th0, src, padded, ni, nj, nk, theta);
  GM1_LAUNCH_CHECK();
  return fill_family(padded, 0, sides, ni, nj, nk, s);
}

int gm1hip_rk_compose(const float* a0, const float* held, const float* divterm, const float* adv, float* out,
                      int64_t n, float dttmp, void* stream) {
  hipLaunchKernelGGL(k_compose, dim3(gm1_blocks(n, 256)), dim3(256), 0, (hipStream_t)stream, a0, held, divterm, adv,
                     out, n, dttmp);
  GM1_LAUNCH_CHECK();
  return 0;
}

int gm1hip_seed_held_momentum(const float* uten1, const float* vten1, const float* wten1, float* uten, float* vten,
                              float* wten, int ni, int nj, int nk, void* stream) {
  hipLaunchKernelGGL(k_seed_held, dim3((ni + 1 + 63) / 64, nj + 1, nk), dim3(64), 0, (hipStream_t)stream, uten1,
                     vten1, wten1, uten, vten, wten, ni, nj, nk);
  GM1_LAUNCH_CHECK();
  return 0;
}

int gm1hip_held_compose(const float* th0, const float* pi0, const float* rho, const float* ppi, const float* tha,
                        const float* qa, float* thheld, float* qheld, float* ppten, const float* qpten,
                        const float* qtten, const float* qvten, const float* qcten, int ni, int nj, int nk, float dt,
                        int zero_held, void* stream) {
  hipLaunchKernelGGL(k_held_compose, dim3((ni + 63) / 64, nj, nk), dim3(64), 0, (hipStream_t)stream, th0, pi0, rho,
                     ppi, tha, qa, thheld, qheld, ppten, qpten, qtten, qvten, qcten, ni, nj, nk, dt, zero_held ? 1 : 0);
  GM1_LAUNCH_CHECK();
  return 0;
}

int gm1hip_solve3_pressure(float* ppx, float* pp3d, float* ppi, const float* pp3d_mass, int ni, int nj, int nk,
                           unsigned sides, void* stream) {
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_solve3_pressure, dim3((ni + 63) / 64, nj, nk), dim3(64), 0, s, ppx, pp3d, ppi, pp3d_mass, ni,
                     nj);
  GM1_LAUNCH_CHECK();
  int rc = fill_pressure(ppx, sides, ni, nj, nk, s);
  if (!rc) rc = fill_pressure(pp3d, sides, ni, nj, nk, s);
  if (!rc) rc = fill_pressure(ppi, sides, ni, nj, nk, s);
  return rc;
}

int gm1hip_solve3_density(const float* rho, const float* c1, const float* c2, float* rr, float* rf, int ni, int nj,
                          int nk, void* stream) {
  hipLaunchKernelGGL(k_solve3_density, dim3((ni + 63) / 64, nj, nk + 1), dim3(64), 0, (hipStream_t)stream, rho, c1,
                     c2, rr, rf, ni, nj, nk);
  GM1_LAUNCH_CHECK();
  return 0;
}

}  // extern "C"
