// The flat, uniform, doubly periodic psolver=3 acoustic stage on gfx950.
//
// Layout (cuda_sound3_periodic.py:449-491): the scalar family (nk+2, nj+2,
// ni+2) with a one-cell periodic halo, u (nk+2, nj+2, ni+1), v (nk+2, nj+1,
// ni+2), w (nk+2, nj, ni) without lateral halo; 1-D metrics uf/uh (ni+2),
// vf/vh (nj+2), face weights dtu0 (nj+2, ni+1), dtv0 (nj+1, ni+2).  The
// statement order is cpu_sound3_periodic_stage_inplace's (:672-1064), which
// stores every elimination and tridiagonal diagnostic as an output.
//
// One stage is: prepare (coefficients, carriers, zeroed diagnostics), then per
// small step a face pass (pressure-gradient update of every live u and v face
// with the time average folded in -- the average of step n reads only the
// faces step n wrote), a column pass (divergence, pressure advection, fpk,
// Thomas solve, residual, pressure update and filter) and, between steps, the
// periodic ppd fill; the stage ends with the u/v and (nrk < 3) pp3d fills.
#include <algorithm>

#include "cm1_common.h"

namespace {

using gm1::Dims;

// Pointer-table slots, in the order of cuda_sound3_periodic.ALL_ARRAY_NAMES.
enum Slot {
  MH, RR0, RF0, C1, C2, MF, PIADV, PPTERM, THV, PPTEN, PPI, PP3D, PPX, PPD, QK, PK1, PK2, MM, FPK, DIVH, HADVP,
  ELIM_C, ELIM_D, TRI_AA, TRI_BB, TRI_CC, TRI_DD, TRI_PIVOT, TRI_RESIDUAL, WTEN, WA, W3D, RRW, UA, U3D, UTEN, RRU,
  VA, V3D, VTEN, RRV, DTU0, DTV0, UF, VF, UH, VH, NSLOTS
};

struct Tab {
  float* p[NSLOTS];
};

constexpr int NZEROED = 15;  // QK .. TRI_RESIDUAL

GM1_DEV int64_t si(int k, int j, int i, int ni, int nj) { return ((int64_t)k * (nj + 2) + j) * (ni + 2) + i; }
GM1_DEV int64_t wi(int k, int j, int i, int ni, int nj) { return ((int64_t)k * nj + j) * ni + i; }
GM1_DEV int64_t ui(int k, int j, int f, int ni, int nj) { return ((int64_t)k * (nj + 2) + j) * (ni + 1) + f; }
GM1_DEV int64_t vi(int k, int f, int i, int ni, int nj) { return ((int64_t)k * (nj + 1) + f) * (ni + 2) + i; }

// :706-784.  One thread per word of the scalar family: diagnostics zeroed,
// ppd <- ppx and (nrk != 1) pp3d <- ppi on levels 1..nk, the stage
// coefficients qk, pk1, pk2, mm and the w carriers on physical cells.
__global__ void __launch_bounds__(256) k_sp_prepare(Tab t, Dims d, int nrk, float dts, float rdz, float cp,
                                                    float beta) {
  const int ni = d.ni, nj = d.nj, nk = d.nk;
  const int64_t total = (int64_t)(nk + 2) * (nj + 2) * (ni + 2);
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= total) return;
  const int i = (int)(q % (ni + 2));
  const int j = (int)((q / (ni + 2)) % (nj + 2));
  const int k = (int)(q / ((int64_t)(ni + 2) * (nj + 2)));
  const bool level = k >= 1 && k <= nk;
  const bool phys = level && j >= 1 && j <= nj && i >= 1 && i <= ni;
#pragma unroll
  for (int s = 0; s < NZEROED; ++s) {
    if (phys && s <= MM - QK) continue;  // coefficients below
    t.p[QK + s][q] = 0.0f;
  }
  if (level) {
    t.p[PPD][q] = t.p[PPX][q];
    if (nrk != 1) t.p[PP3D][q] = t.p[PPI][q];
  }
  if (!phys) return;
  const float* mh = t.p[MH];
  const float* rr0 = t.p[RR0];
  const float* rf0 = t.p[RF0];
  const float* c1 = t.p[C1];
  const float* c2 = t.p[C2];
  const float* piadv = t.p[PIADV];
  const float* thv = t.p[THV];
  const int64_t up = si(k + 1, j, i, ni, nj), dn = si(k - 1, j, i, ni, nj);
  const float dr = dts * rdz;
  const float tem = dr * mh[q];
  float r2 = dr * mh[q];
  r2 = r2 * rr0[q];
  r2 = r2 * rf0[up];
  float r1 = dr * mh[q];
  r1 = r1 * rr0[q];
  r1 = r1 * rf0[q];
  const float tem1 = (dr * cp) * t.p[MF][q];
  t.p[QK][q] = tem * t.p[PPTERM][q];
  t.p[PK2][q] = (k == nk) ? 0.0f : r2 * ((-c2[up]) * piadv[up] + (1.0f - c1[up]) * piadv[q]);
  t.p[PK1][q] = (k == 1) ? 0.0f : r1 * (c1[q] * piadv[dn] + (c2[q] - 1.0f) * piadv[q]);
  t.p[MM][q] = tem1 * (c2[q] * thv[q] + c1[q] * thv[dn]);
  const int64_t w = wi(k, j - 1, i - 1, ni, nj);
  const float wa = t.p[WA][w];
  t.p[RRW][w] = wa * beta;
  if (nrk != 1) t.p[W3D][w] = wa;
}

// :722-727.  The face carriers: rru = rrv = 0, and (nrk != 1) u3d <- ua, v3d <- va.
__global__ void __launch_bounds__(256) k_sp_prepare_faces(Tab t, Dims d, int nrk) {
  const int ni = d.ni, nj = d.nj, nk = d.nk;
  const int64_t nu = (int64_t)(nk + 2) * (nj + 2) * (ni + 1);
  const int64_t nv = (int64_t)(nk + 2) * (nj + 1) * (ni + 2);
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < nu) {
    t.p[RRU][q] = 0.0f;
    if (nrk != 1) t.p[U3D][q] = t.p[UA][q];
  }
  if (q < nv) {
    t.p[RRV][q] = 0.0f;
    if (nrk != 1) t.p[V3D][q] = t.p[VA][q];
  }
}

// :787-828 and :1013-1032.  Every live u face (k 1..nk, j 1..nj+1, f 0..ni)
// and v face (k 1..nk, f 0..nj, i 1..ni+1): u += dts*dtu0*(uten - pg), then
// the running sum rru (n < nloop) or the average (rru + u)*(1/nloop).
__global__ void __launch_bounds__(256) k_sp_faces(Tab t, Dims d, float dts, float tem1, float tem2, float tavg,
                                                  int last) {
  const int ni = d.ni, nj = d.nj, nk = d.nk;
  const int64_t nu = (int64_t)nk * (nj + 1) * (ni + 1);
  const int64_t nv = (int64_t)nk * (nj + 1) * (ni + 1);
  int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const float* ppd = t.p[PPD];
  const float* thv = t.p[THV];
  if (q < nu) {
    const int f = (int)(q % (ni + 1));
    const int j = (int)((q / (ni + 1)) % (nj + 1)) + 1;
    const int k = (int)(q / ((int64_t)(ni + 1) * (nj + 1))) + 1;
    float pg = tem1 * (ppd[si(k, j, f + 1, ni, nj)] - ppd[si(k, j, f, ni, nj)]);
    pg = pg * t.p[UF][f + 1];
    pg = pg * (thv[si(k, j, f + 1, ni, nj)] + thv[si(k, j, f, ni, nj)]);
    const int64_t o = ui(k, j, f, ni, nj);
    const float u = t.p[U3D][o] + (dts * t.p[DTU0][(int64_t)j * (ni + 1) + f]) * (t.p[UTEN][o] - pg);
    t.p[U3D][o] = u;
    const float r = t.p[RRU][o] + u;
    t.p[RRU][o] = last ? r * tavg : r;
    return;
  }
  q -= nu;
  if (q < nv) {
    const int i = (int)(q % (ni + 1)) + 1;
    const int f = (int)((q / (ni + 1)) % (nj + 1));
    const int k = (int)(q / ((int64_t)(ni + 1) * (nj + 1))) + 1;
    float pg = tem2 * (ppd[si(k, f + 1, i, ni, nj)] - ppd[si(k, f, i, ni, nj)]);
    pg = pg * t.p[VF][f + 1];
    pg = pg * (thv[si(k, f + 1, i, ni, nj)] + thv[si(k, f, i, ni, nj)]);
    const int64_t o = vi(k, f, i, ni, nj);
    const float v = t.p[V3D][o] + (dts * t.p[DTV0][(int64_t)f * (ni + 2) + i]) * (t.p[VTEN][o] - pg);
    t.p[V3D][o] = v;
    const float r = t.p[RRV][o] + v;
    t.p[RRV][o] = last ? r * tavg : r;
  }
}

struct ColumnScalars {
  float dts, rdx, rdy, temx, temy, alph, beta, asq, kdiv, smeps, tavg;
  int write_ppx, last;
};

// :830-1010.  One thread per physical column; k runs serially.  Lanes of a
// wavefront take consecutive i, so every level's loads are row-contiguous.
__global__ void __launch_bounds__(256) k_sp_column(Tab t, Dims d, ColumnScalars c) {
  const int ni = d.ni, nj = d.nj, nk = d.nk;
  const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
  const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
  if (i > ni || j > nj) return;
  const float* u = t.p[U3D];
  const float* v = t.p[V3D];
  const float* piadv = t.p[PIADV];
  const float* ppterm = t.p[PPTERM];
  const float* ppten = t.p[PPTEN];
  const float* qk = t.p[QK];
  const float* pk1 = t.p[PK1];
  const float* pk2 = t.p[PK2];
  const float* mm = t.p[MM];
  const float* wten = t.p[WTEN];
  float* w = t.p[W3D];
  float* pp3d = t.p[PP3D];
  float* fpk = t.p[FPK];
  float* ec = t.p[ELIM_C];
  float* ed = t.p[ELIM_D];
  const float uh = t.p[UH][i], vh = t.p[VH][j];
  const int64_t plane = (int64_t)(nj + 2) * (ni + 2);
  const int64_t wplane = (int64_t)nj * ni;
  const int64_t s0 = si(0, j, i, ni, nj);
  const int64_t w0 = wi(0, j - 1, i - 1, ni, nj);

  // fpk on every level (:830-880).
  for (int k = 1; k <= nk; ++k) {
    const int64_t s = s0 + k * plane, w_k = w0 + k * wplane;
    const float du = ((u[ui(k, j, i, ni, nj)] - u[ui(k, j, i - 1, ni, nj)]) * c.rdx) * uh;
    const float dv = ((v[vi(k, j, i, ni, nj)] - v[vi(k, j - 1, i, ni, nj)]) * c.rdy) * vh;
    float div = du + dv;
    if (fabsf(div) < c.smeps) div = 0.0f;
    t.p[DIVH][s] = div;
    const float pc = piadv[s];
    const float gxw = pc - piadv[s - 1];
    const float gxe = piadv[s + 1] - pc;
    const float gys = pc - piadv[s - (ni + 2)];
    const float gyn = piadv[s + (ni + 2)] - pc;
    float hx = u[ui(k, j, i - 1, ni, nj)] * gxw + u[ui(k, j, i, ni, nj)] * gxe;
    hx = ((-hx) * c.temx) * uh;
    float hy = v[vi(k, j - 1, i, ni, nj)] * gys + v[vi(k, j, i, ni, nj)] * gyn;
    hy = ((-hy) * c.temy) * vh;
    const float hp = hx + hy;
    t.p[HADVP][s] = hp;
    const float src = c.dts * (ppten[s] - ppterm[s] * div);
    const float vert = w[w_k] * (pk1[s] + qk[s]) + w[w_k + wplane] * (pk2[s] - qk[s]);
    fpk[s] = (src + hp) + c.beta * vert;
  }

  // Forward elimination (:881-953); level k-1's terms are carried.
  float ec_prev = 0.0f, ed_prev = 0.0f;
  float pp_prev = pp3d[s0 + plane], fpk_prev = fpk[s0 + plane];
  float s1_prev = pk1[s0 + plane] + qk[s0 + plane];
  float s2_prev = pk2[s0 + plane] - qk[s0 + plane];
  for (int k = 2; k <= nk; ++k) {
    const int64_t s = s0 + k * plane, w_k = w0 + k * wplane;
    const float m = mm[s];
    const float s1 = pk1[s] + qk[s];
    const float s2 = pk2[s] - qk[s];
    const float aa = (k >= 3) ? (-c.asq) * (m * s1_prev) : 0.0f;
    const float cc = (k < nk) ? (-c.asq) * ((-m) * s2) : 0.0f;
    const float diag = (-s1) + s2_prev;
    const float bb = 1.0f - c.asq * (m * diag);
    const float pk = pp3d[s];
    const float fk = fpk[s];
    const float pd = pk - pp_prev;
    const float fwk = c.dts * wten[w_k] + c.beta * ((-m) * pd);
    const float upd = (pk + fk) - (pp_prev + fpk_prev);
    const float dd = (w[w_k] + fwk) + c.alph * ((-m) * upd);
    const float pivot = (k < 3) ? bb : aa * ec_prev + bb;
    const float inv = 1.0f / pivot;
    t.p[TRI_AA][s] = aa;
    t.p[TRI_BB][s] = bb;
    t.p[TRI_CC][s] = cc;
    t.p[TRI_DD][s] = dd;
    t.p[TRI_PIVOT][s] = pivot;
    if (k < nk) {
      ec_prev = (-cc) * inv;
      ed_prev = ((k < 3) ? dd : dd - aa * ed_prev) * inv;
      ec[s] = ec_prev;
      ed[s] = ed_prev;
    } else {
      w[w_k] = (dd - aa * ed_prev) * inv;
    }
    pp_prev = pk;
    fpk_prev = fk;
    s1_prev = s1;
    s2_prev = s2;
  }
  // Back substitution (:954-959).
  float w_up = w[w0 + nk * wplane];
  for (int k = nk - 1; k >= 2; --k) {
    const int64_t s = s0 + k * plane;
    w_up = ec[s] * w_up + ed[s];
    w[w0 + k * wplane] = w_up;
  }
  // Residual (:960-968).
  for (int k = 2; k <= nk; ++k) {
    const int64_t s = s0 + k * plane, w_k = w0 + k * wplane;
    float lhs = t.p[TRI_AA][s] * w[w_k - wplane];
    lhs = lhs + t.p[TRI_BB][s] * w[w_k];
    lhs = lhs + t.p[TRI_CC][s] * w[w_k + wplane];
    t.p[TRI_RESIDUAL][s] = lhs - t.p[TRI_DD][s];
  }
  // Pressure update, filter, ppx and the w average (:969-1010).
  float* rrw = t.p[RRW];
  for (int k = 1; k <= nk; ++k) {
    const int64_t s = s0 + k * plane, w_k = w0 + k * wplane;
    const float old = pp3d[s];
    const float wk = w[w_k];
    const float vert = wk * (pk1[s] + qk[s]) + w[w_k + wplane] * (pk2[s] - qk[s]);
    float upd = (old + fpk[s]) + c.alph * vert;
    if (fabsf(upd) < c.smeps) upd = 0.0f;
    pp3d[s] = upd;
    const float filt = c.kdiv * (upd - old);
    ec[s] = filt;
    t.p[PPD][s] = upd + filt;
    if (c.write_ppx) t.p[PPX][s] = filt;
    rrw[w_k] = c.last ? (rrw[w_k] + wk * c.alph) * c.tavg : rrw[w_k] + wk;
  }
}

// :506-516.  Halo word (k, j, i) of levels 1..nk takes the interior word the
// x-then-y copy sequence leaves there: i' = ni for i = 0, 1 for i = ni+1, and
// likewise in j (corners therefore read (nj or 1, ni or 1)).
__global__ void __launch_bounds__(256) k_sp_fill_scalar(float* s, Dims d) {
  const int ni = d.ni, nj = d.nj, nk = d.nk;
  const int64_t total = (int64_t)nk * (nj + 2) * (ni + 2);
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= total) return;
  const int i = (int)(q % (ni + 2));
  const int j = (int)((q / (ni + 2)) % (nj + 2));
  const int k = (int)(q / ((int64_t)(ni + 2) * (nj + 2))) + 1;
  if (i >= 1 && i <= ni && j >= 1 && j <= nj) return;
  const int is = (i == 0) ? ni : (i == ni + 1) ? 1 : i;
  const int js = (j == 0) ? nj : (j == nj + 1) ? 1 : j;
  s[si(k, j, i, ni, nj)] = s[si(k, js, is, ni, nj)];
}

// :519-532.  u rows 0 and nj+1, v columns 0 and ni+1, levels 1..nk.
__global__ void __launch_bounds__(256) k_sp_fill_uv(float* u, float* v, Dims d) {
  const int ni = d.ni, nj = d.nj, nk = d.nk;
  const int64_t nu = (int64_t)nk * 2 * (ni + 1);
  const int64_t nv = (int64_t)nk * (nj + 1) * 2;
  int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < nu) {
    const int f = (int)(q % (ni + 1));
    const int side = (int)((q / (ni + 1)) % 2);
    const int k = (int)(q / (2 * (int64_t)(ni + 1))) + 1;
    const int j = side ? nj + 1 : 0, js = side ? 1 : nj;
    u[ui(k, j, f, ni, nj)] = u[ui(k, js, f, ni, nj)];
    return;
  }
  q -= nu;
  if (q < nv) {
    const int side = (int)(q % 2);
    const int f = (int)((q / 2) % (nj + 1));
    const int k = (int)(q / (2 * (int64_t)(nj + 1))) + 1;
    const int i = side ? ni + 1 : 0, is = side ? 1 : ni;
    v[vi(k, f, i, ni, nj)] = v[vi(k, f, is, ni, nj)];
  }
}

unsigned blocks(int64_t n, int t) { return (unsigned)((n + t - 1) / t); }

}  // namespace

extern "C" int gm1hip_sound3_periodic_stage(void* const* arrays, int ni, int nj, int nk, int nrk, int nloop,
                                            float dts, float rdx, float rdy, float rdz, float cp, float alph,
                                            float kdiv, void* stream) {
  if (!arrays || ni < 3 || nj < 3 || nk < 3 || nrk < 1 || nrk > 3 || nloop < 1) return (int)hipErrorInvalidValue;
  Tab t;
  for (int s = 0; s < NSLOTS; ++s) {
    t.p[s] = static_cast<float*>(arrays[s]);
    if (!t.p[s]) return (int)hipErrorInvalidValue;
  }
  const hipStream_t st = static_cast<hipStream_t>(stream);
  const Dims d{ni, nj, nk};
  const float beta = 1.0f - alph;
  const float asq = alph * alph;
  constexpr int T = 256;
  const int64_t scalar_words = (int64_t)(nk + 2) * (nj + 2) * (ni + 2);
  const int64_t face_words =
      std::max((int64_t)(nk + 2) * (nj + 2) * (ni + 1), (int64_t)(nk + 2) * (nj + 1) * (ni + 2));
  const int64_t live_faces = 2 * (int64_t)nk * (nj + 1) * (ni + 1);
  const int64_t halo_words = (int64_t)nk * (nj + 2) * (ni + 2);
  const int64_t uv_halo = (int64_t)nk * 2 * (ni + 1) + (int64_t)nk * (nj + 1) * 2;

  hipLaunchKernelGGL(k_sp_prepare, dim3(blocks(scalar_words, T)), dim3(T), 0, st, t, d, nrk, dts, rdz, cp, beta);
  hipLaunchKernelGGL(k_sp_prepare_faces, dim3(blocks(face_words, T)), dim3(T), 0, st, t, d, nrk);
  const float tem1 = (rdx * cp) * 0.5f;
  const float tem2 = (rdy * cp) * 0.5f;
  ColumnScalars c;
  c.dts = dts;
  c.rdx = rdx;
  c.rdy = rdy;
  c.temx = (dts * 0.5f) * rdx;
  c.temy = (dts * 0.5f) * rdy;
  c.alph = alph;
  c.beta = beta;
  c.asq = asq;
  c.kdiv = kdiv;
  c.smeps = 1.0e-30f;
  c.tavg = 1.0f / (float)nloop;
  const dim3 cblock(64, 4);
  const dim3 cgrid((unsigned)((ni + 63) / 64), (unsigned)((nj + 3) / 4));
  for (int n = 1; n <= nloop; ++n) {
    const int last = n == nloop;
    hipLaunchKernelGGL(k_sp_faces, dim3(blocks(live_faces, T)), dim3(T), 0, st, t, d, dts, tem1, tem2, c.tavg,
                       last);
    c.last = last;
    c.write_ppx = (nrk == 3 && last) ? 1 : 0;
    hipLaunchKernelGGL(k_sp_column, cgrid, cblock, 0, st, t, d, c);
    if (!last) hipLaunchKernelGGL(k_sp_fill_scalar, dim3(blocks(halo_words, T)), dim3(T), 0, st, t.p[PPD], d);
  }
  hipLaunchKernelGGL(k_sp_fill_uv, dim3(blocks(uv_halo, T)), dim3(T), 0, st, t.p[U3D], t.p[V3D], d);
  if (nrk < 3) hipLaunchKernelGGL(k_sp_fill_scalar, dim3(blocks(halo_words, T)), dim3(T), 0, st, t.p[PP3D], d);
  return (int)hipGetLastError();
}
