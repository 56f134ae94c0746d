// Adjoint of the acoustic RK stage (psolver=3 `sound`) on gfx950, for the
// dry flat-terrain, uniform-column, doubly periodic geometry.
//
// For fixed coefficients (thv, ppterm, piadv -> qk, pk1, pk2, mm) one stage
// (cuda_psolver3_terrain_open.py:774-893; per small step :838-863) is linear
// in its state and forcing; this is the transpose of that map, statement by
// statement in reverse:
//   stage end   family fills of u, v, w; bcw_rigid; bcwsfc; bcp(pp3d) (nrk<3)
//   small step  time averages <- bcp(ppd) <- pressure update and filter <-
//               Thomas solve (transposed: a second tridiagonal solve with the
//               off-diagonals exchanged) <- pressure advection <- bcwsfc <-
//               fpk <- pressure gradient on faces 0..n
//   stage start ppd = ppx, pp3d = ppi, rru = rrv = 0, rrw = beta*wa, u/v/w = A
// The smeps clamps are taken as the identity.  Arithmetic is binary32; the
// coefficients are formed with the forward's own expressions.
#include "cm1_common.h"

namespace {

using gm1::ix3;
using gm1::NG;
using gm1::pix;
using gm1::sx;
using gm1::sy;
using gm1::ux;
using gm1::vy;

constexpr float CPF = 1005.7f;

// Slots of the pointer table of gm1hip_psolver3_stage_adjoint.
enum Slot {
  S_THV, S_PPTERM, S_PIADV,                          // pressure family, fixed
  S_RHO0, S_RR0, S_RF0, S_C1, S_C2,                  // metrics (terrain_static_shapes), flat
  S_PROF,                                            // sigma | sigmaf | rds | rdsf (nk+2 each) | ...
  S_QK, S_PK1, S_PK2, S_MM, S_FPKB, S_EC, S_ED,      // scratch, pressure family
  S_UB, S_VB, S_WB, S_PB, S_PDB, S_PPXB, S_RRUB, S_RRVB, S_RRWB,  // adjoint state
  S_UTENB, S_VTENB, S_WTENB, S_PPTENB, S_UAB, S_VAB, S_WAB, S_PPIB,  // input adjoints (+=)
  S_COUNT
};

struct Adj {
  const float* thv; const float* ppterm; const float* pia;
  const float* rho0; const float* rr0; const float* rf0; const float* c1; const float* c2;
  const float* prof;
  float* qk; float* pk1; float* pk2; float* mm; float* fpkb; float* ec; float* ed;
  float* ub; float* vb; float* wb; float* pb; float* pdb; float* ppxb; float* rrub; float* rrvb; float* rrwb;
  float* utenb; float* vtenb; float* wtenb; float* pptenb; float* uab; float* vab; float* wab; float* ppib;
  int ni, nj, nk;
  float dts, rdx, rdy, alph, beta, kdiv, tavg, temx, temy;
};

GM1_DEV const float* rdsf_of(const Adj& a) { return a.prof + 3 * (a.nk + 2); }
GM1_DEV const float* rds_of(const Adj& a) { return a.prof + 2 * (a.nk + 2); }

// qk, pk1, pk2, mm at interior pressure cells (prepare_coefficients, flat metrics)
__global__ __launch_bounds__(64) void k_adj_coef(Adj a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x + 1, j = blockIdx.y + 1, k = blockIdx.z + 1;
  const int ni = a.ni, nj = a.nj, nk = a.nk;
  if (i > ni) return;
  const int64_t p = pix(k, j, i, ni, nj);
  const int64_t h = (int64_t)(j - 1) * ni + (i - 1), plane = (int64_t)nj * ni;
  const float tem = a.dts * rdsf_of(a)[k];
  const float qk = tem * a.ppterm[p];
  const float tem2 = (a.dts * rdsf_of(a)[k]) * a.rr0[(k - 1) * plane + h];
  const float b2 = (-a.c2[k * plane + h]) * a.pia[pix(k + 1, j, i, ni, nj)] + (1.0f - a.c1[k * plane + h]) * a.pia[p];
  float pk2 = (tem2 * a.rf0[k * plane + h]) * b2;
  const float b1 = a.c1[(k - 1) * plane + h] * a.pia[pix(k - 1, j, i, ni, nj)] +
                   (a.c2[(k - 1) * plane + h] - 1.0f) * a.pia[p];
  float pk1 = (tem2 * a.rf0[(k - 1) * plane + h]) * b1;
  float mm = 0.0f;
  if (k >= 2) {
    const float tem1 = (a.dts * rds_of(a)[k]) * CPF;
    const float th = a.c2[(k - 1) * plane + h] * a.thv[p] + a.c1[(k - 1) * plane + h] * a.thv[pix(k - 1, j, i, ni, nj)];
    mm = tem1 * th;
  }
  if (k == 1) pk1 = 0.0f;
  if (k == nk) pk2 = 0.0f;
  a.qk[p] = qk;
  a.pk1[p] = pk1;
  a.pk2[p] = pk2;
  a.mm[p] = mm;
}

// bcp on periodic sides, reversed: north and south halo rows join their
// source rows (all columns), then east and west halo columns (all rows).
__global__ __launch_bounds__(64) void k_adj_bcp_sn(float* pb, int ni, int nj, int nk) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, k = blockIdx.y + 1;
  if (i > ni + 1) return;
  float* P = pb + (int64_t)k * (nj + 2) * (ni + 2);
  const int w = ni + 2;
  P[1 * w + i] += P[(nj + 1) * w + i];
  P[(nj + 1) * w + i] = 0.0f;
  P[nj * w + i] += P[0 * w + i];
  P[0 * w + i] = 0.0f;
}

__global__ __launch_bounds__(64) void k_adj_bcp_we(float* pb, int ni, int nj, int nk) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x, k = blockIdx.y + 1;
  if (j > nj + 1) return;
  float* R = pb + ((int64_t)k * (nj + 2) + j) * (ni + 2);
  R[1] += R[ni + 1];
  R[ni + 1] = 0.0f;
  R[ni] += R[0];
  R[0] = 0.0f;
}

// Periodic family fill reversed, one direction: every ghost's adjoint joins
// its source, then clears, in the reverse of the forward copy order.  One
// thread owns one line (a row for dir 0, a column for dir 1) of one level, so
// sources shared by a west and an east ghost on narrow grids do not race.
// dir 0: west/east ghost columns over all rows; dir 1: south/north ghost rows
// over all columns.  stag: the array is staggered in that direction.
__global__ __launch_bounds__(64) void k_adj_family(float* ab, int nx, int ny, int n, int kfirst, int dir, int stag) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;  // along the plane
  const int k = kfirst + (int)blockIdx.y;
  if (t >= (dir == 0 ? ny : nx)) return;
  float* A = ab + (int64_t)k * ny * nx;
  auto at = [&](int c) -> float& { return dir == 0 ? A[(int64_t)t * nx + c] : A[(int64_t)c * nx + t]; };
  for (int layer = NG - 1; layer >= 0; --layer) {
    const int de = stag ? n + NG + 1 + layer : n + NG + layer, se = stag ? NG + 1 + layer : NG + layer;
    at(se) += at(de);
    at(de) = 0.0f;
    const int dw = NG - 1 - layer, sw = n + NG - 1 - layer;
    at(sw) += at(dw);
    at(dw) = 0.0f;
  }
}

// w = 0 on the floor (bcwsfc, bcw_rigid) and at the lid (bcw_rigid), reversed
__global__ __launch_bounds__(64) void k_adj_wzero(float* wb, int ni, int nj, int nk) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
  if (i > ni + 1 || j > nj + 1) return;
  const int64_t st = (int64_t)sy(nj) * sx(ni);
  const int64_t c = (int64_t)(j + NG - 1) * sx(ni) + (i + NG - 1);
  wb[st + c] = 0.0f;
  wb[(int64_t)(nk + 1) * st + c] = 0.0f;
}

// time averages reversed over rows ng..ng+nj, columns ng..ng+ni, levels 1..nk
__global__ __launch_bounds__(64) void k_adj_average(Adj a, int last) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x + NG, y = blockIdx.y + NG, k = blockIdx.z + 1;
  const int ni = a.ni, nj = a.nj;
  if (x > ni + NG) return;
  const int64_t pu = ix3(k, y, x, sy(nj), ux(ni)), pv = ix3(k, y, x, vy(nj), sx(ni)), pw = ix3(k, y, x, sy(nj), sx(ni));
  if (last) {
    const float ru = a.rrub[pu] * a.tavg, rv = a.rrvb[pv] * a.tavg, rw = a.rrwb[pw] * a.tavg;
    a.rrub[pu] = ru;
    a.rrvb[pv] = rv;
    a.rrwb[pw] = rw;
    a.ub[pu] += ru;
    a.vb[pv] += rv;
    a.wb[pw] += a.alph * rw;
  } else {
    a.ub[pu] += a.rrub[pu];
    a.vb[pv] += a.rrvb[pv];
    a.wb[pw] += a.rrwb[pw];
  }
}

// One column, reversed: pressure update and filter, the transposed Thomas
// solve, the w and ppten parts of pressure advection and fpk; fpk's adjoint
// is left in fpkb for the face pass.
__global__ __launch_bounds__(64) void k_adj_column(Adj a, int last3) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x + 1, j = blockIdx.y + 1;
  const int ni = a.ni, nj = a.nj, nk = a.nk;
  if (i > ni) return;
  const int64_t wst = (int64_t)sy(nj) * sx(ni), wc = (int64_t)(j - 1 + NG) * sx(ni) + (i - 1 + NG);
  auto P = [&](int k) { return pix(k, j, i, ni, nj); };
  float* wb = a.wb;
  const float alph = a.alph, beta = a.beta, kdiv = a.kdiv, dts = a.dts;
  // A6: update and filter (each chunk of 8 levels' words requested together)
  for (int k0 = 1; k0 <= nk; k0 += 8) {
    float pb_[8], pdb_[8], ppxb_[8], s1_[8], s2_[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int k = k0 + u;
      if (k <= nk) {
        const int64_t p = P(k);
        pb_[u] = a.pb[p];
        pdb_[u] = a.pdb[p];
        ppxb_[u] = last3 ? a.ppxb[p] : 0.0f;
        s1_[u] = a.pk1[p] + a.qk[p];
        s2_[u] = a.pk2[p] - a.qk[p];
      }
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int k = k0 + u;
      if (k <= nk) {
        const int64_t p = P(k);
        float upb = pb_[u] + (1.0f + kdiv) * pdb_[u];
        float oldb = -kdiv * pdb_[u];
        if (last3) {
          upb = upb + kdiv * ppxb_[u];
          oldb = oldb - kdiv * ppxb_[u];
          a.ppxb[p] = 0.0f;
        }
        a.pdb[p] = 0.0f;
        a.pb[p] = oldb + upb;
        a.fpkb[p] = upb;
        if (k >= 2) wb[(int64_t)k * wst + wc] += alph * s1_[u] * upb;
        if (k <= nk - 1) wb[(int64_t)(k + 1) * wst + wc] += alph * s2_[u] * upb;
      }
    }
  }
  // A5: T^T x = wb(2..nk); T's row k: aa(k) below, bb(k) on, cc(k) above the diagonal
  const float asq = alph * alph;
  auto AA = [&](int k) {
    return k >= 3 ? (-asq) * (a.mm[P(k)] * (a.pk1[P(k - 1)] + a.qk[P(k - 1)])) : 0.0f;
  };
  auto CC = [&](int k) { return k < nk ? (-asq) * ((-a.mm[P(k)]) * (a.pk2[P(k)] - a.qk[P(k)])) : 0.0f; };
  auto BB = [&](int k) {
    const float diag = (-(a.pk1[P(k)] + a.qk[P(k)])) + (a.pk2[P(k - 1)] - a.qk[P(k - 1)]);
    return 1.0f - asq * (a.mm[P(k)] * diag);
  };
  // forward sweep of T^T: row k has CC(k-1) below and AA(k+1) above its
  // diagonal.  Each chunk of CH levels has its coefficients and right-hand
  // sides formed into registers before its recurrence runs, so the chunk's
  // loads are in flight together.
  constexpr int CH = 8;
  float cprev = 0.0f, dprev = 0.0f;
  for (int k0 = 2; k0 <= nk; k0 += CH) {
    float lo_[CH], up_[CH], bb_[CH], rhs_[CH];
#pragma unroll
    for (int u = 0; u < CH; ++u) {
      const int k = k0 + u;
      if (k <= nk) {
        lo_[u] = (k == 2) ? 0.0f : CC(k - 1);
        up_[u] = (k < nk) ? AA(k + 1) : 0.0f;
        bb_[u] = BB(k);
        rhs_[u] = wb[(int64_t)k * wst + wc];
      }
    }
#pragma unroll
    for (int u = 0; u < CH; ++u) {
      const int k = k0 + u;
      if (k <= nk) {
        const float piv = bb_[u] - lo_[u] * cprev;
        const float cpk = up_[u] / piv;
        const float dpk = (rhs_[u] - lo_[u] * dprev) / piv;
        a.ec[P(k)] = cpk;
        a.ed[P(k)] = dpk;
        cprev = cpk;
        dprev = dpk;
      }
    }
  }
  // back substitution, CH levels' ec, ed and mm requested at a time
  float xnext = 0.0f;
  for (int k0 = nk; k0 >= 2; k0 -= CH) {
    float ec_[CH], ed_[CH], mm_[CH];
#pragma unroll
    for (int u = 0; u < CH; ++u) {
      const int k = k0 - u;
      if (k >= 2) {
        ec_[u] = a.ec[P(k)];
        ed_[u] = a.ed[P(k)];
        mm_[u] = a.mm[P(k)];
      }
    }
#pragma unroll
    for (int u = 0; u < CH; ++u) {
      const int k = k0 - u;
      if (k >= 2) {
        const float x = (k == nk) ? ed_[u] : ed_[u] - ec_[u] * xnext;
        xnext = x;
        // dd(k) = w(k) + dts*wten(k) + beta*(-mm)*(pp(k)-pp(k-1)) + alph*(-mm)*((pp+fpk)(k) - (pp+fpk)(k-1))
        const float mmk = mm_[u];
        wb[(int64_t)k * wst + wc] = x;
        a.wtenb[(int64_t)k * wst + wc] += dts * x;
        a.pb[P(k)] += (-mmk) * x;
        a.pb[P(k - 1)] -= (-mmk) * x;
        a.fpkb[P(k)] += alph * (-mmk) * x;
        a.fpkb[P(k - 1)] -= alph * (-mmk) * x;
      }
    }
  }
  // A4 (w part), A3, A2 (w part, ppten), 8 levels' words requested together
  for (int k0 = 1; k0 <= nk; k0 += 8) {
    float f_[8], p1_[8], p2_[8];
#pragma unroll
    for (int u = 0; u < 8; ++u)
      if (k0 + u <= nk) {
        f_[u] = a.fpkb[P(k0 + u)];
        p1_[u] = a.pk1[P(k0 + u)];
        p2_[u] = a.pk2[P(k0 + u)];
      }
#pragma unroll
    for (int u = 0; u < 8; ++u)
      if (k0 + u <= nk) {
        const int k = k0 + u;
        wb[(int64_t)k * wst + wc] += beta * p1_[u] * f_[u];
        wb[(int64_t)(k + 1) * wst + wc] += beta * p2_[u] * f_[u];
      }
  }
  wb[wst + wc] = 0.0f;
  for (int k0 = 1; k0 <= nk; k0 += 8) {
    float f_[8], q_[8], t_[8];
#pragma unroll
    for (int u = 0; u < 8; ++u)
      if (k0 + u <= nk) {
        f_[u] = a.fpkb[P(k0 + u)];
        q_[u] = a.qk[P(k0 + u)];
        t_[u] = a.pptenb[P(k0 + u)];
      }
#pragma unroll
    for (int u = 0; u < 8; ++u)
      if (k0 + u <= nk) {
        const int k = k0 + u;
        const float f = f_[u];
        a.pptenb[P(k)] = t_[u] + dts * f;
        if (k < nk) wb[(int64_t)(k + 1) * wst + wc] += -beta * q_[u] * f;
        if (k > 1) wb[(int64_t)k * wst + wc] += beta * q_[u] * f;
      }
  }
}

// u faces 0..ni and v faces 0..nj: fpk's divergence and pressure-advection
// adjoints gathered from the faces' two cells, then the pressure gradient's
// forcing adjoint (uten, vten).
__global__ __launch_bounds__(64) void k_adj_faces(Adj a) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x, r = blockIdx.y, k = blockIdx.z + 1;
  const int ni = a.ni, nj = a.nj;
  const int64_t plane = (int64_t)nj * ni;
  const float rho = a.rho0[(int64_t)(k - 1) * plane];
  auto cell_terms = [&](int ci, int cj, float& divb, float& hzb) {  // interior cell (0-based)
    const int64_t p = pix(k, cj + 1, ci + 1, ni, nj);
    const float f = a.fpkb[p];
    divb = -a.dts * a.ppterm[p] * f;
    hzb = f * a.rr0[(int64_t)(k - 1) * plane + (int64_t)cj * ni + ci];
  };
  // u face f = t of row j = r (t <= ni, r < nj)
  if (t <= ni && r < nj) {
    float add = 0.0f;
    if (t >= 1) {  // east face of cell t-1
      float divb, hzb;
      cell_terms(t - 1, r, divb, hzb);
      const int64_t pc = pix(k, r + 1, t, ni, nj);
      add += a.rdx * divb + (-(a.temx * rho)) * (a.pia[pc + 1] - a.pia[pc]) * hzb;
    }
    if (t < ni) {  // west face of cell t
      float divb, hzb;
      cell_terms(t, r, divb, hzb);
      const int64_t pc = pix(k, r + 1, t + 1, ni, nj);
      add += -a.rdx * divb + (-(a.temx * rho)) * (a.pia[pc] - a.pia[pc - 1]) * hzb;
    }
    const int64_t pu = ix3(k, r + NG, t + NG, sy(nj), ux(ni));
    const float ubt = a.ub[pu] + add;
    a.ub[pu] = ubt;
    a.utenb[pu] += a.dts * ubt;
  }
  // v face g = r of column i = t (r <= nj, t < ni)
  if (t < ni && r <= nj) {
    float add = 0.0f;
    if (r >= 1) {  // north face of cell r-1
      float divb, hzb;
      cell_terms(t, r - 1, divb, hzb);
      const int64_t pc = pix(k, r, t + 1, ni, nj);
      add += a.rdy * divb + (-(a.temy * rho)) * (a.pia[pc + (ni + 2)] - a.pia[pc]) * hzb;
    }
    if (r < nj) {  // south face of cell r
      float divb, hzb;
      cell_terms(t, r, divb, hzb);
      const int64_t pc = pix(k, r + 1, t + 1, ni, nj);
      add += -a.rdy * divb + (-(a.temy * rho)) * (a.pia[pc] - a.pia[pc - (ni + 2)]) * hzb;
    }
    const int64_t pv = ix3(k, r + NG, t + NG, vy(nj), sx(ni));
    const float vbt = a.vb[pv] + add;
    a.vb[pv] = vbt;
    a.vtenb[pv] += a.dts * vbt;
  }
}

// ppd's adjoint from the pressure gradient on faces 0..n: each pressure cell
// (halos included) gathers its (up to four) faces.
__global__ __launch_bounds__(64) void k_adj_ppd(Adj a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y, k = blockIdx.z + 1;
  const int ni = a.ni, nj = a.nj;
  if (i > ni + 1) return;
  const float halfcp = CPF * 0.5f;
  const int64_t p = pix(k, j, i, ni, nj);
  float acc = 0.0f;
  if (j >= 1 && j <= nj) {
    const int y = j - 1 + NG;
    if (i >= 1) {  // right cell of u face f = i-1
      const float cu = (halfcp * (a.thv[p] + a.thv[p - 1])) * a.rdx;
      acc += -a.dts * cu * a.ub[ix3(k, y, i - 1 + NG, sy(nj), ux(ni))];
    }
    if (i <= ni) {  // left cell of u face f = i
      const float cu = (halfcp * (a.thv[p + 1] + a.thv[p])) * a.rdx;
      acc += a.dts * cu * a.ub[ix3(k, y, i + NG, sy(nj), ux(ni))];
    }
  }
  if (i >= 1 && i <= ni) {
    const int x = i - 1 + NG;
    if (j >= 1) {
      const float cv = (halfcp * (a.thv[p] + a.thv[p - (ni + 2)])) * a.rdy;
      acc += -a.dts * cv * a.vb[ix3(k, j - 1 + NG, x, vy(nj), sx(ni))];
    }
    if (j <= nj) {
      const float cv = (halfcp * (a.thv[p + (ni + 2)] + a.thv[p])) * a.rdy;
      acc += a.dts * cv * a.vb[ix3(k, j + NG, x, vy(nj), sx(ni))];
    }
  }
  a.pdb[p] += acc;
}

// the stage-start copies reversed (rows ng-1..ng+nj, columns ng-1..ng+ni for the winds)
__global__ __launch_bounds__(64) void k_adj_start(Adj a, int nrk) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y, k = blockIdx.z + 1;
  const int ni = a.ni, nj = a.nj;
  if (i <= ni + 1 && j <= nj + 1) {
    const int64_t p = pix(k, j, i, ni, nj);
    a.ppxb[p] += a.pdb[p];
    a.pdb[p] = 0.0f;
    if (nrk != 1) {
      a.ppib[p] += a.pb[p];
      a.pb[p] = 0.0f;
    }
    const int y = j + NG - 1, x = i + NG - 1;
    const int64_t pu = ix3(k, y, x, sy(nj), ux(ni)), pv = ix3(k, y, x, vy(nj), sx(ni));
    const int64_t pw = ix3(k, y, x, sy(nj), sx(ni));
    a.wab[pw] += a.beta * a.rrwb[pw];
    a.rrwb[pw] = 0.0f;
    a.rrub[pu] = 0.0f;
    a.rrvb[pv] = 0.0f;
    if (nrk != 1) {
      a.uab[pu] += a.ub[pu];
      a.ub[pu] = 0.0f;
      a.vab[pv] += a.vb[pv];
      a.vb[pv] = 0.0f;
      a.wab[pw] += a.wb[pw];
      a.wb[pw] = 0.0f;
    }
  }
  // pressure halos outside rows/columns 0..n+1 do not exist; nothing else to move
}

int family_adjoint(float* ab, int kind, int ni, int nj, int nk, hipStream_t s) {
  // kind 1 u, 2 v, 3 w; forward order: v is south/north then west/east, the others west/east then south/north
  const int nx = ni + 2 * NG + (kind == 1 ? 1 : 0);
  const int ny = nj + 2 * NG + (kind == 2 ? 1 : 0);
  const int kfirst = kind == 3 ? 2 : 1, kcount = kind == 3 ? nk - 1 : nk;
  if (kcount <= 0) return 0;
  auto we = [&]() {
    hipLaunchKernelGGL(k_adj_family, dim3((ny + 63) / 64, kcount), dim3(64), 0, s, ab, nx, ny, ni, kfirst, 0,
                       kind == 1 ? 1 : 0);
  };
  auto sn = [&]() {
    hipLaunchKernelGGL(k_adj_family, dim3((nx + 63) / 64, kcount), dim3(64), 0, s, ab, nx, ny, nj, kfirst, 1,
                       kind == 2 ? 1 : 0);
  };
  if (kind == 2) { we(); sn(); }
  else { sn(); we(); }
  GM1_LAUNCH_CHECK();
  return 0;
}

}  // namespace

extern "C" {

int gm1hip_psolver3_stage_adjoint(void* const* arrays, int ni, int nj, int nk, int nrk, int nloop, float dts,
                                  float dx, float dy, float alph, float kdiv, void* stream) {
  if (!arrays || ni < 3 || nj < 3 || nk < 3 || nloop < 1 || nrk < 1 || nrk > 3) return (int)hipErrorInvalidValue;
  for (int n = 0; n < S_COUNT; ++n)
    if (!arrays[n]) return (int)hipErrorInvalidValue;
  hipStream_t s = (hipStream_t)stream;
  auto F = [&](int slot) { return (float*)arrays[slot]; };
  Adj a;
  a.thv = F(S_THV); a.ppterm = F(S_PPTERM); a.pia = F(S_PIADV);
  a.rho0 = F(S_RHO0); a.rr0 = F(S_RR0); a.rf0 = F(S_RF0); a.c1 = F(S_C1); a.c2 = F(S_C2); a.prof = F(S_PROF);
  a.qk = F(S_QK); a.pk1 = F(S_PK1); a.pk2 = F(S_PK2); a.mm = F(S_MM); a.fpkb = F(S_FPKB); a.ec = F(S_EC);
  a.ed = F(S_ED);
  a.ub = F(S_UB); a.vb = F(S_VB); a.wb = F(S_WB); a.pb = F(S_PB); a.pdb = F(S_PDB); a.ppxb = F(S_PPXB);
  a.rrub = F(S_RRUB); a.rrvb = F(S_RRVB); a.rrwb = F(S_RRWB);
  a.utenb = F(S_UTENB); a.vtenb = F(S_VTENB); a.wtenb = F(S_WTENB); a.pptenb = F(S_PPTENB);
  a.uab = F(S_UAB); a.vab = F(S_VAB); a.wab = F(S_WAB); a.ppib = F(S_PPIB);
  a.ni = ni; a.nj = nj; a.nk = nk;
  a.dts = dts; a.rdx = 1.0f / dx; a.rdy = 1.0f / dy; a.alph = alph; a.beta = 1.0f - alph; a.kdiv = kdiv;
  a.tavg = 1.0f / (float)nloop;
  a.temx = (dts * 0.5f) * a.rdx;
  a.temy = (dts * 0.5f) * a.rdy;
  const dim3 b64(64);
  hipLaunchKernelGGL(k_adj_coef, dim3((ni + 63) / 64, nj, nk), b64, 0, s, a);
  // stage end
  if (nrk < 3) {
    hipLaunchKernelGGL(k_adj_bcp_sn, dim3((ni + 2 + 63) / 64, nk), b64, 0, s, a.pb, ni, nj, nk);
    hipLaunchKernelGGL(k_adj_bcp_we, dim3((nj + 2 + 63) / 64, nk), b64, 0, s, a.pb, ni, nj, nk);
  }
  hipLaunchKernelGGL(k_adj_wzero, dim3((ni + 2 + 63) / 64, nj + 2), b64, 0, s, a.wb, ni, nj, nk);
  GM1_LAUNCH_CHECK();
  int rc = family_adjoint(a.wb, 3, ni, nj, nk, s);
  if (!rc) rc = family_adjoint(a.vb, 2, ni, nj, nk, s);
  if (!rc) rc = family_adjoint(a.ub, 1, ni, nj, nk, s);
  if (rc) return rc;
  for (int n = nloop; n >= 1; --n) {
    const int last = n == nloop;
    hipLaunchKernelGGL(k_adj_average, dim3((ni + 1 + 63) / 64, nj + 1, nk), b64, 0, s, a, last);
    if (n < nloop) {
      hipLaunchKernelGGL(k_adj_bcp_sn, dim3((ni + 2 + 63) / 64, nk), b64, 0, s, a.pdb, ni, nj, nk);
      hipLaunchKernelGGL(k_adj_bcp_we, dim3((nj + 2 + 63) / 64, nk), b64, 0, s, a.pdb, ni, nj, nk);
    }
    hipLaunchKernelGGL(k_adj_column, dim3((ni + 63) / 64, nj), b64, 0, s, a, (nrk == 3 && last) ? 1 : 0);
    hipLaunchKernelGGL(k_adj_faces, dim3((ni + 1 + 63) / 64, nj + 1, nk), b64, 0, s, a);
    hipLaunchKernelGGL(k_adj_ppd, dim3((ni + 2 + 63) / 64, nj + 2, nk), b64, 0, s, a);
    GM1_LAUNCH_CHECK();
  }
  hipLaunchKernelGGL(k_adj_start, dim3((ni + 2 + 63) / 64, nj + 2, nk), b64, 0, s, a, nrk);
  GM1_LAUNCH_CHECK();
  return 0;
}

}  // extern "C"
