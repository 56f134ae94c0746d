// Dry large step and its adjoint on halo-free periodic fields (gfx950).
//
// Geometry: flat terrain, horizontally uniform base state, CM1's (1,1,1,1)
// lateral sides.  Fields carry no halo: u, v and cell fields are
// (nk, nj, ni), w is (nk+1, nj, ni) holding levels 1..nk+1; u face f is the
// west face of cell f, v face g the south face of row g; every lateral
// neighbour is a periodic index.  The forward stage restates one RK stage
// of the large step (cuda_gm1_forecast.py:1241-1345 with the acoustic stage
// cuda_psolver3_terrain_open.py:774-893 and the scalar stage
// cuda_scalar_moist_rk.py:283-349) in the reference's association order, so
// it produces the words of the padded device step; it keeps every acoustic
// substep's state (u, v, w, pp, ppd) as the adjoint's trajectory.  The
// adjoint stage is its transpose at that trajectory with the upwind choices,
// WENO5 weights and positive-definite limiters frozen at their forward
// values and the smeps clamps taken as the identity.
#include <algorithm>

#include "cm1_common.h"

namespace {

using gm1::flx3;
using gm1::flx5;
using gm1::upstrpd;
using gm1::weno5;

// Pointer-table slots of gm1hip_dry_stage_forward / gm1hip_dry_stage_adjoint.
enum Slot {
  COL,                                    // packed columns (see Cols)
  XU, XV, XW, XTH, XPP,                   // stage input state
  AU, AV, AW, ATH, API,                   // large-step start state (ua va wa tha ppi)
  PPX,                                    // ppx entering the stage
  THV, PPTERM, PIADV, DUM8, THTERM,       // stage terms
  RRU, RRV, RRW, DIVX, UTEN, VTEN, WTEN,  // momentum mass fluxes, divx, tendencies
  QK, PK1, PK2, MM, FPK, EC, ED,          // acoustic coefficients and column scratch
  SU, SV, SW, SPP, SPPD,                  // substep states, (nloop+1) planes each; plane n = after substep n
  ARU, ARV, ARW,                          // stage-averaged winds
  PPXO,                                   // ppx leaving the stage (nrk 3: after solve3)
  DIV, TDIV, DIVX2, THO,                  // post-acoustic divergence, theta_div, divx, theta out
  // adjoint
  B_OU, B_OV, B_OW, B_OTH, B_OPP,         // in: adjoints of the stage outputs
  B_PPX,                                  // in: adjoint of PPXO; out: adjoint of PPX
  B_XU, B_XV, B_XW, B_XTH, B_XPP,         // out: adjoints of the stage inputs X
  B_AU, B_AV, B_AW, B_ATH, B_API,         // accumulated adjoints of A
  B_THV, B_PPTERM, B_PIADV, B_DUM8, B_THTERM,
  B_RRU, B_RRV, B_RRW, B_DIVX, B_UTEN, B_VTEN, B_WTEN,
  B_QK, B_PK1, B_PK2, B_MM, B_U, B_V, B_W, B_PP, B_PPD, B_ARU, B_ARV, B_ARW,
  B_DIV, B_XA, B_TOT, B_MU, B_MV, B_MW, B_DIVX2, B_SC, B_EC, B_ED, B_FPKF,
  RA, RB, RC, RD,                         // tridiagonal rows of the stage (aa, bb, cc) and a substep's right side
  G0, G1, G2, G3, G4, G5,                 // momentum transposes: per-flux adjoints (upwind/downwind parts, mass)
  NSLOT
};

// Scalars of one stage, formed on the host in binary32 as the restatement forms them.
struct Consts {
  float rdx, rdy, rdz, dts, dttmp, alph, beta, kdiv, halfcp, cp, tavg, temx, temy;
  float g, rovcp, reps, repsm1, cpm, cvm, kterm, asq, weps_f;
  float exact;  // 1: WENO5 weights and limiters differentiated; 0: frozen at their forward values
};
constexpr int NCONST = 23;

struct Cols {
  const float *rho, *rr0, *rf0, *rrf0, *c1, *c2, *th0, *pi0, *qv0, *rth, *rds, *rdsf;
};

struct W {
  float* p[NSLOT];
  int ni, nj, nk, nrk, nloop;
  double weps;
  Consts c;
  Cols col;
  // index of (level l, row j, column i), levels 1-based
  __device__ __forceinline__ int64_t at(int l, int j, int i) const {
    return ((int64_t)(l - 1) * nj + j) * ni + i;
  }
  __device__ __forceinline__ int X(int i, int o) const { return (i + o + 4 * ni) % ni; }
  __device__ __forceinline__ int Y(int j, int o) const { return (j + o + 4 * nj) % nj; }
  __device__ __forceinline__ int64_t plane() const { return (int64_t)nk * nj * ni; }
  __device__ __forceinline__ int64_t wplane() const { return (int64_t)(nk + 1) * nj * ni; }
  // substep snapshot n of a cell/u/v array or of w
  __device__ __forceinline__ float* snap(int s, int n) const { return p[s] + (int64_t)n * plane(); }
  __device__ __forceinline__ float* wsnap(int n) const { return p[SW] + (int64_t)n * wplane(); }
};

Cols make_cols(const float* col, int nk) {
  Cols c;
  c.rho = col;
  c.rr0 = col + nk;
  c.rf0 = col + 2 * nk;
  c.rrf0 = col + 3 * nk + 1;
  c.c1 = col + 4 * nk + 2;
  c.c2 = col + 5 * nk + 3;
  c.th0 = col + 6 * nk + 4;
  c.pi0 = col + 7 * nk + 4;
  c.qv0 = col + 8 * nk + 4;
  c.rth = col + 9 * nk + 4;
  c.rds = col + 10 * nk + 4;
  c.rdsf = col + 11 * nk + 6;
  return c;
}

#define CELL_INDEX                                               \
  const int i = blockIdx.x * blockDim.x + threadIdx.x;           \
  const int j = blockIdx.y, l = blockIdx.z + 1;                  \
  if (i >= w.ni) return;

// ------------------------------------------------------------------ forward

__global__ __launch_bounds__(64) void k_terms(W w) {
  CELL_INDEX
  const Consts& c = w.c;
  const int64_t q = w.at(l, j, i);
  const float th = w.p[XTH][q], pp = w.p[XPP][q];
  const float theta = w.col.th0[l - 1] + th;
  float b = th * w.col.rth[l - 1] + c.repsm1 * (0.0f - w.col.qv0[l - 1]);
  b = b - ((0.0f - 0.0f) - 0.0f);
  w.p[THV][q] = (theta * (1.0f + c.reps * 0.0f)) / ((1.0f + 0.0f) + 0.0f);
  w.p[THTERM][q] = (theta * c.kterm) * c.cvm;
  const float pi = w.col.pi0[l - 1] + pp;
  w.p[PPTERM][q] = ((pi * c.rovcp) * c.cpm) * c.cvm;
  w.p[PIADV][q] = w.col.pi0[l - 1] + pp;
  w.p[DUM8][q] = c.g * b;
}

__global__ __launch_bounds__(64) void k_mass(W w) {
  CELL_INDEX
  const int64_t q = w.at(l, j, i);
  const float den = 0.5f * (w.col.rho[l - 1] + w.col.rho[l - 1]);
  w.p[RRU][q] = den * w.p[XU][q];
  w.p[RRV][q] = den * w.p[XV][q];
  w.p[RRW][q] = (l >= 2) ? w.col.rf0[l - 1] * w.p[XW][q] : 0.0f;
  if (l == w.nk) w.p[RRW][w.at(w.nk + 1, j, i)] = 0.0f;
}

__global__ __launch_bounds__(64) void k_divx(W w) {
  CELL_INDEX
  const Consts& c = w.c;
  const float* ru = w.p[RRU];
  const float* rv = w.p[RRV];
  const float* rw = w.p[RRW];
  float d = (ru[w.at(l, j, w.X(i, 1))] - ru[w.at(l, j, i)]) * c.rdx + (rv[w.at(l, w.Y(j, 1), i)] - rv[w.at(l, j, i)]) * c.rdy;
  d = d + (rw[w.at(l + 1, j, i)] - rw[w.at(l, j, i)]) * c.rdz;
  w.p[DIVX][w.at(l, j, i)] = fabsf(d) < 1.0e-20f ? 0.0f : d;
}

// Vertical stencil kinds of the u/v and w vertical fluxes and of the scalar z flux.
enum Kind { LIN = 0, HALF2 = 1, F3 = 2, F5 = 3, UPD = 4, WENO = 5 };
struct Stencil {
  Kind kind;
  int lev[5];
};

__device__ __forceinline__ void uvz_branch(int kf, int nk, Stencil& pos, Stencil& neg) {
  if (kf == 2) {
    pos = {LIN, {1, 2}};
    neg = {F3, {3, 2, 1}};
  } else if (kf == nk) {
    pos = {F3, {nk - 2, nk - 1, nk}};
    neg = {LIN, {nk - 1, nk}};
  } else if (kf == 3) {
    pos = {F3, {1, 2, 3}};
    neg = {F5, {5, 4, 3, 2, 1}};
  } else if (kf == nk - 1) {
    pos = {F5, {kf - 3, kf - 2, kf - 1, kf, kf + 1}};
    neg = {F3, {kf + 1, kf, kf - 1}};
  } else {
    pos = {F5, {kf - 3, kf - 2, kf - 1, kf, kf + 1}};
    neg = {F5, {kf + 2, kf + 1, kf, kf - 1, kf - 2}};
  }
}

__device__ __forceinline__ void wz_branch(int k, int nk, Stencil& pos, Stencil& neg) {
  if (k == 1) {
    pos = {HALF2, {1, 2}};
    neg = {F3, {3, 2, 1}};
  } else if (k == 2) {
    pos = {F3, {1, 2, 3}};
    neg = {F5, {5, 4, 3, 2, 1}};
  } else if (k == nk - 1) {
    pos = {F5, {k - 2, k - 1, k, k + 1, k + 2}};
    neg = {F3, {k + 2, k + 1, k}};
  } else if (k == nk) {
    pos = {F3, {k - 1, k, k + 1}};
    neg = {HALF2, {k, k + 1}};
  } else {
    pos = {F5, {k - 2, k - 1, k, k + 1, k + 2}};
    neg = {F5, {k + 3, k + 2, k + 1, k, k - 1}};
  }
}

// scalar z flux at face k: (pos, neg, strict '>' selection)
__device__ __forceinline__ bool zflux_branch(int k, int nk, bool weno, Stencil& pos, Stencil& neg) {
  const Kind up3 = weno ? UPD : F3;
  const Kind r5 = weno ? WENO : F5;
  if (k == 2) {
    pos = {LIN, {1, 2}};
    neg = {up3, {3, 2, 1}};
    return false;
  }
  if (k == nk) {
    pos = {up3, {nk - 2, nk - 1, nk}};
    neg = {LIN, {nk - 1, nk}};
    return weno;
  }
  if (k == 3) {
    pos = {up3, {1, 2, 3}};
    neg = {r5, {5, 4, 3, 2, 1}};
    return false;
  }
  if (k == nk - 1) {
    pos = {r5, {k - 3, k - 2, k - 1, k, k + 1}};
    neg = {up3, {k + 1, k, k - 1}};
    return false;
  }
  pos = {r5, {k - 3, k - 2, k - 1, k, k + 1}};
  neg = {r5, {k + 2, k + 1, k, k - 1, k - 2}};
  return false;
}

template <typename Q>
__device__ __forceinline__ float eval_branch(const Stencil& s, Q q, float c2h, float c1lin, double weps, float wf) {
  switch (s.kind) {
    case LIN: return c1lin * q(s.lev[0]) + c2h * q(s.lev[1]);
    case HALF2: return 0.5f * (q(s.lev[0]) + q(s.lev[1]));
    case F3: return flx3(q(s.lev[0]), q(s.lev[1]), q(s.lev[2]));
    case F5: return flx5(q(s.lev[0]), q(s.lev[1]), q(s.lev[2]), q(s.lev[3]), q(s.lev[4]));
    case UPD: return upstrpd(q(s.lev[0]), q(s.lev[1]), q(s.lev[2]), wf);
    default: return weno5(q(s.lev[0]), q(s.lev[1]), q(s.lev[2]), q(s.lev[3]), q(s.lev[4]), weps);
  }
}

// Coefficients of the stencil's values (frozen WENO5 weights and limiter) in branch order.
template <typename Q>
__device__ __forceinline__ int branch_coefs(const Stencil& s, Q q, float c2h, float c1lin, float wf, float* cf,
                                            bool exact = false) {
  switch (s.kind) {
    case LIN: cf[0] = c1lin; cf[1] = c2h; return 2;
    case HALF2: cf[0] = 0.5f; cf[1] = 0.5f; return 2;
    case F3: cf[0] = gm1::CM1_6; cf[1] = gm1::C5_6; cf[2] = gm1::C2_6; return 3;
    case F5:
      cf[0] = gm1::C2_60; cf[1] = gm1::CM13_60; cf[2] = gm1::C47_60; cf[3] = gm1::C27_60; cf[4] = gm1::CM3_60;
      return 5;
    case UPD: {
      const float s1 = q(s.lev[0]), s2 = q(s.lev[1]), s3 = q(s.lev[2]);
      const float dd = s2 - s1;
      const float se = copysignf(fabsf(wf), dd);
      const float rr = (s3 - s2) / (se + dd);
      const float cand = gm1::C1_6 + gm1::C2_6 * rr;
      const float phi = fmaxf(0.0f, fminf(rr, fminf(cand, 1.0f)));
      if (!exact) {
        cf[0] = -phi; cf[1] = 1.0f + phi; cf[2] = 0.0f;
        return 3;
      }
      // the limiter's active branch differentiated: phi = rr (scale 1), 1/6 + rr/3 (scale 1/3), 0 or 1 (0)
      const float inner = fminf(cand, 1.0f);
      const float mn = fminf(rr, inner);
      const float scale = !(mn > 0.0f) ? 0.0f : (rr <= inner ? 1.0f : (cand <= 1.0f ? gm1::C2_6 : 0.0f));
      const double den = (double)se + (double)dd;
      const double g1 = ((double)s3 - s2) / (den * den), g2 = -1.0 / den - ((double)s3 - s2) / (den * den);
      const double g3 = 1.0 / den;
      cf[0] = (float)(-phi + dd * (scale * g1));
      cf[1] = (float)(1.0f + phi + dd * (scale * g2));
      cf[2] = (float)(dd * (scale * g3));
      return 3;
    }
    default: return -1;
  }
}

// Frozen WENO5 coefficients of s1..s5 (weights as weno5 forms them).
__device__ __forceinline__ void weno_coefs(float s1, float s2, float s3, float s4, float s5, double weps, float* cf) {
  const float d1 = (s1 - 2.0f * s2) + s3;
  const float e1 = (s1 - 4.0f * s2) + 3.0f * s3;
  const float b1r = gm1::C13_12 * (d1 * d1) + 0.25f * (e1 * e1);
  const float d2 = (s2 - 2.0f * s3) + s4;
  const float e2 = s2 - s4;
  const float b2r = gm1::C13_12 * (d2 * d2) + 0.25f * (e2 * e2);
  const float d3 = (s3 - 2.0f * s4) + s5;
  const float e3 = (3.0f * s3 - 4.0f * s4) + s5;
  const float b3r = gm1::C13_12 * (d3 * d3) + 0.25f * (e3 * e3);
  const double b1 = b1r, b2 = b2r, b3 = b3r;
  const double tau = fabs(b1 - b3);
  const double q1 = fmin(1.0e30, tau / (b1 + weps));
  const double q2 = fmin(1.0e30, tau / (b2 + weps));
  const double q3 = fmin(1.0e30, tau / (b3 + weps));
  const double w1 = (double)0.1f * (1.0 + q1 * q1);
  const double w2 = (double)0.6f * (1.0 + q2 * q2);
  const double w3 = (double)0.3f * (1.0 + q3 * q3);
  const double den = (w1 + w2) + w3;
  const double o1 = w1 / den, o2 = w2 / den, o3 = w3 / den;
  cf[0] = (float)(o1 * (double)gm1::C2_6);
  cf[1] = (float)(o1 * (double)gm1::CM7_6 + o2 * (double)gm1::CM1_6);
  cf[2] = (float)((o1 * (double)gm1::C11_6 + o2 * (double)gm1::C5_6) + o3 * (double)gm1::C2_6);
  cf[3] = (float)(o2 * (double)gm1::C2_6 + o3 * (double)gm1::C5_6);
  cf[4] = (float)(o3 * (double)gm1::CM1_6);
}

// d(WENO5 reconstruction)/d(s1..s5) with the WENO-Z weights differentiated (binary64; zero where q is clamped).
__device__ __forceinline__ void weno_jacobian(float f1, float f2, float f3, float f4, float f5, double weps,
                                              float* cf) {
  const double s1 = f1, s2 = f2, s3 = f3, s4 = f4, s5 = f5;
  const double c13 = (double)gm1::C13_12;
  const double d1 = (s1 - 2.0 * s2) + s3, e1 = (s1 - 4.0 * s2) + 3.0 * s3;
  const double d2 = (s2 - 2.0 * s3) + s4, e2 = s2 - s4;
  const double d3 = (s3 - 2.0 * s4) + s5, e3 = (3.0 * s3 - 4.0 * s4) + s5;
  const double b[3] = {c13 * (d1 * d1) + 0.25 * (e1 * e1), c13 * (d2 * d2) + 0.25 * (e2 * e2),
                       c13 * (d3 * d3) + 0.25 * (e3 * e3)};
  const double db[3][5] = {
      {2 * c13 * d1 + 0.5 * e1, -4 * c13 * d1 - 2.0 * e1, 2 * c13 * d1 + 1.5 * e1, 0.0, 0.0},
      {0.0, 2 * c13 * d2 + 0.5 * e2, -4 * c13 * d2, 2 * c13 * d2 - 0.5 * e2, 0.0},
      {0.0, 0.0, 2 * c13 * d3 + 1.5 * e3, -4 * c13 * d3 - 2.0 * e3, 2 * c13 * d3 + 0.5 * e3}};
  const double diff = b[0] - b[2];
  const double tau = fabs(diff);
  const double sg = diff > 0.0 ? 1.0 : (diff < 0.0 ? -1.0 : 0.0);
  const double a[3] = {(double)0.1f, (double)0.6f, (double)0.3f};
  double wv[3], dw[3][5];
  for (int m = 0; m < 3; ++m) {
    const double den = b[m] + weps;
    double q = tau / den;
    const bool live = q < 1.0e30;
    q = fmin(1.0e30, q);
    wv[m] = a[m] * (1.0 + q * q);
    for (int n = 0; n < 5; ++n) {
      const double dtau = sg * (db[0][n] - db[2][n]);
      const double dq = live ? dtau / den - tau * db[m][n] / (den * den) : 0.0;
      dw[m][n] = 2.0 * a[m] * q * dq;
    }
  }
  const double W = (wv[0] + wv[1]) + wv[2];
  const double om[3] = {wv[0] / W, wv[1] / W, wv[2] / W};
  const double p[3] = {((double)gm1::C2_6 * s1 + (double)gm1::CM7_6 * s2) + (double)gm1::C11_6 * s3,
                       ((double)gm1::CM1_6 * s2 + (double)gm1::C5_6 * s3) + (double)gm1::C2_6 * s4,
                       ((double)gm1::C2_6 * s3 + (double)gm1::C5_6 * s4) + (double)gm1::CM1_6 * s5};
  const double dp[3][5] = {{gm1::C2_6, gm1::CM7_6, gm1::C11_6, 0.0, 0.0},
                           {0.0, gm1::CM1_6, gm1::C5_6, gm1::C2_6, 0.0},
                           {0.0, 0.0, gm1::C2_6, gm1::C5_6, gm1::CM1_6}};
  for (int n = 0; n < 5; ++n) {
    const double dW = (dw[0][n] + dw[1][n]) + dw[2][n];
    double g = 0.0;
    for (int m = 0; m < 3; ++m) g += om[m] * dp[m][n] + p[m] * ((dw[m][n] - om[m] * dW) / W);
    cf[n] = (float)g;
  }
}

template <typename Q>
__device__ __forceinline__ int stencil_coefs(const Stencil& s, Q q, float c2h, float c1lin, double weps, float wf,
                                             float* cf, bool exact) {
  if (s.kind == WENO) {
    if (exact) weno_jacobian(q(s.lev[0]), q(s.lev[1]), q(s.lev[2]), q(s.lev[3]), q(s.lev[4]), weps, cf);
    else weno_coefs(q(s.lev[0]), q(s.lev[1]), q(s.lev[2]), q(s.lev[3]), q(s.lev[4]), weps, cf);
    return 5;
  }
  return branch_coefs(s, q, c2h, c1lin, wf, cf, exact);
}

// u tendency at (l, j, face i)
__global__ __launch_bounds__(64) void k_advu(W w) {
  CELL_INDEX
  const Consts& c = w.c;
  const float *u = w.p[XU], *ru = w.p[RRU], *rv = w.p[RRV], *rw = w.p[RRW], *dx = w.p[DIVX];
  auto U = [&](int ll, int jj, int ii) { return u[w.at(ll, w.Y(jj, 0), w.X(ii, 0))]; };
  auto fx = [&](int cc) {
    const float m = 0.5f * (ru[w.at(l, j, w.X(cc, 0))] + ru[w.at(l, j, w.X(cc, 1))]);
    const float pos = flx5(U(l, j, cc - 2), U(l, j, cc - 1), U(l, j, cc), U(l, j, cc + 1), U(l, j, cc + 2));
    const float neg = flx5(U(l, j, cc + 3), U(l, j, cc + 2), U(l, j, cc + 1), U(l, j, cc), U(l, j, cc - 1));
    return m * (m >= 0.0f ? pos : neg);
  };
  auto fy = [&](int pp) {
    const float m = 0.5f * (rv[w.at(l, w.Y(pp, 0), i)] + rv[w.at(l, w.Y(pp, 0), w.X(i, -1))]);
    const float pos = flx5(U(l, pp - 3, i), U(l, pp - 2, i), U(l, pp - 1, i), U(l, pp, i), U(l, pp + 1, i));
    const float neg = flx5(U(l, pp + 2, i), U(l, pp + 1, i), U(l, pp, i), U(l, pp - 1, i), U(l, pp - 2, i));
    return m * (m >= 0.0f ? pos : neg);
  };
  auto fz = [&](int kf) {
    const float pair = rw[w.at(kf, j, i)] + rw[w.at(kf, j, w.X(i, -1))];
    const float m = 0.5f * pair;
    const float c2h = 0.5f * (w.col.c2[kf - 1] + w.col.c2[kf - 1]);
    Stencil ps, ng;
    uvz_branch(kf, w.nk, ps, ng);
    auto q = [&](int ll) { return U(ll, j, i); };
    const float r = m >= 0.0f ? eval_branch(ps, q, c2h, 1.0f - c2h, 0.0, 0.0f)
                              : eval_branch(ng, q, c2h, 1.0f - c2h, 0.0, 0.0f);
    return m * r;
  };
  const float advx = (-(fx(i) - fx(i - 1))) * c.rdx;
  const float advy = (-(fy(j + 1) - fy(j))) * c.rdy;
  float advz;
  if (l == 1) advz = (-fz(2)) * c.rdz;
  else if (l == w.nk) advz = fz(w.nk) * c.rdz;
  else advz = (-(fz(l + 1) - fz(l))) * c.rdz;
  float total = (advx + advy) + advz;
  total = total + U(l, j, i) * (0.5f * (dx[w.at(l, j, i)] + dx[w.at(l, j, w.X(i, -1))]));
  const float rho = w.col.rho[l - 1];
  w.p[UTEN][w.at(l, j, i)] = total * (1.0f / (0.5f * (rho + rho)));
}

// v tendency at (l, face j, i)
__global__ __launch_bounds__(64) void k_advv(W w) {
  CELL_INDEX
  const Consts& c = w.c;
  const float *v = w.p[XV], *ru = w.p[RRU], *rv = w.p[RRV], *rw = w.p[RRW], *dx = w.p[DIVX];
  auto V = [&](int ll, int jj, int ii) { return v[w.at(ll, w.Y(jj, 0), w.X(ii, 0))]; };
  auto fx = [&](int pp) {
    const float m = 0.5f * (ru[w.at(l, j, w.X(pp, 0))] + ru[w.at(l, w.Y(j, -1), w.X(pp, 0))]);
    const float pos = flx5(V(l, j, pp - 3), V(l, j, pp - 2), V(l, j, pp - 1), V(l, j, pp), V(l, j, pp + 1));
    const float neg = flx5(V(l, j, pp + 2), V(l, j, pp + 1), V(l, j, pp), V(l, j, pp - 1), V(l, j, pp - 2));
    return m * (m >= 0.0f ? pos : neg);
  };
  auto fy = [&](int cc) {
    const float m = 0.5f * (rv[w.at(l, w.Y(cc, 0), i)] + rv[w.at(l, w.Y(cc, 1), i)]);
    const float pos = flx5(V(l, cc - 2, i), V(l, cc - 1, i), V(l, cc, i), V(l, cc + 1, i), V(l, cc + 2, i));
    const float neg = flx5(V(l, cc + 3, i), V(l, cc + 2, i), V(l, cc + 1, i), V(l, cc, i), V(l, cc - 1, i));
    return m * (m >= 0.0f ? pos : neg);
  };
  auto fz = [&](int kf) {
    const float pair = rw[w.at(kf, j, i)] + rw[w.at(kf, w.Y(j, -1), i)];
    const float m = 0.5f * pair;
    const float c2h = 0.5f * (w.col.c2[kf - 1] + w.col.c2[kf - 1]);
    Stencil ps, ng;
    uvz_branch(kf, w.nk, ps, ng);
    auto q = [&](int ll) { return V(ll, j, i); };
    const float r = m >= 0.0f ? eval_branch(ps, q, c2h, 1.0f - c2h, 0.0, 0.0f)
                              : eval_branch(ng, q, c2h, 1.0f - c2h, 0.0, 0.0f);
    return m * r;
  };
  const float advx = (-(fx(i + 1) - fx(i))) * c.rdx;
  const float advy = (-(fy(j) - fy(j - 1))) * c.rdy;
  float advz;
  if (l == 1) advz = (-fz(2)) * c.rdz;
  else if (l == w.nk) advz = fz(w.nk) * c.rdz;
  else advz = (-(fz(l + 1) - fz(l))) * c.rdz;
  float total = (advx + advy) + advz;
  total = total + V(l, j, i) * (0.5f * (dx[w.at(l, j, i)] + dx[w.at(l, w.Y(j, -1), i)]));
  const float rho = w.col.rho[l - 1];
  w.p[VTEN][w.at(l, j, i)] = total * (1.0f / (0.5f * (rho + rho)));
}

// w tendency (with buoyancy) at levels 2..nk; levels 1 and nk+1 are zero
__global__ __launch_bounds__(64) void k_advw(W w) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y, kf = blockIdx.z + 1;
  if (i >= w.ni) return;
  const Consts& c = w.c;
  const int nk = w.nk;
  if (kf == 1 || kf == nk + 1) {
    w.p[WTEN][w.at(kf, j, i)] = 0.0f;
    return;
  }
  const float *ww = w.p[XW], *ru = w.p[RRU], *rv = w.p[RRV], *rw = w.p[RRW], *dx = w.p[DIVX];
  auto Wv = [&](int ll, int jj, int ii) { return ww[w.at(ll, w.Y(jj, 0), w.X(ii, 0))]; };
  const float c2h = 0.5f * (w.col.c2[kf - 1] + w.col.c2[kf - 1]);
  const float c1h = 1.0f - c2h;
  auto fx = [&](int pp) {
    const float m = c2h * ru[w.at(kf, j, w.X(pp, 0))] + c1h * ru[w.at(kf - 1, j, w.X(pp, 0))];
    const float pos = flx5(Wv(kf, j, pp - 3), Wv(kf, j, pp - 2), Wv(kf, j, pp - 1), Wv(kf, j, pp), Wv(kf, j, pp + 1));
    const float neg = flx5(Wv(kf, j, pp + 2), Wv(kf, j, pp + 1), Wv(kf, j, pp), Wv(kf, j, pp - 1), Wv(kf, j, pp - 2));
    return m * (m >= 0.0f ? pos : neg);
  };
  auto fy = [&](int pp) {
    const float m = c2h * rv[w.at(kf, w.Y(pp, 0), i)] + c1h * rv[w.at(kf - 1, w.Y(pp, 0), i)];
    const float pos = flx5(Wv(kf, pp - 3, i), Wv(kf, pp - 2, i), Wv(kf, pp - 1, i), Wv(kf, pp, i), Wv(kf, pp + 1, i));
    const float neg = flx5(Wv(kf, pp + 2, i), Wv(kf, pp + 1, i), Wv(kf, pp, i), Wv(kf, pp - 1, i), Wv(kf, pp - 2, i));
    return m * (m >= 0.0f ? pos : neg);
  };
  auto wz = [&](int k) {
    const float m = 0.5f * (rw[w.at(k, j, i)] + rw[w.at(k + 1, j, i)]);
    Stencil ps, ng;
    wz_branch(k, nk, ps, ng);
    auto q = [&](int ll) { return Wv(ll, j, i); };
    const float r = m >= 0.0f ? eval_branch(ps, q, 0.0f, 0.0f, 0.0, 0.0f) : eval_branch(ng, q, 0.0f, 0.0f, 0.0, 0.0f);
    return m * r;
  };
  const float advx = (-(fx(i + 1) - fx(i))) * (c.rdx * 1.0f);
  const float advy = (-(fy(j + 1) - fy(j))) * (c.rdy * 1.0f);
  const float advz = (-(wz(kf) - wz(kf - 1))) * c.rdz;
  const float cc2 = w.col.c2[kf - 1];
  const float cc1 = 1.0f - cc2;
  float total = (advx + advy) + advz;
  total = total + Wv(kf, j, i) * (cc2 * dx[w.at(kf, j, i)] + cc1 * dx[w.at(kf - 1, j, i)]);
  const float adv = total * (w.col.rrf0[kf - 1] * 1.0f);
  const float* d8 = w.p[DUM8];
  const float add = w.col.c1[kf - 1] * d8[w.at(kf - 1, j, i)] + w.col.c2[kf - 1] * d8[w.at(kf, j, i)];
  w.p[WTEN][w.at(kf, j, i)] = adv + add;
}

// ------------------------------------------------------------------ acoustic stage (forward)

__global__ __launch_bounds__(64) void k_coef(W w) {
  CELL_INDEX
  const Consts& c = w.c;
  const int nk = w.nk;
  const int64_t q = w.at(l, j, i);
  const float tem = c.dts * w.col.rdsf[l];
  w.p[QK][q] = (tem * w.p[PPTERM][q]) * 1.0f;
  const float tem2 = ((c.dts * 1.0f) * w.col.rdsf[l]) * w.col.rr0[l - 1];
  const float* pia = w.p[PIADV];
  float pk2 = 0.0f, pk1 = 0.0f, mm = 0.0f;
  if (l < nk) {
    const float b2 = (-w.col.c2[l]) * pia[w.at(l + 1, j, i)] + (1.0f - w.col.c1[l]) * pia[q];
    pk2 = (tem2 * w.col.rf0[l]) * b2;
  }
  if (l > 1) {
    const float b1 = w.col.c1[l - 1] * pia[w.at(l - 1, j, i)] + (w.col.c2[l - 1] - 1.0f) * pia[q];
    pk1 = (tem2 * w.col.rf0[l - 1]) * b1;
    const float tem1 = (c.dts * w.col.rds[l]) * c.cp;
    const float th = w.col.c2[l - 1] * w.p[THV][q] + w.col.c1[l - 1] * w.p[THV][w.at(l - 1, j, i)];
    mm = (tem1 * th) * 1.0f;
  }
  w.p[PK1][q] = pk1;
  w.p[PK2][q] = pk2;
  w.p[MM][q] = mm;
}

// rows k = 2..nk of the stage's tridiagonal systems (they depend on the coefficients only)
__global__ __launch_bounds__(64) void k_rows(W w);

__global__ __launch_bounds__(64) void k_ac_start(W w) {
  CELL_INDEX
  const int64_t q = w.at(l, j, i);
  const bool first = w.nrk == 1;
  w.snap(SPPD, 0)[q] = w.p[PPX][q];
  w.snap(SPP, 0)[q] = first ? w.p[XPP][q] : w.p[API][q];
  w.snap(SU, 0)[q] = first ? w.p[XU][q] : w.p[AU][q];
  w.snap(SV, 0)[q] = first ? w.p[XV][q] : w.p[AV][q];
  w.wsnap(0)[q] = first ? w.p[XW][q] : w.p[AW][q];
  w.p[ARU][q] = 0.0f;
  w.p[ARV][q] = 0.0f;
  w.p[ARW][q] = w.p[AW][q] * w.c.beta;
  if (l == w.nk) {
    const int64_t top = w.at(w.nk + 1, j, i);
    w.wsnap(0)[top] = w.p[XW][top];
    w.p[ARW][top] = 0.0f;
  }
}

// substep n: pressure gradient on u and v faces, running sums of the winds
__global__ __launch_bounds__(64) void k_ss_pgrad(W w, int n) {
  CELL_INDEX
  const Consts& c = w.c;
  const bool last = n == w.nloop;
  const float* thv = w.p[THV];
  const float* ppd = w.snap(SPPD, n - 1);
  const int64_t q = w.at(l, j, i);
  {
    const int64_t qw = w.at(l, j, w.X(i, -1));
    const float pg = (c.halfcp * (thv[q] + thv[qw])) * ((ppd[q] - ppd[qw]) * c.rdx);
    const float u1 = w.snap(SU, n - 1)[q] + c.dts * (w.p[UTEN][q] - pg);
    w.snap(SU, n)[q] = u1;
    w.p[ARU][q] = last ? (w.p[ARU][q] + u1) * c.tavg : w.p[ARU][q] + u1;
  }
  {
    const int64_t qs = w.at(l, w.Y(j, -1), i);
    const float pg = (c.halfcp * (thv[q] + thv[qs])) * ((ppd[q] - ppd[qs]) * c.rdy);
    const float v1 = w.snap(SV, n - 1)[q] + c.dts * (w.p[VTEN][q] - pg);
    w.snap(SV, n)[q] = v1;
    w.p[ARV][q] = last ? (w.p[ARV][q] + v1) * c.tavg : w.p[ARV][q] + v1;
  }
}

// fpk of one column at level l from the substep's u1, v1 and starting w (sound.F's fpk and pressure advection)
__device__ __forceinline__ float column_fpk(const W& w, int n, int l, int j, int i, float* div_out) {
  const Consts& c = w.c;
  const int nk = w.nk;
  const float* u1 = w.snap(SU, n);
  const float* v1 = w.snap(SV, n);
  const float* ws = w.wsnap(n - 1);
  const float* pia = w.p[PIADV];
  const int64_t q = w.at(l, j, i);
  const int64_t qe = w.at(l, j, w.X(i, 1)), qn = w.at(l, w.Y(j, 1), i);
  const float hz = (u1[qe] - u1[q]) * c.rdx + (v1[qn] - v1[q]) * c.rdy;
  float div = 1.0f * (hz + 0.0f);
  div = fabsf(div) < 1.0e-20f ? 0.0f : div;
  if (div_out) *div_out = div;
  const float wlo = l == 1 ? 0.0f : ws[q];
  const float whi = l == nk ? 0.0f : ws[w.at(l + 1, j, i)];
  const float src = c.dts * (0.0f - w.p[PPTERM][q] * div);
  const float qp = (-w.p[QK][q]) * (whi - wlo);
  float fpk = src + c.beta * qp;
  const float rho = w.col.rho[l - 1];
  const float uu2 = c.temx * (rho * u1[qe]), uu1 = c.temx * (rho * u1[q]);
  const float vv2 = c.temy * (rho * v1[qn]), vv1 = c.temy * (rho * v1[q]);
  const float ww2 = 0.0f + c.beta * ws[w.at(l + 1, j, i)];
  const float ww1 = 0.0f + c.beta * wlo;
  const float pc = pia[q];
  const float xa = uu2 * (pia[qe] - pc) + uu1 * (pc - pia[w.at(l, j, w.X(i, -1))]);
  const float ya = vv2 * (pia[qn] - pc) + vv1 * (pc - pia[w.at(l, w.Y(j, -1), i)]);
  float hz2 = (-(xa + ya)) * w.col.rr0[l - 1];
  hz2 = hz2 * 1.0f;
  const float vt = ww1 * w.p[PK1][q] + ww2 * w.p[PK2][q];
  return fpk + (hz2 + vt);
}

// Thomas rows of level k (2..nk) of one column
struct Row {
  float aa, bb, cc;
};
__device__ __forceinline__ Row thomas_row(const W& w, int k, int j, int i) {
  const int nk = w.nk;
  const float asq = w.c.asq;
  const float *qk = w.p[QK], *pk1 = w.p[PK1], *pk2 = w.p[PK2], *mm = w.p[MM];
  const int64_t q = w.at(k, j, i), qd = w.at(k - 1, j, i);
  Row r;
  r.aa = k >= 3 ? (-asq) * (mm[q] * (pk1[qd] + qk[qd])) : 0.0f;
  r.cc = k < nk ? (-asq) * ((-mm[q]) * (pk2[q] - qk[q])) : 0.0f;
  const float diag = (-(pk1[q] + qk[q])) + (pk2[qd] - qk[qd]);
  r.bb = 1.0f - asq * (mm[q] * diag);
  return r;
}

// substep n, every cell: fpk
__global__ __launch_bounds__(64) void k_rows(W w) {
  CELL_INDEX
  if (l < 2) return;
  const Row r = thomas_row(w, l, j, i);
  const int64_t q = w.at(l, j, i);
  w.p[RA][q] = r.aa;
  w.p[RB][q] = r.bb;
  w.p[RC][q] = r.cc;
}

// substep n, levels 2..nk: the right side of the tridiagonal system (fpk of levels l and l-1 from FPK)
__global__ __launch_bounds__(64) void k_ss_rhs(W w, int n) {
  CELL_INDEX
  if (l < 2) return;
  const Consts& c = w.c;
  const float* ws = w.wsnap(n - 1);
  const float* pp = w.snap(SPP, n - 1);
  const float* fpk = w.p[FPK];
  const int64_t q = w.at(l, j, i), qd = w.at(l - 1, j, i);
  const float mmk = w.p[MM][q];
  const float pdelta = pp[q] - pp[qd];
  const float fwk = c.dts * w.p[WTEN][q] + c.beta * ((-mmk) * pdelta);
  const float udelta = (pp[q] + fpk[q]) - (pp[qd] + fpk[qd]);
  w.p[RD][q] = (ws[q] + fwk) + c.alph * ((-mmk) * udelta);
}

__global__ __launch_bounds__(64) void k_ss_fpk(W w, int n) {
  CELL_INDEX
  w.p[FPK][w.at(l, j, i)] = column_fpk(w, n, l, j, i, nullptr);
}

// substep n, one column: the Thomas elimination of rows 2..nk and back substitution (w at levels 2..nk).
// Rows are fetched TH_CHUNK levels at a time into registers (independent loads in flight together), then the
// recurrence runs over them.
constexpr int TH_CHUNK = 8;

__global__ __launch_bounds__(64) void k_ss_thomas(W w, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
  if (i >= w.ni) return;
  const int nk = w.nk;
  const float* __restrict__ ws = w.wsnap(n - 1);
  float* __restrict__ wn = w.wsnap(n);
  const float* __restrict__ ra = w.p[RA];
  const float* __restrict__ rb = w.p[RB];
  const float* __restrict__ rc = w.p[RC];
  const float* __restrict__ rd = w.p[RD];
  float* __restrict__ ec = w.p[EC];
  float* __restrict__ ed = w.p[ED];
  float ecd = 0.0f, edd = 0.0f;
  for (int k0 = 2; k0 <= nk; k0 += TH_CHUNK) {
    float a[TH_CHUNK], bb[TH_CHUNK], c[TH_CHUNK], d[TH_CHUNK];
#pragma unroll
    for (int u = 0; u < TH_CHUNK; ++u) {
      const int k = k0 + u;
      const int64_t q = w.at(k <= nk ? k : nk, j, i);
      a[u] = ra[q];
      bb[u] = rb[q];
      c[u] = rc[q];
      d[u] = rd[q];
    }
#pragma unroll
    for (int u = 0; u < TH_CHUNK; ++u) {
      const int k = k0 + u;
      if (k > nk) break;
      const int64_t q = w.at(k, j, i);
      const float pivot = k == 2 ? bb[u] : a[u] * ecd + bb[u];
      const float inv = 1.0f / pivot;
      if (k < nk) {
        ecd = (-c[u]) * inv;
        edd = (k == 2 ? d[u] : d[u] - a[u] * edd) * inv;
        ec[q] = ecd;
        ed[q] = edd;
      } else {
        wn[q] = (d[u] - a[u] * edd) * inv;
        edd = wn[q];
      }
    }
  }
  float wnext = edd;  // w(nk)
  for (int k1 = nk - 1; k1 >= 2; k1 -= TH_CHUNK) {
    float e[TH_CHUNK], f[TH_CHUNK];
#pragma unroll
    for (int u = 0; u < TH_CHUNK; ++u) {
      const int k = k1 - u;
      const int64_t q = w.at(k >= 2 ? k : 2, j, i);
      e[u] = ec[q];
      f[u] = ed[q];
    }
#pragma unroll
    for (int u = 0; u < TH_CHUNK; ++u) {
      const int k = k1 - u;
      if (k < 2) break;
      wnext = e[u] * wnext + f[u];
      wn[w.at(k, j, i)] = wnext;
    }
  }
  wn[w.at(1, j, i)] = 0.0f;
  wn[w.at(nk + 1, j, i)] = ws[w.at(nk + 1, j, i)];
}

// substep n, every cell: pressure update, filter, ppx, w running sum
__global__ __launch_bounds__(64) void k_ss_update(W w, int n) {
  CELL_INDEX
  const Consts& c = w.c;
  const int nk = w.nk;
  const bool last = n == w.nloop;
  const float* wn = w.wsnap(n);
  const float* pp = w.snap(SPP, n - 1);
  const int64_t q = w.at(l, j, i);
  const float wlo = l == 1 ? 0.0f : wn[q];
  const float whi = l == nk ? 0.0f : wn[w.at(l + 1, j, i)];
  const float vt = wlo * (w.p[PK1][q] + w.p[QK][q]) + whi * (w.p[PK2][q] - w.p[QK][q]);
  float up = (pp[q] + w.p[FPK][q]) + c.alph * vt;
  up = fabsf(up) < 1.0e-20f ? 0.0f : up;
  const float tend = c.kdiv * (up - pp[q]);
  w.snap(SPP, n)[q] = up;
  w.snap(SPPD, n)[q] = up + tend;
  if (w.nrk == 3 && last) w.p[PPXO][q] = tend;
  float* arw = w.p[ARW];
  arw[q] = last ? (arw[q] + wn[q] * c.alph) * c.tavg : arw[q] + wn[q];
}

// ------------------------------------------------------------------ post-acoustic and theta (forward)

__device__ __forceinline__ float mw_at(const W& w, int l, int j, int i) {
  return (l >= 2 && l <= w.nk) ? w.col.rf0[l - 1] * w.p[ARW][w.at(l, j, i)] + 0.0f : 0.0f;
}
__device__ __forceinline__ float mu_at(const W& w, int l, int j, int i) {
  const float rho = w.col.rho[l - 1];
  return ((0.5f * (rho + rho)) * w.p[ARU][w.at(l, j, i)]) * 1.0f;
}
__device__ __forceinline__ float mv_at(const W& w, int l, int j, int i) {
  const float rho = w.col.rho[l - 1];
  return ((0.5f * (rho + rho)) * w.p[ARV][w.at(l, j, i)]) * 1.0f;
}

__global__ __launch_bounds__(64) void k_post(W w) {
  CELL_INDEX
  const Consts& c = w.c;
  const int nk = w.nk;
  const int64_t q = w.at(l, j, i);
  const int ie = w.X(i, 1), jn = w.Y(j, 1);
  const float *aru = w.p[ARU], *arv = w.p[ARV], *arw = w.p[ARW];
  // getdiv
  auto wt = [&](int ll) { return (ll >= 2 && ll <= nk) ? arw[w.at(ll, j, i)] + 0.0f : 0.0f; };
  float d = (aru[w.at(l, j, ie)] * 1.0f - aru[q] * 1.0f) * c.rdx + (arv[w.at(l, jn, i)] * 1.0f - arv[q] * 1.0f) * c.rdy;
  d = d + (wt(l + 1) - wt(l)) * c.rdz;
  d = fabsf(d) < 1.0e-20f ? 0.0f : d;
  w.p[DIV][q] = d;
  w.p[TDIV][q] = d * w.p[THTERM][q];
  // getdivx
  float dx = (mu_at(w, l, j, ie) - mu_at(w, l, j, i)) * c.rdx + (mv_at(w, l, jn, i) - mv_at(w, l, j, i)) * c.rdy;
  dx = dx + (mw_at(w, l + 1, j, i) - mw_at(w, l, j, i)) * c.rdz;
  w.p[DIVX2][q] = fabsf(dx) < 1.0e-20f ? 0.0f : dx;
}

__device__ __forceinline__ float sc_at(const W& w, int l, int j, int i) {
  return (w.col.th0[l - 1] - 300.0f) + w.p[XTH][w.at(l, w.Y(j, 0), w.X(i, 0))];
}

__global__ __launch_bounds__(64) void k_theta(W w) {
  CELL_INDEX
  const Consts& c = w.c;
  const int nk = w.nk;
  const bool weno = w.nrk == 3;
  const double weps = w.weps;
  auto S = [&](int ll, int jj, int ii) { return sc_at(w, ll, jj, ii); };
  auto recon = [&](float a, float b, float cc, float d, float e) {
    return weno ? weno5(a, b, cc, d, e, weps) : flx5(a, b, cc, d, e);
  };
  auto fx = [&](int f) {
    const float m = mu_at(w, l, j, w.X(f, 0));
    const float pos = recon(S(l, j, f - 3), S(l, j, f - 2), S(l, j, f - 1), S(l, j, f), S(l, j, f + 1));
    const float neg = recon(S(l, j, f + 2), S(l, j, f + 1), S(l, j, f), S(l, j, f - 1), S(l, j, f - 2));
    return m * (m >= 0.0f ? pos : neg);
  };
  auto fy = [&](int g) {
    const float m = mv_at(w, l, w.Y(g, 0), i);
    const float pos = recon(S(l, g - 3, i), S(l, g - 2, i), S(l, g - 1, i), S(l, g, i), S(l, g + 1, i));
    const float neg = recon(S(l, g + 2, i), S(l, g + 1, i), S(l, g, i), S(l, g - 1, i), S(l, g - 2, i));
    return m * (m >= 0.0f ? pos : neg);
  };
  auto fz = [&](int k) {
    const float m = mw_at(w, k, j, i);
    Stencil ps, ng;
    const bool strict = zflux_branch(k, nk, weno, ps, ng);
    auto q = [&](int ll) { return S(ll, j, i); };
    const float c1 = w.col.c1[k - 1], c2 = w.col.c2[k - 1];
    const bool up = strict ? (m > 0.0f) : (m >= 0.0f);
    const float r = up ? eval_branch(ps, q, c2, c1, weps, c.weps_f) : eval_branch(ng, q, c2, c1, weps, c.weps_f);
    return m * r;
  };
  const float ax = (fx(i) - fx(i + 1)) * c.rdx;
  const float ay = (fy(j) - fy(j + 1)) * c.rdy;
  float az;
  if (l == 1) az = (-fz(2)) * c.rdz;
  else if (l == nk) az = fz(nk) * c.rdz;
  else az = (fz(l) - fz(l + 1)) * c.rdz;
  const int64_t q = w.at(l, j, i);
  float total = (ax + ay) + az;
  total = total + S(l, j, i) * w.p[DIVX2][q];
  const float adv = (total * w.col.rr0[l - 1]) * 1.0f;
  float out = w.p[ATH][q] + c.dttmp * ((0.0f - w.p[TDIV][q]) + adv);
  w.p[THO][q] = fabsf(out) < 1.0e-30f ? 0.0f : out;
}

__global__ __launch_bounds__(64) void k_solve3(W w) {
  CELL_INDEX
  const int64_t q = w.at(l, j, i);
  w.p[PPXO][q] = w.snap(SPP, w.nloop)[q] + w.p[PPXO][q];
}

// ------------------------------------------------------------------ adjoint

__device__ __forceinline__ void aadd(float* p, float v) { atomicAdd(p, v); }

struct FillList {
  float* p[24];
  int64_t n[24];
  int count;
};
// zero every listed array (blockIdx.y selects the array)
__global__ __launch_bounds__(256) void k_b_fill(FillList f) {
  const int a = blockIdx.y;
  if (a >= f.count) return;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < f.n[a]; t += (int64_t)gridDim.x * blockDim.x)
    f.p[a][t] = 0.0f;
}

// theta compose and the cell parts of the scalar tendency
__global__ __launch_bounds__(64) void k_b_theta_cell(W w) {
  CELL_INDEX
  const Consts& c = w.c;
  const int64_t q = w.at(l, j, i);
  const float thb = w.p[B_OTH][q];
  w.p[B_ATH][q] += thb;
  const float tdb = -c.dttmp * thb;
  const float advb = c.dttmp * thb;
  const float totb = advb * w.col.rr0[l - 1];
  w.p[B_TOT][q] = totb;
  w.p[B_DIV][q] = tdb * w.p[THTERM][q];
  w.p[B_THTERM][q] = tdb * w.p[DIV][q];
  w.p[B_SC][q] = totb * w.p[DIVX2][q];
  w.p[B_DIVX2][q] = totb * sc_at(w, l, j, i);
}

// scalar x (dir 0) and y (dir 1) face fluxes: face index i (dir 0) or j (dir 1)
__global__ __launch_bounds__(64) void k_b_theta_xy(W w, int dir) {
  CELL_INDEX
  const Consts& c = w.c;
  const bool weno = w.nrk == 3;
  const float* tot = w.p[B_TOT];
  float fb, m;
  if (dir == 0) {
    fb = tot[w.at(l, j, i)] * c.rdx - tot[w.at(l, j, w.X(i, -1))] * c.rdx;
    m = mu_at(w, l, j, i);
  } else {
    fb = tot[w.at(l, j, i)] * c.rdy - tot[w.at(l, w.Y(j, -1), i)] * c.rdy;
    m = mv_at(w, l, j, i);
  }
  auto S = [&](int o) { return dir == 0 ? sc_at(w, l, j, i + o) : sc_at(w, l, j + o, i); };
  const bool pos = m >= 0.0f;
  const int o0 = pos ? -3 : 2, step = pos ? 1 : -1;
  float s[5], cf[5];
  for (int n = 0; n < 5; ++n) s[n] = S(o0 + step * n);
  float r;
  if (weno) {
    r = weno5(s[0], s[1], s[2], s[3], s[4], w.weps);
    if (c.exact != 0.0f) weno_jacobian(s[0], s[1], s[2], s[3], s[4], w.weps, cf);
    else weno_coefs(s[0], s[1], s[2], s[3], s[4], w.weps, cf);
  } else {
    r = flx5(s[0], s[1], s[2], s[3], s[4]);
    cf[0] = gm1::C2_60; cf[1] = gm1::CM13_60; cf[2] = gm1::C47_60; cf[3] = gm1::C27_60; cf[4] = gm1::CM3_60;
  }
  w.p[dir == 0 ? B_MU : B_MV][w.at(l, j, i)] = fb * r;
  const float rb = fb * m;
  for (int n = 0; n < 5; ++n) {
    const int o = o0 + step * n;
    const int64_t t = dir == 0 ? w.at(l, j, w.X(i, o)) : w.at(l, w.Y(j, o), i);
    aadd(&w.p[B_SC][t], cf[n] * rb);
  }
}

// scalar z face fluxes at face levels k = 2..nk (grid level l = k)
__global__ __launch_bounds__(64) void k_b_theta_z(W w) {
  CELL_INDEX
  const Consts& c = w.c;
  const int nk = w.nk, k = l;
  if (k < 2) {
    w.p[B_MW][w.at(1, j, i)] = 0.0f;
    w.p[B_MW][w.at(nk + 1, j, i)] = 0.0f;
    return;
  }
  const bool weno = w.nrk == 3;
  const float* tot = w.p[B_TOT];
  const float fb = tot[w.at(k, j, i)] * c.rdz - tot[w.at(k - 1, j, i)] * c.rdz;
  const float m = mw_at(w, k, j, i);
  Stencil ps, ng;
  const bool strict = zflux_branch(k, nk, weno, ps, ng);
  const bool up = strict ? (m > 0.0f) : (m >= 0.0f);
  const Stencil& st = up ? ps : ng;
  auto q = [&](int ll) { return sc_at(w, ll, j, i); };
  const float c1 = w.col.c1[k - 1], c2 = w.col.c2[k - 1];
  const float r = eval_branch(st, q, c2, c1, w.weps, c.weps_f);
  w.p[B_MW][w.at(k, j, i)] = fb * r;
  float cf[5];
  const int nc = stencil_coefs(st, q, c2, c1, w.weps, c.weps_f, cf, c.exact != 0.0f);
  const float rb = fb * m;
  for (int n = 0; n < nc; ++n) aadd(&w.p[B_SC][w.at(st.lev[n], j, i)], cf[n] * rb);
}

// getdivx and getdiv transposed: adjoints of the stage-averaged winds (w levels on grid level l = 1..nk+1)
__global__ __launch_bounds__(64) void k_b_post(W w) {
  CELL_INDEX
  const Consts& c = w.c;
  const int nk = w.nk;
  const float *dx2 = w.p[B_DIVX2], *db = w.p[B_DIV];
  if (l <= nk) {
    const float rho = w.col.rho[l - 1];
    const float den = 0.5f * (rho + rho);
    const int64_t q = w.at(l, j, i), qw = w.at(l, j, w.X(i, -1)), qs = w.at(l, w.Y(j, -1), i);
    w.p[B_ARU][q] = den * (w.p[B_MU][q] + c.rdx * (dx2[qw] - dx2[q])) + c.rdx * (db[qw] - db[q]);
    w.p[B_ARV][q] = den * (w.p[B_MV][q] + c.rdy * (dx2[qs] - dx2[q])) + c.rdy * (db[qs] - db[q]);
  }
  const int64_t qw = w.at(l, j, i);
  if (l >= 2 && l <= nk) {
    const int64_t lo = w.at(l - 1, j, i);
    w.p[B_ARW][qw] = w.col.rf0[l - 1] * (w.p[B_MW][qw] + c.rdz * (dx2[lo] - dx2[qw])) + c.rdz * (db[lo] - db[qw]);
  } else {
    w.p[B_ARW][qw] = 0.0f;
  }
}

// acoustic adjoint state from the stage-output adjoints (grid level l = 1..nk+1)
__global__ __launch_bounds__(64) void k_b_ac_init(W w) {
  CELL_INDEX
  const int64_t q = w.at(l, j, i);
  w.p[B_W][q] = w.p[B_OW][q];
  if (l > w.nk) return;
  w.p[B_U][q] = w.p[B_OU][q];
  w.p[B_V][q] = w.p[B_OV][q];
  w.p[B_PP][q] = w.p[B_OPP][q] + (w.nrk == 3 ? w.p[B_PPX][q] : 0.0f);
  w.p[B_PPD][q] = 0.0f;
}

// The column adjoint of substep n in three passes.  A (every cell): forward fpk and div recomputed, the w
// averages, the pressure update's adjoint (fpk's part kept in B_FPKF), vt's coefficient adjoints and the w
// adjoint of vt gathered from levels l (w(k)) and l-1 (w(k+1)).  B_PP, B_PPD and B_PPX are only read here.
__global__ __launch_bounds__(64) void k_b_ssA(W w, int n) {
  CELL_INDEX
  const Consts& c = w.c;
  const int nk = w.nk;
  const bool last = n == w.nloop;
  const bool last3 = last && w.nrk == 3;
  const float* wn = w.wsnap(n);
  const int64_t q = w.at(l, j, i);
  float dv;
  w.p[FPK][q] = column_fpk(w, n, l, j, i, &dv);
  w.p[DIV][q] = dv;
  auto upb_at = [&](int64_t qq) {
    float u = w.p[B_PP][qq] + (1.0f + c.kdiv) * w.p[B_PPD][qq];
    if (last3) u = u + c.kdiv * w.p[B_PPX][qq];
    return u;
  };
  float bw = w.p[B_W][q];
  if (last) {
    bw += (w.p[B_ARW][q] * c.tavg) * c.alph;
    w.p[B_ARW][q] *= c.tavg;
  } else {
    bw += w.p[B_ARW][q];
  }
  const float upb = upb_at(q);
  w.p[B_FPKF][q] = upb;
  const float vtb = c.alph * upb;
  const float *qk = w.p[QK], *pk1 = w.p[PK1], *pk2 = w.p[PK2];
  const float wlo = l == 1 ? 0.0f : wn[q];
  const float whi = l == nk ? 0.0f : wn[w.at(l + 1, j, i)];
  w.p[B_PK1][q] += vtb * wlo;
  w.p[B_QK][q] += vtb * (wlo - whi);
  w.p[B_PK2][q] += vtb * whi;
  if (l >= 2) {
    const int64_t qd = w.at(l - 1, j, i);
    bw += vtb * (pk1[q] + qk[q]) + (c.alph * upb_at(qd)) * (pk2[qd] - qk[qd]);
  }
  w.p[B_W][q] = bw;
}

// B (one column): the transposed Thomas solve T^T lam = wbar(2..nk); lam is left in B_ED.  Row k of T^T
// has cc(k-1) below and aa(k+1) above its diagonal bb(k); rows are fetched TH_CHUNK levels at a time.
__global__ __launch_bounds__(64) void k_b_ssB(W w) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
  if (i >= w.ni) return;
  const int nk = w.nk;
  float* __restrict__ bec = w.p[B_EC];
  float* __restrict__ bed = w.p[B_ED];
  const float* __restrict__ bw = w.p[B_W];
  const float* __restrict__ ra = w.p[RA];
  const float* __restrict__ rb = w.p[RB];
  const float* __restrict__ rc = w.p[RC];
  float cprev = 0.0f, dprev = 0.0f;
  for (int k0 = 2; k0 <= nk; k0 += TH_CHUNK) {
    float lo[TH_CHUNK], up[TH_CHUNK], dg[TH_CHUNK], rh[TH_CHUNK];
#pragma unroll
    for (int u = 0; u < TH_CHUNK; ++u) {
      const int k = k0 + u;
      const int kk = k <= nk ? k : nk;
      lo[u] = kk > 2 ? rc[w.at(kk - 1, j, i)] : 0.0f;
      up[u] = kk < nk ? ra[w.at(kk + 1, j, i)] : 0.0f;
      dg[u] = rb[w.at(kk, j, i)];
      rh[u] = bw[w.at(kk, j, i)];
    }
#pragma unroll
    for (int u = 0; u < TH_CHUNK; ++u) {
      const int k = k0 + u;
      if (k > nk) break;
      const int64_t q = w.at(k, j, i);
      const float piv = dg[u] - lo[u] * cprev;
      const float cpk = up[u] / piv;
      const float dpk = (rh[u] - lo[u] * dprev) / piv;
      bec[q] = cpk;
      bed[q] = dpk;
      cprev = cpk;
      dprev = dpk;
    }
  }
  float lamnext = dprev;  // lam(nk)
  for (int k1 = nk - 1; k1 >= 2; k1 -= TH_CHUNK) {
    float e[TH_CHUNK], f[TH_CHUNK];
#pragma unroll
    for (int u = 0; u < TH_CHUNK; ++u) {
      const int64_t q = w.at(k1 - u >= 2 ? k1 - u : 2, j, i);
      e[u] = bec[q];
      f[u] = bed[q];
    }
#pragma unroll
    for (int u = 0; u < TH_CHUNK; ++u) {
      const int k = k1 - u;
      if (k < 2) break;
      lamnext = f[u] - e[u] * lamnext;
      bed[w.at(k, j, i)] = lamnext;
    }
  }
}

// C (every cell, w levels 1..nk+1): adjoints of the Thomas rows' right-hand sides and coefficients, of the
// pressure advection and fpk's w terms; the start-w adjoint, fpk's adjoint as divergence and advection parts
__global__ __launch_bounds__(64) void k_b_ssC(W w, int n) {
  CELL_INDEX
  const Consts& c = w.c;
  const int nk = w.nk;
  const int64_t q = w.at(l, j, i);
  if (l == nk + 1) return;  // the lid level's adjoint passes through (pk2(nk) = 0)
  const float* ws = w.wsnap(n - 1);
  const float* wn = w.wsnap(n);
  const float* pp = w.snap(SPP, n - 1);
  const float *qk = w.p[QK], *pk1 = w.p[PK1], *pk2 = w.p[PK2], *mm = w.p[MM], *fpk = w.p[FPK];
  const float* lam = w.p[B_ED];
  auto LAM = [&](int k) { return (k >= 2 && k <= nk) ? lam[w.at(k, j, i)] : 0.0f; };
  auto PP = [&](int k) { return pp[w.at(k, j, i)]; };
  auto FP = [&](int k) { return fpk[w.at(k, j, i)]; };
  // udb(k), pdb(k) of row k (2..nk)
  auto udb = [&](int k) { return (k >= 2 && k <= nk) ? c.alph * (-mm[w.at(k, j, i)]) * LAM(k) : 0.0f; };
  auto pdb = [&](int k) { return (k >= 2 && k <= nk) ? c.beta * (-mm[w.at(k, j, i)]) * LAM(k) : 0.0f; };
  // fpk's full adjoint at level k (1..nk)
  auto F = [&](int k) { return w.p[B_FPKF][w.at(k, j, i)] + udb(k) - udb(k + 1); };
  // bb's and aa's, cc's adjoints of row k
  auto diagb = [&](int k) {
    if (k < 2 || k > nk) return 0.0f;
    const int64_t qq = w.at(k, j, i);
    return (-c.asq * mm[qq]) * (-LAM(k) * wn[qq]);
  };
  auto aab = [&](int k) { return (k >= 3 && k <= nk) ? -LAM(k) * wn[w.at(k - 1, j, i)] : 0.0f; };
  auto ccb = [&](int k) { return (k >= 2 && k < nk) ? -LAM(k) * wn[w.at(k + 1, j, i)] : 0.0f; };
  const float f = F(l);
  // start-w adjoint of level l
  float bw = LAM(l);
  if (l >= 2) bw += c.beta * (f * pk1[q]) + qk[q] * (c.beta * f);
  if (l >= 2) {
    const int64_t qd = w.at(l - 1, j, i);
    const float fd = F(l - 1);
    bw += c.beta * (fd * pk2[qd]);
    if (l - 1 <= nk - 1) bw += -qk[qd] * (c.beta * fd);
  }
  w.p[B_W][q] = l == 1 ? 0.0f : bw;
  if (l >= 2) {
    const float lk = LAM(l);
    w.p[B_WTEN][q] += c.dts * lk;
    const int64_t qd = w.at(l - 1, j, i);
    const float pdelta = PP(l) - PP(l - 1);
    const float udelta = (PP(l) + FP(l)) - (PP(l - 1) + FP(l - 1));
    float mmb = (-(c.beta * pdelta + c.alph * udelta)) * lk;
    const float diag = (-(pk1[q] + qk[q])) + (pk2[qd] - qk[qd]);
    mmb += (-c.asq * diag) * (-lk * wn[q]);
    if (l >= 3) mmb += (-c.asq) * (pk1[qd] + qk[qd]) * aab(l);
    if (l < nk) mmb += c.asq * (pk2[q] - qk[q]) * ccb(l);
    w.p[B_MM][q] += mmb;
  }
  // pressure adjoint: the update's (from the step's output adjoints) and rows l and l+1's
  {
    const bool last3 = n == w.nloop && w.nrk == 3;
    float ppb = -c.kdiv * w.p[B_PPD][q];
    if (last3) {
      ppb = ppb - c.kdiv * w.p[B_PPX][q];
      w.p[B_PPX][q] = 0.0f;
    }
    w.p[B_PP][q] = (ppb + w.p[B_FPKF][q]) + ((pdb(l) + udb(l)) - (pdb(l + 1) + udb(l + 1)));
  }
  // coefficient adjoints of level l
  const float wlo = l == 1 ? 0.0f : ws[q];
  const float wup = ws[w.at(l + 1, j, i)];
  const float whi = l == nk ? 0.0f : wup;
  float bpk1 = -diagb(l) + f * (c.beta * wlo);
  float bqk = -diagb(l) - diagb(l + 1) - (c.beta * f) * (whi - wlo);
  float bpk2 = diagb(l + 1) + f * (c.beta * wup);
  if (l + 1 <= nk && l + 1 >= 3) {
    const float m1 = mm[w.at(l + 1, j, i)];
    bpk1 += (-c.asq) * m1 * aab(l + 1);
    bqk += (-c.asq) * m1 * aab(l + 1);
  }
  if (l < nk && l >= 2) {
    bpk2 += c.asq * mm[q] * ccb(l);
    bqk -= c.asq * mm[q] * ccb(l);
  }
  w.p[B_PK1][q] += bpk1;
  w.p[B_QK][q] += bqk;
  w.p[B_PK2][q] += bpk2;
  w.p[B_PPTERM][q] -= c.dts * (w.p[DIV][q] * f);
  w.p[B_DIV][q] = -c.dts * (w.p[PPTERM][q] * f);
  w.p[B_XA][q] = -(f * w.col.rr0[l - 1]);
}

// u and v faces of substep n: averages, pressure advection and divergence adjoints, forcing
__global__ __launch_bounds__(64) void k_b_ss_faces(W w, int n) {
  CELL_INDEX
  const Consts& c = w.c;
  const bool last = n == w.nloop;
  const float* pia = w.p[PIADV];
  const float *xa = w.p[B_XA], *db = w.p[B_DIV];
  const float rho = w.col.rho[l - 1];
  const int64_t q = w.at(l, j, i);
  {
    const int64_t qw = w.at(l, j, w.X(i, -1));
    float ub = w.p[B_U][q];
    if (last) {
      ub += w.p[B_ARU][q] * c.tavg;
      w.p[B_ARU][q] *= c.tavg;
    } else {
      ub += w.p[B_ARU][q];
    }
    const float dp = pia[q] - pia[qw];
    const float xub = c.temx * (xa[qw] * dp) + c.temx * (xa[q] * dp);
    ub += rho * xub;
    ub += db[qw] * c.rdx - db[q] * c.rdx;
    w.p[B_UTEN][q] += c.dts * ub;
    w.p[B_U][q] = ub;
  }
  {
    const int64_t qs = w.at(l, w.Y(j, -1), i);
    float vb = w.p[B_V][q];
    if (last) {
      vb += w.p[B_ARV][q] * c.tavg;
      w.p[B_ARV][q] *= c.tavg;
    } else {
      vb += w.p[B_ARV][q];
    }
    const float dp = pia[q] - pia[qs];
    const float yvb = c.temy * (xa[qs] * dp) + c.temy * (xa[q] * dp);
    vb += rho * yvb;
    vb += db[qs] * c.rdy - db[q] * c.rdy;
    w.p[B_VTEN][q] += c.dts * vb;
    w.p[B_V][q] = vb;
  }
}

// cells of substep n: ppd (start) and thv adjoints of the pressure gradient, piadv's of the pressure advection
__global__ __launch_bounds__(64) void k_b_ss_cells(W w, int n) {
  CELL_INDEX
  const Consts& c = w.c;
  const float* thv = w.p[THV];
  const float* ppd = w.snap(SPPD, n - 1);
  const float* u1 = w.snap(SU, n);
  const float* v1 = w.snap(SV, n);
  const float *bu = w.p[B_U], *bv = w.p[B_V], *xa = w.p[B_XA];
  const float rho = w.col.rho[l - 1];
  const int64_t q = w.at(l, j, i);
  const int64_t qe = w.at(l, j, w.X(i, 1)), qw = w.at(l, j, w.X(i, -1));
  const int64_t qn = w.at(l, w.Y(j, 1), i), qs = w.at(l, w.Y(j, -1), i);
  // u faces i (cell i is the right cell) and i+1 (left cell)
  const float pgx0 = -c.dts * bu[q], pgx1 = -c.dts * bu[qe];
  const float A0 = c.halfcp * (thv[q] + thv[qw]), A1 = c.halfcp * (thv[qe] + thv[q]);
  const float B0 = (ppd[q] - ppd[qw]) * c.rdx, B1 = (ppd[qe] - ppd[q]) * c.rdx;
  const float pgy0 = -c.dts * bv[q], pgy1 = -c.dts * bv[qn];
  const float Ay0 = c.halfcp * (thv[q] + thv[qs]), Ay1 = c.halfcp * (thv[qn] + thv[q]);
  const float By0 = (ppd[q] - ppd[qs]) * c.rdy, By1 = (ppd[qn] - ppd[q]) * c.rdy;
  w.p[B_PPD][q] = (c.rdx * (A0 * pgx0) - c.rdx * (A1 * pgx1)) + (c.rdy * (Ay0 * pgy0) - c.rdy * (Ay1 * pgy1));
  float tb = c.halfcp * (pgx0 * B0) + c.halfcp * (pgx1 * B1);
  tb += c.halfcp * (pgy0 * By0) + c.halfcp * (pgy1 * By1);
  w.p[B_THV][q] += tb;
  // pressure advection: t2(c) = xa(c)*temx*rho*u1(c+1), t1(c) = xa(c)*temx*rho*u1(c)
  auto t2x = [&](int64_t cell, int64_t east) { return xa[cell] * (c.temx * (rho * u1[east])); };
  auto t1x = [&](int64_t cell) { return xa[cell] * (c.temx * (rho * u1[cell])); };
  auto t2y = [&](int64_t cell, int64_t north) { return xa[cell] * (c.temy * (rho * v1[north])); };
  auto t1y = [&](int64_t cell) { return xa[cell] * (c.temy * (rho * v1[cell])); };
  float pb = (t2x(qw, q) - t2x(q, qe)) + (t1x(q) - t1x(qe));
  pb += (t2y(qs, q) - t2y(q, qn)) + (t1y(q) - t1y(qn));
  w.p[B_PIADV][q] += pb;
}

// coefficient adjoints -> thv, ppterm, piadv, gathered per cell from the coefficients that read it
__global__ __launch_bounds__(64) void k_b_coef(W w) {
  CELL_INDEX
  const Consts& c = w.c;
  const int nk = w.nk;
  const int64_t q = w.at(l, j, i);
  auto tem2 = [&](int k) { return ((c.dts * 1.0f) * w.col.rdsf[k]) * w.col.rr0[k - 1]; };
  auto b2b = [&](int k) { return (tem2(k) * w.col.rf0[k]) * w.p[B_PK2][w.at(k, j, i)]; };      // k < nk
  auto b1b = [&](int k) { return (tem2(k) * w.col.rf0[k - 1]) * w.p[B_PK1][w.at(k, j, i)]; };  // k > 1
  auto thb = [&](int k) { return ((c.dts * w.col.rds[k]) * c.cp) * w.p[B_MM][w.at(k, j, i)]; };  // k >= 2
  w.p[B_PPTERM][q] += (c.dts * w.col.rdsf[l]) * w.p[B_QK][q];
  float pia = 0.0f;
  if (l < nk) pia += (1.0f - w.col.c1[l]) * b2b(l);
  if (l >= 2 && l - 1 < nk) pia += (-w.col.c2[l - 1]) * b2b(l - 1);
  if (l > 1) pia += (w.col.c2[l - 1] - 1.0f) * b1b(l);
  if (l + 1 <= nk) pia += w.col.c1[l] * b1b(l + 1);
  w.p[B_PIADV][q] += pia;
  float tv = 0.0f;
  if (l >= 2) tv += w.col.c2[l - 1] * thb(l);
  if (l + 1 <= nk) tv += w.col.c1[l] * thb(l + 1);
  w.p[B_THV][q] += tv;
}

// stage start transposed (grid level l = 1..nk+1)
__global__ __launch_bounds__(64) void k_b_ac_end(W w) {
  CELL_INDEX
  const int nk = w.nk;
  const int64_t q = w.at(l, j, i);
  if (l <= nk) {
    w.p[B_PPX][q] = w.p[B_PPD][q] + w.p[B_PPX][q];
    w.p[B_AW][q] += w.c.beta * w.p[B_ARW][q];
    if (w.nrk != 1) {
      w.p[B_API][q] += w.p[B_PP][q];
      w.p[B_AU][q] += w.p[B_U][q];
      w.p[B_AV][q] += w.p[B_V][q];
      w.p[B_AW][q] += w.p[B_W][q];
    } else {
      w.p[B_XU][q] += w.p[B_U][q];
      w.p[B_XV][q] += w.p[B_V][q];
      w.p[B_XW][q] += w.p[B_W][q];
      w.p[B_XPP][q] += w.p[B_PP][q];
    }
  } else {
    w.p[B_XW][q] += w.p[B_W][q];
  }
}

// momentum: local terms (u*divx, w*divx), buoyancy (grid level l = 1..nk+1)
__global__ __launch_bounds__(64) void k_b_adv_local(W w) {
  CELL_INDEX
  const int nk = w.nk;
  const float *dx = w.p[DIVX], *btu = w.p[B_UTEN], *btv = w.p[B_VTEN], *btw = w.p[B_WTEN];
  auto totu = [&](int64_t qq, int ll) {
    const float rho = w.col.rho[ll - 1];
    return btu[qq] * (1.0f / (0.5f * (rho + rho)));
  };
  auto totv = [&](int64_t qq, int ll) {
    const float rho = w.col.rho[ll - 1];
    return btv[qq] * (1.0f / (0.5f * (rho + rho)));
  };
  auto totw = [&](int kf, int64_t qq) { return btw[qq] * (w.col.rrf0[kf - 1] * 1.0f); };
  const int64_t q = w.at(l, j, i);
  if (l <= nk) {
    const int64_t qw = w.at(l, j, w.X(i, -1)), qe = w.at(l, j, w.X(i, 1));
    const int64_t qs = w.at(l, w.Y(j, -1), i), qn = w.at(l, w.Y(j, 1), i);
    const float tu = totu(q, l), tv = totv(q, l);
    w.p[B_XU][q] += tu * (0.5f * (dx[q] + dx[qw]));
    w.p[B_XV][q] += tv * (0.5f * (dx[q] + dx[qs]));
    const float* u = w.p[XU];
    const float* v = w.p[XV];
    const float* ww = w.p[XW];
    float g = 0.5f * (tu * u[q]) + 0.5f * (totu(qe, l) * u[qe]);
    g += 0.5f * (tv * v[q]) + 0.5f * (totv(qn, l) * v[qn]);
    if (l >= 2) g += w.col.c2[l - 1] * (totw(l, q) * ww[q]);
    if (l + 1 <= nk) {
      const int64_t qu = w.at(l + 1, j, i);
      g += (1.0f - w.col.c2[l]) * (totw(l + 1, qu) * ww[qu]);
    }
    w.p[B_DIVX][q] += g;
    // buoyancy: dum8 level l from w levels l (c2) and l+1 (c1)
    float d8 = 0.0f;
    if (l >= 2) d8 += w.col.c2[l - 1] * btw[q];
    if (l + 1 <= nk) d8 += w.col.c1[l] * btw[w.at(l + 1, j, i)];
    w.p[B_DUM8][q] = d8;
  }
  if (l >= 2 && l <= nk) {
    const float cc2 = w.col.c2[l - 1], cc1 = 1.0f - cc2;
    w.p[B_XW][q] += totw(l, q) * (cc2 * dx[q] + cc1 * dx[w.at(l - 1, j, i)]);
  }
}

// Horizontal momentum flux transposes without atomics.  Pass 1 stores, per flux, fb*m on the side its
// upwind choice took (G0/G3 where m >= 0, G1/G4 where m < 0) and fb*r (G2/G5), for the x (G0-G2) and y
// (G3-G5) fluxes of one component; pass 2 gathers them at each grid point.  comp 0 u, 1 v, 2 w (levels 2..nk).
__global__ __launch_bounds__(64) void k_b_advh_flux(W w, int comp) {
  CELL_INDEX
  const Consts& c = w.c;
  const int nk = w.nk;
  const int64_t q = w.at(l, j, i);
  const float *ru = w.p[RRU], *rv = w.p[RRV];
  float *g0 = w.p[G0], *g1 = w.p[G1], *g2 = w.p[G2], *g3 = w.p[G3], *g4 = w.p[G4], *g5 = w.p[G5];
  if (comp == 2 && (l < 2 || l > nk)) {
    g0[q] = g1[q] = g2[q] = g3[q] = g4[q] = g5[q] = 0.0f;
    return;
  }
  const float* fld = comp == 0 ? w.p[XU] : (comp == 1 ? w.p[XV] : w.p[XW]);
  const float* bt = comp == 0 ? w.p[B_UTEN] : (comp == 1 ? w.p[B_VTEN] : w.p[B_WTEN]);
  auto F = [&](int jj, int ii) { return fld[w.at(l, w.Y(jj, 0), w.X(ii, 0))]; };
  auto tot = [&](int jj, int ii) {
    const float v = bt[w.at(l, w.Y(jj, 0), w.X(ii, 0))];
    if (comp == 2) return v * (w.col.rrf0[l - 1] * 1.0f);
    const float rho = w.col.rho[l - 1];
    return v * (1.0f / (0.5f * (rho + rho)));
  };
  float fbx, mx, posx, negx, fby, my, posy, negy;
  if (comp == 0) {  // x flux at the cell centre i, y flux at row face j (face column i)
    fbx = c.rdx * tot(j, i + 1) - c.rdx * tot(j, i);
    mx = 0.5f * (ru[q] + ru[w.at(l, j, w.X(i, 1))]);
    posx = flx5(F(j, i - 2), F(j, i - 1), F(j, i), F(j, i + 1), F(j, i + 2));
    negx = flx5(F(j, i + 3), F(j, i + 2), F(j, i + 1), F(j, i), F(j, i - 1));
    fby = c.rdy * tot(j, i) - c.rdy * tot(j - 1, i);
    my = 0.5f * (rv[q] + rv[w.at(l, j, w.X(i, -1))]);
    posy = flx5(F(j - 3, i), F(j - 2, i), F(j - 1, i), F(j, i), F(j + 1, i));
    negy = flx5(F(j + 2, i), F(j + 1, i), F(j, i), F(j - 1, i), F(j - 2, i));
  } else if (comp == 1) {  // x flux at face column i of row j, y flux at the cell-centre row j
    fbx = c.rdx * tot(j, i) - c.rdx * tot(j, i - 1);
    mx = 0.5f * (ru[q] + ru[w.at(l, w.Y(j, -1), i)]);
    posx = flx5(F(j, i - 3), F(j, i - 2), F(j, i - 1), F(j, i), F(j, i + 1));
    negx = flx5(F(j, i + 2), F(j, i + 1), F(j, i), F(j, i - 1), F(j, i - 2));
    fby = c.rdy * tot(j + 1, i) - c.rdy * tot(j, i);
    my = 0.5f * (rv[q] + rv[w.at(l, w.Y(j, 1), i)]);
    posy = flx5(F(j - 2, i), F(j - 1, i), F(j, i), F(j + 1, i), F(j + 2, i));
    negy = flx5(F(j + 3, i), F(j + 2, i), F(j + 1, i), F(j, i), F(j - 1, i));
  } else {  // w at level l: x flux at face i, y flux at face j
    const float c2h = 0.5f * (w.col.c2[l - 1] + w.col.c2[l - 1]);
    const float c1h = 1.0f - c2h;
    fbx = c.rdx * tot(j, i) - c.rdx * tot(j, i - 1);
    mx = c2h * ru[w.at(l, j, i)] + c1h * ru[w.at(l - 1, j, i)];
    posx = flx5(F(j, i - 3), F(j, i - 2), F(j, i - 1), F(j, i), F(j, i + 1));
    negx = flx5(F(j, i + 2), F(j, i + 1), F(j, i), F(j, i - 1), F(j, i - 2));
    fby = c.rdy * tot(j, i) - c.rdy * tot(j - 1, i);
    my = c2h * rv[w.at(l, j, i)] + c1h * rv[w.at(l - 1, j, i)];
    posy = flx5(F(j - 3, i), F(j - 2, i), F(j - 1, i), F(j, i), F(j + 1, i));
    negy = flx5(F(j + 2, i), F(j + 1, i), F(j, i), F(j - 1, i), F(j - 2, i));
  }
  const bool ux = mx >= 0.0f, uy = my >= 0.0f;
  g0[q] = ux ? fbx * mx : 0.0f;
  g1[q] = ux ? 0.0f : fbx * mx;
  g2[q] = fbx * (ux ? posx : negx);
  g3[q] = uy ? fby * my : 0.0f;
  g4[q] = uy ? 0.0f : fby * my;
  g5[q] = fby * (uy ? posy : negy);
}

// pass 2: the field adjoint gathered from the fluxes whose stencils hold the point, and the mass-flux adjoints
__global__ __launch_bounds__(64) void k_b_advh_gather(W w, int comp) {
  CELL_INDEX
  const int nk = w.nk;
  const int64_t q = w.at(l, j, i);
  const float *g0 = w.p[G0], *g1 = w.p[G1], *g2 = w.p[G2], *g3 = w.p[G3], *g4 = w.p[G4], *g5 = w.p[G5];
  const float cf[5] = {gm1::C2_60, gm1::CM13_60, gm1::C47_60, gm1::C27_60, gm1::CM3_60};
  auto at = [&](int ll, int jj, int ii) { return w.at(ll, w.Y(jj, 0), w.X(ii, 0)); };
  // stencil offsets: a flux at index p reads points p + o0 + n (upwind side, o0 = -2 or -3) or p + o1 - n
  const int xo_pos = comp == 0 ? -2 : -3, xo_neg = comp == 0 ? 3 : 2;
  const int yo_pos = comp == 1 ? -2 : -3, yo_neg = comp == 1 ? 3 : 2;
  const bool wlevel = comp != 2 || (l >= 2 && l <= nk);
  if (wlevel) {
    float acc = 0.0f;
    for (int n = 0; n < 5; ++n) {
      acc += cf[n] * g0[at(l, j, i - xo_pos - n)];
      acc += cf[n] * g1[at(l, j, i - xo_neg + n)];
      acc += cf[n] * g3[at(l, j - yo_pos - n, i)];
      acc += cf[n] * g4[at(l, j - yo_neg + n, i)];
    }
    float* bf = comp == 0 ? w.p[B_XU] : (comp == 1 ? w.p[B_XV] : w.p[B_XW]);
    bf[q] += acc;
  }
  if (l > nk) return;
  // mass-flux adjoints at the rru / rrv point (l, j, i)
  if (comp == 0) {
    w.p[B_RRU][q] += 0.5f * g2[q] + 0.5f * g2[at(l, j, i - 1)];
    w.p[B_RRV][q] += 0.5f * g5[q] + 0.5f * g5[at(l, j, i + 1)];
  } else if (comp == 1) {
    w.p[B_RRU][q] += 0.5f * g2[q] + 0.5f * g2[at(l, j + 1, i)];
    w.p[B_RRV][q] += 0.5f * g5[q] + 0.5f * g5[at(l, j - 1, i)];
  } else {
    float gu = 0.0f, gv = 0.0f;
    if (l >= 2) {
      const float c2h = 0.5f * (w.col.c2[l - 1] + w.col.c2[l - 1]);
      gu += c2h * g2[q];
      gv += c2h * g5[q];
    }
    if (l + 1 <= nk) {
      const float c2h = 0.5f * (w.col.c2[l] + w.col.c2[l]);
      const int64_t qu = w.at(l + 1, j, i);
      gu += (1.0f - c2h) * g2[qu];
      gv += (1.0f - c2h) * g5[qu];
    }
    w.p[B_RRU][q] += gu;
    w.p[B_RRV][q] += gv;
  }
}

// Vertical momentum flux transposes, the same two passes: pass 1 per flux (u/v: w levels kf = 2..nk,
// w: scalar levels k = 1..nk) stores G0/G1 (fb*m by upwind side) and G2 (fb*r); pass 2 gathers per point.
__global__ __launch_bounds__(64) void k_b_advz_flux(W w, int comp) {
  CELL_INDEX
  const Consts& c = w.c;
  const int nk = w.nk;
  const int64_t q = w.at(l, j, i);
  float *g0 = w.p[G0], *g1 = w.p[G1], *g2 = w.p[G2];
  const float* rw = w.p[RRW];
  const bool live = comp == 2 ? (l <= nk) : (l >= 2 && l <= nk);
  if (!live) {
    g0[q] = g1[q] = g2[q] = 0.0f;
    return;
  }
  const float* fld = comp == 0 ? w.p[XU] : (comp == 1 ? w.p[XV] : w.p[XW]);
  auto Q = [&](int ll) { return fld[w.at(ll, j, i)]; };
  float fb, m, r;
  if (comp < 2) {
    const float* bt = comp == 0 ? w.p[B_UTEN] : w.p[B_VTEN];
    auto tot = [&](int ll) {
      const float rho = w.col.rho[ll - 1];
      return bt[w.at(ll, j, i)] * (1.0f / (0.5f * (rho + rho)));
    };
    const int kf = l;
    fb = c.rdz * tot(kf) - c.rdz * tot(kf - 1);
    const int64_t qo = comp == 0 ? w.at(kf, j, w.X(i, -1)) : w.at(kf, w.Y(j, -1), i);
    m = 0.5f * (rw[q] + rw[qo]);
    const float c2h = 0.5f * (w.col.c2[kf - 1] + w.col.c2[kf - 1]);
    Stencil ps, ng;
    uvz_branch(kf, nk, ps, ng);
    r = m >= 0.0f ? eval_branch(ps, Q, c2h, 1.0f - c2h, 0.0, 0.0f) : eval_branch(ng, Q, c2h, 1.0f - c2h, 0.0, 0.0f);
  } else {
    const float* bt = w.p[B_WTEN];
    auto tot = [&](int kf) {
      if (kf < 2 || kf > nk) return 0.0f;
      return bt[w.at(kf, j, i)] * (w.col.rrf0[kf - 1] * 1.0f);
    };
    const int k = l;
    fb = c.rdz * tot(k + 1) - c.rdz * tot(k);
    m = 0.5f * (rw[w.at(k, j, i)] + rw[w.at(k + 1, j, i)]);
    Stencil ps, ng;
    wz_branch(k, nk, ps, ng);
    r = m >= 0.0f ? eval_branch(ps, Q, 0.0f, 0.0f, 0.0, 0.0f) : eval_branch(ng, Q, 0.0f, 0.0f, 0.0, 0.0f);
  }
  const bool up = m >= 0.0f;
  g0[q] = up ? fb * m : 0.0f;
  g1[q] = up ? 0.0f : fb * m;
  g2[q] = fb * r;
}

__global__ __launch_bounds__(64) void k_b_advz_gather(W w, int comp) {
  CELL_INDEX
  const int nk = w.nk;
  const int64_t q = w.at(l, j, i);
  const float *g0 = w.p[G0], *g1 = w.p[G1], *g2 = w.p[G2];
  auto Q0 = [&](int) { return 0.0f; };
  float acc = 0.0f;
  const float cf5[5] = {gm1::C2_60, gm1::CM13_60, gm1::C47_60, gm1::C27_60, gm1::CM3_60};
  if (comp < 2 && l >= 6 && l <= nk - 5) {
    // every flux kf = l-2..l+3 uses 5th-order stencils: upwind levels kf-3..kf+1, downwind kf+2..kf-2
#pragma unroll
    for (int d = -1; d <= 3; ++d) acc += cf5[3 - d] * g0[w.at(l + d, j, i)];
#pragma unroll
    for (int d = -2; d <= 2; ++d) acc += cf5[2 + d] * g1[w.at(l + d, j, i)];
    (comp == 0 ? w.p[B_XU] : w.p[B_XV])[q] += acc;
    const int64_t qn = comp == 0 ? w.at(l, j, w.X(i, 1)) : w.at(l, w.Y(j, 1), i);
    w.p[B_RRW][q] += 0.5f * g2[q] + 0.5f * g2[qn];
    return;
  }
  if (comp == 2 && l >= 6 && l <= nk - 4) {
    // every flux k = l-3..l+2 uses 5th-order stencils: upwind levels k-2..k+2, downwind k+3..k-1
#pragma unroll
    for (int d = -2; d <= 2; ++d) acc += cf5[2 - d] * g0[w.at(l + d, j, i)];
#pragma unroll
    for (int d = -3; d <= 1; ++d) acc += cf5[3 + d] * g1[w.at(l + d, j, i)];
    w.p[B_XW][q] += acc;
    w.p[B_RRW][q] += 0.5f * g2[q] + 0.5f * g2[w.at(l - 1, j, i)];
    return;
  }
  if (comp < 2) {
    if (l <= nk) {
      const int lo = l - 2 < 2 ? 2 : l - 2, hi = l + 3 > nk ? nk : l + 3;
      for (int kf = lo; kf <= hi; ++kf) {
        const float c2h = 0.5f * (w.col.c2[kf - 1] + w.col.c2[kf - 1]);
        Stencil ps, ng;
        uvz_branch(kf, nk, ps, ng);
        float cf[5];
        const int64_t qf = w.at(kf, j, i);
        int nc = branch_coefs(ps, Q0, c2h, 1.0f - c2h, 0.0f, cf);
        for (int n = 0; n < nc; ++n)
          if (ps.lev[n] == l) acc += cf[n] * g0[qf];
        nc = branch_coefs(ng, Q0, c2h, 1.0f - c2h, 0.0f, cf);
        for (int n = 0; n < nc; ++n)
          if (ng.lev[n] == l) acc += cf[n] * g1[qf];
      }
      (comp == 0 ? w.p[B_XU] : w.p[B_XV])[q] += acc;
    }
    // rrw at w level l (2..nk) from the fluxes at this point and its +x / +y neighbour
    if (l >= 2 && l <= nk) {
      const int64_t qn = comp == 0 ? w.at(l, j, w.X(i, 1)) : w.at(l, w.Y(j, 1), i);
      w.p[B_RRW][q] += 0.5f * g2[q] + 0.5f * g2[qn];
    }
  } else {
    const int lo = l - 3 < 1 ? 1 : l - 3, hi = l + 2 > nk ? nk : l + 2;
    for (int k = lo; k <= hi; ++k) {
      Stencil ps, ng;
      wz_branch(k, nk, ps, ng);
      float cf[5];
      const int64_t qf = w.at(k, j, i);
      int nc = branch_coefs(ps, Q0, 0.0f, 0.0f, 0.0f, cf);
      for (int n = 0; n < nc; ++n)
        if (ps.lev[n] == l) acc += cf[n] * g0[qf];
      nc = branch_coefs(ng, Q0, 0.0f, 0.0f, 0.0f, cf);
      for (int n = 0; n < nc; ++n)
        if (ng.lev[n] == l) acc += cf[n] * g1[qf];
    }
    w.p[B_XW][q] += acc;
    float mb = 0.0f;
    if (l <= nk) mb += 0.5f * g2[q];
    if (l >= 2) mb += 0.5f * g2[w.at(l - 1, j, i)];
    w.p[B_RRW][q] += mb;
  }
}

// mass fluxes and divx transposed, then the stage terms (grid level l = 1..nk+1)
__global__ __launch_bounds__(64) void k_b_mass_terms(W w) {
  CELL_INDEX
  const Consts& c = w.c;
  const int nk = w.nk;
  const float* dxb = w.p[B_DIVX];
  const int64_t q = w.at(l, j, i);
  if (l <= nk) {
    const float rho = w.col.rho[l - 1];
    const float den = 0.5f * (rho + rho);
    const int64_t qw = w.at(l, j, w.X(i, -1)), qs = w.at(l, w.Y(j, -1), i);
    w.p[B_XU][q] += den * (w.p[B_RRU][q] + (dxb[qw] * c.rdx - dxb[q] * c.rdx));
    w.p[B_XV][q] += den * (w.p[B_RRV][q] + (dxb[qs] * c.rdy - dxb[q] * c.rdy));
    w.p[B_XTH][q] += ((w.p[B_THV][q] * 1.0f + w.p[B_DUM8][q] * (c.g * w.col.rth[l - 1])) +
                      w.p[B_THTERM][q] * (c.kterm * c.cvm)) + w.p[B_SC][q];
    w.p[B_XPP][q] += w.p[B_PPTERM][q] * ((c.rovcp * c.cpm) * c.cvm) + w.p[B_PIADV][q];
  }
  if (l >= 2 && l <= nk) {
    const int64_t qd = w.at(l - 1, j, i);
    w.p[B_XW][q] += w.col.rf0[l - 1] * (w.p[B_RRW][q] + (dxb[qd] * c.rdz - dxb[q] * c.rdz));
  }
}

W make_w(void* const* table, const float* consts, int ni, int nj, int nk, int nrk, int nloop, double weps) {
  W w;
  for (int s = 0; s < NSLOT; ++s) w.p[s] = (float*)table[s];
  w.ni = ni;
  w.nj = nj;
  w.nk = nk;
  w.nrk = nrk;
  w.nloop = nloop;
  w.weps = weps;
  const float* cs = consts;
  float* dst = (float*)&w.c;
  for (int n = 0; n < NCONST; ++n) dst[n] = cs[n];
  w.col = make_cols(w.p[COL], nk);
  return w;
}

}  // namespace

extern "C" {

int gm1hip_dry_stage_forward(void* const* table, const float* consts, int ni, int nj, int nk, int nrk, int nloop,
                             double weps, void* stream) {
  if (!table || !consts || ni < 6 || nj < 6 || nk < 6 || nrk < 1 || nrk > 3 || nloop < 1)
    return (int)hipErrorInvalidValue;
  hipStream_t s = (hipStream_t)stream;
  const W w = make_w(table, consts, ni, nj, nk, nrk, nloop, weps);
  const dim3 b(64), g((ni + 63) / 64, nj, nk), gw((ni + 63) / 64, nj, nk + 1), gc((ni + 63) / 64, nj);
  hipLaunchKernelGGL(k_terms, g, b, 0, s, w);
  hipLaunchKernelGGL(k_mass, g, b, 0, s, w);
  hipLaunchKernelGGL(k_divx, g, b, 0, s, w);
  hipLaunchKernelGGL(k_advu, g, b, 0, s, w);
  hipLaunchKernelGGL(k_advv, g, b, 0, s, w);
  hipLaunchKernelGGL(k_advw, gw, b, 0, s, w);
  hipLaunchKernelGGL(k_coef, g, b, 0, s, w);
  hipLaunchKernelGGL(k_rows, g, b, 0, s, w);
  hipLaunchKernelGGL(k_ac_start, g, b, 0, s, w);
  GM1_LAUNCH_CHECK();
  for (int n = 1; n <= nloop; ++n) {
    hipLaunchKernelGGL(k_ss_pgrad, g, b, 0, s, w, n);
    hipLaunchKernelGGL(k_ss_fpk, g, b, 0, s, w, n);
    hipLaunchKernelGGL(k_ss_rhs, g, b, 0, s, w, n);
    hipLaunchKernelGGL(k_ss_thomas, gc, b, 0, s, w, n);
    hipLaunchKernelGGL(k_ss_update, g, b, 0, s, w, n);
  }
  hipLaunchKernelGGL(k_post, g, b, 0, s, w);
  hipLaunchKernelGGL(k_theta, g, b, 0, s, w);
  if (nrk == 3) hipLaunchKernelGGL(k_solve3, g, b, 0, s, w);
  GM1_LAUNCH_CHECK();
  return 0;
}

int gm1hip_dry_stage_adjoint(void* const* table, const float* consts, int ni, int nj, int nk, int nrk, int nloop,
                             double weps, void* stream) {
  if (!table || !consts || ni < 6 || nj < 6 || nk < 6 || nrk < 1 || nrk > 3 || nloop < 1)
    return (int)hipErrorInvalidValue;
  hipStream_t s = (hipStream_t)stream;
  const W w = make_w(table, consts, ni, nj, nk, nrk, nloop, weps);
  const int64_t cells = (int64_t)ni * nj * nk, wcells = (int64_t)ni * nj * (nk + 1);
  const int zero_cell[] = {B_XU, B_XV, B_XTH, B_XPP, B_THV, B_PPTERM, B_PIADV, B_RRU, B_RRV, B_DIVX, B_UTEN,
                           B_VTEN, B_QK, B_PK1, B_PK2, B_MM};
  const int zero_w[] = {B_XW, B_RRW, B_WTEN};
  FillList fl;
  fl.count = 0;
  for (int z : zero_cell) {
    fl.p[fl.count] = w.p[z];
    fl.n[fl.count++] = cells;
  }
  for (int z : zero_w) {
    fl.p[fl.count] = w.p[z];
    fl.n[fl.count++] = wcells;
  }
  const int fill_blocks = (int)std::min<int64_t>(gm1_blocks(wcells, 256), 2048);
  hipLaunchKernelGGL(k_b_fill, dim3(fill_blocks, fl.count), dim3(256), 0, s, fl);
  GM1_LAUNCH_CHECK();
  const dim3 b(64), g((ni + 63) / 64, nj, nk), gw((ni + 63) / 64, nj, nk + 1), gc((ni + 63) / 64, nj);
  // theta and the post-acoustic divergences
  hipLaunchKernelGGL(k_b_theta_cell, g, b, 0, s, w);
  hipLaunchKernelGGL(k_b_theta_xy, g, b, 0, s, w, 0);
  hipLaunchKernelGGL(k_b_theta_xy, g, b, 0, s, w, 1);
  hipLaunchKernelGGL(k_b_theta_z, g, b, 0, s, w);
  hipLaunchKernelGGL(k_b_post, gw, b, 0, s, w);
  // acoustic stage
  hipLaunchKernelGGL(k_b_ac_init, gw, b, 0, s, w);
  GM1_LAUNCH_CHECK();
  for (int n = nloop; n >= 1; --n) {
    hipLaunchKernelGGL(k_b_ssA, g, b, 0, s, w, n);
    hipLaunchKernelGGL(k_b_ssB, gc, b, 0, s, w);
    hipLaunchKernelGGL(k_b_ssC, g, b, 0, s, w, n);
    hipLaunchKernelGGL(k_b_ss_faces, g, b, 0, s, w, n);
    hipLaunchKernelGGL(k_b_ss_cells, g, b, 0, s, w, n);
  }
  hipLaunchKernelGGL(k_b_coef, g, b, 0, s, w);
  hipLaunchKernelGGL(k_b_ac_end, gw, b, 0, s, w);
  // momentum advection, buoyancy, mass fluxes, stage terms
  hipLaunchKernelGGL(k_b_adv_local, gw, b, 0, s, w);
  for (int comp = 0; comp < 3; ++comp) {
    hipLaunchKernelGGL(k_b_advh_flux, comp == 2 ? gw : g, b, 0, s, w, comp);
    hipLaunchKernelGGL(k_b_advh_gather, comp == 2 ? gw : g, b, 0, s, w, comp);
  }
  for (int comp = 0; comp < 3; ++comp) {
    hipLaunchKernelGGL(k_b_advz_flux, gw, b, 0, s, w, comp);
    hipLaunchKernelGGL(k_b_advz_gather, gw, b, 0, s, w, comp);
  }
  hipLaunchKernelGGL(k_b_mass_terms, gw, b, 0, s, w);
  GM1_LAUNCH_CHECK();
  return 0;
}

}  // extern "C"
