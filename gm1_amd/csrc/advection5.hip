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
constexpr int TY = 8;
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
// max |s| over a five-point stencil: the WENO gate of the positive-definite
// fluxes (flux = 0 where every stencil value is <= weps).
__device__ __forceinline__ float max5(float a, float b, float c, float d, float e) {
  return fmaxf(fmaxf(fabsf(a), fabsf(b)), fmaxf(fabsf(c), fmaxf(fabsf(d), fabsf(e))));
}

template <bool WENO, bool GATE>
__device__ __forceinline__ float recon_gated(float m, float s1, float s2, float s3, float s4, float s5, double weps) {
  if (GATE && !((double)max5(s1, s2, s3, s4, s5) > weps)) return 0.0f;
  return m * recon5(s1, s2, s3, s4, s5, WENO, weps);
}

// z-face flux at face k (2..nk) of one column with the lower-order closures.
// S(level) reads the column's scalar at a level.
template <bool WENO, class Get>
__device__ __forceinline__ float zflux_g(Get S, float m, float cc1, float cc2, int k, int nk, double weps) {
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

// z-face flux of the positive-definite chain (cm1_rk_pdef_vertical_flux_v1):
// the lower closures of vadv_weno5 / vadv_lwr_weno_35(_b) and, with
// pdefweno, the magnitude gate (a gated face stores +0, never m*0).
template <class Get>
__device__ __forceinline__ float zflux_pdef_g(Get S, float m, float cc1, float cc2, int k, int nk, double weps,
                                              bool pdefweno) {
  const float wf = (float)weps;
  float r;
  if (k == 2) {
    r = (m >= 0.0f) ? cc1 * S(k - 1) + cc2 * S(k) : upstrpd(S(k + 1), S(k), S(k - 1), wf);
  } else if (k == nk) {
    r = (m > 0.0f) ? upstrpd(S(k - 2), S(k - 1), S(k), wf) : cc1 * S(k - 1) + cc2 * S(k);
  } else {
    const bool low = (k == 3 && (pdefweno ? m > 0.0f : m >= 0.0f)) || (k == nk - 1 && m < 0.0f);
    if (low) {
      r = (k == 3) ? upstrpd(S(k - 2), S(k - 1), S(k), wf) : upstrpd(S(k + 1), S(k), S(k - 1), wf);
    } else {
      float s1, s2, s3, s4, s5;
      if (m >= 0.0f) { s1 = S(k - 3); s2 = S(k - 2); s3 = S(k - 1); s4 = S(k); s5 = S(k + 1); }
      else { s1 = S(k + 2); s2 = S(k + 1); s3 = S(k); s4 = S(k - 1); s5 = S(k - 2); }
      if (pdefweno && !((double)max5(s1, s2, s3, s4, s5) > weps)) return 0.0f;
      r = weno5(s1, s2, s3, s4, s5, weps);
    }
  }
  return m * r;
}

// Output modes of the fused tile kernel.
constexpr int MODE_TENDENCY = 0;  // tend = advection tendency
constexpr int MODE_COMPOSE = 1;   // out = a0 + dttmp*((held - divterm) + tendency), |out|<1e-30 -> 0
constexpr int MODE_PDEF = 2;      // advx/advy/advz = directional tendencies of the pdef chain

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
  // MODE_COMPOSE
  const float* a0;
  const float* held;
  const float* divterm;  // may be null
  float* out;
  float dttmp;
  // MODE_PDEF
  float* advx;
  float* advy;
  float* advz;
  int pdefweno;
  int ni, nj, nk, kchunk, nchunks;
  float rdx, rdy, rdsf;  // rdsf: 1/dz, or rdz of the metric reference on a flat stretched column
  double weps;
  unsigned obc;
  // Flat stretched column (terrain_flag=.false., adv.F:458-478, 555-572,
  // 646-656): mh (nk+2, CM1-indexed) or null.  With it the horizontal
  // differences are -(fe-fw)*rdx, the vertical ((d*rdz)*mh(k)) and the
  // tendency is scaled by rr0(1,1,k) alone.
  const float* mh;
  // Unpadded source (nk, nj, ni) read in place of the padded scalar s, or
  // null; th0 (nk) turns it into the packed theta (th0 - 300) + theta'.
  const float* src;
  const float* th0;
  // the caller states c1 and c2 hold one word per level (every column's the
  // same): the z faces read the first column's
  int c12col;
};

// Lateral-family word with its flag bits split off: bit 16 of a caller's obc
// states that c1 and c2 are horizontally uniform.
constexpr unsigned OBC_C12_COLUMN = 1u << 16;
inline void set_obc(ScalarAdvArgs& a, unsigned obc) {
  a.obc = obc & 0xFFFFu;
  a.c12col = (obc & OBC_C12_COLUMN) ? 1 : 0;
}

// bc.F's halo cell c (-3..n+2) of a row of n cells as a physical index.
GM1_DEV int family_cell(int c, int n, unsigned lo, unsigned hi) {
  if (c < 0) return lo == 1u ? c + n : (lo == 2u ? 0 : -c - 1);
  if (c >= n) return hi == 1u ? c - n : (hi == 2u ? n - 1 : 2 * n - 1 - c);
  return c;
}


// The fused tile kernel, one lane per face.  Each lane forms its own cell's
// west x face and south y face from its own mass-flux words (wavefront TY-1
// also the tile's east faces, wavefront TY-2 its north faces), so no slot
// index is formed per level; the level plane's slots keep their byte offsets within
// a level (from an unpadded carrier, the lateral family's image of each halo
// cell) in LDS, formed once per workgroup.  A lane whose face lies outside the
// domain reads word 0 and its flux is never consumed.  FLAT: the flat
// stretched column's text (a.mh set); SRC: the scalar read from its unpadded
// carrier (a.src) rather than a padded array (a.s).
constexpr int LQ = TX + 6;  // the level plane's row pitch
static_assert(TX == 64, "a wavefront is one tile row");

template <bool WENO, bool GATE>
__device__ __forceinline__ float face_flux(float m, float q0, float q1, float q2, float q3, float q4, float q5,
                                           double weps) {
  const bool up = m >= 0.0f;  // the five cells upwind of the face: q0..q4, or q5..q1 against the flow
  return recon_gated<WENO, GATE>(m, up ? q0 : q5, up ? q1 : q4, up ? q2 : q3, up ? q3 : q2, up ? q4 : q1, weps);
}

template <bool WENO, int MODE, bool FLAT, bool SRC>
__device__ __forceinline__ void scalar_adv5_body_lanes(const ScalarAdvArgs& a) {
  constexpr bool GATE = (MODE == MODE_PDEF);
  constexpr int PS = ((TY + 6) * LQ + NT - 1) / NT;
  __shared__ float lfx[TY][TX + 1];
  __shared__ float lfy[TY + 1][TX];
  __shared__ float lp[PS * NT];       // the level plane, TY+6 rows of LQ (slots past them unread)
  __shared__ uint32_t lpo[PS * NT];   // each plane slot's byte offset within a level
  const int ni = a.ni, nj = a.nj, nk = a.nk;
  const int tx = threadIdx.x % TX, ty = threadIdx.x / TX;
  // tiles ordered i fastest, then j, then column chunk; each XCD sweeps a
  // contiguous range so the y-halo rows a tile shares stay in one L2
  const int gxt = (ni + TX - 1) / TX, gyt = (nj + TY - 1) / TY;
  const int tile = xcd_tile((int)blockIdx.x, gxt * gyt * a.nchunks);
  const int i0 = (tile % gxt) * TX, j0 = ((tile / gxt) % gyt) * TY;
  const int i = i0 + tx, j = j0 + ty;
  const int kb = 1 + (tile / (gxt * gyt)) * a.kchunk;
  const int ke = min(nk, kb + a.kchunk - 1);
  const int nxs = sx(ni), nys = sy(nj), nxu = ux(ni), nyv = vy(nj);
  const bool live = (i < ni) && (j < nj);
  const unsigned obc = a.obc;
  const bool owned = live && (obc != 0u || (i >= 3 && i + 3 < ni && j >= 3 && j + 3 < nj));
  // an owned cell on an open side reads its four faces' mass fluxes for the outflow closure
  const bool edge = owned && obc != 0u && (i == 0 || i + 1 == ni || j == 0 || j + 1 == nj);
  const int x = i + NG, y = j + NG;
  const int64_t st = (int64_t)nys * nxs, ust = (int64_t)nys * nxu, vst = (int64_t)nyv * nxs;
  const int64_t mst = (int64_t)nj * ni;
  const uint32_t o_s = (uint32_t)(y * nxs + x) * 4u;  // padded scalar, w and v (south face) families
  const uint32_t o_u = (uint32_t)(y * nxu + x) * 4u;  // u family (west face)
  const uint32_t o_m = (uint32_t)(j * ni + i) * 4u;   // unpadded cells
  const uint32_t o_fu = (i <= ni && j < nj) ? o_u : 0u;  // the own west x face
  const uint32_t o_fv = (i < ni && j <= nj) ? o_s : 0u;  // the own south y face
  const bool xe = (ty == TY - 1), yn = (ty == TY - 2);   // the wavefronts forming the east / north faces
  uint32_t o_ex = 0u;
  if (xe && tx < TY && i0 + TX <= ni && j0 + tx < nj) o_ex = (uint32_t)((j0 + tx + NG) * nxu + i0 + TX + NG) * 4u;
  if (yn && j0 + TY <= nj && i0 + tx < ni) o_ex = (uint32_t)((j0 + TY + NG) * nxs + i0 + tx + NG) * 4u;
  // the plane's rows j0-3..j0+TY+2, columns i0-3..i0+TX+2 (padded indices
  // j0..j0+TY+5, i0..i0+TX+5, clipped to the padded array)
#pragma unroll
  for (int u = 0; u < PS; ++u) {
    const int slot = threadIdx.x + u * NT;
    const int r = slot / LQ, c = slot % LQ;
    const int yy = j0 + r, xx = i0 + c;
    uint32_t off = 0u;
    if (slot < (TY + 6) * LQ && yy < nys && xx < nxs) {
      if constexpr (SRC) {
        const int pj = family_cell(yy - NG, nj, side_of(obc, 2), side_of(obc, 3));
        const int pi = family_cell(xx - NG, ni, side_of(obc, 0), side_of(obc, 1));
        off = (uint32_t)(pj * ni + pi) * 4u;
      } else {
        off = (uint32_t)(yy * nxs + xx) * 4u;
      }
    }
    lpo[slot] = off;  // read back by this lane alone
  }

  // The column's scalar S(k-3..k+3) at level k rides in registers: win[q]
  // holds level k-3+q (0 outside the padded column; those words are never
  // read by a face's stencil, which stays within levels 1..nk).
  float win[7];
  auto colword = [&](int lev) {  // the column's word at a level, 0 where no stencil reads
    if constexpr (SRC) {
      // the own column is a physical cell, so its unpadded word needs no lateral image
      if (lev < 1 || lev > nk) return 0.0f;
      const float v = ldo(a.src + (int64_t)(lev - 1) * mst, o_m);
      return a.th0 ? (a.th0[lev - 1] - 300.0f) + v : v;
    }
    return (lev >= 0 && lev <= nk + 1) ? ldo(a.s + (int64_t)lev * st, o_s) : 0.0f;
  };
#pragma unroll
  for (int q = 0; q < 7; ++q) win[q] = live ? colword(kb - 3 + q) : 0.0f;
  // z-face flux at face kf from the window b, b[0] = S(kf-3).
  auto zf = [&](int kf, const float* b) {
    const float m = ldo(a.rrw + (int64_t)kf * st, o_s);
    const float* p1 = a.c1 + (int64_t)(kf - 1) * mst;
    const float* p2 = a.c2 + (int64_t)(kf - 1) * mst;
    const float cc1 = a.c12col ? p1[0] : ldo(p1, o_m), cc2 = a.c12col ? p2[0] : ldo(p2, o_m);
    auto S = [&](int level) { return b[level - (kf - 3)]; };
    if (MODE == MODE_PDEF) return zflux_pdef_g(S, m, cc1, cc2, kf, nk, a.weps, a.pdefweno != 0);
    return zflux_g<WENO>(S, m, cc1, cc2, kf, nk, a.weps);
  };
  float fzb = 0.0f;
  if (live && kb >= 2) fzb = zf(kb, win);  // S(kb-3..)
  const float gzc = (!FLAT && owned && MODE != MODE_PDEF) ? ldo(a.gz, o_m) : 0.0f;
  // A level's plane slots and own faces' mass fluxes are requested one level
  // ahead into registers, and its cell words at the top of the level, so they
  // arrive while the level before it (or its fluxes) is formed.
  struct Lv {
    float plane[PS];
    float mu, mv, mx;
    float snext;
  };
  auto fetch = [&](int k) {
    Lv v;
    v.snext = live ? colword(k + 4) : 0.0f;  // the level entering the window after level k
    const float* pl = SRC ? a.src + (int64_t)(k - 1) * mst : a.s + (int64_t)k * st;
    const float pack = (SRC && a.th0) ? a.th0[k - 1] - 300.0f : 0.0f;
#pragma unroll
    for (int u = 0; u < PS; ++u) {
      const float w = ldo(pl, lpo[threadIdx.x + u * NT]);
      v.plane[u] = (SRC && a.th0) ? pack + w : w;
    }
    const float* pu = a.rru + (int64_t)k * ust;
    const float* pv = a.rrv + (int64_t)k * vst;
    v.mu = ldo(pu, o_fu);
    v.mv = ldo(pv, o_fv);
    v.mx = xe ? ldo(pu, o_ex) : (yn ? ldo(pv, o_ex) : 0.0f);
    return v;
  };
  Lv nx = fetch(kb);
  for (int k = kb; k <= ke; ++k) {
    const Lv cu = nx;
    if (k < ke) nx = fetch(k + 1);
    const float snext = cu.snext;
    const int64_t lm = (int64_t)(k - 1) * mst;  // the unpadded families' level k
    // this level's cell words, consumed after its fluxes
    float c_uw = 0.0f, c_ue = 0.0f, c_vs = 0.0f, c_vn = 0.0f, c_divx = 0.0f, c_rr0 = 0.0f, c_a0 = 0.0f,
          c_held = 0.0f, c_divterm = 0.0f;
    if (owned) {
      if (MODE != MODE_PDEF) {
        c_divx = ldo(a.divx + lm, o_m);
        c_rr0 = FLAT ? a.rr0[lm] : ldo(a.rr0 + lm, o_m);
        if (MODE == MODE_COMPOSE) {
          c_a0 = ldo(a.a0 + lm, o_m);
          c_held = ldo(a.held + lm, o_m);
          if (a.divterm) c_divterm = ldo(a.divterm + lm, o_m);
        }
      }
      if (edge) {
        const float* pu = a.rru + (int64_t)k * ust;
        const float* pv = a.rrv + (int64_t)k * vst;
        c_uw = ldo(pu, o_u);
        c_ue = ldo(pu, o_u + 4u);
        c_vs = ldo(pv, o_s);
        c_vn = ldo(pv, o_s + (uint32_t)nxs * 4u);
      }
    }
#pragma unroll
    for (int u = 0; u < PS; ++u) lp[threadIdx.x + u * NT] = cu.plane[u];
    __syncthreads();
    {  // the own faces: x face i0+tx of row j0+ty, y face j0+ty of column i0+tx
      const float* q = lp + (ty + 3) * LQ + tx;
      const float fx = face_flux<WENO, GATE>(cu.mu, q[0], q[1], q[2], q[3], q[4], q[5], a.weps);
      lfx[ty][tx] = fx;
      const float* p = lp + ty * LQ + tx + 3;
      const float fy = face_flux<WENO, GATE>(cu.mv, p[0], p[LQ], p[2 * LQ], p[3 * LQ], p[4 * LQ], p[5 * LQ], a.weps);
      lfy[ty][tx] = fy;
      if (MODE == MODE_TENDENCY) {
        if (a.fx && i <= ni && j < nj) sto(a.fx + (int64_t)k * ust, o_u, fx);
        if (a.fy && i < ni && j <= nj) sto(a.fy + (int64_t)k * vst, o_s, fy);
      }
    }
    if (xe) {  // face i0+TX of row j0+tx
      if (tx < TY) {
        const float* q = lp + (tx + 3) * LQ + TX;
        const float fx = face_flux<WENO, GATE>(cu.mx, q[0], q[1], q[2], q[3], q[4], q[5], a.weps);
        lfx[tx][TX] = fx;
        // the domain's east face, which no tile forms as its own
        if (MODE == MODE_TENDENCY && a.fx && i0 + TX == ni && j0 + tx < nj) sto(a.fx + (int64_t)k * ust, o_ex, fx);
      }
    } else if (yn) {  // face j0+TY of column i0+tx
      const float* p = lp + TY * LQ + tx + 3;
      const float fy = face_flux<WENO, GATE>(cu.mx, p[0], p[LQ], p[2 * LQ], p[3 * LQ], p[4 * LQ], p[5 * LQ], a.weps);
      lfy[TY][tx] = fy;
      if (MODE == MODE_TENDENCY && a.fy && j0 + TY == nj && i0 + tx < ni) sto(a.fy + (int64_t)k * vst, o_ex, fy);
    }
    float fzt = 0.0f;
    if (live && k + 1 <= nk) fzt = zf(k + 1, win + 1);  // S(k-2..)
    __syncthreads();
    if (MODE == MODE_TENDENCY && live && a.fz) {
      if (k == 1) sto(a.fz + st, o_s, 0.0f);
      sto(a.fz + (int64_t)(k + 1) * st, o_s, (k + 1 <= nk) ? fzt : 0.0f);
      if (k == kb && kb >= 2) sto(a.fz + (int64_t)kb * st, o_s, fzb);
    }
    if (owned) {
      float fw = lfx[ty][tx], fe = lfx[ty][tx + 1];
      float fs = lfy[ty][tx], fn = lfy[ty + 1][tx];
      const float sv = lp[(ty + 3) * LQ + tx + 3];
      float ax, ay, az;
      if (obc != 0u) {
        const float uw = c_uw, ue = c_ue, vs = c_vs, vn = c_vn;  // zero off the edges, where no closure applies
        const bool iw = (i == 0) && side_of(obc, 0) == 2u && uw >= 0.0f;
        const bool ie = (i + 1 == ni) && side_of(obc, 1) == 2u && ue <= 0.0f;
        const bool js = (j == 0) && side_of(obc, 2) == 2u && vs >= 0.0f;
        const bool jn = (j + 1 == nj) && side_of(obc, 3) == 2u && vn <= 0.0f;
        if (iw) fw = fe;
        if (ie) fe = fw;
        if (js) fs = fn;
        if (jn) fn = fs;
        ax = FLAT ? (-(fe - fw)) * a.rdx : (fw - fe) * a.rdx;
        ay = FLAT ? (-(fn - fs)) * a.rdy : (fs - fn) * a.rdy;
        if (iw) ax = ax - (sv * (ue - uw)) * a.rdx;
        if (ie) ax = ax - (sv * (ue - uw)) * a.rdx;
        if (js) ay = ay - (sv * (vn - vs)) * a.rdy;
        if (jn) ay = ay - (sv * (vn - vs)) * a.rdy;
      } else {
        ax = FLAT ? (-(fe - fw)) * a.rdx : (fw - fe) * a.rdx;
        ay = FLAT ? (-(fn - fs)) * a.rdy : (fs - fn) * a.rdy;
      }
      if constexpr (FLAT) {
        if (k == 1) az = ((-fzt) * a.rdsf) * a.mh[k];
        else if (k == nk) az = (fzb * a.rdsf) * a.mh[k];
        else az = ((-(fzt - fzb)) * a.rdsf) * a.mh[k];
      } else {
        if (k == 1) az = (-fzt) * a.rdsf;
        else if (k == nk) az = fzb * a.rdsf;
        else az = (fzb - fzt) * a.rdsf;
      }
      if (MODE == MODE_PDEF) {
        sto(a.advx + lm, o_m, ax);
        sto(a.advy + lm, o_m, ay);
        sto(a.advz + lm, o_m, az);
      } else {
        float total = (ax + ay) + az;
        total = total + sv * c_divx;
        const float adv = FLAT ? total * c_rr0 : (total * c_rr0) * gzc;
        if (MODE == MODE_TENDENCY) {
          sto(a.tend + lm, o_m, adv);
        } else {
          const float seed = a.divterm ? c_held - c_divterm : c_held;
          const float v = c_a0 + a.dttmp * (seed + adv);
          sto(a.out + lm, o_m, (fabsf(v) < 1.0e-30f) ? 0.0f : v);
        }
      }
    }
    fzb = fzt;
#pragma unroll
    for (int q = 0; q < 6; ++q) win[q] = win[q + 1];
    win[6] = snext;
    __syncthreads();
  }
}


// Held to six waves per SIMD (three 512-lane workgroups per CU): more
// waves in flight hide more of each level's load latency.  Measured at
// 640x768x80 against the unconstrained register budget (four waves): the
// pdef tendencies 641 -> 585 us and the WENO compose 1025 -> 964 us, though
// both spill a few registers.
template <bool WENO, int MODE, bool FLAT, bool SRC>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(6))) void k_scalar_adv5(ScalarAdvArgs a) {
  scalar_adv5_body_lanes<WENO, MODE, FLAT, SRC>(a);
}

// The fifth-order compose of NS scalars read from their unpadded carriers in
// one pass: the mass fluxes, divx and the addressing are shared, each scalar
// keeps its own level plane, flux planes and vertical window.  Every scalar's
// words are those of its own single-scalar compose (the same statements, in
// the same order).
constexpr int NSMAX = 4;
struct ScalarBatch {
  const float* src[NSMAX];
  const float* th0[NSMAX];      // packs theta' as (th0 - 300) + theta', or null
  const float* a0[NSMAX];
  const float* held[NSMAX];
  const float* divterm[NSMAX];  // may be null
  float* out[NSMAX];
};

template <int NS, bool FLAT>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(4))) void k_scalar_compose_batch(ScalarAdvArgs a, ScalarBatch b) {
  constexpr int PS = ((TY + 6) * LQ + NT - 1) / NT;
  __shared__ float lfx[NS][TY][TX + 1];
  __shared__ float lfy[NS][TY + 1][TX];
  __shared__ float lp[NS][PS * NT];
  __shared__ uint32_t lpo[PS * NT];
  const int ni = a.ni, nj = a.nj, nk = a.nk;
  const int tx = threadIdx.x % TX, ty = threadIdx.x / TX;
  const int gxt = (ni + TX - 1) / TX, gyt = (nj + TY - 1) / TY;
  const int tile = xcd_tile((int)blockIdx.x, gxt * gyt * a.nchunks);
  const int i0 = (tile % gxt) * TX, j0 = ((tile / gxt) % gyt) * TY;
  const int i = i0 + tx, j = j0 + ty;
  const int kb = 1 + (tile / (gxt * gyt)) * a.kchunk;
  const int ke = min(nk, kb + a.kchunk - 1);
  const int nxs = sx(ni), nys = sy(nj), nxu = ux(ni), nyv = vy(nj);
  const bool live = (i < ni) && (j < nj);
  const unsigned obc = a.obc;
  const bool owned = live && (obc != 0u || (i >= 3 && i + 3 < ni && j >= 3 && j + 3 < nj));
  const bool edge = owned && obc != 0u && (i == 0 || i + 1 == ni || j == 0 || j + 1 == nj);
  const int x = i + NG, y = j + NG;
  const int64_t st = (int64_t)nys * nxs, ust = (int64_t)nys * nxu, vst = (int64_t)nyv * nxs;
  const int64_t mst = (int64_t)nj * ni;
  const uint32_t o_s = (uint32_t)(y * nxs + x) * 4u;
  const uint32_t o_u = (uint32_t)(y * nxu + x) * 4u;
  const uint32_t o_m = (uint32_t)(j * ni + i) * 4u;
  const uint32_t o_fu = (i <= ni && j < nj) ? o_u : 0u;
  const uint32_t o_fv = (i < ni && j <= nj) ? o_s : 0u;
  const bool xe = (ty == TY - 1), yn = (ty == TY - 2);
  uint32_t o_ex = 0u;
  if (xe && tx < TY && i0 + TX <= ni && j0 + tx < nj) o_ex = (uint32_t)((j0 + tx + NG) * nxu + i0 + TX + NG) * 4u;
  if (yn && j0 + TY <= nj && i0 + tx < ni) o_ex = (uint32_t)((j0 + TY + NG) * nxs + i0 + tx + NG) * 4u;
#pragma unroll
  for (int u = 0; u < PS; ++u) {
    const int slot = threadIdx.x + u * NT;
    const int r = slot / LQ, c = slot % LQ;
    const int yy = j0 + r, xx = i0 + c;
    uint32_t off = 0u;
    if (slot < (TY + 6) * LQ && yy < nys && xx < nxs) {
      const int pj = family_cell(yy - NG, nj, side_of(obc, 2), side_of(obc, 3));
      const int pi = family_cell(xx - NG, ni, side_of(obc, 0), side_of(obc, 1));
      off = (uint32_t)(pj * ni + pi) * 4u;
    }
    lpo[slot] = off;
  }
  auto colword = [&](int q, int lev) {  // scalar q's own column at a level, 0 where no stencil reads
    if (lev < 1 || lev > nk) return 0.0f;
    const float v = ldo(b.src[q] + (int64_t)(lev - 1) * mst, o_m);
    return b.th0[q] ? (b.th0[q][lev - 1] - 300.0f) + v : v;
  };
  float win[NS][7];
#pragma unroll
  for (int q = 0; q < NS; ++q)
#pragma unroll
    for (int l = 0; l < 7; ++l) win[q][l] = live ? colword(q, kb - 3 + l) : 0.0f;
  // the z face kf's mass flux and weights, shared by the scalars
  auto zmass = [&](int kf, float& m, float& cc1, float& cc2) {
    m = ldo(a.rrw + (int64_t)kf * st, o_s);
    const float* p1 = a.c1 + (int64_t)(kf - 1) * mst;
    const float* p2 = a.c2 + (int64_t)(kf - 1) * mst;
    cc1 = a.c12col ? p1[0] : ldo(p1, o_m);
    cc2 = a.c12col ? p2[0] : ldo(p2, o_m);
  };
  float fzb[NS];
#pragma unroll
  for (int q = 0; q < NS; ++q) fzb[q] = 0.0f;
  if (live && kb >= 2) {
    float m, cc1, cc2;
    zmass(kb, m, cc1, cc2);
#pragma unroll
    for (int q = 0; q < NS; ++q) {
      const float* w = win[q];
      fzb[q] = zflux_g<false>([&](int level) { return w[level - (kb - 3)]; }, m, cc1, cc2, kb, nk, a.weps);
    }
  }
  const float gzc = (!FLAT && owned) ? ldo(a.gz, o_m) : 0.0f;
  struct Lv {
    float plane[NS][PS];
    float snext[NS];
    float mu, mv, mx;
  };
  auto fetch = [&](int k) {
    Lv v;
#pragma unroll
    for (int q = 0; q < NS; ++q) {
      v.snext[q] = live ? colword(q, k + 4) : 0.0f;
      const float* pl = b.src[q] + (int64_t)(k - 1) * mst;
      const float pack = b.th0[q] ? b.th0[q][k - 1] - 300.0f : 0.0f;
#pragma unroll
      for (int u = 0; u < PS; ++u) {
        const float w = ldo(pl, lpo[threadIdx.x + u * NT]);
        v.plane[q][u] = b.th0[q] ? pack + w : w;
      }
    }
    const float* pu = a.rru + (int64_t)k * ust;
    const float* pv = a.rrv + (int64_t)k * vst;
    v.mu = ldo(pu, o_fu);
    v.mv = ldo(pv, o_fv);
    v.mx = xe ? ldo(pu, o_ex) : (yn ? ldo(pv, o_ex) : 0.0f);
    return v;
  };
  Lv nx = fetch(kb);
  for (int k = kb; k <= ke; ++k) {
    const Lv cu = nx;
    if (k < ke) nx = fetch(k + 1);
    const int64_t lm = (int64_t)(k - 1) * mst;
    float c_uw = 0.0f, c_ue = 0.0f, c_vs = 0.0f, c_vn = 0.0f, c_divx = 0.0f, c_rr0 = 0.0f;
    float c_a0[NS], c_held[NS], c_divterm[NS];
#pragma unroll
    for (int q = 0; q < NS; ++q) c_a0[q] = c_held[q] = c_divterm[q] = 0.0f;
    if (owned) {
      c_divx = ldo(a.divx + lm, o_m);
      c_rr0 = FLAT ? a.rr0[lm] : ldo(a.rr0 + lm, o_m);
#pragma unroll
      for (int q = 0; q < NS; ++q) {
        c_a0[q] = ldo(b.a0[q] + lm, o_m);
        c_held[q] = ldo(b.held[q] + lm, o_m);
        if (b.divterm[q]) c_divterm[q] = ldo(b.divterm[q] + lm, o_m);
      }
      if (edge) {
        const float* pu = a.rru + (int64_t)k * ust;
        const float* pv = a.rrv + (int64_t)k * vst;
        c_uw = ldo(pu, o_u);
        c_ue = ldo(pu, o_u + 4u);
        c_vs = ldo(pv, o_s);
        c_vn = ldo(pv, o_s + (uint32_t)nxs * 4u);
      }
    }
#pragma unroll
    for (int q = 0; q < NS; ++q)
#pragma unroll
      for (int u = 0; u < PS; ++u) lp[q][threadIdx.x + u * NT] = cu.plane[q][u];
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NS; ++q) {
      const float* P = lp[q];
      const float* xq = P + (ty + 3) * LQ + tx;
      lfx[q][ty][tx] = face_flux<false, false>(cu.mu, xq[0], xq[1], xq[2], xq[3], xq[4], xq[5], a.weps);
      const float* yq = P + ty * LQ + tx + 3;
      lfy[q][ty][tx] =
          face_flux<false, false>(cu.mv, yq[0], yq[LQ], yq[2 * LQ], yq[3 * LQ], yq[4 * LQ], yq[5 * LQ], a.weps);
      if (xe) {
        if (tx < TY) {
          const float* e = P + (tx + 3) * LQ + TX;
          lfx[q][tx][TX] = face_flux<false, false>(cu.mx, e[0], e[1], e[2], e[3], e[4], e[5], a.weps);
        }
      } else if (yn) {
        const float* e = P + TY * LQ + tx + 3;
        lfy[q][TY][tx] =
            face_flux<false, false>(cu.mx, e[0], e[LQ], e[2 * LQ], e[3 * LQ], e[4 * LQ], e[5 * LQ], a.weps);
      }
    }
    float fzt[NS];
#pragma unroll
    for (int q = 0; q < NS; ++q) fzt[q] = 0.0f;
    if (live && k + 1 <= nk) {
      float m, cc1, cc2;
      zmass(k + 1, m, cc1, cc2);
#pragma unroll
      for (int q = 0; q < NS; ++q) {
        const float* w = win[q] + 1;  // S(k-2..)
        fzt[q] = zflux_g<false>([&](int level) { return w[level - (k - 2)]; }, m, cc1, cc2, k + 1, nk, a.weps);
      }
    }
    __syncthreads();
    if (owned) {
#pragma unroll
      for (int q = 0; q < NS; ++q) {
        float fw = lfx[q][ty][tx], fe = lfx[q][ty][tx + 1];
        float fs = lfy[q][ty][tx], fn = lfy[q][ty + 1][tx];
        const float sv = lp[q][(ty + 3) * LQ + tx + 3];
        float ax, ay, az;
        if (obc != 0u) {
          const float uw = c_uw, ue = c_ue, vs = c_vs, vn = c_vn;  // zero off the edges, where no closure applies
          const bool iw = (i == 0) && side_of(obc, 0) == 2u && uw >= 0.0f;
          const bool ie = (i + 1 == ni) && side_of(obc, 1) == 2u && ue <= 0.0f;
          const bool js = (j == 0) && side_of(obc, 2) == 2u && vs >= 0.0f;
          const bool jn = (j + 1 == nj) && side_of(obc, 3) == 2u && vn <= 0.0f;
          if (iw) fw = fe;
          if (ie) fe = fw;
          if (js) fs = fn;
          if (jn) fn = fs;
          ax = FLAT ? (-(fe - fw)) * a.rdx : (fw - fe) * a.rdx;
          ay = FLAT ? (-(fn - fs)) * a.rdy : (fs - fn) * a.rdy;
          if (iw) ax = ax - (sv * (ue - uw)) * a.rdx;
          if (ie) ax = ax - (sv * (ue - uw)) * a.rdx;
          if (js) ay = ay - (sv * (vn - vs)) * a.rdy;
          if (jn) ay = ay - (sv * (vn - vs)) * a.rdy;
        } else {
          ax = FLAT ? (-(fe - fw)) * a.rdx : (fw - fe) * a.rdx;
          ay = FLAT ? (-(fn - fs)) * a.rdy : (fs - fn) * a.rdy;
        }
        if constexpr (FLAT) {
          if (k == 1) az = ((-fzt[q]) * a.rdsf) * a.mh[k];
          else if (k == nk) az = (fzb[q] * a.rdsf) * a.mh[k];
          else az = ((-(fzt[q] - fzb[q])) * a.rdsf) * a.mh[k];
        } else {
          if (k == 1) az = (-fzt[q]) * a.rdsf;
          else if (k == nk) az = fzb[q] * a.rdsf;
          else az = (fzb[q] - fzt[q]) * a.rdsf;
        }
        float total = (ax + ay) + az;
        total = total + sv * c_divx;
        const float adv = FLAT ? total * c_rr0 : (total * c_rr0) * gzc;
        const float seed = b.divterm[q] ? c_held[q] - c_divterm[q] : c_held[q];
        const float v = c_a0[q] + a.dttmp * (seed + adv);
        sto(b.out[q] + lm, o_m, (fabsf(v) < 1.0e-30f) ? 0.0f : v);
      }
    }
#pragma unroll
    for (int q = 0; q < NS; ++q) {
      fzb[q] = fzt[q];
#pragma unroll
      for (int l = 0; l < 6; ++l) win[q][l] = win[q][l + 1];
      win[q][6] = cu.snext[q];
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// positive-definite repair of the final-stage moisture advection
// (cm1_rk_pdef_candidate_{x,y,z}_v1 / correct_{x,y,z}_v1 / total_v1, pdef.F)
// fused with the A-anchored compose.  A 64x8 tile marches up a column chunk;
// per level the x candidates of the tile's rows (+-2 cells, the row extended
// past the domain by the lateral family) and the y candidates of its columns
// (+-2 rows) are formed once into LDS; the z candidates of each column slide
// through a five-level register window (zero beyond the column).
// ---------------------------------------------------------------------------
constexpr int RX = 64;
constexpr int RY = 8;
constexpr int RT = RX * RY;

struct PdefArgs {
  const float* s;  // padded scalar (current values)
  const float* rru; const float* rrv; const float* rrw;
  const float* rho0; const float* rr0; const float* gz; const float* rgz;
  const float* divx;
  const float* qa; const float* held;
  const float* advx; const float* advy; const float* advz;
  float* out;
  int ni, nj, nk, kchunk, nchunks;
  float dt, rdt, rdx, rdy, rdz;
  unsigned obc;
  // Flat stretched column (pdef.F:37-47, 263-275, 477-486, 548-557 with
  // terrain_flag=.false.): mh (nk+2) or null; candidates rho0(1,1,k)*s0 +
  // dt*(adv + div) without gz, corrections without rgz, vertical
  // ((s*d)*rdz)*mh(k), the total scaled by rr0(1,1,k).
  const float* mh;
  const float* src;  // unpadded source (nk, nj, ni) in place of s, or null
};

// A candidate of the directional repair from its cell's words (pdef.F's
// x/y candidate): rho0*qa, and unless an open edge's face pair points
// inward, + dt*(adv + s*(m_hi - m_lo)*rd) (times gz on the terrain family).
template <bool FLAT>
GM1_DEV float pdef_cand(float r0a, float adv, float sv, float mlo, float mhi, float rd, float dt, float gzv,
                        bool lo_edge, bool hi_edge) {
  if ((lo_edge && mlo >= 0.0f) || (hi_edge && mhi <= 0.0f)) return r0a;
  const float d = mhi - mlo;
  if constexpr (FLAT) return r0a + dt * (adv + (sv * d) * rd);
  return r0a + (dt * (adv + (sv * d) * rd)) * gzv;
}

// ratio of the donor repair at the middle of three consecutive values
GM1_DEV float donor_ratio(float lo, float c, float hi) {
  if (c >= 0.0f) return 0.0f;
  const float sum = fmaxf(0.0f, lo) + fmaxf(0.0f, hi);
  return (sum > 1.0e-30f) ? fmaxf(c, -sum) / sum : 0.0f;
}

// mass change and flag of the centre of five consecutive values v[0..4]
GM1_DEV void row_mass(const float v[5], float& mass, bool& flag) {
  mass = 0.0f;
  flag = false;
  const float c = v[2];
  if (c < 0.0f) {
    const float r = donor_ratio(v[1], c, v[3]);
    if (r != 0.0f) {
      const float sum = fmaxf(0.0f, v[1]) + fmaxf(0.0f, v[3]);
      mass = 0.0f - sum * r;
      flag = true;
    }
  } else if (c > 1.0e-30f) {
    const float rl = donor_ratio(v[0], v[1], v[2]), rr = donor_ratio(v[2], v[3], v[4]);
    if (rl != 0.0f) { mass = mass + c * rl; flag = true; }
    if (rr != 0.0f) { mass = mass + c * rr; flag = true; }
  }
}

// The repair pass: each lane forms the x and y candidates of its own cell
// (a lane in the two columns east of the domain or the two rows north of it,
// those of the lateral family's image), wavefronts 0-3 the tile's south and
// north halo rows' y candidates and wavefront 4 the west and east halo
// columns' x candidates; the z candidates ride in a five-level register
// window.  Every word a level needs is requested one level ahead.
template <bool FLAT>
__global__ __launch_bounds__(RT) void k_pdef_repair(PdefArgs a) {
  __shared__ float lcx[RY][RX + 4];
  __shared__ float lcy[RY + 4][RX];
  const int ni = a.ni, nj = a.nj, nk = a.nk;
  const int tx = threadIdx.x % RX, ty = threadIdx.x / RX;
  const int gxt = (ni + RX - 1) / RX, gyt = (nj + RY - 1) / RY;
  const int tile = xcd_tile((int)blockIdx.x, gxt * gyt * a.nchunks);
  const int i0 = (tile % gxt) * RX, j0 = ((tile / gxt) % gyt) * RY;
  const int i = i0 + tx, j = j0 + ty;
  const bool live = (i < ni) && (j < nj);
  const int kb = 1 + (tile / (gxt * gyt)) * a.kchunk;
  const int ke = min(nk, kb + a.kchunk - 1);
  const unsigned wbc = side_of(a.obc, 0), ebc = side_of(a.obc, 1), sbc = side_of(a.obc, 2), nbc = side_of(a.obc, 3);
  const int nxs = sx(ni), nys = sy(nj), nxu = ux(ni), nyv = vy(nj);
  const int64_t st = (int64_t)nys * nxs, ust = (int64_t)nys * nxu, vst = (int64_t)nyv * nxs, mst = (int64_t)nj * ni;
  // a cell's words: unpadded (m), padded scalar / v / w families (s), u family (u)
  struct Cell {
    uint32_t m, s, u;
  };
  auto cell = [&](int ci, int cj) {
    Cell c;
    c.m = (uint32_t)(cj * ni + ci) * 4u;
    c.s = (uint32_t)((cj + NG) * nxs + ci + NG) * 4u;
    c.u = (uint32_t)((cj + NG) * nxu + ci + NG) * 4u;
    return c;
  };
  auto word = [&](int k, const Cell& c) {  // the scalar's current value
    return a.src ? ldo(a.src + (int64_t)(k - 1) * mst, c.m) : ldo(a.s + (int64_t)k * st, c.s);
  };
  auto rho_at = [&](int k, const Cell& c) {
    return FLAT ? a.rho0[(int64_t)(k - 1) * mst] : ldo(a.rho0 + (int64_t)(k - 1) * mst, c.m);
  };
  // the lane's candidate cell
  const bool xc = (j < nj) && (i <= ni + 1);  // x candidate at lcx[ty][tx + 2]
  const bool yc = (i < ni) && (j <= nj + 1);  // y candidate at lcy[ty + 2][tx]
  const int ci = (i < ni) ? i : family_cell(i, ni, wbc, ebc);
  const int cj = (j < nj) ? j : family_cell(j, nj, sbc, nbc);
  const Cell oc = cell((xc || yc) ? ci : 0, (xc || yc) ? cj : 0);
  const bool x_lo = ci == 0 && wbc == 2u, x_hi = ci + 1 == ni && ebc == 2u;
  const bool y_lo = cj == 0 && sbc == 2u, y_hi = cj + 1 == nj && nbc == 2u;
  const float gzc = (!FLAT && (xc || yc)) ? ldo(a.gz, oc.m) : 1.0f;
  // the lane's halo candidate: hk 1 (x, lcx[hr][ht]) or 2 (y, lcy[hr][ht])
  int hk = 0, hr = 0, ht = 0, hi = 0, hj = 0;
  if (ty == 4 && tx < 4 * RY) {
    hr = tx >> 2;
    ht = (tx & 3) < 2 ? (tx & 3) : RX + (tx & 3);
    const int c = i0 - 2 + ht, jj = j0 + hr;
    if (jj < nj && c >= -2 && c <= ni + 1) {
      hk = 1;
      hi = family_cell(c, ni, wbc, ebc);
      hj = jj;
    }
  } else if (ty < 4) {
    hr = ty < 2 ? ty : RY + ty;
    ht = tx;
    const int c = j0 - 2 + hr, ii = i0 + tx;
    if (ii < ni && c >= -2 && c <= nj + 1) {
      hk = 2;
      hi = ii;
      hj = family_cell(c, nj, sbc, nbc);
    }
  }
  const Cell hc = cell(hi, hj);
  const bool h_lo = hk == 1 ? (hi == 0 && wbc == 2u) : (hj == 0 && sbc == 2u);
  const bool h_hi = hk == 1 ? (hi + 1 == ni && ebc == 2u) : (hj + 1 == nj && nbc == 2u);
  const float hgz = (!FLAT && hk) ? ldo(a.gz, hc.m) : 1.0f;
  // z candidate of the own column at level k (0 outside 1..nk) from its words
  auto zcand = [&](int k, float rho, float qa, float az, float sv, float wb, float wt) {
    if (k < 1 || k > nk) return 0.0f;
    const float d = wt - wb;
    if constexpr (FLAT) return rho * qa + a.dt * (az + ((sv * d) * a.rdz) * a.mh[k]);
    return rho * qa + (a.dt * (az + (sv * d) * a.rdz)) * gzc;
  };
  auto zwords = [&](int k) {  // the own column's z candidate at level k from memory
    if (!live || k < 1 || k > nk) return 0.0f;
    const int64_t lm = (int64_t)(k - 1) * mst;
    return zcand(k, rho_at(k, oc), ldo(a.qa + lm, oc.m), ldo(a.advz + lm, oc.m), word(k, oc),
                 ldo(a.rrw + (int64_t)k * st, oc.s), ldo(a.rrw + (int64_t)(k + 1) * st, oc.s));
  };
  float cz[5] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f};  // candidates at k-2..k+2
  for (int d = 0; d < 4; ++d) cz[d + 1] = zwords(kb - 2 + d);
  struct Lv {
    float qa, ax, ay, az, sv, divx, held, uw, ue, vs, vn, wb, wt, rho, rr0;  // the own cell at level k
    float zqa, zaz, zsv, zwb, zwt, zrho;                                      // the own column at level k+2
    float hqa, hadv, hsv, hm0, hm1, hrho;                                     // the halo cell at level k
  };
  auto fetch = [&](int k) {
    Lv v{};
    const int64_t lm = (int64_t)(k - 1) * mst;
    if (xc || yc) {
      v.qa = ldo(a.qa + lm, oc.m);
      v.sv = word(k, oc);
      v.rho = rho_at(k, oc);
      if (xc) {
        v.ax = ldo(a.advx + lm, oc.m);
        v.uw = ldo(a.rru + (int64_t)k * ust, oc.u);
        v.ue = ldo(a.rru + (int64_t)k * ust, oc.u + 4u);
      }
      if (yc) {
        v.ay = ldo(a.advy + lm, oc.m);
        v.vs = ldo(a.rrv + (int64_t)k * vst, oc.s);
        v.vn = ldo(a.rrv + (int64_t)k * vst, oc.s + (uint32_t)nxs * 4u);
      }
    }
    if (live) {
      v.az = ldo(a.advz + lm, oc.m);
      v.divx = ldo(a.divx + lm, oc.m);
      v.held = ldo(a.held + lm, oc.m);
      v.wb = ldo(a.rrw + (int64_t)k * st, oc.s);
      v.wt = ldo(a.rrw + (int64_t)(k + 1) * st, oc.s);
      v.rr0 = FLAT ? a.rr0[lm] : ldo(a.rr0 + lm, oc.m);
      const int kz = k + 2;
      if (kz <= nk) {
        const int64_t lz = (int64_t)(kz - 1) * mst;
        v.zqa = ldo(a.qa + lz, oc.m);
        v.zaz = ldo(a.advz + lz, oc.m);
        v.zsv = word(kz, oc);
        v.zwb = ldo(a.rrw + (int64_t)kz * st, oc.s);
        v.zwt = ldo(a.rrw + (int64_t)(kz + 1) * st, oc.s);
        v.zrho = rho_at(kz, oc);
      }
    }
    if (hk) {
      v.hqa = ldo(a.qa + lm, hc.m);
      v.hsv = word(k, hc);
      v.hrho = rho_at(k, hc);
      if (hk == 1) {
        v.hadv = ldo(a.advx + lm, hc.m);
        v.hm0 = ldo(a.rru + (int64_t)k * ust, hc.u);
        v.hm1 = ldo(a.rru + (int64_t)k * ust, hc.u + 4u);
      } else {
        v.hadv = ldo(a.advy + lm, hc.m);
        v.hm0 = ldo(a.rrv + (int64_t)k * vst, hc.s);
        v.hm1 = ldo(a.rrv + (int64_t)k * vst, hc.s + (uint32_t)nxs * 4u);
      }
    }
    return v;
  };
  Lv nx = fetch(kb);
  for (int k = kb; k <= ke; ++k) {
    const Lv cu = nx;
    if (k < ke) nx = fetch(k + 1);
    if (xc) lcx[ty][tx + 2] = pdef_cand<FLAT>(cu.rho * cu.qa, cu.ax, cu.sv, cu.uw, cu.ue, a.rdx, a.dt, gzc, x_lo, x_hi);
    if (yc) lcy[ty + 2][tx] = pdef_cand<FLAT>(cu.rho * cu.qa, cu.ay, cu.sv, cu.vs, cu.vn, a.rdy, a.dt, gzc, y_lo, y_hi);
    if (hk) {
      const float c = pdef_cand<FLAT>(cu.hrho * cu.hqa, cu.hadv, cu.hsv, cu.hm0, cu.hm1, hk == 1 ? a.rdx : a.rdy, a.dt,
                                      hgz, h_lo, h_hi);
      if (hk == 1) lcx[hr][ht] = c;
      else lcy[hr][ht] = c;
    }
    for (int d = 0; d < 4; ++d) cz[d] = cz[d + 1];
    cz[4] = live ? zcand(k + 2, cu.zrho, cu.zqa, cu.zaz, cu.zsv, cu.zwb, cu.zwt) : 0.0f;
    __syncthreads();
    if (live) {
      const int64_t lm = (int64_t)(k - 1) * mst;
      const float sv = cu.sv;
      constexpr bool flat = FLAT;
      const float r0a = cu.rho * cu.qa;
      const float rg = flat ? 1.0f : ldo(a.rgz, oc.m);
      float ax = cu.ax, ay = cu.ay, az = cu.az;
      float v[5], mass;
      bool flag;
      for (int d = 0; d < 5; ++d) v[d] = lcx[ty][tx + d];
      row_mass(v, mass, flag);
      if (flag) {
        const float du = cu.ue - cu.uw;
        const float first = flat ? ((v[2] + mass) - r0a) * a.rdt : (((v[2] + mass) - r0a) * a.rdt) * rg;
        ax = first - (sv * du) * a.rdx;
      }
      for (int d = 0; d < 5; ++d) v[d] = lcy[ty + d][tx];
      row_mass(v, mass, flag);
      if (flag) {
        const float dv = cu.vn - cu.vs;
        const float first = flat ? ((v[2] + mass) - r0a) * a.rdt : (((v[2] + mass) - r0a) * a.rdt) * rg;
        ay = first - (sv * dv) * a.rdy;
      }
      row_mass(cz, mass, flag);
      if (flag) {
        const float dw = cu.wt - cu.wb;
        if (flat) az = ((cz[2] + mass) - r0a) * a.rdt - ((sv * dw) * a.rdz) * a.mh[k];
        else az = (((cz[2] + mass) - r0a) * a.rdt) * rg - (sv * dw) * a.rdz;
      }
      float t = (ax + ay) + az;
      t = t + sv * cu.divx;
      const float adv = flat ? t * cu.rr0 : (t * cu.rr0) * gzc;
      const float o = cu.qa + a.dt * (cu.held + adv);
      sto(a.out + lm, oc.m, (fabsf(o) < 1.0e-30f) ? 0.0f : o);
    }
    __syncthreads();
  }
}

int chunks_for(int tiles, int nk, int* kchunk) {
  // enough column chunks to put >= 1024 tiles in flight on 256 CUs (measured at
  // 640x768x80 against 2048 and 4096: the paired compose 764 -> 744 us, the
  // pdef repair 521 -> 508 us)
  int chunks = (1024 + tiles - 1) / tiles;
  if (chunks > nk) chunks = nk;
  if (chunks < 1) chunks = 1;
  *kchunk = (nk + chunks - 1) / chunks;
  return (nk + *kchunk - 1) / *kchunk;
}

template <bool FLAT, bool SRC>
void launch_scalar_variant(const ScalarAdvArgs& a, int use_weno, int mode, dim3 grid, hipStream_t s) {
  if (mode == MODE_TENDENCY) {
    if (use_weno)
      GM1_TIMED(TK_SCALAR_ADV5, s,
                hipLaunchKernelGGL((k_scalar_adv5<true, MODE_TENDENCY, FLAT, SRC>), grid, dim3(NT), 0, s, a));
    else
      GM1_TIMED(TK_SCALAR_ADV5, s,
                hipLaunchKernelGGL((k_scalar_adv5<false, MODE_TENDENCY, FLAT, SRC>), grid, dim3(NT), 0, s, a));
  } else if (mode == MODE_COMPOSE) {
    if (use_weno)
      GM1_TIMED(TK_SCALAR_ADV5, s,
                hipLaunchKernelGGL((k_scalar_adv5<true, MODE_COMPOSE, FLAT, SRC>), grid, dim3(NT), 0, s, a));
    else
      GM1_TIMED(TK_SCALAR_ADV5, s,
                hipLaunchKernelGGL((k_scalar_adv5<false, MODE_COMPOSE, FLAT, SRC>), grid, dim3(NT), 0, s, a));
  } else {
    GM1_TIMED(TK_SCALAR_ADV5, s,
              hipLaunchKernelGGL((k_scalar_adv5<true, MODE_PDEF, FLAT, SRC>), grid, dim3(NT), 0, s, a));
  }
}

int launch_scalar_tile(ScalarAdvArgs& a, int use_weno, int mode, hipStream_t s) {
  const int tiles = ((a.ni + TX - 1) / TX) * ((a.nj + TY - 1) / TY);
  const int chunks = chunks_for(tiles, a.nk, &a.kchunk);
  a.nchunks = chunks;
  dim3 grid(tiles * chunks);
  if (a.mh) {
    if (a.src) launch_scalar_variant<true, true>(a, use_weno, mode, grid, s);
    else launch_scalar_variant<true, false>(a, use_weno, mode, grid, s);
  } else {
    if (a.src) launch_scalar_variant<false, true>(a, use_weno, mode, grid, s);
    else launch_scalar_variant<false, false>(a, use_weno, mode, grid, s);
  }
  GM1_LAUNCH_CHECK();
  return 0;
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
  ScalarAdvArgs a{};
  a.s = scalar; a.rru = rru; a.rrv = rrv; a.rrw = rrw; a.c1 = c1; a.c2 = c2;
  a.divx = divx; a.rr0 = rr0; a.gz = gz; a.tend = tendency;
  a.fx = flux_x; a.fy = flux_y; a.fz = flux_z;
  a.ni = ni; a.nj = nj; a.nk = nk;
  a.rdx = 1.0f / dx; a.rdy = 1.0f / dy; a.rdsf = 1.0f / dz;
  a.weps = weps; set_obc(a, obc);
  return launch_scalar_tile(a, use_weno, MODE_TENDENCY, (hipStream_t)stream);
}

int gm1hip_scalar_rk_stage(const float* scalar, const float* rru, const float* rrv, const float* rrw,
                           const float* c1, const float* c2, const float* rr0, const float* gz, const float* divx,
                           const float* a0, const float* held, const float* divterm, float* out, int ni, int nj,
                           int nk, int use_weno, double weps, float dx, float dy, float dz, unsigned obc,
                           float dttmp, void* stream) {
  if ((obc & 0xFFFFu) == 0u) return (int)hipErrorInvalidValue;
  ScalarAdvArgs a{};
  a.s = scalar; a.rru = rru; a.rrv = rrv; a.rrw = rrw; a.c1 = c1; a.c2 = c2;
  a.divx = divx; a.rr0 = rr0; a.gz = gz;
  a.a0 = a0; a.held = held; a.divterm = divterm; a.out = out; a.dttmp = dttmp;
  a.ni = ni; a.nj = nj; a.nk = nk;
  a.rdx = 1.0f / dx; a.rdy = 1.0f / dy; a.rdsf = 1.0f / dz;
  a.weps = weps; set_obc(a, obc);
  return launch_scalar_tile(a, use_weno, MODE_COMPOSE, (hipStream_t)stream);
}

int gm1hip_scalar_rk_stage_flat(const float* scalar, const float* rru, const float* rrv, const float* rrw,
                                const float* c1, const float* c2, const float* rr0, const float* flat_column,
                                const float* divx, const float* a0, const float* held, const float* divterm, float* out,
                                int ni, int nj, int nk, int use_weno, double weps, float dx, float dy, float rdz,
                                unsigned obc, float dttmp, void* stream) {
  if ((obc & 0xFFFFu) == 0u || !flat_column) return (int)hipErrorInvalidValue;
  ScalarAdvArgs a{};
  a.s = scalar; a.rru = rru; a.rrv = rrv; a.rrw = rrw; a.c1 = c1; a.c2 = c2;
  a.divx = divx; a.rr0 = rr0; a.gz = nullptr; a.mh = flat_column;
  a.a0 = a0; a.held = held; a.divterm = divterm; a.out = out; a.dttmp = dttmp;
  a.ni = ni; a.nj = nj; a.nk = nk;
  a.rdx = 1.0f / dx; a.rdy = 1.0f / dy; a.rdsf = rdz;
  a.weps = weps; set_obc(a, obc);
  return launch_scalar_tile(a, use_weno, MODE_COMPOSE, (hipStream_t)stream);
}

static int pdef_stage(const float* scalar, const float* src, const float* rru, const float* rrv, const float* rrw,
                      const float* c1, const float* c2, const float* rho0_3d, const float* rr0, const float* gz,
                      const float* rgz, const float* mh, const float* divx, const float* qa, const float* held, float* out,
                      float* advx, float* advy, float* advz, int ni, int nj, int nk, int pdefweno, double weps,
                      float dx, float dy, float rdzv, unsigned obc, float dttmp, void* stream) {
  if ((obc & 0xFFFFu) == 0u || nk < 4) return (int)hipErrorInvalidValue;
  hipStream_t s = (hipStream_t)stream;
  ScalarAdvArgs a{};
  a.s = scalar; a.src = src; a.rru = rru; a.rrv = rrv; a.rrw = rrw; a.c1 = c1; a.c2 = c2;
  a.divx = divx; a.rr0 = rr0; a.gz = gz; a.mh = mh;
  a.advx = advx; a.advy = advy; a.advz = advz; a.pdefweno = pdefweno;
  a.ni = ni; a.nj = nj; a.nk = nk;
  a.rdx = 1.0f / dx; a.rdy = 1.0f / dy; a.rdsf = rdzv;
  a.weps = weps; set_obc(a, obc);
  int rc = launch_scalar_tile(a, 1, MODE_PDEF, s);
  if (rc) return rc;
  PdefArgs p{};
  p.s = scalar; p.src = src; p.rru = rru; p.rrv = rrv; p.rrw = rrw;
  p.rho0 = rho0_3d; p.rr0 = rr0; p.gz = gz; p.rgz = rgz; p.divx = divx; p.mh = mh;
  p.qa = qa; p.held = held; p.advx = advx; p.advy = advy; p.advz = advz; p.out = out;
  p.ni = ni; p.nj = nj; p.nk = nk;
  p.dt = dttmp; p.rdt = 1.0f / dttmp; p.rdx = 1.0f / dx; p.rdy = 1.0f / dy; p.rdz = rdzv;
  p.obc = obc & 0xFFFFu;
  const int tiles = ((ni + RX - 1) / RX) * ((nj + RY - 1) / RY);
  const int chunks = chunks_for(tiles, nk, &p.kchunk);
  p.nchunks = chunks;
  if (p.mh) hipLaunchKernelGGL(k_pdef_repair<true>, dim3(tiles * chunks), dim3(RT), 0, s, p);
  else hipLaunchKernelGGL(k_pdef_repair<false>, dim3(tiles * chunks), dim3(RT), 0, s, p);
  GM1_LAUNCH_CHECK();
  return 0;
}

int gm1hip_pdef_moisture_stage(const float* scalar, const float* rru, const float* rrv, const float* rrw,
                               const float* c1, const float* c2, const float* rho0_3d, const float* rr0,
                               const float* gz, const float* rgz, const float* divx, const float* qa,
                               const float* held, float* out, float* advx, float* advy, float* advz, int ni,
                               int nj, int nk, int pdefweno, double weps, float dx, float dy, float dz,
                               unsigned obc, float dttmp, void* stream) {
  return pdef_stage(scalar, nullptr, rru, rrv, rrw, c1, c2, rho0_3d, rr0, gz, rgz, nullptr, divx, qa, held, out,
                    advx, advy, advz, ni, nj, nk, pdefweno, weps, dx, dy, 1.0f / dz, obc, dttmp, stream);
}

int gm1hip_pdef_moisture_stage_flat(const float* scalar, const float* rru, const float* rrv, const float* rrw,
                                    const float* c1, const float* c2, const float* rho0_3d, const float* rr0,
                                    const float* flat_column, const float* divx, const float* qa, const float* held,
                                    float* out, float* advx, float* advy, float* advz, int ni, int nj, int nk,
                                    int pdefweno, double weps, float dx, float dy, float rdz, unsigned obc,
                                    float dttmp, void* stream) {
  if (!flat_column) return (int)hipErrorInvalidValue;
  return pdef_stage(scalar, nullptr, rru, rrv, rrw, c1, c2, rho0_3d, rr0, nullptr, nullptr, flat_column, divx, qa,
                    held, out, advx, advy, advz, ni, nj, nk, pdefweno, weps, dx, dy, rdz, obc, dttmp, stream);
}

// The two calls above reading the scalar from its unpadded carrier (theta'
// with th0, or one moisture species) with the pack and the lateral family
// fill folded into the tiles' loads.  Terrain columns pass gz (and rgz) and
// rdz = 1/dz; flat stretched columns pass flat_column and the metric rdz.
int gm1hip_scalar_rk_stage_src(const float* src, const float* th0, const float* rru, const float* rrv,
                               const float* rrw, const float* c1, const float* c2, const float* rr0,
                               const float* gz, const float* flat_column, const float* divx, const float* a0,
                               const float* held, const float* divterm, float* out, int ni, int nj, int nk,
                               int use_weno, double weps, float dx, float dy, float rdz, unsigned obc, float dttmp,
                               void* stream) {
  if ((obc & 0xFFFFu) == 0u || !src || (gz == nullptr) == (flat_column == nullptr)) return (int)hipErrorInvalidValue;
  ScalarAdvArgs a{};
  a.src = src; a.th0 = th0; a.rru = rru; a.rrv = rrv; a.rrw = rrw; a.c1 = c1; a.c2 = c2;
  a.divx = divx; a.rr0 = rr0; a.gz = gz; a.mh = flat_column;
  a.a0 = a0; a.held = held; a.divterm = divterm; a.out = out; a.dttmp = dttmp;
  a.ni = ni; a.nj = nj; a.nk = nk;
  a.rdx = 1.0f / dx; a.rdy = 1.0f / dy; a.rdsf = rdz;
  a.weps = weps; set_obc(a, obc);
  return launch_scalar_tile(a, use_weno, MODE_COMPOSE, (hipStream_t)stream);
}

int gm1hip_scalar_rk_stage_src_batch(int ns, const float* const* srcs, const float* const* th0s, const float* rru,
                                     const float* rrv, const float* rrw, const float* c1, const float* c2,
                                     const float* rr0, const float* gz, const float* flat_column, const float* divx,
                                     const float* const* a0s, const float* const* helds,
                                     const float* const* divterms, float* const* outs, int ni, int nj, int nk,
                                     float dx, float dy, float rdz, unsigned obc, float dttmp, void* stream) {
  if (ns < 1 || ns > NSMAX || !srcs || !a0s || !helds || !outs || (obc & 0xFFFFu) == 0u ||
      (gz == nullptr) == (flat_column == nullptr))
    return (int)hipErrorInvalidValue;
  ScalarAdvArgs a{};
  a.rru = rru; a.rrv = rrv; a.rrw = rrw; a.c1 = c1; a.c2 = c2;
  a.divx = divx; a.rr0 = rr0; a.gz = gz; a.mh = flat_column; a.dttmp = dttmp;
  a.ni = ni; a.nj = nj; a.nk = nk;
  a.rdx = 1.0f / dx; a.rdy = 1.0f / dy; a.rdsf = rdz;
  a.weps = 0.0; set_obc(a, obc);
  ScalarBatch b{};
  for (int q = 0; q < ns; ++q) {
    if (!srcs[q] || !a0s[q] || !helds[q] || !outs[q]) return (int)hipErrorInvalidValue;
    b.src[q] = srcs[q]; b.th0[q] = th0s ? th0s[q] : nullptr; b.a0[q] = a0s[q]; b.held[q] = helds[q];
    b.divterm[q] = divterms ? divterms[q] : nullptr; b.out[q] = outs[q];
  }
  const int tiles = ((ni + TX - 1) / TX) * ((nj + TY - 1) / TY);
  a.nchunks = chunks_for(tiles, nk, &a.kchunk);
  const dim3 grid(tiles * a.nchunks);
  hipStream_t s = (hipStream_t)stream;
  const bool flat = flat_column != nullptr;
#define GM1_BATCH(N)                                                                                           \
  do {                                                                                                         \
    if (flat)                                                                                                  \
      GM1_TIMED(TK_SCALAR_ADV5, s, hipLaunchKernelGGL((k_scalar_compose_batch<N, true>), grid, dim3(NT), 0, s, a, b)); \
    else                                                                                                       \
      GM1_TIMED(TK_SCALAR_ADV5, s, hipLaunchKernelGGL((k_scalar_compose_batch<N, false>), grid, dim3(NT), 0, s, a, b)); \
  } while (0)
  switch (ns) {
    case 1: GM1_BATCH(1); break;
    case 2: GM1_BATCH(2); break;
    case 3: GM1_BATCH(3); break;
    default: GM1_BATCH(4); break;
  }
#undef GM1_BATCH
  GM1_LAUNCH_CHECK();
  return 0;
}

int gm1hip_pdef_moisture_stage_src(const float* src, const float* rru, const float* rrv, const float* rrw,
                                   const float* c1, const float* c2, const float* rho0_3d, const float* rr0,
                                   const float* gz, const float* rgz, const float* flat_column, const float* divx,
                                   const float* qa, const float* held, float* out, float* advx, float* advy,
                                   float* advz, int ni, int nj, int nk, int pdefweno, double weps, float dx, float dy,
                                   float rdz, unsigned obc, float dttmp, void* stream) {
  if (!src || (gz == nullptr) == (flat_column == nullptr) || (gz != nullptr) != (rgz != nullptr))
    return (int)hipErrorInvalidValue;
  return pdef_stage(nullptr, src, rru, rrv, rrw, c1, c2, rho0_3d, rr0, gz, rgz, flat_column, divx, qa, held, out,
                    advx, advy, advz, ni, nj, nk, pdefweno, weps, dx, dy, rdz, obc, dttmp, stream);
}

}  // extern "C"
