// Kessler warm rain (CM1 ptype=1, efall=0 or 1) and its large-step transition on gfx950.
//
// Replaces cuda_gm1_kessler_driver.py / cuda_gm1_kessler_local.py /
// cuda_gm1_kessler_large_step.py / cuda_apmasscon.py.  Design:
//   * k_kes_column: ONE thread per column does everything that is serial in
//     k: pdefq of qv, qc, qr (binary64 column sums in k order, so the column
//     rescale and its budget are the reference's words), the fall speed, and
//     the flux-form fallout in nfall substeps with a carried flux.  (The
//     reference runs pre-pdefq and fallout one thread per j-ROW.)  With
//     efall=1 the energy fallout (getcvm, getefall, the guarded temperature /
//     pressure / Exner / theta update, cuda_gm1_kessler_efall.py:201-256) runs
//     in the same column pass between the surface-energy budget and the fallout,
//     and the held temperature it leaves is the local conversion's.
//   * k_kes_cell: ONE thread per cell for autoconversion, accretion, rain
//     evaporation and the saturation adjustment (eqtset 2 iterates to
//     |dT| <= 2e-4 K, at most 100 times); the eqtset=2 held recapture of the
//     large-step transition is fused in.  Per-wave binary64 partial budgets
//     are formed with a fixed shuffle tree.
//   * k_kes_budgets: one block folds the column and wave partials into asq
//     and qbudget in a fixed order (rows in i order, then j order, as the
//     reference; levels in k order).
// Transcendentals are binary64 exp/pow rounded once to binary32.  Budgets and
// status words stay on the device; nothing synchronises with the host.
#include "cm1_common.h"

#include <mutex>

using namespace gm1;

namespace {

constexpr float RD = 287.04f, CP = 1005.7f, CPV = 1870.0f, RV = 461.5f, XLV = 2501000.0f, CPL = 4190.0f;

struct KesConst {
  float cv, cvv, rovcp, eps, reps, lv2, lv1, rp00, p00, cpdcv, rddcv, cpdrd;
};

GM1_DEV KesConst kes_const() {
  KesConst c;
  c.cv = CP - RD;
  c.cvv = CPV - RV;
  c.rovcp = RD / CP;
  c.eps = RD / RV;
  c.reps = RV / RD;
  c.lv2 = CPL - CPV;
  c.lv1 = XLV + c.lv2 * 273.15f;
  c.rp00 = 1.0f / 1.0e5f;
  c.p00 = 1.0e5f;
  c.cpdcv = CP / c.cv;
  c.rddcv = RD / c.cv;
  c.cpdrd = CP / RD;
  return c;
}

GM1_DEV float exp_r(float x) { return (float)exp((double)x); }
GM1_DEV float pow_r(float b, float e) { return (float)pow((double)b, (double)e); }

GM1_DEV float sat_qv(const KesConst& c, float t, float prs_cap, float pden) {
  float esl = 611.2f * exp_r((17.67f * (t - 273.15f)) / (t - 29.65f));
  esl = fminf(esl, prs_cap * 0.5f);
  return (c.eps * esl) / (pden - esl);
}

struct KesArgs {
  const float* pi0; const float* th0; const float* ruh; const float* rvh; const float* rmh; const float* mh;
  float* rho; float* rr; float* pp3d; float* th3d; float* prs; float* qv; float* qc; float* qr; float* vt;
  float* rain; float* prate;
  double* colbud;   // [5][nj][ni]: pdefq qv, qc, qr, fallout, surface energy
  double* wavebud;  // [5][nk][nj][nseg]: autoconversion, accretion, evaporation, condensation, cloud evaporation
  int* wavestat;    // [nk][nj][nseg][2]: max satadj iterations, failures
  // per-cell branch decisions (null: not written).  Bits 0..6: qc > 0, autoconversion
  // (qc > 0.001), accretion, evaporation (qr > 0 and qv < qvs), evaporation capped at
  // saturation, the ar + cr > qc limiter, the theta/pi' update (er > 1e-7); bit 7: the
  // saturation adjustment ran; bits 8..15: its iteration count; bit 16: it did not converge.
  unsigned* decisions;
  // held rates of the large-step transition (eqtset 2; null when not recapturing)
  float* qpten; float* qtten; float* qvten; float* qcten;
  // efall=1: mf (nk+1, nj, ni) and the held temperature (nk, nj, ni) the column
  // pass writes and the local conversion reads; both null for efall=0
  const float* mf; float* temp;
  int ni, nj, nk, nseg, eqtset;
  float dx, dy, dz, dt;
};

// ------------------------------------------------------------------ column phase
// Column loops load KB levels at a time (independent loads in flight) and
// then run the level arithmetic in k order, so every binary64 sum and every
// carried value is the sequential one.
constexpr int KB = 8;

__global__ __launch_bounds__(64) void k_kes_column(KesArgs a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  const int ni = a.ni, nj = a.nj, nk = a.nk;
  if (i >= ni) return;
  const KesConst c = kes_const();
  const int64_t plane = (int64_t)nj * ni, h = (int64_t)j * ni + i;
  const float* __restrict__ rho = a.rho;
  const float* __restrict__ rmh = a.rmh;
  const float* __restrict__ mh = a.mh;
  float* __restrict__ vt = a.vt;
  float* __restrict__ qr = a.qr;
  const float ru = a.ruh[i], rv = a.rvh[j];
  float* qs[3] = {a.qv, a.qc, a.qr};
  const float rmax[3] = {0.0f, 1.0e-14f, 1.0e-14f};
  // temperature at the lowest level before any phase (pdefq leaves theta, pi alone)
  const float t0 = (a.th0[h] + a.th3d[h]) * (a.pi0[h] + a.pp3d[h]);
  const float rdz = 1.0f / a.dz;
  // pdefq of the three species in two sweeps up the column (each species' sums
  // in level order, as species by species): the first forms the masses before
  // and after the small values are cleared, the second scales each species by
  // its mass ratio -- re-clearing the same small values rather than reading
  // back cleared words -- and forms the fall speed and its Courant maximum
  double t1[3] = {0.0, 0.0, 0.0}, t2[3] = {0.0, 0.0, 0.0}, a1[3] = {0.0, 0.0, 0.0}, a2[3] = {0.0, 0.0, 0.0};
  for (int k0 = 0; k0 < nk; k0 += KB) {
    float r[KB], m[KB], v[3][KB];
#pragma unroll
    for (int u = 0; u < KB; ++u)
      if (k0 + u < nk) {
        const int64_t p = (k0 + u) * plane + h;
        r[u] = rho[p];
        m[u] = rmh[p];
#pragma unroll
        for (int sp = 0; sp < 3; ++sp) v[sp][u] = qs[sp][p];
      }
#pragma unroll
    for (int u = 0; u < KB; ++u)
      if (k0 + u < nk) {
#pragma unroll
        for (int sp = 0; sp < 3; ++sp) {
          const float x = v[sp][u];
          t1[sp] += (double)(r[u] * x);
          a1[sp] += (double)((((r[u] * x) * ru) * rv) * m[u]);
          const float xz = (x < rmax[sp]) ? 0.0f : x;
          t2[sp] += (double)(r[u] * xz);
        }
      }
  }
  double t3[3];
#pragma unroll
  for (int sp = 0; sp < 3; ++sp) {
    t3[sp] = (t1[sp] + 1.0e-20) / (t2[sp] + 1.0e-20);
    if (t3[sp] < 0.0) t3[sp] = 1.0;
  }
  float crmax = 0.0f;
  for (int k0 = 0; k0 < nk; k0 += KB) {
    float r[KB], m[KB], mf[KB], v[3][KB];
#pragma unroll
    for (int u = 0; u < KB; ++u)
      if (k0 + u < nk) {
        const int64_t p = (k0 + u) * plane + h;
        r[u] = rho[p];
        m[u] = rmh[p];
        mf[u] = mh[p];
#pragma unroll
        for (int sp = 0; sp < 3; ++sp) v[sp][u] = qs[sp][p];
      }
#pragma unroll
    for (int u = 0; u < KB; ++u)
      if (k0 + u < nk) {
        const int64_t p = (k0 + u) * plane + h;
        float qrn = 0.0f;
#pragma unroll
        for (int sp = 0; sp < 3; ++sp) {
          const float x = (v[sp][u] < rmax[sp]) ? 0.0f : v[sp][u];
          const float nv = (float)(t3[sp] * (double)x);
          qs[sp][p] = nv;
          a2[sp] += (double)((((r[u] * nv) * ru) * rv) * m[u]);
          qrn = nv;  // the last species is rain
        }
        float vf = 0.0f;
        if (qrn > 1.0e-12f) vf = (14.34f * pow_r(r[u] * qrn, 0.1346f)) * sqrtf(1.15f / r[u]);
        vt[p] = vf;
        crmax = fmaxf(crmax, ((vf * a.dt) * rdz) * mf[u]);
      }
  }
#pragma unroll
  for (int sp = 0; sp < 3; ++sp) a.colbud[(int64_t)sp * plane + h] = a2[sp] - a1[sp];
  {  // surface energy flux of falling rain
    float term = a.dt * vt[h];
    term = term * a.rr[h];
    term = term * qr[h];
    term = term * (CPL * t0 - c.lv1);
    term = term * ru;
    term = term * rv;
    a.colbud[4 * plane + h] = (double)term;
  }
  if (a.temp) {  // efall=1: energy fallout of every level below the top, bottom up
    // level k's tendency reads level k+1's temperature before level k+1 is updated
    const float rdzf = 1.0f / a.dz;
    float tk = t0;
    for (int k = 0; k < nk; ++k) {
      const int64_t p = (int64_t)k * plane + h;
      float tnew = tk, tup = 0.0f;
      if (k + 1 < nk) {
        const int64_t pu = p + plane;
        tup = (a.th0[pu] + a.th3d[pu]) * (a.pi0[pu] + a.pp3d[pu]);
        const float qvk = a.qv[p], qrk = qr[p];
        const float cvm = (a.eqtset <= 1) ? c.cv : ((c.cv + c.cvv * qvk) + CPL * a.qc[p]) + CPL * qrk;
        float vertical = CPL * (tup - tk);
        vertical = vertical * rdzf;
        vertical = vertical * a.mf[(int64_t)(k + 1) * plane + h];
        const float tend = ((qrk * vt[p]) * (vertical + 9.81f)) / cvm;
        const float inc = a.dt * tend;
        if (fabsf(inc) >= 0.0001f) {
          tnew = tk + inc;
          const float pr = ((rho[p] * RD) * tnew) * (1.0f + qvk * c.reps);
          const float pp = pow_r(pr * c.rp00, c.rovcp) - a.pi0[p];
          a.prs[p] = pr;
          a.pp3d[p] = pp;
          a.th3d[p] = tnew / (a.pi0[p] + pp) - a.th0[p];
        }
      }
      a.temp[p] = tnew;
      tk = tup;
    }
  }
  // fallout (crmax from the second pdefq sweep)
  int nfall = (int)(crmax + 1.0f);
  if (nfall < 1) nfall = 1;
  {  // levels with little rain take the speed of the level above, top down
    float above = vt[(int64_t)(nk - 1) * plane + h];
    for (int k1 = nk - 2; k1 >= 0; k1 -= KB) {
      float q[KB], v[KB];
#pragma unroll
      for (int u = 0; u < KB; ++u)
        if (k1 - u >= 0) {
          q[u] = qr[(k1 - u) * plane + h];
          v[u] = vt[(k1 - u) * plane + h];
        }
#pragma unroll
      for (int u = 0; u < KB; ++u)
        if (k1 - u >= 0) {
          if (q[u] < 1.0e-8f) {
            v[u] = above;
            vt[(k1 - u) * plane + h] = above;
          }
          above = v[u];
        }
    }
  }
  const float dtfall = a.dt / (float)nfall;
  double bud = 0.0;
  for (int sub = 0; sub < nfall; ++sub) {
    float rqk = (rho[h] * vt[h]) * fmaxf(0.0f, qr[h]);
    const float rq0 = rqk;
    for (int k0 = 0; k0 < nk; k0 += KB) {
      // old qr of levels k0..k0+KB (the update of level k reads only level k+1's old value)
      float r[KB + 1], v[KB + 1], q[KB + 1], m[KB];
#pragma unroll
      for (int u = 0; u <= KB; ++u)
        if (k0 + u < nk) {
          const int64_t p = (k0 + u) * plane + h;
          r[u] = rho[p];
          v[u] = vt[p];
          q[u] = qr[p];
          if (u < KB) m[u] = mh[p];
        }
#pragma unroll
      for (int u = 0; u < KB; ++u)
        if (k0 + u < nk) {
          const float rqk1 = (k0 + u + 1 < nk) ? (r[u + 1] * v[u + 1]) * fmaxf(0.0f, q[u + 1]) : 0.0f;
          const float delta = (((dtfall * (rqk1 - rqk)) * rdz) * m[u]) / r[u];
          qr[(k0 + u) * plane + h] = q[u] + delta;
          rqk = rqk1;
        }
    }
    a.rain[h] = a.rain[h] + (0.1f * dtfall) * rq0;
    a.prate[h] = rq0;
    float term = dtfall * a.rr[h];
    term = term * vt[h];
    term = term * fmaxf(0.0f, qr[h]);
    term = term * ru;
    term = term * rv;
    bud += (double)term;
  }
  a.colbud[3 * plane + h] = bud;
}

// ------------------------------------------------------------------ cell phase
GM1_DEV double wave_sum(double v) {
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

GM1_DEV int wave_max(int v) {
  for (int off = 32; off >= 1; off >>= 1) v = max(v, __shfl_xor(v, off, 64));
  return v;
}

__global__ __launch_bounds__(64) void k_kes_cell(KesArgs a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y, k = blockIdx.z;
  const int ni = a.ni, nj = a.nj;
  const bool live = i < ni;
  double bud[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  int iters = 0, fail = 0;
  if (live) {
    const KesConst c = kes_const();
    const int64_t p = ((int64_t)k * nj + j) * ni + i;
    const float pi0 = a.pi0[p], th0 = a.th0[p];
    float pp = a.pp3d[p], th = a.th3d[p], prs = a.prs[p], rho = a.rho[p];
    const float rr = a.rr[p];
    float qv = a.qv[p], qc = a.qc[p], qr = a.qr[p];
    const float dt = a.dt;
    // ---- local conversion (cuda_gm1_kessler_local.py:290-416)
    float pisum = pi0 + pp;
    const float tmp = a.temp ? a.temp[p] : (th0 + th) * pisum;  // efall=1: the held temperature
    float ar = 0.0f, cr = 0.0f, er = 0.0f;
    unsigned word = 0u;
    if (qc > 0.0f) {
      word |= 1u;
      if (qc - 0.001f > 0.0f) word |= 2u;
      ar = fmaxf(0.001f * (qc - 0.001f), 0.0f);
      ar = ar * dt;
    }
    if (qc > 0.0f && qr > 0.0f) {
      word |= 4u;
      cr = ((2.2f * qc) * pow_r(qr, 0.875f)) * dt;
    }
    if (qr > 0.0f) {
      const float qvs = sat_qv(c, tmp, prs, prs);
      if (qv < qvs) {
        word |= 8u;
        const float rhoqr = rho * qr;
        const float vent = 1.6f + 30.3922f * pow_r(rhoqr, 0.2046f);
        const float deficit = 1.0f - qv / qvs;
        const float num = (vent * deficit) * pow_r(rhoqr, 0.525f);
        const float den = (2.03e4f + 9.584e6f / (qvs * prs)) * rho;
        er = fminf((num / den) * dt, qr);
        if (qv + er > qvs) {
          word |= 16u;
          er = qvs - qv;
        }
      }
    }
    if (ar + cr > qc) {
      word |= 32u;
      const float term1 = ar + cr;
      ar = (qc * ar) / term1;
      cr = (qc * cr) / term1;
    }
    qv = qv + er;
    qc = qc - (ar + cr);
    qr = qr + ((ar + cr) - er);
    if (er > 1.0e-7f) word |= 64u;
    if (er > 1.0e-7f) {
      if (a.eqtset == 2) {
        const float cpml = (CP + CPV * qv) + CPL * (qc + qr);
        const float cvml = (c.cv + c.cvv * qv) + CPL * (qc + qr);
        const float rm = RD + RV * qv;
        const float latent = c.lv1 - c.lv2 * tmp;
        const float first = latent / ((c.cpdcv * cvml) * pisum);
        const float second = ((th0 + th) * (RV / cvml)) * (1.0f - (c.rovcp * cpml) / rm);
        th = th - er * (first - second);
        const float base = ((rho * rm) * (th0 + th)) * c.rp00;
        pp = pow_r(base, c.rddcv) - pi0;
        pisum = pi0 + pp;
        prs = c.p00 * pow_r(pisum, c.cpdrd);
      } else {
        const float latent = c.lv1 - c.lv2 * tmp;
        th = th - er * (latent / (CP * pisum));
        const float den = ((RD * (th0 + th)) * pisum) * (1.0f + qv * c.reps);
        rho = prs / den;
      }
    }
    const float tem = (a.ruh[i] * a.rvh[j]) * a.rmh[p];
    bud[0] = (double)((rr * ar) * tem);
    bud[1] = (double)((rr * cr) * tem);
    bud[2] = (double)((rr * er) * tem);
    // ---- saturation adjustment (cuda_gm1_kessler_driver.py:337-428)
    const float oqv = qv, oqc = qc;
    float qvn = qv, qcn = qc;
    float tnew = (th0 + th) * (pi0 + pp);
    float qvs = sat_qv(c, tnew, prs, prs);
    if (a.eqtset == 2) {
      if (oqc > 1.0e-12f || oqv > qvs) {
        const float ql = fmaxf(oqc + qr, 0.0f);
        const float cvml = (c.cv + c.cvv * qvn) + CPL * ql;
        const float lhv = c.lv1 - c.lv2 * tnew;
        const float t1 = (lhv - RV * tnew) / cvml;
        const float d1 = (t1 * 17.67f) * 243.5f;
        const float converge = 2.0f * 0.0001f;
        float tlast = tnew, pnew = prs;
        int n = 0;
        while (true) {
          n += 1;
          const float sq = (tnew - 29.65f) * (tnew - 29.65f);
          const float denom = 1.0f + (d1 * qvs) / sq;
          float dqv = (qvs - qvn) / denom;
          dqv = fminf(dqv, qcn);
          if (qvn + dqv < 1.0e-20f) dqv = 1.0e-20f - qvn;
          qvn = qvn + dqv;
          qcn = qcn - dqv;
          tnew = tnew - dqv * t1;
          pnew = (rho * (RD + RV * qvn)) * tnew;
          const bool conv = fabs((double)(tnew - tlast)) <= (double)converge;
          if (!conv) {
            tlast = tnew;
            qvs = sat_qv(c, tnew, prs, pnew);
          }
          if (n == 100) {
            fail = conv ? 0 : 1;
            break;
          }
          if (conv) break;
        }
        iters = n;
        prs = pnew;
        pp = pow_r(pnew * c.rp00, c.rovcp) - pi0;
        th = tnew / (pi0 + pp) - th0;
        qv = qvn;
        qc = qcn;
      }
    } else {
      const float lhv = c.lv1 - c.lv2 * tnew;
      const float sq = (tnew - 29.65f) * (tnew - 29.65f);
      const float num = ((lhv * qvs) * 17.67f) * 243.5f;
      float dqv = (qvs - qvn) / (1.0f + num / (CP * sq));
      if (qvn + dqv < 1.0e-20f) dqv = 1.0e-20f - qvn;
      dqv = fminf(dqv, fmaxf(0.0f, qcn));
      qvn = qvn + dqv;
      qcn = qcn - dqv;
      iters = 1;
      th = th - dqv * (lhv / (CP * (pi0 + pp)));
      qv = qvn;
      qc = qcn;
      const float den = ((RD * (th0 + th)) * (pi0 + pp)) * (1.0f + qvn * c.reps);
      rho = prs / den;
    }
    const float cond = fmaxf(qcn - oqc, 0.0f), evac = fmaxf(qvn - oqv, 0.0f);
    bud[3] = (double)((rr * cond) * tem);
    bud[4] = (double)((rr * evac) * tem);
    a.pp3d[p] = pp;
    a.th3d[p] = th;
    a.prs[p] = prs;
    a.rho[p] = rho;
    a.qv[p] = qv;
    a.qc[p] = qc;
    a.qr[p] = qr;
    if (iters > 0) word |= 128u | ((unsigned)(iters & 0xFF) << 8) | (fail ? (1u << 16) : 0u);
    if (a.decisions) a.decisions[p] = word;
    if (a.qpten) {  // eqtset=2 held recapture (cuda_gm1_kessler_large_step.py:281-285)
      a.qpten[p] = (pp - a.qpten[p]) / dt;
      a.qtten[p] = (th - a.qtten[p]) / dt;
      a.qvten[p] = (qv - a.qvten[p]) / dt;
      a.qcten[p] = (qc - a.qcten[p]) / dt;
    }
  }
  // per-wave budget partials and status (fixed shuffle tree: deterministic)
  const int64_t slot = ((int64_t)k * nj + j) * a.nseg + blockIdx.x;
  const int64_t stride = (int64_t)a.nk * nj * a.nseg;
  for (int b = 0; b < 5; ++b) {
    const double v = wave_sum(bud[b]);
    if (threadIdx.x == 0) a.wavebud[b * stride + slot] = v;
  }
  const int mi = wave_max(iters), mf = wave_max(fail);
  if (threadIdx.x == 0) {
    a.wavestat[2 * slot] = mi;
    a.wavestat[2 * slot + 1] = mf;
  }
}

// ------------------------------------------------------------------ budgets
// Sum of n doubles in index order, eight loads in flight at a time (the
// additions stay in order, so the result is the sequential sum).
GM1_DEV double ordered_sum(const double* __restrict__ v, int64_t n, int64_t stride) {
  double s = 0.0;
  int64_t q = 0;
  for (; q + 8 <= n; q += 8) {
    double b[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) b[u] = v[(q + u) * stride];
#pragma unroll
    for (int u = 0; u < 8; ++u) s += b[u];
  }
  for (; q < n; ++q) s += v[q * stride];
  return s;
}

// Sum of n doubles by one workgroup of BS threads in a fixed order: thread t
// adds elements t, t + BS, ... in index order (four loads requested ahead),
// then a fixed shared-memory tree.  Deterministic; every thread returns it.
constexpr int BS = 256;
GM1_DEV double block_sum(const double* __restrict__ v, int64_t n, int64_t stride) {
  __shared__ double sh[BS];
  const int t = threadIdx.x;
  double acc = 0.0;
  int64_t q = t;
  for (; q + 3 * BS < n; q += 4 * BS) {
    double b[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) b[u] = v[(q + u * BS) * stride];
#pragma unroll
    for (int u = 0; u < 4; ++u) acc += b[u];
  }
  for (; q < n; q += BS) acc += v[q * stride];
  sh[t] = acc;
  __syncthreads();
  for (int w = BS / 2; w > 0; w >>= 1) {
    if (t < w) sh[t] += sh[t + w];
    __syncthreads();
  }
  return sh[0];
}

// rows: colbud rows (5 x nj, i order), one wavefront each: the wave stages a
// chunk of the row in LDS with coalesced loads and its first lane adds the
// chunk in index order (the sequential sum); then STAT_PARTS strided partial
// maxima of the satadj status words, one thread each, in the blocks after.
constexpr int STAT_PARTS = 2048;
constexpr int ROW_WAVES = 4, ROW_CHUNK = 512;

__global__ __launch_bounds__(64 * ROW_WAVES) void k_kes_rows(const double* __restrict__ colbud,
                                                             const int* __restrict__ wavestat,
                                                             double* __restrict__ rowscr, int* __restrict__ statscr,
                                                             int ni, int nj, int nk, int nseg, int row_blocks) {
  __shared__ double chunk[ROW_WAVES][ROW_CHUNK];
  const int nrow = 5 * nj;
  if ((int)blockIdx.x < row_blocks) {
    const int lane = threadIdx.x % 64, wv = threadIdx.x / 64;
    const int r = blockIdx.x * ROW_WAVES + wv;
    const double* row = colbud + (int64_t)(r < nrow ? r / nj : 0) * nj * ni + (int64_t)(r < nrow ? r % nj : 0) * ni;
    double acc = 0.0;
    for (int c0 = 0; c0 < ni; c0 += ROW_CHUNK) {  // the same trip count in every wave of the block
      const int len = min(ROW_CHUNK, ni - c0);
#pragma unroll
      for (int u = 0; u < ROW_CHUNK / 64; ++u) {
        const int q = u * 64 + lane;
        if (r < nrow && q < len) chunk[wv][q] = row[c0 + q];
      }
      __syncthreads();
      if (r < nrow && lane == 0) {
        int q = 0;
        for (; q + 8 <= len; q += 8) {  // eight words read ahead, added in index order
          double b[8];
#pragma unroll
          for (int u = 0; u < 8; ++u) b[u] = chunk[wv][q + u];
#pragma unroll
          for (int u = 0; u < 8; ++u) acc += b[u];
        }
        for (; q < len; ++q) acc += chunk[wv][q];
      }
      __syncthreads();
    }
    if (r < nrow && lane == 0) rowscr[r] = acc;
    return;
  }
  const int part = (blockIdx.x - row_blocks) * blockDim.x + threadIdx.x;
  if (part < STAT_PARTS) {
    const int64_t n = (int64_t)nk * nj * nseg;
    int mi = 0, mf = 0;  // (an integer maximum and count: any order gives the same words)
    for (int64_t q0 = part; q0 < n; q0 += 8 * (int64_t)STAT_PARTS) {
      int2 w[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int64_t q = q0 + (int64_t)u * STAT_PARTS;
        w[u] = (q < n) ? reinterpret_cast<const int2*>(wavestat)[q] : make_int2(0, 0);
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        mi = max(mi, w[u].x);
        mf += w[u].y;
      }
    }
    statscr[2 * part] = mi;
    statscr[2 * part + 1] = mf;
  }
}

// Level sums of the wave partials (5 x nk), one workgroup each, into rowscr
// after the 5 x nj row sums.
__global__ __launch_bounds__(BS) void k_kes_levels(const double* __restrict__ wavebud, double* __restrict__ rowscr,
                                                   int nj, int nk, int nseg) {
  const int q = blockIdx.x, b = q / nk, k = q % nk;
  const int64_t lstride = (int64_t)nj * nseg;
  const double v = block_sum(wavebud + (int64_t)b * nk * lstride + (int64_t)k * lstride, lstride, 1);
  if (threadIdx.x == 0) rowscr[5 * nj + q] = v;
}

// The folds into asq/qbudget in the reference's order (j rows, k levels) and
// the satadj status maxima.
__global__ __launch_bounds__(256) void k_kes_budgets(const double* __restrict__ rowscr_g,
                                                     const int* __restrict__ statscr, double* __restrict__ asq,
                                                     double* __restrict__ qbudget, int* __restrict__ status, int nj,
                                                     int nk, float dx, float dy, float dz, int staged) {
  __shared__ int smax[256];
  __shared__ int sfail[256];
  extern __shared__ double lrows[];  // the row and level sums, when staged
  const int t = threadIdx.x;
  if (staged)
    for (int q = t; q < 5 * nj + 5 * nk; q += blockDim.x) lrows[q] = rowscr_g[q];
  const double* __restrict__ rowscr = staged ? lrows : rowscr_g;
  {
    int mi = 0, mf = 0;
    for (int q = t; q < STAT_PARTS; q += blockDim.x) {
      mi = max(mi, statscr[2 * q]);
      mf += statscr[2 * q + 1];
    }
    smax[t] = mi;
    sfail[t] = mf;
  }
  __syncthreads();
  const double volume = (double)((dx * dy) * dz);
  const double area = (double)(dx * dy);
  // each fold on a wavefront of its own (lanes of one wave running the same
  // fold's code in parallel), its words read eight ahead and added in order
  auto fold = [&](double acc, const double* v, int n, double f1, double f2, bool two) {
    int q = 0;
    for (; q + 8 <= n; q += 8) {
      double b[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) b[u] = v[q + u];
#pragma unroll
      for (int u = 0; u < 8; ++u) acc = two ? acc + (b[u] * f1) * f2 : acc + b[u] * f1;
    }
    for (; q < n; ++q) acc = two ? acc + (v[q] * f1) * f2 : acc + v[q] * f1;
    return acc;
  };
  if (t < 3) {  // pdefq asq[t]
    asq[t] = fold(asq[t], rowscr + t * nj, nj, volume, 1.0, false);
  } else if (t == 64) {  // fallout -> qbudget[5]
    qbudget[5] = fold(qbudget[5], rowscr + 3 * nj, nj, area, 1.0, false);
  } else if (t == 128) {  // surface energy -> qbudget[6]
    qbudget[6] = fold(qbudget[6], rowscr + 4 * nj, nj, (double)dx, (double)dy, true);
  } else if (t >= 192 && t < 197) {  // level budgets -> qbudget[2,3,4,0,1]
    const int b = t - 192;
    const int slot = (b < 3) ? 2 + b : b - 3;
    qbudget[slot] = fold(qbudget[slot], rowscr + 5 * nj + b * nk, nk, volume, 1.0, false);
  } else if (t == 200) {
    int mi = 0, mf = 0;
    for (int q = 0; q < (int)blockDim.x; ++q) {
      mi = max(mi, smax[q]);
      mf += sfail[q];
    }
    status[0] = max(status[0], mi);
    status[1] = status[1] + mf;
  }
}

// ------------------------------------------------------------------ transition
// eqtset=2 held remove/snapshot, then the pre-physics EOS; pp3d is taken from
// the dynamics' pressure family (one-cell halo) into the mass array.
// One cell of the pre-physics transition: the eqtset=2 held remove/snapshot,
// pp3d/th3d/qv/qc written back, and the equation of state (prs, rho, and rr
// when copy_rr); returns pp and rho.
GM1_DEV void kes_pre_cell(int64_t p, const float* __restrict__ pp_family, int64_t pf, float* __restrict__ pp3d,
                          float* __restrict__ th3d, float* __restrict__ qv, float* __restrict__ qc,
                          float* __restrict__ qpten, float* __restrict__ qtten, float* __restrict__ qvten,
                          float* __restrict__ qcten, const float* __restrict__ pi0, const float* __restrict__ th0,
                          float* __restrict__ prs, float* __restrict__ rho, float* __restrict__ rr, float dt,
                          int eqtset, int copy_rr, float& pp_out, float& rho_out) {
  float pp = pp_family ? pp_family[pf] : pp3d[p];
  float th = th3d[p], v = qv[p], cc = qc[p];
  if (eqtset == 2) {
    pp = pp - qpten[p] * dt;
    qpten[p] = pp;
    th = th - qtten[p] * dt;
    qtten[p] = th;
    v = v - qvten[p] * dt;
    qvten[p] = v;
    cc = cc - qcten[p] * dt;
    qcten[p] = cc;
  }
  pp3d[p] = pp;
  th3d[p] = th;
  qv[p] = v;
  qc[p] = cc;
  const float work = pi0[p] + pp;
  const float pr = pow_r(work, CP / RD) * 1.0e5f;
  prs[p] = pr;
  float r = (th0[p] + th) * work;
  const float gas = fmaxf(v, 0.0f) * RV + RD;
  r = pr / (r * gas);
  rho[p] = r;
  if (copy_rr) rr[p] = r;
  pp_out = pp;
  rho_out = r;
}

__global__ __launch_bounds__(256) void k_kes_transition_pre(
    const float* __restrict__ pp_family, float* __restrict__ pp3d, float* __restrict__ th3d, float* __restrict__ qv,
    float* __restrict__ qc, float* __restrict__ qpten, float* __restrict__ qtten, float* __restrict__ qvten,
    float* __restrict__ qcten, const float* __restrict__ pi0, const float* __restrict__ th0, float* __restrict__ prs,
    float* __restrict__ rho, float* __restrict__ rr, int ni, int nj, int nk, float dt, int eqtset, int copy_rr) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y, k = blockIdx.z;
  if (i >= ni) return;
  const int64_t p = ((int64_t)k * nj + j) * ni + i;
  float pp, r;
  kes_pre_cell(p, pp_family, pix(k + 1, j + 1, i + 1, ni, nj), pp3d, th3d, qv, qc, qpten, qtten, qvten, qcten, pi0,
               th0, prs, rho, rr, dt, eqtset, copy_rr, pp, r);
}

// dry-air mass ledger (cuda_apmasscon.py): per-wave partials of
// rho*ruh*rvh*rmh and pi0+pp3d, a fixed-order fold, the Exner correction.
__global__ __launch_bounds__(64) void k_mass_partials(const float* __restrict__ rho, const float* __restrict__ ruh,
                                                      const float* __restrict__ rvh, const float* __restrict__ rmh,
                                                      const float* __restrict__ pi0, const float* __restrict__ pp3d,
                                                      double* __restrict__ part, int ni, int nj, int nk, int nseg,
                                                      int with_pressure) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y, k = blockIdx.z;
  double d = 0.0, pr = 0.0;
  if (i < ni) {
    const int64_t p = ((int64_t)k * nj + j) * ni + i;
    d = (double)(((rho[p] * ruh[i]) * rvh[j]) * rmh[p]);
    if (with_pressure) pr = (double)(pi0[p] + pp3d[p]);
  }
  d = wave_sum(d);
  pr = wave_sum(pr);
  const int64_t slot = ((int64_t)k * nj + j) * nseg + blockIdx.x;
  if (threadIdx.x == 0) {
    part[2 * slot] = d;
    part[2 * slot + 1] = pr;
  }
}

// The transition's cells with the dry-air mass partials of the rho and pp3d
// they write (the words k_mass_partials would read back), one wave per
// 64-cell row segment in k_mass_partials' layout.
__global__ __launch_bounds__(64) void k_kes_transition_pre_mass(
    const float* __restrict__ pp_family, float* __restrict__ pp3d, float* __restrict__ th3d, float* __restrict__ qv,
    float* __restrict__ qc, float* __restrict__ qpten, float* __restrict__ qtten, float* __restrict__ qvten,
    float* __restrict__ qcten, const float* __restrict__ pi0, const float* __restrict__ th0, float* __restrict__ prs,
    float* __restrict__ rho, const float* __restrict__ ruh, const float* __restrict__ rvh,
    const float* __restrict__ rmh, double* __restrict__ part, int ni, int nj, int nk, int nseg, float dt,
    int eqtset) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y, k = blockIdx.z;
  double d = 0.0, pr = 0.0;
  if (i < ni) {
    const int64_t p = ((int64_t)k * nj + j) * ni + i;
    float pp, r;
    kes_pre_cell(p, pp_family, pix(k + 1, j + 1, i + 1, ni, nj), pp3d, th3d, qv, qc, qpten, qtten, qvten, qcten,
                 pi0, th0, prs, rho, nullptr, dt, eqtset, 0, pp, r);
    d = (double)(((r * ruh[i]) * rvh[j]) * rmh[p]);
    pr = (double)(pi0[p] + pp);
  }
  d = wave_sum(d);
  pr = wave_sum(pr);
  const int64_t slot = ((int64_t)k * nj + j) * nseg + blockIdx.x;
  if (threadIdx.x == 0) {
    part[2 * slot] = d;
    part[2 * slot + 1] = pr;
  }
}

// ledger[0] = dry mass (levels in k order); with mass1 > 0 also the f32
// Exner correction into corr[0].
// One workgroup per level sum (dry mass, then the pressure sum).
__global__ __launch_bounds__(BS) void k_mass_levels(const double* __restrict__ part, double* __restrict__ levels,
                                                    int nk, int nj, int nseg) {
  const int r = blockIdx.x;
  const int which = r / nk, k = r % nk;
  const int64_t lstride = (int64_t)nj * nseg;
  const double v = block_sum(part + 2 * (int64_t)k * lstride + which, lstride, 2);
  if (threadIdx.x == 0) levels[r] = v;
}

__global__ __launch_bounds__(64) void k_mass_fold(const double* __restrict__ levels, double* __restrict__ ledger,
                                                  float* __restrict__ corr, int nk, double mass1, int ncells) {
  const int t = threadIdx.x;
  if (t == 0) {
    double m2 = 0.0, ps = 0.0;
    for (int k = 0; k < nk; ++k) m2 += levels[k];
    for (int k = 0; k < nk; ++k) ps += levels[nk + k];
    ledger[0] = m2;
    ledger[1] = ps;
    if (mass1 > 0.0) {
      const double e = (double)(RD / (CP - RD));
      double c = pow(mass1 / m2, e) - 1.0;
      c = c * ps;
      c = c / (double)ncells;
      corr[0] = (float)c;
    }
  }
}

__global__ __launch_bounds__(256) void k_mass_apply(const float* __restrict__ corr, float* __restrict__ pp3d,
                                                    const float* __restrict__ pi0, const float* __restrict__ th0,
                                                    const float* __restrict__ th3d, const float* __restrict__ qv,
                                                    float* __restrict__ prs, float* __restrict__ rho,
                                                    float* __restrict__ rr, int64_t n) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const float pp = pp3d[p] + corr[0];
  pp3d[p] = pp;
  const float pi = pi0[p] + pp;
  const float pr = 1.0e5f * pow_r(pi, CP / RD);
  const float theta = th0[p] + th3d[p];
  const float gas = RD + fmaxf(0.0f, qv[p]) * RV;
  const float r = pr / ((theta * pi) * gas);
  prs[p] = pr;
  rho[p] = r;
  if (rr) rr[p] = r;
}

}  // namespace

// ------------------------------------------------------------------ the no-allocation driver
// File-scope device workspace of gm1hip_kessler_driver_noalloc, for callers that
// cannot allocate (code under HIP graph capture, tight serving loops).  Each GPU
// holds its own copy for the life of the process, at its own address: host code
// uses the per-device address from hipGetSymbolAddress, never this name.  Every
// call on one GPU shares it; each call writes every word it reads, so nothing
// needs clearing, and calls that overlap on one GPU race.
constexpr int64_t KESSLER_NOALLOC_MAX_BYTES = 64LL << 20;  // 640x768x80 needs 49.6 MB
constexpr int KESSLER_NOALLOC_MAX_DEVICES = 64;
static __device__ __align__(256) unsigned char scratch[KESSLER_NOALLOC_MAX_BYTES];

namespace {
std::mutex g_scratch_mutex;
void* g_scratch_address[KESSLER_NOALLOC_MAX_DEVICES] = {};

// The current device's address of `scratch`, looked up once per device.
int scratch_address(void** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return (int)e;
  if (dev < 0 || dev >= KESSLER_NOALLOC_MAX_DEVICES) return (int)hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (g_scratch_address[dev] == nullptr) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return (int)e;
    if (p == nullptr) return (int)hipErrorInvalidSymbol;
    g_scratch_address[dev] = p;
  }
  *out = g_scratch_address[dev];
  return 0;
}
}  // namespace

extern "C" {

// Scratch the driver needs (bytes), for ni x nj x nk.
int64_t gm1hip_kessler_scratch_bytes(int ni, int nj, int nk) {
  const int64_t nseg = (ni + 63) / 64;
  const int64_t col = 5LL * nj * ni * 8;
  const int64_t wave = 5LL * nk * nj * nseg * 8;
  const int64_t stat = 2LL * nk * nj * nseg * 4;
  const int64_t rows = (5LL * nj + 5LL * nk) * 8;
  return col + wave + stat + rows + 8LL * 2048 + 64;
}

}  // extern "C"

namespace {
int kessler_driver(const float* ruh, const float* rvh, const float* rmh, const float* mh, const float* mf,
                   const float* pi0, const float* th0, float* rho, float* rr, float* pp3d, float* th3d, float* prs,
                   float* qv, float* qc, float* qr, float* fall_velocity, float* rain, float* prate,
                   float* temperature, double* asq, double* qbudget, int* status, float* qpten, float* qtten,
                   float* qvten, float* qcten, unsigned* decisions, void* work, int64_t work_bytes, int ni, int nj,
                   int nk, float dx, float dy, float dz, float dt, int eqtset, void* stream) {
  if (eqtset != 1 && eqtset != 2) return (int)hipErrorInvalidValue;
  if (work_bytes < gm1hip_kessler_scratch_bytes(ni, nj, nk)) return (int)hipErrorInvalidValue;
  if ((mf == nullptr) != (temperature == nullptr) || (mf && nk < 2)) return (int)hipErrorInvalidValue;
  hipStream_t s = (hipStream_t)stream;
  KesArgs a{};
  a.mf = mf; a.temp = temperature;
  a.pi0 = pi0; a.th0 = th0; a.ruh = ruh; a.rvh = rvh; a.rmh = rmh; a.mh = mh;
  a.rho = rho; a.rr = rr; a.pp3d = pp3d; a.th3d = th3d; a.prs = prs; a.qv = qv; a.qc = qc; a.qr = qr;
  a.vt = fall_velocity; a.rain = rain; a.prate = prate;
  a.qpten = qpten; a.qtten = qtten; a.qvten = qvten; a.qcten = qcten;
  a.decisions = decisions;
  a.ni = ni; a.nj = nj; a.nk = nk; a.nseg = (ni + 63) / 64; a.eqtset = eqtset;
  a.dx = dx; a.dy = dy; a.dz = dz; a.dt = dt;
  char* base = (char*)work;
  a.colbud = (double*)base;
  base += 5LL * nj * ni * 8;
  a.wavebud = (double*)base;
  base += 5LL * nk * nj * a.nseg * 8;
  double* rows = (double*)base;
  base += (5LL * nj + 5LL * nk) * 8;
  a.wavestat = (int*)base;
  hipLaunchKernelGGL(k_kes_column, dim3((ni + 63) / 64, nj), dim3(64), 0, s, a);
  hipLaunchKernelGGL(k_kes_cell, dim3(a.nseg, nj, nk), dim3(64), 0, s, a);
  int* statscr = a.wavestat + 2LL * nk * nj * a.nseg;
  const int row_blocks = (5 * nj + ROW_WAVES - 1) / ROW_WAVES;
  const int stat_blocks = (STAT_PARTS + 64 * ROW_WAVES - 1) / (64 * ROW_WAVES);
  hipLaunchKernelGGL(k_kes_rows, dim3(row_blocks + stat_blocks), dim3(64 * ROW_WAVES), 0, s, a.colbud, a.wavestat,
                     rows, statscr, ni, nj, nk, a.nseg, row_blocks);
  hipLaunchKernelGGL(k_kes_levels, dim3(5 * nk), dim3(BS), 0, s, a.wavebud, rows, nj, nk, a.nseg);
  // the folds read the row and level sums from LDS when they fit in 48 KiB
  const int64_t stage_bytes = (5LL * nj + 5LL * nk) * 8;
  const int staged = stage_bytes <= 48 * 1024 ? 1 : 0;
  hipLaunchKernelGGL(k_kes_budgets, dim3(1), dim3(256), staged ? (size_t)stage_bytes : 0, s, rows, statscr, asq,
                     qbudget, status, nj, nk, dx, dy, dz, staged);
  GM1_LAUNCH_CHECK();
  return 0;
}
}  // namespace

extern "C" {

int gm1hip_kessler_driver_decisions(const float* ruh, const float* rvh, const float* rmh, const float* mh,
                                    const float* pi0, const float* th0, float* rho, float* rr, float* pp3d,
                                    float* th3d, float* prs, float* qv, float* qc, float* qr, float* fall_velocity,
                                    float* rain, float* prate, double* asq, double* qbudget, int* status,
                                    float* qpten, float* qtten, float* qvten, float* qcten, unsigned* decisions,
                                    void* work, int64_t work_bytes, int ni, int nj, int nk, float dx, float dy,
                                    float dz, float dt, int eqtset, void* stream) {
  return kessler_driver(ruh, rvh, rmh, mh, nullptr, pi0, th0, rho, rr, pp3d, th3d, prs, qv, qc, qr, fall_velocity,
                        rain, prate, nullptr, asq, qbudget, status, qpten, qtten, qvten, qcten, decisions, work,
                        work_bytes, ni, nj, nk, dx, dy, dz, dt, eqtset, stream);
}

int gm1hip_kessler_driver_efall(const float* ruh, const float* rvh, const float* rmh, const float* mh,
                                const float* mf, const float* pi0, const float* th0, float* rho, float* rr,
                                float* pp3d, float* th3d, float* prs, float* qv, float* qc, float* qr,
                                float* fall_velocity, float* rain, float* prate, float* temperature, double* asq,
                                double* qbudget, int* status, float* qpten, float* qtten, float* qvten, float* qcten,
                                unsigned* decisions, void* work, int64_t work_bytes, int ni, int nj, int nk, float dx,
                                float dy, float dz, float dt, int eqtset, void* stream) {
  if (!mf || !temperature) return (int)hipErrorInvalidValue;
  return kessler_driver(ruh, rvh, rmh, mh, mf, pi0, th0, rho, rr, pp3d, th3d, prs, qv, qc, qr, fall_velocity, rain,
                        prate, temperature, asq, qbudget, status, qpten, qtten, qvten, qcten, decisions, work,
                        work_bytes, ni, nj, nk, dx, dy, dz, dt, eqtset, stream);
}

int gm1hip_kessler_driver(const float* ruh, const float* rvh, const float* rmh, const float* mh, const float* pi0,
                          const float* th0, float* rho, float* rr, float* pp3d, float* th3d, float* prs, float* qv,
                          float* qc, float* qr, float* fall_velocity, float* rain, float* prate, double* asq,
                          double* qbudget, int* status, float* qpten, float* qtten, float* qvten, float* qcten,
                          void* work, int64_t work_bytes, int ni, int nj, int nk, float dx, float dy, float dz,
                          float dt, int eqtset, void* stream) {
  return gm1hip_kessler_driver_decisions(ruh, rvh, rmh, mh, pi0, th0, rho, rr, pp3d, th3d, prs, qv, qc, qr,
                                         fall_velocity, rain, prate, asq, qbudget, status, qpten, qtten, qvten, qcten,
                                         nullptr, work, work_bytes, ni, nj, nk, dx, dy, dz, dt, eqtset, stream);
}

int gm1hip_kessler_driver_noalloc(const float* ruh, const float* rvh, const float* rmh, const float* mh,
                                  const float* pi0, const float* th0, float* rho, float* rr, float* pp3d, float* th3d,
                                  float* prs, float* qv, float* qc, float* qr, float* fall_velocity, float* rain,
                                  float* prate, double* asq, double* qbudget, int* status, float* qpten,
                                  float* qtten, float* qvten, float* qcten, int ni, int nj, int nk, float dx,
                                  float dy, float dz, float dt, int eqtset, void* stream) {
  if (ni <= 0 || nj <= 0 || nk <= 0) return (int)hipErrorInvalidValue;
  if (gm1hip_kessler_scratch_bytes(ni, nj, nk) > KESSLER_NOALLOC_MAX_BYTES) return (int)hipErrorInvalidValue;
  void* work = nullptr;
  const int rc = scratch_address(&work);
  if (rc != 0) return rc;
  return gm1hip_kessler_driver_decisions(ruh, rvh, rmh, mh, pi0, th0, rho, rr, pp3d, th3d, prs, qv, qc, qr,
                                         fall_velocity, rain, prate, asq, qbudget, status, qpten, qtten, qvten, qcten,
                                         nullptr, work, KESSLER_NOALLOC_MAX_BYTES, ni, nj, nk, dx, dy, dz, dt, eqtset,
                                         stream);
}

int gm1hip_kessler_transition_pre(const float* pp_family, float* pp3d, float* th3d, float* qv, float* qc, float* qpten,
                                  float* qtten, float* qvten, float* qcten, const float* pi0, const float* th0,
                                  float* prs, float* rho, float* rr, int ni, int nj, int nk, float dt, int eqtset,
                                  int copy_rr, void* stream) {
  hipLaunchKernelGGL(k_kes_transition_pre, dim3((ni + 255) / 256, nj, nk), dim3(256), 0, (hipStream_t)stream,
                     pp_family, pp3d, th3d, qv, qc, qpten, qtten, qvten, qcten, pi0, th0, prs, rho, rr, ni, nj, nk,
                     dt, eqtset, copy_rr);
  GM1_LAUNCH_CHECK();
  return 0;
}

int64_t gm1hip_dry_mass_scratch_bytes(int ni, int nj, int nk) {
  const int64_t nseg = (ni + 63) / 64;
  return 2LL * nk * nj * nseg * 8 + 2LL * nk * 8 + 64;
}

// Dry-air mass ledger; with mass1 > 0 also the apmasscon correction applied
// to pp3d/prs/rho (and rr when given).  ledger[0] = mass before, ledger[1] =
// sum(pi0+pp3d) before, corr[0] = the f32 correction.
int gm1hip_apmasscon(const float* ruh, const float* rvh, const float* rmh, const float* pi0, const float* th0,
                     const float* th3d, const float* qv, float* pp3d, float* prs, float* rho, float* rr,
                     double* ledger, float* corr, double mass1, void* work, int64_t work_bytes, int ni, int nj,
                     int nk, void* stream) {
  if (work_bytes < gm1hip_dry_mass_scratch_bytes(ni, nj, nk)) return (int)hipErrorInvalidValue;
  hipStream_t s = (hipStream_t)stream;
  const int nseg = (ni + 63) / 64;
  double* part = (double*)work;
  double* levels = part + 2LL * nk * nj * nseg;
  hipLaunchKernelGGL(k_mass_partials, dim3(nseg, nj, nk), dim3(64), 0, s, rho, ruh, rvh, rmh, pi0, pp3d, part, ni,
                     nj, nk, nseg, mass1 > 0.0 ? 1 : 0);
  hipLaunchKernelGGL(k_mass_levels, dim3(2 * nk), dim3(BS), 0, s, part, levels, nk, nj, nseg);
  hipLaunchKernelGGL(k_mass_fold, dim3(1), dim3(64), 0, s, levels, ledger, corr, nk, mass1, ni * nj * nk);
  if (mass1 > 0.0) {
    const int64_t n = (int64_t)ni * nj * nk;
    hipLaunchKernelGGL(k_mass_apply, dim3(gm1_blocks(n, 256)), dim3(256), 0, s, corr, pp3d, pi0, th0, th3d, qv, prs,
                       rho, rr, n);
  }
  GM1_LAUNCH_CHECK();
  return 0;
}

int gm1hip_kessler_transition_pre_apmasscon(const float* pp_family, float* pp3d, float* th3d, float* qv, float* qc,
                                            float* qpten, float* qtten, float* qvten, float* qcten, const float* pi0,
                                            const float* th0, float* prs, float* rho, float* rr, const float* ruh,
                                            const float* rvh, const float* rmh, double* ledger, float* corr,
                                            double mass1, void* work, int64_t work_bytes, int ni, int nj, int nk,
                                            float dt, int eqtset, void* stream) {
  if ((eqtset != 1 && eqtset != 2) || !(mass1 > 0.0) || work_bytes < gm1hip_dry_mass_scratch_bytes(ni, nj, nk))
    return (int)hipErrorInvalidValue;
  hipStream_t s = (hipStream_t)stream;
  const int nseg = (ni + 63) / 64;
  double* part = (double*)work;
  double* levels = part + 2LL * nk * nj * nseg;
  hipLaunchKernelGGL(k_kes_transition_pre_mass, dim3(nseg, nj, nk), dim3(64), 0, s, pp_family, pp3d, th3d, qv, qc,
                     qpten, qtten, qvten, qcten, pi0, th0, prs, rho, ruh, rvh, rmh, part, ni, nj, nk, nseg, dt, eqtset);
  hipLaunchKernelGGL(k_mass_levels, dim3(2 * nk), dim3(BS), 0, s, part, levels, nk, nj, nseg);
  hipLaunchKernelGGL(k_mass_fold, dim3(1), dim3(64), 0, s, levels, ledger, corr, nk, mass1, ni * nj * nk);
  const int64_t n = (int64_t)ni * nj * nk;
  hipLaunchKernelGGL(k_mass_apply, dim3(gm1_blocks(n, 256)), dim3(256), 0, s, corr, pp3d, pi0, th0, th3d, qv, prs, rho,
                     rr, n);
  GM1_LAUNCH_CHECK();
  return 0;
}

}  // extern "C"
