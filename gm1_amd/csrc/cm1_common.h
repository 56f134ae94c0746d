// Shared device helpers for the CM1 large-step kernels (gfx950).
//
// Arithmetic contract: every float operation is one IEEE binary32 operation
// rounded to nearest, in the association order the CM1 transcription spells.
// The translation units are compiled with -ffp-contract=off and without fast
// math, so `a*b + c` is a rounded product followed by a rounded sum and never
// a fused multiply-add; division and sqrt are the correctly rounded forms.
// Denormals are preserved (no flush to zero).
#pragma once
#pragma clang fp contract(off)

#include <hip/hip_runtime.h>
#include <stdint.h>

#define GM1_DEV __device__ __forceinline__

// Field layout (z, y, x), C order, i fastest.  Winds and scalars carry a
// three-cell lateral halo (ng = 3); the pressure family carries one cell.
namespace gm1 {

constexpr int NG = 3;

struct Dims {
  int ni, nj, nk;
};

// Index of (k, j, i) in an array with row length nx and nz*ny*nx words.
GM1_DEV int64_t ix3(int k, int j, int i, int ny, int nx) {
  return ((int64_t)k * ny + j) * (int64_t)nx + i;
}

// Padded-array extents for the four staggerings.
GM1_DEV int sx(int ni) { return ni + 2 * NG; }      // scalar / w row length
GM1_DEV int sy(int nj) { return nj + 2 * NG; }      // scalar / w / u rows
GM1_DEV int ux(int ni) { return ni + 1 + 2 * NG; }  // u row length
GM1_DEV int vy(int nj) { return nj + 1 + 2 * NG; }  // v rows

// Tile index for linear workgroup id `linear` of `n`, chosen so that each of
// the eight XCDs -- which receive workgroups round-robin in dispatch order --
// works on one contiguous range of tiles, keeping neighbouring rows in the
// same L2.
GM1_DEV int xcd_tile(int linear, int n) {
  constexpr int X = 8;
  const int per = n / X, rem = n % X;
  const int x = linear % X, r = linear / X;
  return (x < rem) ? x * (per + 1) + r : rem * (per + 1) + (x - rem) * per + r;
}

// The word at a 32-bit byte offset from a base the whole workgroup shares:
// the load takes the base in scalar registers and the offset in one vector
// register (the hardware adds them), so a thread's level-k address costs no
// 64-bit vector arithmetic.  The offset must stay below 4 GiB.
GM1_DEV float ldo(const float* base, uint32_t byte_offset) {
  return *reinterpret_cast<const float*>(reinterpret_cast<const char*>(base) + byte_offset);
}
GM1_DEV void sto(float* base, uint32_t byte_offset, float v) {
  *reinterpret_cast<float*>(reinterpret_cast<char*>(base) + byte_offset) = v;
}

// Pressure family (nk+2, nj+2, ni+2).
GM1_DEV int64_t pix(int k, int j, int i, int ni, int nj) {
  return ((int64_t)k * (nj + 2) + j) * (int64_t)(ni + 2) + i;
}

// Constant weights: each is the binary64 quotient rounded once to binary32,
// which is how the CM1 transcription forms them.
constexpr float C2_60 = (float)(2.0 / 60.0);
constexpr float CM13_60 = (float)(-13.0 / 60.0);
constexpr float C47_60 = (float)(47.0 / 60.0);
constexpr float C27_60 = (float)(27.0 / 60.0);
constexpr float CM3_60 = (float)(-3.0 / 60.0);
constexpr float CM1_6 = (float)(-1.0 / 6.0);
constexpr float C1_6 = (float)(1.0 / 6.0);
constexpr float C2_6 = (float)(2.0 / 6.0);
constexpr float C5_6 = (float)(5.0 / 6.0);
constexpr float CM7_6 = (float)(-7.0 / 6.0);
constexpr float C11_6 = (float)(11.0 / 6.0);
constexpr float C13_12 = (float)(13.0 / 12.0);

// Third-order upwind face value (adv_routines.F flx3).
GM1_DEV float flx3(float s1, float s2, float s3) {
  float r = CM1_6 * s1;
  r = r + C5_6 * s2;
  return r + C2_6 * s3;
}

// Fifth-order upwind-biased face value (adv_routines.F flx5).
GM1_DEV float flx5(float s1, float s2, float s3, float s4, float s5) {
  float r = C2_60 * s1;
  r = r + CM13_60 * s2;
  r = r + C47_60 * s3;
  r = r + C27_60 * s4;
  return r + CM3_60 * s5;
}

// WENO-Z (siform 2) face value: binary32 smoothness indicators and
// candidate stencils, binary64 nonlinear weights (adv_routines.F weno5).
GM1_DEV float weno5(float s1, float s2, float s3, float s4, float s5,
                    double weps) {
  float d1 = (s1 - 2.0f * s2) + s3;
  float e1 = (s1 - 4.0f * s2) + 3.0f * s3;
  float b1r = C13_12 * (d1 * d1) + 0.25f * (e1 * e1);
  float d2 = (s2 - 2.0f * s3) + s4;
  float e2 = s2 - s4;
  float b2r = C13_12 * (d2 * d2) + 0.25f * (e2 * e2);
  float d3 = (s3 - 2.0f * s4) + s5;
  float e3 = (3.0f * s3 - 4.0f * s4) + s5;
  float b3r = C13_12 * (d3 * d3) + 0.25f * (e3 * e3);
  double b1 = (double)b1r, b2 = (double)b2r, b3 = (double)b3r;
  double tau = fabs(b1 - b3);
  double q1 = fmin(1.0e30, tau / (b1 + weps));
  double q2 = fmin(1.0e30, tau / (b2 + weps));
  double q3 = fmin(1.0e30, tau / (b3 + weps));
  double w1 = (double)0.1f * (1.0 + q1 * q1);
  double w2 = (double)0.6f * (1.0 + q2 * q2);
  double w3 = (double)0.3f * (1.0 + q3 * q3);
  float p1 = (C2_6 * s1 + CM7_6 * s2) + C11_6 * s3;
  float p2 = (CM1_6 * s2 + C5_6 * s3) + C2_6 * s4;
  float p3 = (C2_6 * s3 + C5_6 * s4) + CM1_6 * s5;
  double num = (w1 * (double)p1 + w2 * (double)p2) + w3 * (double)p3;
  return (float)(num / ((w1 + w2) + w3));
}

// Positive-definite third-order limiter used at the WENO lower closures
// (adv_routines.F upstrpd).
GM1_DEV float upstrpd(float s1, float s2, float s3, float weps) {
  float dd = s2 - s1;
  float se = copysignf(fabsf(weps), dd);
  float rr = (s3 - s2) / (se + dd);
  float cand = C1_6 + C2_6 * rr;
  float phi = fmaxf(0.0f, fminf(rr, fminf(cand, 1.0f)));
  return s2 + phi * (s2 - s1);
}

GM1_DEV float recon5(float s1, float s2, float s3, float s4, float s5,
                     bool use_weno, double weps) {
  return use_weno ? weno5(s1, s2, s3, s4, s5, weps)
                  : flx5(s1, s2, s3, s4, s5);
}

// CM1's lateral family word: wbc | ebc<<4 | sbc<<8 | nbc<<12 with each side
// 1 periodic, 2 open-radiative, 3/4 rigid.  Zero is a tile-local launch.
__host__ __device__ inline unsigned side_of(unsigned obc, unsigned side) {
  return (obc >> (4u * side)) & 15u;
}

}  // namespace gm1

// Status handling shared by every C entry point.
#define GM1_LAUNCH_CHECK()                          \
  do {                                              \
    hipError_t e_ = hipGetLastError();              \
    if (e_ != hipSuccess) return (int)e_;           \
  } while (0)

inline int gm1_blocks(int64_t n, int threads) {
  int64_t b = (n + threads - 1) / threads;
  return (int)(b < 1 ? 1 : b);
}
