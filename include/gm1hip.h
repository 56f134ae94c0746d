// C ABI of libgm1hip.so: the CM1 dynamical-core large step on gfx950.
//
// Every entry point enqueues its kernels on `stream` (a hipStream_t passed
// as void*; NULL is the default stream), allocates nothing, never
// synchronises, and returns a hipError_t code (0 on success).  Arrays are
// caller-owned float32 device pointers in CM1's (z, y, x) C-order layout:
//
//   u family   (nk+2, nj+6, ni+7)      v family   (nk+2, nj+7, ni+6)
//   w family   (nk+3, nj+6, ni+6)      scalars    (nk+2, nj+6, ni+6)
//   pressure   (nk+2, nj+2, ni+2)      cells      (nk, nj, ni)
//   metrics    terrain_static_shapes (cuda_terrain_open_boundary.py:429-460)
//
// Padded index i (0-based) holds CM1's Fortran column i-2 (ngxy = 3).
// `obc`/`sides` packs CM1's lateral families (wbc, ebc, sbc, nbc) as
// wbc | ebc<<4 | sbc<<8 | nbc<<12 with 1 periodic, 2 open-radiative,
// 3/4 rigid; 0 (scalar advection only) is a tile-local launch.  The scalar
// advection and RK stage entry points also read bit 16 of obc: set, the
// caller states that c1 and c2 hold one word per level at every column (flat
// terrain), and the vertical fluxes read the first column's words -- the same
// words, with one read per level in place of one per cell.
//
// Arithmetic: IEEE binary32 with no contraction, IEEE denormals and
// correctly rounded division and square root; WENO weights in binary64.
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

// ----------------------------------------------------------------- advection
// Replaces launch_cuda_terrain_mass_flux5 (cuda_advection5.py:849-945):
// rru, rrv, rrw mass fluxes and divx (|d| < 1e-20 -> 0).
int gm1hip_terrain_mass_flux5(const float* u3d, const float* v3d, const float* w3d, float* rru, float* rrv,
                              float* rrw, const float* rho0_3d, const float* rf0, const float* rgzu,
                              const float* rgzv, const float* dzdx, const float* dzdy, const float* gz, float* divx,
                              int ni, int nj, int nk, float dx, float dy, float dz, void* stream);

// Replaces launch_cuda_scalar_advection5 (cuda_advection5.py:948-1049):
// 5th-order (use_weno=0) or WENO5 (use_weno=1) face fluxes and the scalar
// tendency in one LDS-tiled pass.  flux_x/y/z may be NULL (fluxes then stay
// on chip); otherwise they receive the face fluxes as the reference's
// scratch does.
int gm1hip_scalar_advection5(const float* scalar, const float* rru, const float* rrv, const float* rrw,
                             const float* c1, const float* c2, const float* rr0, const float* gz, const float* divx,
                             float* tendency, float* flux_x, float* flux_y, float* flux_z, int ni, int nj, int nk,
                             int use_weno, double weps, float dx, float dy, float dz, unsigned obc, void* stream);

// One scalar's RK stage of cuda_scalar_moist_rk.py (:1041-1499) with the
// advection fused into the A-anchored compose (:283-349):
// out = a0 + dttmp*((held - divterm) + advection), |out| < 1e-30 -> 0;
// divterm may be NULL (moisture rows).  WENO5 when use_weno.  obc must name
// the lateral families (non-zero).
int gm1hip_scalar_rk_stage(const float* scalar, const float* rru, const float* rrv, const float* rrw,
                           const float* c1, const float* c2, const float* rr0, const float* gz, const float* divx,
                           const float* a0, const float* held, const float* divterm, float* out, int ni, int nj,
                           int nk, int use_weno, double weps, float dx, float dy, float dz, unsigned obc,
                           float dttmp, void* stream);

// Final-stage moisture row (cuda_scalar_moist_rk.py:1376-1499, pdef.F):
// gated WENO5 fluxes (pdefweno: the vertical gate, off for qv), directional
// tendencies into advx/advy/advz (cell-shaped scratch), the donor repair in
// x, y and z (cm1_rk_pdef_candidate_*/correct_*), the total and the compose
// out = qa + dttmp*(held + advection).
int gm1hip_pdef_moisture_stage(const float* scalar, const float* rru, const float* rrv, const float* rrw,
                               const float* c1, const float* c2, const float* rho0_3d, const float* rr0,
                               const float* gz, const float* rgz, const float* divx, const float* qa,
                               const float* held, float* out, float* advx, float* advy, float* advz, int ni,
                               int nj, int nk, int pdefweno, double weps, float dx, float dy, float dz,
                               unsigned obc, float dttmp, void* stream);

// Replaces launch_cuda_terrain_open_momentum_advection5
// (cuda_momentum_advection5.py:1173-1330): mass fluxes, divx and the
// advu/advv/advw increments added into uten/vten/wten.  With
// stop_after_divergence=1 only the mass fluxes and divx are produced (the
// reference's rrw_boundary_hook / post_divergence_hook seam); the caller
// then finishes with gm1hip_momentum_transport5.
int gm1hip_momentum_advection5(const float* u3d, const float* v3d, const float* w3d, float* uten, float* vten,
                               float* wten, float* rru, float* rrv, float* rrw, const float* uh, const float* vh,
                               const float* rho0_3d, const float* rf0, const float* rrf0, const float* c2,
                               const float* gz, const float* gzu, const float* gzv, const float* rgzu,
                               const float* rgzv, const float* dzdx, const float* dzdy, float* divx, int ni, int nj,
                               int nk, float dx, float dy, float dz, unsigned obc, int stop_after_divergence,
                               void* stream);

// Second half of the above (cuda_momentum_advection5.py:871-1011,
// _cpu_momentum_transport5) from given mass fluxes and divx.
int gm1hip_momentum_transport5(const float* u3d, const float* v3d, const float* w3d, float* uten, float* vten,
                               float* wten, const float* rru, const float* rrv, const float* rrw, const float* uh,
                               const float* vh, const float* rho0_3d, const float* rrf0, const float* c2,
                               const float* gz, const float* gzu, const float* gzv, const float* divx, int ni, int nj,
                               int nk, float dx, float dy, float dz, unsigned obc, void* stream);

// ----------------------------------------------------------------- acoustic
// Replaces launch_cuda_psolver3_terrain_open_stage
// (cuda_psolver3_terrain_open.py:1436-1507) for one RK stage `nrk` with
// `nloop` small steps of length `dts`: coefficient preparation, per small
// step [ssopen on open sides], the forward u/v step (periodic boundary faces
// included, cuda_boundary_families.py:1640), one fused column pass
// (divergence, fpk, bcwsfc, pressure advection, Thomas solve for w, pressure
// update and filter, time averages), the ppd halo; then the radiation
// history (flags & 1) and the stage-end family fills (flags & 2).
// `profiles` packs sigma | sigmaf | rds | rdsf (nk+2 each) | cgs[3] | cgt[3].
// fpk, elim_c and elim_d are the column solve's scratch (pressure shape).
int gm1hip_psolver3_stage(
    float* ppi, float* pp3d, float* ppx, float* ppd, const float* piadv, const float* ppterm, const float* ppten,
    const float* thv, float* qk, float* pk1, float* pk2, float* mm, float* fpk, float* elim_c, float* elim_d,
    const float* ua, float* u3d, const float* uten, float* rru, const float* va, float* v3d, const float* vten,
    float* rrv, const float* wa, float* w3d, const float* wten, float* rrw, float* radbcw, float* radbce,
    float* radbcs, float* radbcn, const float* uh, const float* vh, const float* ruf, const float* rvf,
    float* scalar, const float* rho0_3d, const float* rr0, const float* rf0, const float* rrf0, const float* c1,
    const float* c2, const float* gz, const float* rgzu, const float* rgzv, const float* dzdx, const float* dzdy,
    const float* gx, const float* gy, const float* gxu, const float* gyv, const float* profiles, float zt,
    int ni, int nj, int nk, int nrk, int nloop, float dts, float dttmp, float dx, float dy, float alph,
    float kdiv, unsigned sides, int flags, void* stream);

// Replaces one stage of launch_cuda_sound3_periodic_three_stage
// (cuda_sound3_periodic.py:1216-1460; CPU form cpu_sound3_periodic_stage_inplace
// :672-1064): the flat, uniform, doubly periodic stage on 1-cell halos.
// `arrays` is a host table of 47 device pointers in the order of
// cuda_sound3_periodic.ALL_ARRAY_NAMES (:59-94): mh rr0 rf0 c1 c2 mf piadv
// ppterm thv ppten ppi pp3d ppx ppd qk pk1 pk2 mm fpk divh hadvp elim_c elim_d
// tri_aa tri_bb tri_cc tri_dd tri_pivot tri_residual | wten wa w3d rrw | ua u3d
// uten rru | va v3d vten rrv | dtu0 dtv0 uf vf uh vh, shaped per array_shapes
// (:449-491).  Every elimination and tridiagonal diagnostic is written.
int gm1hip_sound3_periodic_stage(void* const* arrays, int ni, int nj, int nk, int nrk, int nloop, float dts,
                                 float rdx, float rdy, float rdz, float cp, float alph, float kdiv, void* stream);

// Adjoint of one gm1hip_psolver3_stage call on flat terrain metrics, a
// horizontally uniform base state and doubly periodic sides (no counterpart
// in the reference, which has no adjoint; the forward it transposes is
// cuda_psolver3_terrain_open.py:774-893).  For the fixed thv, ppterm and
// piadv the forward ran with, it maps the adjoints of the stage's outputs to
// those of its inputs, in binary32.  `arrays` is a host table of 33 device
// pointers:
//   thv ppterm piadv | rho0_3d rr0 rf0 c1 c2 (terrain_static_shapes) |
//   profiles (sigma|sigmaf|rds|rdsf..., nk+2 each) |
//   qk pk1 pk2 mm fpkb elim_c elim_d (pressure-shaped scratch) |
//   u3d v3d w3d pp3d ppd ppx rru rrv rrw (in: output adjoints; out: u3d v3d
//   w3d pp3d ppx hold input adjoints, the rest is consumed) |
//   uten vten wten ppten ua va wa ppi (input adjoints, accumulated: zero them)
// with the forward's shapes.  ua, va and ppi receive nothing at nrk = 1.
int gm1hip_psolver3_stage_adjoint(void* const* arrays, int ni, int nj, int nk, int nrk, int nloop, float dts,
                                  float dx, float dy, float alph, float kdiv, void* stream);

// One RK stage of the dry large step on halo-free periodic fields (flat
// terrain, uniform base column, CM1's (1,1,1,1) sides): no counterpart in the
// reference's interface (it restates cuda_gm1_forecast.py:1241-1345 for the
// adjoint's trajectory; the reference has no adjoint).  `table` holds the
// device pointers of the workspace in the order of gm1_amd/adjoint_step.py
// SLOTS, `consts` the stage's 23 binary32 scalars (stage_consts).  Fields are
// (nk, nj, ni), w-type fields (nk+1, nj, ni), substep states (nloop+1)
// planes.  ni, nj, nk >= 6.
int gm1hip_dry_stage_forward(void* const* table, const float* consts, int ni, int nj, int nk, int nrk, int nloop,
                             double weps, void* stream);

// Adjoint of one gm1hip_dry_stage_forward call (no counterpart in the
// reference, which has no adjoint).  Run it right after the forward call of
// the same stage on the same workspace (it reads that call's intermediates
// and substep states).  In: B_OU B_OV B_OW B_OTH B_OPP (adjoints of the
// stage's u v w theta' pp'), B_PPX (adjoint of the stage's ppx; at nrk 3 of
// solve3's pp + ppx).  Out: B_XU B_XV B_XW B_XTH B_XPP (adjoints of the
// stage's input state), B_PPX (adjoint of the ppx it read); B_AU B_AV B_AW
// B_ATH B_API (adjoints of the step-start state) are accumulated.  Upwind
// choices are held at their forward values; the last of `consts` selects
// whether the theta WENO5 weights and limiters are differentiated (1) or
// frozen at their forward values (0).
int gm1hip_dry_stage_adjoint(void* const* table, const float* consts, int ni, int nj, int nk, int nrk, int nloop,
                             double weps, void* stream);

// ----------------------------------------------------------------- boundaries
// Family halo fill of one padded field (cuda_boundary_families.py:887-940,
// cpu_family_we_inplace / cpu_family_sn_inplace); kind 0 scalar, 1 u, 2 v, 3 w.
int gm1hip_fill_family(float* a, int kind, unsigned sides, int ni, int nj, int nk, void* stream);
// One-cell pressure halo (cuda_boundary_families.py:941-964, cpu_pressure_we/sn_inplace).
int gm1hip_fill_pressure(float* p, unsigned sides, int ni, int nj, int nk, void* stream);
// Rigid floor and lid, w = 0 at k = 1 and nk+1 (cuda_sound3_open_boundary.py:315, flag 1).
int gm1hip_w_rigid(float* w, int ni, int nj, int nk, void* stream);
// w(k=1) from the terrain slope (cuda_terrain_open_boundary.py:641-665).
int gm1hip_bcwsfc(const float* u3d, const float* v3d, const float* dzdx, const float* dzdy, const float* gz,
                  float* w3d, int ni, int nj, void* stream);
// Stage-entry w(k=1) halo in W, E, S, N order (cuda_momentum_rk.py:274-290).
int gm1hip_bc2d_open(float* w3d, int ni, int nj, void* stream);

// ----------------------------------------------------------------- RK stage bridge
// eqtset=2 stage terms (cuda_moist_stage_bridge.py:378-430 and
// cpu_eqtset2_cell_f32 :62-128): thv, ppterm, piadv (pressure family),
// dum8 (padded scalar, g*b), thterm (cells); thv and dum8 halos filled.
// q3d is (numq, nk, nj, ni) in the order qv, qc, qr, qi, qs, qg (numq 0..6;
// missing species are zero); qc0/qi0 may be NULL (zero profiles).
int gm1hip_stage_terms(const float* th0, const float* qv0, const float* pi0, const float* qc0, const float* qi0,
                       const float* th3d, const float* q3d, int numq, const float* pp3d, float* thv, float* ppterm,
                       float* piadv, float* dum8, float* thterm, int ni, int nj, int nk, unsigned sides,
                       void* stream);
// Buoyancy added into wten/uten/vten (cuda_moist_stage_bridge.py:433-453).
int gm1hip_buoyancy(const float* dum8, const float* c1, const float* c2, const float* gxu, const float* gyv,
                    float* uten, float* vten, float* wten, int ni, int nj, int nk, void* stream);
// Post-acoustic bridge (cuda_moist_stage_bridge.py:456-520): divergence of
// the time-averaged winds, theta_div = div*thterm, then rru/rrv/rrw rebuilt
// in place as mass fluxes and divx.
int gm1hip_post_acoustic(float* rru, float* rrv, float* rrw, const float* rho0_3d, const float* rf0,
                         const float* rgzu, const float* rgzv, const float* dzdx, const float* dzdy, const float* gz,
                         const float* thterm, float* divergence, float* theta_div, float* divx, int ni, int nj,
                         int nk, float dx, float dy, float dz, void* stream);
// Scalar packing into the padded carrier, theta as (th0 - 300) + th3d when
// theta != 0 (cuda_scalar_moist_rk.py:255-280), then its family halo.
int gm1hip_pack_scalar(const float* th0, const float* src, float* padded, int ni, int nj, int nk, int theta,
                       unsigned sides, void* stream);
// A-anchored RK compose out = a0 + dttmp*((held - divterm) + adv) with
// |out| < 1e-30 -> 0 (cuda_scalar_moist_rk.py:283-349); divterm may be NULL.
int gm1hip_rk_compose(const float* a0, const float* held, const float* divterm, const float* adv, float* out,
                      int64_t n, float dttmp, void* stream);
// Held momentum tendencies copied over i=1:ni+1, j=1:nj+1, k=1:nk
// (cuda_momentum_rk.py:242-271).
int gm1hip_seed_held_momentum(const float* uten1, const float* vten1, const float* wten1, float* uten, float* vten,
                              float* wten, int ni, int nj, int nk, void* stream);
// The held seed and the terrain momentum advection of one RK stage
// (cuda_momentum_rk.py:303-372, statements 1 and 3): every word equals
// gm1hip_seed_held_momentum followed by gm1hip_momentum_advection5, the seed
// folded into the advection updates.  All three held tendencies null: they
// are +0 (a step without a solve1 contribution) and are not read; every word
// equals the seed of zeroed fields followed by the advection.
int gm1hip_momentum_stage(const float* u3d, const float* v3d, const float* w3d, const float* uten1,
                          const float* vten1, const float* wten1, float* uten, float* vten, float* wten, float* rru,
                          float* rrv, float* rrw, const float* uh, const float* vh, const float* rho0_3d,
                          const float* rf0, const float* rrf0, const float* c2, const float* gz, const float* gzu,
                          const float* gzv, const float* rgzu, const float* rgzv, const float* dzdx,
                          const float* dzdy, float* divx, int ni, int nj, int nk, float dx, float dy, float dz,
                          unsigned obc, void* stream);
// solve1 held compose: ppten from held theta/qv and the microphysics held
// rates (kernel gm1_predict_and_compose_held_v1, cuda_gm1_forecast.py:555-566).
// zero_held != 0 states that thheld and qheld hold +0 (a step without a solve1
// contribution): they are not read, every word written is the same as with
// zeroed fields, and qheld's species beyond qc are left as they are (the
// caller keeps them +0).
int gm1hip_held_compose(const float* th0, const float* pi0, const float* rho, const float* ppi, const float* tha,
                        const float* qa, float* thheld, float* qheld, float* ppten, const float* qpten,
                        const float* qtten, const float* qvten, const float* qcten, int ni, int nj, int nk, float dt,
                        int zero_held, void* stream);
// solve3 pressure: ppx = pp3d + ppx, ppi = pp3d and their halos
// (cuda_gm1_forecast.py:1414-1435, _finish_solve3).  With pp3d_mass (the
// microphysics' (nk,nj,ni) pp3d, may be NULL) the new pp3d comes from it and
// is copied into the pressure family first.
int gm1hip_solve3_pressure(float* ppx, float* pp3d, float* ppi, const float* pp3d_mass, int ni, int nj, int nk,
                           unsigned sides, void* stream);
// solve3 density: rr = 1/rho, rf by c1/c2 with 3-point end extrapolation
// (kernel gm1_solve3_density_v1, cuda_gm1_forecast.py:568-577).
int gm1hip_solve3_density(const float* rho, const float* c1, const float* c2, float* rr, float* rf, int ni, int nj,
                          int nk, void* stream);

// ----------------------------------------------------------------- Kessler
// Device scratch (bytes) gm1hip_kessler_driver needs at ni x nj x nk.
int64_t gm1hip_kessler_scratch_bytes(int ni, int nj, int nk);

// Replaces launch_cuda_gm1_kessler_driver (cuda_gm1_kessler_driver.py:907-1087,
// ptype 1, efall 0, eqtset 1 or 2): pdefq of qv/qc/qr, fall speed, fallout in
// nfall substeps, autoconversion/accretion/evaporation
// (cuda_gm1_kessler_local.py:251-426) and saturation adjustment (:337-428).
// Mass arrays are (nk,nj,ni); rain (1,nj,ni), prate (nj,ni); asq (3) and
// qbudget (10) are binary64 and accumulate.  status[0] keeps the largest
// satadj iteration count, status[1] counts cells that did not converge in
// 100 iterations.  With qpten..qcten non-NULL (holding the eqtset=2
// snapshot) the held rates are recaptured, (state - snapshot)/dt.
int gm1hip_kessler_driver(const float* ruh, const float* rvh, const float* rmh, const float* mh, const float* pi0,
                          const float* th0, float* rho, float* rr, float* pp3d, float* th3d, float* prs, float* qv,
                          float* qc, float* qr, float* fall_velocity, float* rain, float* prate, double* asq,
                          double* qbudget, int* status, float* qpten, float* qtten, float* qvten, float* qcten,
                          void* scratch, int64_t scratch_bytes, int ni, int nj, int nk, float dx, float dy, float dz,
                          float dt, int eqtset, void* stream);

// gm1hip_kessler_driver, plus one uint32 word per cell (decisions, (nk,nj,ni),
// may be NULL) recording the branches the cell took: bit 0 qc > 0, bit 1
// autoconversion (qc > 0.001), bit 2 accretion, bit 3 evaporation (qr > 0 and
// qv < qvs), bit 4 evaporation capped at saturation, bit 5 the ar + cr > qc
// limiter, bit 6 the theta/pi' update (er > 1e-7), bit 7 the saturation
// adjustment ran, bits 8..15 its iteration count, bit 16 it did not converge.
// The reference's driver takes these branches inside
// cpu_gm1_kessler_local_inplace (cuda_gm1_kessler_local.py:290-416) and
// _cpu_satadj (cuda_gm1_kessler_driver.py:337-428); the word lets a caller
// check them cell by cell.
int gm1hip_kessler_driver_decisions(const float* ruh, const float* rvh, const float* rmh, const float* mh,
                                    const float* pi0, const float* th0, float* rho, float* rr, float* pp3d,
                                    float* th3d, float* prs, float* qv, float* qc, float* qr, float* fall_velocity,
                                    float* rain, float* prate, double* asq, double* qbudget, int* status,
                                    float* qpten, float* qtten, float* qvten, float* qcten, unsigned* decisions,
                                    void* scratch, int64_t scratch_bytes, int ni, int nj, int nk, float dx, float dy,
                                    float dz, float dt, int eqtset, void* stream);

// The driver with efall=1 (cuda_gm1_kessler_driver.py:482-551 with
// launch_cuda_gm1_kessler_efall, cuda_gm1_kessler_efall.py:380-431): the
// energy fallout (getcvm, getefall, the temperature / prs / pp3d / th3d
// update where |dt * tendency| >= 1e-4, every level below the top) between
// the surface-energy budget and the mass fallout, and the local conversion
// reading the held temperature.  mf is (nk+1, nj, ni); temperature (nk, nj,
// ni) receives the held temperature; nk >= 2.  decisions may be NULL.
int gm1hip_kessler_driver_efall(const float* ruh, const float* rvh, const float* rmh, const float* mh,
                                const float* mf, const float* pi0, const float* th0, float* rho, float* rr,
                                float* pp3d, float* th3d, float* prs, float* qv, float* qc, float* qr,
                                float* fall_velocity, float* rain, float* prate, float* temperature, double* asq,
                                double* qbudget, int* status, float* qpten, float* qtten, float* qvten, float* qcten,
                                unsigned* decisions, void* scratch, int64_t scratch_bytes, int ni, int nj, int nk,
                                float dx, float dy, float dz, float dt, int eqtset, void* stream);

// No counterpart in the reference: gm1hip_kessler_driver for callers that
// cannot allocate (code under HIP graph capture, tight serving loops, harnesses
// that forbid hipMalloc).  The same kernels and results, with no workspace
// argument: the driver's scratch is a file-scope device array of 64 MiB
// (enough for 640x768x80; gm1hip_kessler_scratch_bytes gives the need), which
// each GPU holds until the process ends.  Inputs whose need exceeds it return
// hipErrorInvalidValue before any launch.  The call allocates nothing.  Every
// call on one GPU shares the array: calls that overlap on one GPU race.
int gm1hip_kessler_driver_noalloc(const float* ruh, const float* rvh, const float* rmh, const float* mh,
                                  const float* pi0, const float* th0, float* rho, float* rr, float* pp3d, float* th3d,
                                  float* prs, float* qv, float* qc, float* qr, float* fall_velocity, float* rain,
                                  float* prate, double* asq, double* qbudget, int* status, float* qpten,
                                  float* qtten, float* qvten, float* qcten, int ni, int nj, int nk, float dx,
                                  float dy, float dz, float dt, int eqtset, void* stream);

// First half of cpu_gm1_kessler_large_step_inplace
// (cuda_gm1_kessler_large_step.py:238-268): eqtset=2 held remove/snapshot
// (state -= rate*dt, rate <- state) and the pre-physics equation of state.
// pp_family (pressure family, may be NULL) supplies pp3d; rr <- rho when
// copy_rr (no dry-mass hook follows).
int gm1hip_kessler_transition_pre(const float* pp_family, float* pp3d, float* th3d, float* qv, float* qc,
                                  float* qpten, float* qtten, float* qvten, float* qcten, const float* pi0,
                                  const float* th0, float* prs, float* rho, float* rr, int ni, int nj, int nk,
                                  float dt, int eqtset, int copy_rr, void* stream);

// Device scratch (bytes) gm1hip_apmasscon needs.
int64_t gm1hip_dry_mass_scratch_bytes(int ni, int nj, int nk);

// Dry-air mass ledger and, with mass1 > 0, the Exner correction of
// cpu_apmasscon_inplace (cuda_apmasscon.py:256-327) applied to
// pp3d/prs/rho (rr <- rho when rr is non-NULL).  ledger[0] = mass before,
// ledger[1] = sum(pi0 + pp3d) before (binary64, fixed reduction order),
// corr[0] = the binary32 correction.
int gm1hip_apmasscon(const float* ruh, const float* rvh, const float* rmh, const float* pi0, const float* th0,
                     const float* th3d, const float* qv, float* pp3d, float* prs, float* rho, float* rr,
                     double* ledger, float* corr, double mass1, void* scratch, int64_t scratch_bytes, int ni, int nj,
                     int nk, void* stream);

// gm1hip_kessler_transition_pre (copy_rr = 0) followed by gm1hip_apmasscon
// with mass1 > 0, the ledger's partials formed in the transition's own pass
// from the rho and pp3d it writes: the transition's dry-mass hook
// (cuda_gm1_kessler_large_step.py:183-292, apmasscon between the EOS and
// the driver) with the same words as the two calls.
int gm1hip_kessler_transition_pre_apmasscon(const float* pp_family, float* pp3d, float* th3d, float* qv, float* qc,
                                            float* qpten, float* qtten, float* qvten, float* qcten, const float* pi0,
                                            const float* th0, float* prs, float* rho, float* rr, const float* ruh,
                                            const float* rvh, const float* rmh, double* ledger, float* corr,
                                            double mass1, void* scratch, int64_t scratch_bytes, int ni, int nj,
                                            int nk, float dt, int eqtset, void* stream);

// ----------------------------------------------------------------- flat stretched column
// The stretched Cartesian lane (cuda_cartesian_dynamics.py:1070-1262): CM1's
// terrain_flag=.false. statements on a stretch_z column.  `flat_column` is a
// device buffer of mh (nk+2) then mf (nk+3), CM1-indexed
// (cm1_vertical_grid.py:616-652); rdz = 1/dz of the metric reference
// (param.F:6327).  The base state is read at the first physical column.

// Replaces launch_cuda_gm1_child_stretched_momentum5 on a flat column
// (cuda_gm1_child_stretched_momentum.py:430-757; texts :215-318): rru, rrv,
// rrw, divx (smeps 1e-30) and the advu/advv/advw increments.
int gm1hip_momentum_advection5_flat(const float* u3d, const float* v3d, const float* w3d, float* uten, float* vten,
                                    float* wten, float* rru, float* rrv, float* rrw, const float* uh,
                                    const float* vh, const float* rho0_3d, const float* rf0, const float* rr0,
                                    const float* rrf0, const float* c2, const float* flat_column, float* divx,
                                    int ni, int nj, int nk, float dx, float dy, float rdz, unsigned obc,
                                    void* stream);

// Replaces launch_cuda_gm1_flat_stretched_buoyancy
// (cuda_gm1_child_stretched_moist.py:240-285): wten += c1*dum8(k-1) + c2*dum8(k).
int gm1hip_buoyancy_flat(const float* dum8, const float* c1, const float* c2, float* wten, int ni, int nj, int nk,
                         void* stream);

// One flat-column RK stage's momentum statements in the composer's order
// (cuda_cartesian_dynamics.py:1070-1262: the held seed of cuda_momentum_rk.py:242-271,
// launch_cuda_gm1_child_stretched_momentum5, launch_cuda_gm1_flat_stretched_buoyancy) in
// three passes: every word equals gm1hip_seed_held_momentum, then
// gm1hip_momentum_advection5_flat, then gm1hip_buoyancy_flat.  Null uten1,
// vten1 and wten1 as in gm1hip_momentum_stage.
int gm1hip_momentum_stage_flat(const float* u3d, const float* v3d, const float* w3d, const float* uten1,
                               const float* vten1, const float* wten1, float* uten, float* vten, float* wten,
                               float* rru, float* rrv, float* rrw, const float* uh, const float* vh,
                               const float* rho0_3d, const float* rf0, const float* rr0, const float* rrf0,
                               const float* c1, const float* c2, const float* dum8, const float* flat_column,
                               float* divx, int ni, int nj, int nk, float dx, float dy, float rdz, unsigned obc,
                               void* stream);

// Replaces launch_cuda_gm1_child_stretched_post_acoustic on a flat column
// (cuda_gm1_child_stretched_moist.py:286-372; misclibs.F's flat getdiv/getdivx).
int gm1hip_post_acoustic_flat(float* rru, float* rrv, float* rrw, const float* rho0_3d, const float* rf0,
                              const float* flat_column, const float* thterm, float* divergence, float* theta_div,
                              float* divx, int ni, int nj, int nk, float dx, float dy, float rdz, void* stream);

// The flat column's scalar advection fused with the A-anchored compose
// (cuda_gm1_child_stretched_scalar.py:262-352, adv.F:458-478, 555-572,
// 646-656), as gm1hip_scalar_rk_stage for the terrain family.
int gm1hip_scalar_rk_stage_flat(const float* scalar, const float* rru, const float* rrv, const float* rrw,
                                const float* c1, const float* c2, const float* rr0, const float* flat_column,
                                const float* divx, const float* a0, const float* held, const float* divterm,
                                float* out, int ni, int nj, int nk, int use_weno, double weps, float dx, float dy,
                                float rdz, unsigned obc, float dttmp, void* stream);

// The flat column's final-stage positive-definite moisture chain
// (launch_cuda_gm1_child_stretched_pdef_moisture, cuda_gm1_child_stretched_scalar.py:851;
// texts :355-448), as gm1hip_pdef_moisture_stage for the terrain family.
int gm1hip_pdef_moisture_stage_flat(const float* scalar, const float* rru, const float* rrv, const float* rrw,
                                    const float* c1, const float* c2, const float* rho0_3d, const float* rr0,
                                    const float* flat_column, const float* divx, const float* qa, const float* held,
                                    float* out, float* advx, float* advy, float* advz, int ni, int nj, int nk,
                                    int pdefweno, double weps, float dx, float dy, float rdz, unsigned obc,
                                    float dttmp, void* stream);

// gm1hip_scalar_rk_stage (terrain: gz, rdz = 1/dz) and
// gm1hip_scalar_rk_stage_flat (flat_column, metric rdz) reading the scalar
// from its unpadded carrier `src` (nk, nj, ni) -- theta' with th0 (nk) for the
// packed theta (th0 - 300) + theta' of cuda_scalar_moist_rk.py:255-280, or a
// moisture species with th0 = NULL -- with the pack and the lateral family
// fill (cuda_boundary_families.py:887-938) folded into the loads.  Exactly one
// of gz and flat_column is non-NULL.  The large step's own call: the
// reference packs into a padded carrier first (cuda_scalar_moist_rk.py:1041-1160).
int gm1hip_scalar_rk_stage_src(const float* src, const float* th0, const float* rru, const float* rrv,
                               const float* rrw, const float* c1, const float* c2, const float* rr0,
                               const float* gz, const float* flat_column, const float* divx, const float* a0,
                               const float* held, const float* divterm, float* out, int ni, int nj, int nk,
                               int use_weno, double weps, float dx, float dy, float rdz, unsigned obc, float dttmp,
                               void* stream);

// gm1hip_scalar_rk_stage_src with use_weno = 0 for ns (1..4) scalars in one
// pass: host arrays of ns device pointers (th0s and divterms may be NULL, or
// hold NULL entries), every other argument shared.  Each out[q] receives the
// words gm1hip_scalar_rk_stage_src writes for scalar q; the reference advects
// theta and then each species in passes of their own within one stage
// (cuda_scalar_moist_rk.py:1296 and :1314-1340), this form shares the mass
// fluxes, divx and addressing among them.
int gm1hip_scalar_rk_stage_src_batch(int ns, const float* const* srcs, const float* const* th0s, const float* rru,
                                     const float* rrv, const float* rrw, const float* c1, const float* c2,
                                     const float* rr0, const float* gz, const float* flat_column, const float* divx,
                                     const float* const* a0s, const float* const* helds,
                                     const float* const* divterms, float* const* outs, int ni, int nj, int nk,
                                     float dx, float dy, float rdz, unsigned obc, float dttmp, void* stream);

// gm1hip_pdef_moisture_stage / _flat on an unpadded moisture carrier `src`,
// as gm1hip_scalar_rk_stage_src (terrain: gz and rgz; flat: flat_column).
int gm1hip_pdef_moisture_stage_src(const float* src, const float* rru, const float* rrv, const float* rrw,
                                   const float* c1, const float* c2, const float* rho0_3d, const float* rr0,
                                   const float* gz, const float* rgz, const float* flat_column, const float* divx,
                                   const float* qa, const float* held, float* out, float* advx, float* advy,
                                   float* advz, int ni, int nj, int nk, int pdefweno, double weps, float dx, float dy,
                                   float rdz, unsigned obc, float dttmp, void* stream);

// Replaces launch_cuda_gm1_flat_acoustic_stage for the all-open Durran-Klemp
// family (cuda_gm1_flat_lateral_acoustic.py:117; cuda_gm1_child_psolver3_open.py
// :340-535, 1618-1792): flat coefficients, flat ssopenbc, the pressure
// gradient on every live face, the flat fpk, the Thomas solve, the update
// with w(k), w(k+1) as they are, the time averages, and (flags & 2) the stage
// fills without bcwsfc.  `profiles` is the packed terrain profile buffer
// (unused by the flat statements; kept for the shared argument block).
int gm1hip_psolver3_stage_flat(
    float* ppi, float* pp3d, float* ppx, float* ppd, const float* piadv, const float* ppterm, const float* ppten,
    const float* thv, float* qk, float* pk1, float* pk2, float* mm, float* fpk, float* elim_c, float* elim_d,
    const float* ua, float* u3d, const float* uten, float* rru, const float* va, float* v3d, const float* vten,
    float* rrv, const float* wa, float* w3d, const float* wten, float* rrw, float* radbcw, float* radbce,
    float* radbcs, float* radbcn, const float* uh, const float* vh, const float* ruf, const float* rvf,
    float* scalar, const float* rho0_3d, const float* rr0, const float* rf0, const float* c1, const float* c2,
    const float* flat_column, float rdz, const float* profiles, int ni, int nj, int nk, int nrk, int nloop,
    float dts, float dttmp, float dx, float dy, float alph, float kdiv, unsigned sides, int flags, void* stream);

// The same stage as gm1hip_psolver3_stage_flat, word for word, with the
// pressure-gradient step of each small step (cuda_gm1_child_psolver3_open.py
// :387-399, launched at :1701-1711) inside the pipelined column pass: one pass per
// small step instead of two.  The reference has no fused form; it replaces the
// same launch_cuda_gm1_flat_acoustic_stage.  u_alt and v_alt (the u3d and v3d
// shapes) and ppd_alt (the ppd shape) are workspace distinct from those
// arrays: the small steps alternate between them so that no pass reads a word
// another workgroup of the pass writes, and the stage ends in the caller's
// arrays (nothing is kept in the workspace between calls).  flags bit 8
// selects 16-column workgroups; the pipelined and fused bits are implied.
// Refused (hipErrorInvalidValue, nothing launched): a missing or aliased
// alternate, a column too tall for the fused pass's LDS, or an array of
// 4 GiB or more (the pass addresses by 32-bit byte offsets).
int gm1hip_psolver3_stage_flat_fused(
    float* ppi, float* pp3d, float* ppx, float* ppd, const float* piadv, const float* ppterm, const float* ppten,
    const float* thv, float* qk, float* pk1, float* pk2, float* mm, float* fpk, float* elim_c, float* elim_d,
    const float* ua, float* u3d, const float* uten, float* rru, const float* va, float* v3d, const float* vten,
    float* rrv, const float* wa, float* w3d, const float* wten, float* rrw, float* radbcw, float* radbce,
    float* radbcs, float* radbcn, const float* uh, const float* vh, const float* ruf, const float* rvf,
    float* scalar, const float* rho0_3d, const float* rr0, const float* rf0, const float* c1, const float* c2,
    const float* flat_column, float rdz, const float* profiles, float* u_alt, float* v_alt, float* ppd_alt, int ni,
    int nj, int nk, int nrk, int nloop, float dts, float dttmp, float dx, float dy, float alph, float kdiv,
    unsigned sides, int flags, void* stream);

// ----------------------------------------------------------------- measurement
// No counterpart in the reference: live per-kernel GPU time for throughput
// reports.  arm() creates `capacity` event pairs (before a timed region) and
// brackets every launch of kernel id 1 k_snd_column, 2 k_snd_pgrad,
// 3 k_scalar_adv5, 4/5/6 k_mom_advu/advv/advw, or 7 one acoustic small step's
// pressure-gradient and column work (one bracket per small step, around one
// pass in the fused forms and two in the others); 0 disarms.  collect()
// waits, sums and disarms; it fails if more launches happened than were armed.
int gm1hip_kernel_timer_arm(int kernel, int capacity);
int gm1hip_kernel_timer_collect(double* total_ms, int* launches);

#ifdef __cplusplus
}
#endif
