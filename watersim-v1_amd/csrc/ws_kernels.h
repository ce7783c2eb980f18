// Kernel declarations and the per-call parameter block of the substep.
#pragma once
#include "ws_internal.h"
#include "ws_stage.h"
#include "ws_world.h"

namespace ws {

// particle-phase tunables (solver.js:56-75, 145-193) as floats
struct Par {
  float pic, maxSpeed, gravity, oceanR, atmosphereH, domainR, coreR, W, H, D;
  float sunX, sunY, sunZ;  // unit direction to the sun, (1,0,0) without one
  float rainTe, rainLiftAge;
  int ceilReflect;
};

// thermal tick, evaporation, currents and phase changes
struct ThermalPar {
  float dt, sunX, sunY, sunZ, sunPow, shadeCool, Tamb, alpha;
  float Rlo2, Rhi2, Rhi, coreR2, inv_dv;
  int hasSun, hasTerrain, tn, conduct;
};
// evaporation law inputs as the reference holds them (doubles): the
// candidate tests and the escape threshold are evaluated in double
struct EvapPar {
  double dt, rate, freezeT;
  double dx, cx, cy, cz;
};
struct CurrPar {
  int gnx, gny, gnz, kMax;
  float sx, invS, t, amp, relax, oceanR, shell;
  float drift[3];
  float wave[3][3][5];  // octave, potential component: kx, ky, kz, A, phase
};
struct PhasePar {
  float snowT, meltT, rainT, cloudT, cloudP, rainLiftAge, cloudBase, dt, damp;
  uint32_t tick;
  int spinOn;
};

__global__ void k_begin_substep(DevScalars *s, double dt_host, int use_law, double cfl_dx, double dt_base,
                                float gravity);
__global__ void k_end_substep(DevScalars *s);
__global__ void k_init_cells(const uint8_t *st, uint8_t *ct, unsigned *cnt, int nC);
__global__ void k_mark(Geo g, World w, int nP, const float *px, const float *py, const float *pz, uint8_t *pflag,
                       uint8_t *ct, unsigned *cnt, int *pcell, unsigned *prank, DevScalars *s);
__global__ void k_scatter(int nP, const int *pcell, const unsigned *prank, const unsigned *start,
                          unsigned *sorted);
template <int TX, int TY, int TZ>
__global__ void k_p2g_tile(Geo g, int tilesX, int tilesY, const unsigned *start, const unsigned *sorted,
                           const float *px, const float *py, const float *pz, const float *pvx, const float *pvy,
                           const float *pvz, float *vel, float *wgt);
__global__ void k_bc_snapshot(Geo g, const uint8_t *ct, float *vel, float *velO, const float *wgt, uint8_t *valid,
                              int fuse_gravity, const float *gvec, const DevScalars *s, int write_valid,
                              const float *sv);
__global__ void k_valid(Geo g, const uint8_t *ct, const float *wgt, uint8_t *valid);
__global__ void k_gravity_bc(Geo g, const uint8_t *ct, float *vel, const float *gvec, const DevScalars *s,
                             const float *sv);
__global__ void k_visc_iter(Geo g, const uint8_t *ct, const float *old, const float *cur, float *out, float nu_dx2,
                            const DevScalars *s, int first);
__global__ void k_assemble(Geo g, const uint8_t *ct, const float *vel, float *q, float *q2, float *rhs, float *pw,
                           uint8_t *nbm, uint8_t *dg, float omega);
__global__ void k_sor_tile(Geo g, int tilesX, int tilesY, const uint8_t *active, const float *qin, float *qout,
                           const float *rhs, const float *pw, const uint8_t *nbm, float om1);
__global__ void k_sor_tiles_mark(Geo g, int tilesX, int tilesY, int nTiles, const uint8_t *nbm, uint8_t *active);
__global__ void k_sor(Geo g, float *q, const float *rhs, const float *pw, const uint8_t *nbm, int color,
                      float om1);
__global__ void k_residual(Geo g, const float *q, const float *rhs, const uint8_t *nbm, const uint8_t *dg,
                           double *partials);
__global__ void k_sor_count(Geo g, const uint8_t *nbm, unsigned long long *key);
__global__ void k_sor_compact(Geo g, const uint8_t *nbm, const uint8_t *dg, const float *rhs, const float *pw,
                              const unsigned long long *off, long half, unsigned *code, float *lrhs, float *lpw,
                              uint8_t *ldg, DevScalars *ds);
__global__ void k_sor_list(Geo g, int n, const unsigned *code, const float *lrhs, const float *lpw, float *q,
                           float om1);
__global__ void k_residual_list(Geo g, const DevScalars *ds, long half, const unsigned *code, const float *lrhs,
                                const uint8_t *ldg, const float *q, double *partials);
template <int S>
__global__ void k_sor_tb(Geo g, int tilesX, int tilesY, const uint8_t *active, const float *qin, float *qout,
                         const unsigned *pk, const float *lrhs, const int *cnt, float omega, float om1);
__global__ void k_sor_tb_build(Geo g, int tilesX, int tilesY, const uint8_t *active, const uint8_t *nbm,
                               const uint8_t *dg, const float *rhs, unsigned *pk, float *lrhs, int *cnt);
__global__ void k_sor_tb_mark(Geo g, int tilesX, int tilesY, int nTiles, const uint8_t *nbm, uint8_t *active);
__global__ void k_residual_final(const double *partials, int n, DevScalars *s, DevScalars *hs, unsigned *seqWord,
                                 unsigned seq);
template <int NT>
__global__ void k_sor_solve(SorJob J);
template <int NT>
__global__ void k_sor_solve_batch(const SorJob *jobs, const int2 *blk);
template <int NT>
__global__ void k_sor_solve_multi(SorJob J);
__global__ void k_sor_arm(int n, unsigned *arrive);
template <int NT>
__global__ void k_sor_flow(SorJob J, int X);
__global__ void k_flow_arm(int G, unsigned *flow, unsigned *arrive);
template <int NT, int U, int H>
__global__ void k_sor_xcd(XcdSweep J);
__global__ void k_exf_rows(Geo g, const uint8_t *valid, unsigned long long *fin, ExRun R);
__global__ void k_exf_k(Geo g, unsigned long long *fin);
__global__ void k_extrap_rows(Geo g, float *vel, uint8_t *valid, const unsigned long long *fin, ExRun R,
                              DevScalars *ds);
__global__ void k_extrap_check(Geo g, ExRun R, DevScalars *ds);
__global__ void k_xcd_arm(int G, unsigned *flow);
__global__ void k_project(Geo g, const uint8_t *ct, const float *phi, float *vel, int do_clamp, float vcap,
                          DevScalars *s);
__global__ void k_clamp(long F, float *vel, float vcap, DevScalars *s);
__global__ void k_vort_center(Geo g, const float *vel, float *vc);
__global__ void k_vort_omega(Geo g, const float *vc, float *om, float inv2);
__global__ void k_vort_force(Geo g, const uint8_t *ct, const float *om, float *fc, float eps,
                             const DevScalars *s);
__global__ void k_vort_splat(Geo g, const uint8_t *ct, const float *fc, float *vel);
__global__ void k_extrap_diag(Geo g, ExDiag D, float *vel, uint8_t *valid);
__global__ void k_extrap_diag_batch(const ExPlanet *planets, const int4 *tiles, const int *planetOf, int off);

__global__ void k_g2p_advect(Geo g, World w, Par P, int nP, int mode, const float *vel, const float *velO,
                             const uint8_t *ct, float *px, float *py, float *pz, float *pvx, float *pvy, float *pvz,
                             uint8_t *pflag, uint8_t *pcool, float *pAir, float *advU, float *advV, float *advW,
                             DevScalars *s);
__global__ void k_vapor_a(Geo g, World w, Par P, int nP, float *px, float *py, float *pz, float *pvx, float *pvy,
                          float *pvz, const uint8_t *pflag, float *pAir, float *scr, unsigned *need,
                          const DevScalars *s);
__global__ void k_vapor_b(Geo g, Par P, int nP, const float *px, const float *py, const float *pz, float *pvx,
                          float *pvy, float *pvz, const uint8_t *pflag, const float *scr, const unsigned *needA,
                          const unsigned *offA, unsigned *needB, const DevScalars *s, const uint32_t *rng);
__global__ void k_vapor_c(Geo g, World w, Par P, int nP, float *px, float *py, float *pz, float *pvx, float *pvy,
                          float *pvz, uint8_t *pflag, float *pAir, const float *pT, const uint8_t *ct,
                          const float *scr, const unsigned *needB, const unsigned *offB, const DevScalars *s,
                          const uint32_t *rng);
__global__ void k_rng_advance(uint32_t *rng, int which, const unsigned *need, const unsigned *off, int nP);
__global__ void k_sample_vel(Geo g, long n, const float *xyz, const float *vel, float *out);

// the substep's stages as launches over many planets (ws_stage.h); the
// P2G ops are indexed like P2G_BINS
extern StageOps k_begin_substep_ops, k_end_substep_ops, k_init_cells_ops, k_mark_ops, k_scatter_ops,
    k_bc_snapshot_ops, k_gravity_bc_ops, k_visc_iter_ops, k_assemble_ops, k_sor_count_ops, k_sor_compact_ops,
    k_project_ops, k_clamp_ops, k_vort_center_ops, k_vort_omega_ops, k_vort_force_ops, k_vort_splat_ops,
    k_g2p_advect_ops, k_p2g_tile_ops0, k_p2g_tile_ops1, k_p2g_tile_ops2, k_p2g_tile_ops3;

// surface
__global__ void k_splat_density(Geo g, World w, int nP, const unsigned *start, const unsigned *sorted,
                                const float *px, const float *py, const float *pz, float *dens, float r2);
__global__ void k_mt_count(int nx, int ny, int nz, const float *dens, double iso, unsigned *cnt);
__global__ void k_mt_emit(int nx, int ny, int nz, double dx, const float *dens, double iso, const unsigned *off,
                          float *pos, float *nrm, long cap, double *vpart);
__global__ void k_mt_volume(int nx, int ny, int nz, double dx, const float *dens, double iso, double *vpart,
                            unsigned *cnt);
__global__ void k_sum_f64(const double *part, int n, double *out);
__global__ void k_kinetic(int nP, const float *pvx, const float *pvy, const float *pvz, double *part);

__global__ void k_heat_particle(Geo g, ThermalPar T, int nP, const float *px, const float *py, const float *pz,
                                float *pT, const uint8_t *pflag, float *pLight, uint8_t *pDepth, const uint8_t *ct,
                                const float *dens, const uint8_t *tsolid, const float *thr, unsigned long long *csum,
                                unsigned *ccnt);
__global__ void k_heat_cells(Geo g, float alpha, const unsigned long long *csum, const unsigned *ccnt, float *mean,
                             float *dT);
__global__ void k_heat_apply(Geo g, ThermalPar T, int nP, const float *px, const float *py, const float *pz,
                             float *pT, const uint8_t *pflag, float *pvx, float *pvy, float *pvz, const float *mean,
                             const float *dT, const float *thr);
__global__ void k_evap_candidates(Geo g, EvapPar E, int nP, const float *px, const float *py, const float *pz,
                                  const float *pT, const uint8_t *pflag, const uint8_t *ct, const float *thr,
                                  unsigned *cand, unsigned *thresh);
__global__ void k_evap_compact(int nP, const unsigned *cand, const unsigned *off, const unsigned *thresh,
                               unsigned *list, unsigned *thrC);
__global__ void k_evap_chooser(const unsigned *thrC, const unsigned *off, const unsigned *cand,
                               int nP, const uint32_t *jmpA, const uint32_t *jmpC, uint32_t *rng, unsigned *evList,
                               unsigned *evState, unsigned *evCount);
__global__ void k_evap_apply(Geo g, EvapPar E, const unsigned *list, const unsigned *evList, const unsigned *evState,
                             const unsigned *evCount, float *px, float *py, float *pz, float *pvx, float *pvy,
                             float *pvz, const float *pT, uint8_t *pflag, float *pAir, const float *vel,
                             const float *thr);
__global__ void k_currents_lattice(CurrPar Q, float *cf);
__global__ void k_currents_relax(Geo g, CurrPar Q, int nP, const float *px, const float *py, const float *pz,
                                 float *pvx, float *pvy, float *pvz, const uint8_t *pflag, const float *cf);
__global__ void k_steam_count(Geo g, int nP, const float *px, const float *py, const float *pz, const uint8_t *pflag,
                              const float *thr, unsigned *cnt);
__global__ void k_push_surface(Geo g, World w, int nP, float domainR, float skyHi, float outer, float *px, float *py,
                               float *pz, float *pvx, float *pvy, float *pvz, const uint8_t *pflag);
__global__ void k_phase(Geo g, PhasePar P, int nP, const float *px, const float *py, const float *pz, float *pvx,
                        float *pvy, float *pvz, const float *pT, uint8_t *pflag, float *pAir, float *pWx, float *pWy,
                        float *pWz, float *pPh, const uint8_t *ct, const unsigned *steam, const float *thr);

}  // namespace ws
