/*
 * CPU oracle for the FLIP substep of johnnybuggy/watersim-v1 (js/solver.js,
 * js/surface.js). TEST INFRASTRUCTURE ONLY: the parity tests, the smoke check
 * and the CPU-baseline leg of the benchmark use it as the checker. The product
 * (the HIP library) never links or calls it.
 *
 * Arithmetic model: the reference is JavaScript, so every local is an IEEE
 * double and every store into a Float32Array rounds to float. The oracle keeps
 * that model (double locals, OC_REAL storage, expression trees in the
 * reference's evaluation order, built with -ffp-contract=off) so that with
 * OC_REAL=float it reproduces the reference bit for bit on the stages that use
 * only + - * / sqrt floor. Built with OC_REAL=double it is the FP64 run that
 * GPU errors are reported against.
 *
 * Only the ocean-planet ("sphere") world mode is restated; the box-pool mode
 * is not on the path being rebuilt.
 */
#ifndef WS_ORACLE_H
#define WS_ORACLE_H
#include <stdint.h>

#ifndef OC_REAL
#define OC_REAL float
#endif
typedef OC_REAL oc_real;

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oc_ctx oc_ctx;

/* a ball (solver.js:3910-3912): JavaScript numbers, so doubles */
#define OC_MAX_BALLS 64
typedef struct { double x, y, z, vx, vy, vz, r, rho, submerged; } oc_ball;

oc_ctx *oc_create(int nx, int ny, int nz, double dx);
void oc_destroy(oc_ctx *c);
/* name-based binding keeps the Python mirror free of struct layouts */
int oc_set_ptr(oc_ctx *c, const char *name, void *ptr);
int oc_set_f(oc_ctx *c, const char *name, double v);
int oc_set_i(oc_ctx *c, const char *name, long long v);
double oc_get_f(oc_ctx *c, const char *name);
long long oc_get_i(oc_ctx *c, const char *name);
int oc_real_size(void);

/* stages (reference file:line in oracle.c) */
void oc_rasterize_static(oc_ctx *c);
void oc_mark_fluid(oc_ctx *c, int p0, int p1);
void oc_clear_faces(oc_ctx *c);
void oc_p2g(oc_ctx *c);
void oc_apply_bc(oc_ctx *c);
void oc_snapshot(oc_ctx *c);
void oc_viscosity(oc_ctx *c, double dt);
void oc_apply_gravity(oc_ctx *c, double dt);
int oc_pressure_assemble(oc_ctx *c);
void oc_sor_lex(oc_ctx *c, int iters);
void oc_sor_redblack(oc_ctx *c, int iters);
double oc_residual(oc_ctx *c, double *rhs_norm);
void oc_pressure_finish(oc_ctx *c);
void oc_pressure_solve(oc_ctx *c);
void oc_project(oc_ctx *c);
void oc_vorticity(oc_ctx *c, double dt);
double oc_clamp_umax(oc_ctx *c);
void oc_extrapolate(oc_ctx *c);
int oc_g2p(oc_ctx *c, double dt);
void oc_advect(oc_ctx *c, double dt);
void oc_substep(oc_ctx *c, double dt);
double oc_dt_law(oc_ctx *c);

/* frame tick */
void oc_update_heat(oc_ctx *c, double dt);
void oc_update_evaporation(oc_ctx *c, double dt);
void oc_update_currents(oc_ctx *c, double dt);
void oc_update_phase(oc_ctx *c, double dt);
/* frame-end passes (solver.js:3290-3587) */
void oc_push_surface(oc_ctx *c);
void oc_vapor_collisions(oc_ctx *c);
void oc_ice_contacts(oc_ctx *c);

/* balls (solver.js:918-938, 1005-1026, 2294-2564); state as 9 doubles per
 * ball: x y z vx vy vz r rho submerged */
int oc_set_balls(oc_ctx *c, int n, const double *s9);
int oc_get_balls(oc_ctx *c, int n, double *s9);
void oc_push_balls(oc_ctx *c);
void oc_update_balls(oc_ctx *c, double dt);
const float *oc_ball_hist(oc_ctx *c);
const float *oc_ball_surf(oc_ctx *c);

/* samplers and terrain predicates */
double oc_sample_face(oc_ctx *c, const oc_real *arr, int comp, double x, double y, double z);
void oc_sample_vel3(oc_ctx *c, double x, double y, double z, double *out);
double oc_terrain_radius_at(oc_ctx *c, double x, double y, double z);
int oc_rock_cell_at(oc_ctx *c, double x, double y, double z);
int oc_voxel_rock_at(oc_ctx *c, double x, double y, double z);

/* surface */
void oc_splat_density(oc_ctx *c);
/* marching tetrahedra over a corner field; pos/nrm may be NULL (count only) */
long oc_mt_build(const oc_real *dens, int nx, int ny, int nz, double dx, double iso,
                 float *pos, float *nrm, long cap, double *volume);
double oc_autotune_iso(oc_ctx *c);
/* Metaballs.build (surface.js:178-284): lattice, kernel radius, stencil
 * reach and contour level of a tension / detail setting, and the field */
void oc_metaball_params(int nx, double dx, double spacing, double tension, double detail, int *m, double *mdx,
                        double *R, int *rad, double *iso);
void oc_metaball_field(const oc_real *px, const oc_real *py, const oc_real *pz, const uint8_t *pflag, int nP, int m,
                       double mdx, double R, int rad, oc_real *field);
double oc_kinetic_energy(oc_ctx *c);

#ifdef __cplusplus
}
#endif
#endif
