/*
 * Host-side planet construction for the watersim library (not on the
 * substep path). Replaces the world setup of the reference FluidSolver:
 * generateTerrain (js/solver.js:272-508), _rasterizeStatic's static raster
 * (:865-916), _buildPlanetGravity (:1177-1210) and the seeded water fill of
 * _resetWaterSphere (:651-731). The arrays it returns are the inputs that
 * ws_upload() takes ("terrainR", "staticTypes", "gu"/"gv"/"gw", particles).
 */
#ifndef WATERSIM_WORLD_H
#define WATERSIM_WORLD_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int nx, ny, nz;       /* MAC grid cells */
  double dx;            /* cell size */
  double coreR;         /* solid core radius (opts.coreR) */
  double depth;         /* resetWater(depth): sea level = coreR + depth */
  double bumpiness;     /* opts.bumpiness, default 1 */
  uint32_t seed;        /* opts.terrainSeed */
  int targetParticles;  /* opts.targetParticles */
} ws_world_params;

typedef struct {
  int tn;               /* voxel lattice size (tn^3), spacing dv = dx/2 */
  double dv;
  uint8_t *solid;       /* tn^3 rock voxels */
  float *R;             /* tn^3 terrain radius */
  uint8_t *staticTypes; /* nx*ny*nz SOLID(2)/AIR(0) */
  float *gravity;       /* gu | gv | gw, unit-g radial gravity per face */
  long nP, capacity;    /* particles placed; capacity = target + 8192 */
  float *px, *py, *pz;
  double spacing, particleVolume, oceanR, coreR, Rlo, Rhi, landFrac;
} ws_world;

/* Builds the planet and its ocean fill into newly allocated host arrays. */
int ws_world_build(const ws_world_params *p, ws_world *w);
/* Releases the arrays of a built world. */
void ws_world_free(ws_world *w);
/* _buildPlanetGravity, solver.js:1177-1210: the unit-strength radial
 * gravity -oceanR^2 e/|e|^3 of every face (e from the domain centre), written
 * as gu | gv | gw into g (3 (nx+1) ny nz floats for a cubic grid). Returns 0,
 * or -1 on bad arguments. */
int ws_planet_gravity(int nx, int ny, int nz, double dx, double oceanR, float *g);

#ifdef __cplusplus
}
#endif
#endif
