/*
 * CPU oracle: a restatement of the reference substep (js/solver.js) and the
 * marching-tetrahedra mesher (js/surface.js) of johnnybuggy/watersim-v1.
 * TEST INFRASTRUCTURE ONLY (see oracle.h): the checker for the HIP library,
 * never part of it.
 *
 * Every stage cites the reference lines it follows. Locals are double and
 * array stores round to OC_REAL, reproducing JavaScript's Float32Array
 * semantics when OC_REAL is float. Sphere-world mode only.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

enum { AIR = 0, FLUID = 1, SOLID = 2 };

/* Math.PI */
static const double JS_PI = 3.141592653589793;

struct oc_ctx {
  int nx, ny, nz;
  double dx, W, H, D, cx, cy, cz, domainR;
  long nC, uN, vN, wN;
  /* tunables (solver.js:56-75, 145-193) */
  double cfl, pic, gravity, viscosity, omega, maxSpeed, vorticity;
  int pressureIters;
  /* solve variant used by the GPU build: red-black order, and optionally
   * iterate to a relative residual checked every residualEvery sweeps */
  int sorRedBlack, residualEvery, maxPressureIters, itersDone;
  double pressureTol, residual;
  double coreR, oceanR, atmosphereH, spacing, particleVolume;
  double gRamp, dtLast;
  double rainT, cloudT, rainLiftAge;
  int ceilReflect; /* 0 linear, 1 quadratic, 2 exponential */
  int hasSun;
  double sunX, sunY, sunZ;
  uint32_t cfS, bfS, evS;
  /* terrain (solver.js:499-504) */
  int tn;
  double dv;
  const uint8_t *tsolid;
  const float *tR;
  const uint8_t *staticTypes;
  const oc_real *gu, *gv, *gw;
  /* grid */
  uint8_t *cellType;
  oc_real *u, *v, *w, *uW, *vW, *wW, *uO, *vO, *wO;
  uint8_t *validU, *validV, *validW;
  oc_real *q, *rhs, *diag, *cellPhi, *dens;
  /* particles */
  int nP;
  oc_real *px, *py, *pz, *pvx, *pvy, *pvz, *pT, *pAir, *pWx, *pWy, *pWz;
  uint8_t *pflag, *pcool;
  double *advU, *advV, *advW;
  /* frame tick (solver.js:137-193, 2715-3281) */
  double heatK, Tamb, shadeCool, sunActivity, evapIntensity, snowT, meltT, cloudP, currents;
  double Rlo, Rhi, simTime;
  int spinOn, evapCount, vaporCount, nightVaporCount;
  oc_real *pLight, *pPh;
  uint8_t *pDepth;
  /* outputs */
  int airborne, nFluid;
  double umax;
  /* balls (solver.js:115, 3910-3912): the reference keeps them as objects
   * of doubles; sv* are the moving-solid cell velocities (solver.js:95-97) */
  int nBalls, hadSolidVelocity;
  oc_ball balls[OC_MAX_BALLS];
  oc_real *svx, *svy, *svz;
  float surfY[OC_MAX_BALLS], hist[OC_MAX_BALLS * 24];
  int histC[OC_MAX_BALLS];
  /* private scratch */
  int *fluidList, *pStart, *pAdj;
  float *pW;
  float *vcx, *vcy, *vcz, *omg, *omx, *omy, *omz;
};

static double clampd(double v, double a, double b) { return v < a ? a : (v > b ? b : v); }
static int clampi(int v, int a, int b) { return v < a ? a : (v > b ? b : v); }

oc_ctx *oc_create(int nx, int ny, int nz, double dx) {
  oc_ctx *c = (oc_ctx *)calloc(1, sizeof(oc_ctx));
  /* solver.js:50-91 */
  c->nx = nx; c->ny = ny; c->nz = nz; c->dx = dx;
  c->W = nx * dx; c->H = ny * dx; c->D = nz * dx;
  c->cx = c->W / 2; c->cy = c->H / 2; c->cz = c->D / 2;
  {
    double m = c->W < c->H ? c->W : c->H;
    m = m < c->D ? m : c->D;
    c->domainR = 0.5 * m - dx;
  }
  c->nC = (long)nx * ny * nz;
  c->uN = (long)(nx + 1) * ny * nz;
  c->vN = (long)nx * (ny + 1) * nz;
  c->wN = (long)nx * ny * (nz + 1);
  c->cfl = 3.0; c->pic = 0.15; c->gravity = 9.81; c->viscosity = 5e-4;
  c->pressureIters = 20; c->omega = 1.55; c->maxSpeed = 3.2;
  c->atmosphereH = 0.35; c->spacing = dx * 0.5;
  c->particleVolume = c->spacing * c->spacing * c->spacing;
  c->gRamp = 1e9;
  c->rainT = 0.22; c->cloudT = 0.34; c->rainLiftAge = 0.4;
  c->cfS = 0x2f6e2b1u; c->bfS = 0x7f4a7c15u; c->evS = 0x51ab3c77u;
  c->heatK = 0.6; c->Tamb = 0.32; c->shadeCool = 0.035; c->evapIntensity = 0.05;
  c->snowT = 0.10; c->meltT = 0.105; c->cloudP = 0.35;
  c->fluidList = (int *)malloc(sizeof(int) * c->nC);
  c->pStart = (int *)malloc(sizeof(int) * (c->nC + 1));
  c->pAdj = (int *)malloc(sizeof(int) * 6 * c->nC);
  c->pW = (float *)malloc(sizeof(float) * c->nC);
  c->vcx = (float *)malloc(sizeof(float) * c->nC * 7);
  c->vcy = c->vcx + c->nC; c->vcz = c->vcy + c->nC; c->omg = c->vcz + c->nC;
  c->omx = c->omg + c->nC; c->omy = c->omx + c->nC; c->omz = c->omy + c->nC;
  return c;
}

void oc_destroy(oc_ctx *c) {
  if (!c) return;
  free(c->fluidList); free(c->pStart); free(c->pAdj); free(c->pW); free(c->vcx);
  free(c);
}

int oc_real_size(void) { return (int)sizeof(oc_real); }

#define PTR(nm, ty) if (!strcmp(name, #nm)) { c->nm = (ty)ptr; return 0; }
int oc_set_ptr(oc_ctx *c, const char *name, void *ptr) {
  PTR(tsolid, const uint8_t *) PTR(tR, const float *) PTR(staticTypes, const uint8_t *)
  PTR(gu, const oc_real *) PTR(gv, const oc_real *) PTR(gw, const oc_real *)
  PTR(cellType, uint8_t *) PTR(u, oc_real *) PTR(v, oc_real *) PTR(w, oc_real *)
  PTR(uW, oc_real *) PTR(vW, oc_real *) PTR(wW, oc_real *)
  PTR(uO, oc_real *) PTR(vO, oc_real *) PTR(wO, oc_real *)
  PTR(validU, uint8_t *) PTR(validV, uint8_t *) PTR(validW, uint8_t *)
  PTR(q, oc_real *) PTR(rhs, oc_real *) PTR(diag, oc_real *) PTR(cellPhi, oc_real *)
  PTR(dens, oc_real *)
  PTR(px, oc_real *) PTR(py, oc_real *) PTR(pz, oc_real *)
  PTR(pvx, oc_real *) PTR(pvy, oc_real *) PTR(pvz, oc_real *)
  PTR(pT, oc_real *) PTR(pAir, oc_real *)
  PTR(pWx, oc_real *) PTR(pWy, oc_real *) PTR(pWz, oc_real *)
  PTR(pflag, uint8_t *) PTR(pcool, uint8_t *)
  PTR(advU, double *) PTR(advV, double *) PTR(advW, double *)
  PTR(pLight, oc_real *) PTR(pPh, oc_real *) PTR(pDepth, uint8_t *)
  PTR(svx, oc_real *) PTR(svy, oc_real *) PTR(svz, oc_real *)
  return -1;
}
#undef PTR

#define FLD(nm) if (!strcmp(name, #nm)) { c->nm = v; return 0; }
int oc_set_f(oc_ctx *c, const char *name, double v) {
  FLD(cfl) FLD(pic) FLD(gravity) FLD(viscosity) FLD(omega) FLD(maxSpeed) FLD(vorticity)
  FLD(coreR) FLD(oceanR) FLD(atmosphereH) FLD(spacing) FLD(particleVolume)
  FLD(gRamp) FLD(dtLast) FLD(rainT) FLD(cloudT) FLD(rainLiftAge) FLD(dv)
  FLD(sunX) FLD(sunY) FLD(sunZ) FLD(umax) FLD(pressureTol)
  FLD(heatK) FLD(Tamb) FLD(shadeCool) FLD(sunActivity) FLD(evapIntensity) FLD(snowT) FLD(meltT) FLD(cloudP)
  FLD(currents) FLD(Rlo) FLD(Rhi) FLD(simTime) FLD(cloudT) FLD(rainT) FLD(rainLiftAge)
  return -1;
}
int oc_set_i(oc_ctx *c, const char *name, long long v) {
  FLD(pressureIters) FLD(nP) FLD(tn) FLD(ceilReflect) FLD(hasSun)
  FLD(cfS) FLD(bfS) FLD(evS) FLD(sorRedBlack) FLD(residualEvery) FLD(maxPressureIters) FLD(spinOn)
  FLD(hadSolidVelocity)
  return -1;
}
#undef FLD
#define GET(nm) if (!strcmp(name, #nm)) return c->nm;
double oc_get_f(oc_ctx *c, const char *name) {
  GET(umax) GET(gRamp) GET(dtLast) GET(domainR) GET(cx) GET(cy) GET(cz) GET(W)
  GET(oceanR) GET(spacing) GET(particleVolume) GET(residual)
  return NAN;
}
long long oc_get_i(oc_ctx *c, const char *name) {
  GET(airborne) GET(nFluid) GET(nP) GET(cfS) GET(bfS) GET(evS) GET(itersDone) GET(evapCount) GET(vaporCount) GET(nightVaporCount)
  GET(nBalls) GET(hadSolidVelocity)
  return -1;
}
#undef GET

/* ------------------------------------------------------------ samplers
 * Trilinear face sampling, solver.js:750-783 (same clamps as _faceCoords
 * :737-748 and the fused _sampleVel3 :820-854). Offsets per component:
 * u (0,.5,.5), v (.5,0,.5), w (.5,.5,0). The i clamp is nx-1 for every
 * component and j,k clamp at n-2, exactly as the reference writes them. */
typedef struct { long idx, sj, sk; double s0, s1, t0, t1, r0, r1; } oc_stencil;

static void face_stencil(const oc_ctx *c, int comp, double x, double y, double z, oc_stencil *st) {
  double dx = c->dx, a, b, cc, tx, ty, tz;
  int i0, j0, k0, nx = c->nx, ny = c->ny, nz = c->nz;
  if (comp == 0) { a = x / dx; b = y / dx - 0.5; cc = z / dx - 0.5; }
  else if (comp == 1) { a = x / dx - 0.5; b = y / dx; cc = z / dx - 0.5; }
  else { a = x / dx - 0.5; b = y / dx - 0.5; cc = z / dx; }
  i0 = (int)floor(a); j0 = (int)floor(b); k0 = (int)floor(cc);
  if (i0 < 0) { i0 = 0; tx = 0; } else if (i0 > nx - 1) { i0 = nx - 1; tx = 1; } else tx = a - i0;
  if (j0 < 0) { j0 = 0; ty = 0; } else if (j0 > ny - 2) { j0 = ny - 2; ty = 1; } else ty = b - j0;
  if (k0 < 0) { k0 = 0; tz = 0; } else if (k0 > nz - 2) { k0 = nz - 2; tz = 1; } else tz = cc - k0;
  if (comp == 0) { st->sj = nx + 1; st->sk = (long)ny * (nx + 1); st->idx = ((long)k0 * ny + j0) * (nx + 1) + i0; }
  else if (comp == 1) { st->sj = nx; st->sk = (long)(ny + 1) * nx; st->idx = ((long)k0 * (ny + 1) + j0) * nx + i0; }
  else { st->sj = nx; st->sk = (long)ny * nx; st->idx = ((long)k0 * ny + j0) * nx + i0; }
  st->s0 = 1 - tx; st->s1 = tx; st->t0 = 1 - ty; st->t1 = ty; st->r0 = 1 - tz; st->r1 = tz;
}

static double stencil_eval(const oc_stencil *st, const oc_real *A) {
  long i = st->idx, sj = st->sj, sk = st->sk;
  double v00 = A[i] * st->s0 + A[i + 1] * st->s1;
  double v10 = A[i + sj] * st->s0 + A[i + sj + 1] * st->s1;
  double v01 = A[i + sk] * st->s0 + A[i + sk + 1] * st->s1;
  double v11 = A[i + sk + sj] * st->s0 + A[i + sk + sj + 1] * st->s1;
  return (v00 * st->t0 + v10 * st->t1) * st->r0 + (v01 * st->t0 + v11 * st->t1) * st->r1;
}

double oc_sample_face(oc_ctx *c, const oc_real *arr, int comp, double x, double y, double z) {
  oc_stencil st;
  face_stencil(c, comp, x, y, z, &st);
  return stencil_eval(&st, arr);
}

void oc_sample_vel3(oc_ctx *c, double x, double y, double z, double *out) {
  out[0] = oc_sample_face(c, c->u, 0, x, y, z);
  out[1] = oc_sample_face(c, c->v, 1, x, y, z);
  out[2] = oc_sample_face(c, c->w, 2, x, y, z);
}

/* ----------------------------------------------------- terrain predicates */
/* terrainRadiusAt, solver.js:512-535: trilinear on the voxel radius lattice */
double oc_terrain_radius_at(oc_ctx *c, double x, double y, double z) {
  if (!c->tR) return c->coreR;
  int n = c->tn;
  double dv = c->dv;
  double gx = x / dv - 0.5, gy = y / dv - 0.5, gz = z / dv - 0.5;
  int i = (int)gx, j = (int)gy, k = (int)gz;
  i = clampi(i, 0, n - 2); j = clampi(j, 0, n - 2); k = clampi(k, 0, n - 2);
  double fx = clampd(gx - i, 0, 1), fy = clampd(gy - j, 0, 1), fz = clampd(gz - k, 0, 1);
  long b = ((long)k * n + j) * n + i, s7 = (long)n * n;
  const float *R = c->tR;
  double v000 = R[b], v100 = R[b + 1], v010 = R[b + n], v110 = R[b + n + 1];
  double v001 = R[b + s7], v101 = R[b + s7 + 1], v011 = R[b + s7 + n], v111 = R[b + s7 + n + 1];
  double a0 = v000 + (v100 - v000) * fx, a1 = v010 + (v110 - v010) * fx;
  double a2 = v001 + (v101 - v001) * fx, a3 = v011 + (v111 - v011) * fx;
  double b0 = a0 + (a1 - a0) * fy, b1 = a2 + (a3 - a2) * fy;
  return b0 + (b1 - b0) * fz;
}

/* _rockCellAt, solver.js:542-559. The cached static raster answers for
 * every interior cell exactly as the voxel-at-centre rule does. */
int oc_rock_cell_at(oc_ctx *c, double x, double y, double z) {
  if (!c->tsolid) return 0;
  double dx = c->dx;
  int i = clampi((int)(x / dx), 1, c->nx - 2);
  int j = clampi((int)(y / dx), 1, c->ny - 2);
  int k = clampi((int)(z / dx), 1, c->nz - 2);
  if (c->staticTypes) return c->staticTypes[((long)k * c->ny + j) * c->nx + i] == SOLID;
  int n = c->tn;
  int vx = clampi((int)((i + 0.5) * dx / c->dv), 0, n - 1);
  int vy = clampi((int)((j + 0.5) * dx / c->dv), 0, n - 1);
  int vz = clampi((int)((k + 0.5) * dx / c->dv), 0, n - 1);
  return c->tsolid[((long)vz * n + vy) * n + vx] == 1;
}

/* _voxelRockAt, solver.js:565-574 */
int oc_voxel_rock_at(oc_ctx *c, double x, double y, double z) {
  if (!c->tsolid) return 0;
  int n = c->tn;
  int i = clampi((int)(x / c->dv), 0, n - 1);
  int j = clampi((int)(y / c->dv), 0, n - 1);
  int k = clampi((int)(z / c->dv), 0, n - 1);
  return c->tsolid[((long)k * n + j) * n + i] == 1;
}

/* _onRockBelow, solver.js:581-587 */
static int on_rock_below(oc_ctx *c, double x, double y, double z) {
  if (!c->tsolid) return 0;
  double ex = x - c->cx, ey = y - c->cy, ez = z - c->cz;
  double el = sqrt(ex * ex + ey * ey + ez * ez);
  if (!(el != 0)) el = 1e-9;
  double s = c->dx * 0.9 / el;
  return oc_rock_cell_at(c, x - ex * s, y - ey * s, z - ez * s);
}

/* ------------------------------------------------------------ rasterize */
static void ball_raster(oc_ctx *c);

/* _rasterizeStatic, solver.js:865-939 (sphere mode): the static raster is
 * built once per world and copied into cellType every substep; balls are
 * then rasterized as moving solids */
void oc_rasterize_static(oc_ctx *c) {
  int nx = c->nx, ny = c->ny, nz = c->nz;
  if (c->staticTypes) { memcpy(c->cellType, c->staticTypes, c->nC); ball_raster(c); return; }
  double dx = c->dx;
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++) {
        int shell = (j == 0 || j == ny - 1 || k == 0 || k == nz - 1 || i == 0 || i == nx - 1);
        int rock = 0;
        if (c->tsolid) {
          int n = c->tn;
          int vx = clampi((int)((i + 0.5) * dx / c->dv), 0, n - 1);
          int vy = clampi((int)((j + 0.5) * dx / c->dv), 0, n - 1);
          int vz = clampi((int)((k + 0.5) * dx / c->dv), 0, n - 1);
          rock = c->tsolid[((long)vz * n + vy) * n + vx] == 1;
        } else {
          double ddx = (i + 0.5) * dx - c->cx, ddy = (j + 0.5) * dx - c->cy, ddz = (k + 0.5) * dx - c->cz;
          rock = ddx * ddx + ddy * ddy + ddz * ddz < c->coreR * c->coreR;
        }
        c->cellType[((long)k * ny + j) * nx + i] = (shell || rock) ? SOLID : AIR;
      }
  ball_raster(c);
}

/* moving solids, solver.js:918-938: the solid velocities are cleared while
 * balls exist (and once after the last one is gone); each ball turns the
 * cells whose centres lie within r + 0.4 dx SOLID, a later ball overwriting
 * an earlier one's velocity */
static void ball_raster(oc_ctx *c) {
  int nx = c->nx, ny = c->ny, nz = c->nz;
  if (c->nBalls || c->hadSolidVelocity) {
    if (c->svx)
      for (long i = 0; i < c->nC; i++) c->svx[i] = c->svy[i] = c->svz[i] = 0;
  }
  c->hadSolidVelocity = c->nBalls > 0;
  double dx = c->dx;
  for (int n = 0; n < c->nBalls; n++) {
    const oc_ball *b = &c->balls[n];
    double r = b->r, rr = (r + 0.4 * dx) * (r + 0.4 * dx);
    int i0 = clampi((int)floor((b->x - r) / dx), 1, nx - 2), i1 = clampi((int)ceil((b->x + r) / dx), 1, nx - 2);
    int j0 = clampi((int)floor((b->y - r) / dx), 1, ny - 2), j1 = clampi((int)ceil((b->y + r) / dx), 1, ny - 2);
    int k0 = clampi((int)floor((b->z - r) / dx), 1, nz - 2), k1 = clampi((int)ceil((b->z + r) / dx), 1, nz - 2);
    for (int k = k0; k <= k1; k++)
      for (int j = j0; j <= j1; j++)
        for (int i = i0; i <= i1; i++) {
          double ex = (i + 0.5) * dx - b->x, ey = (j + 0.5) * dx - b->y, ez = (k + 0.5) * dx - b->z;
          if (ex * ex + ey * ey + ez * ez <= rr) {
            long ci = ((long)k * ny + j) * nx + i;
            c->cellType[ci] = SOLID;
            c->svx[ci] = (oc_real)b->vx; c->svy[ci] = (oc_real)b->vy; c->svz[ci] = (oc_real)b->vz;
          }
        }
  }
}

/* _pushOutOfBalls / _pushBallsChunk, solver.js:1005-1026: every particle
 * (any flag) inside r + 0.002 of a ball moves to that distance along the
 * ball normal and loses its inward relative normal velocity; balls in order */
void oc_push_balls(oc_ctx *c) {
  for (int n = 0; n < c->nBalls; n++) {
    const oc_ball *b = &c->balls[n];
    double rr = b->r + 0.002;
    for (int p = 0; p < c->nP; p++) {
      double ex = c->px[p] - b->x, ey = c->py[p] - b->y, ez = c->pz[p] - b->z;
      double d2 = ex * ex + ey * ey + ez * ez;
      if (d2 < rr * rr) {
        double d = sqrt(d2);
        if (d == 0) d = 1e-9;
        double inv = 1 / d;
        double nxn = ex * inv, nyn = ey * inv, nzn = ez * inv;
        c->px[p] = (oc_real)(b->x + nxn * rr); c->py[p] = (oc_real)(b->y + nyn * rr);
        c->pz[p] = (oc_real)(b->z + nzn * rr);
        double rvx = c->pvx[p] - b->vx, rvy = c->pvy[p] - b->vy, rvz = c->pvz[p] - b->vz;
        double vn = rvx * nxn + rvy * nyn + rvz * nzn;
        if (vn < 0) {
          c->pvx[p] = (oc_real)(c->pvx[p] - nxn * vn);
          c->pvy[p] = (oc_real)(c->pvy[p] - nyn * vn);
          c->pvz[p] = (oc_real)(c->pvz[p] - nzn * vn);
        }
      }
    }
  }
}

/* _markFluidCells, solver.js:945-1000 (sphere mode) */
void oc_mark_fluid(oc_ctx *c, int p0, int p1) {
  double dx = c->dx;
  int nx = c->nx, ny = c->ny, nz = c->nz, jTop = ny - 2;
  double skyAlt = 2 * dx > 0.3 ? 2 * dx : 0.3;
  double lim = (c->oceanR + 0.3) * (c->oceanR + 0.3);
  for (int p = p0; p < p1; p++) {
    if (c->pflag[p] != 0) continue;
    double x = c->px[p], y = c->py[p], z = c->pz[p];
    double rx = x - c->cx, ry = y - c->cy, rz = z - c->cz;
    double r2 = rx * rx + ry * ry + rz * rz;
    if (r2 > lim && sqrt(r2) - oc_terrain_radius_at(c, x, y, z) > skyAlt && !oc_rock_cell_at(c, x, y, z)) {
      c->pflag[p] = 1;
      continue;
    }
    int i = clampi((int)floor(x / dx), 0, nx - 1);
    int j = clampi((int)floor(y / dx), 0, ny - 1);
    int k = clampi((int)floor(z / dx), 0, nz - 1);
    if (j >= jTop) {
      if (sqrt(rx * rx + ry * ry + rz * rz) - oc_terrain_radius_at(c, x, y, z) > 0.3) {
        c->pflag[p] = 1;
        continue;
      }
    }
    long idx = ((long)k * ny + j) * nx + i;
    if (c->cellType[idx] == AIR) c->cellType[idx] = FLUID;
  }
}

/* ------------------------------------------------------------------ P2G */
void oc_clear_faces(oc_ctx *c) {
  /* solver.js:2569-2570 */
  for (long i = 0; i < c->uN; i++) { c->u[i] = 0; c->uW[i] = 0; }
  for (long i = 0; i < c->vN; i++) { c->v[i] = 0; c->vW[i] = 0; }
  for (long i = 0; i < c->wN; i++) { c->w[i] = 0; c->wW[i] = 0; }
}

/* one component of _p2gInto, solver.js:1046-1104: trilinear splat of
 * momentum and weight, accumulated sequentially in particle order */
static void splat_comp(oc_ctx *c, int comp, double x, double y, double z, double vel, oc_real *A, oc_real *Wt) {
  oc_stencil st;
  face_stencil(c, comp, x, y, z, &st);
  double wA = st.s0 * st.t0, wB = st.s1 * st.t0, wC = st.s0 * st.t1, wD = st.s1 * st.t1;
  long i = st.idx, sj = st.sj, sk = st.sk;
  long off[8] = { i, i + 1, i + sj, i + sj + 1, i + sk, i + sk + 1, i + sk + sj, i + sk + sj + 1 };
  double wx[8] = { wA * st.r0, wB * st.r0, wC * st.r0, wD * st.r0, wA * st.r1, wB * st.r1, wC * st.r1, wD * st.r1 };
  for (int e = 0; e < 8; e++) {
    A[off[e]] = (oc_real)(A[off[e]] + wx[e] * vel);
    Wt[off[e]] = (oc_real)(Wt[off[e]] + wx[e]);
  }
}

/* _p2g = _p2gInto + _p2gNormalizeFace, solver.js:1030-1120 */
void oc_p2g(oc_ctx *c) {
  for (int p = 0; p < c->nP; p++) {
    if (c->pflag[p] != 0) continue;
    double x = c->px[p], y = c->py[p], z = c->pz[p];
    splat_comp(c, 0, x, y, z, c->pvx[p], c->u, c->uW);
    splat_comp(c, 1, x, y, z, c->pvy[p], c->v, c->vW);
    splat_comp(c, 2, x, y, z, c->pvz[p], c->w, c->wW);
  }
  for (long i = 0; i < c->uN; i++) if (c->uW[i] > 1e-8) c->u[i] = (oc_real)((double)c->u[i] / c->uW[i]);
  for (long i = 0; i < c->vN; i++) if (c->vW[i] > 1e-8) c->v[i] = (oc_real)((double)c->v[i] / c->vW[i]);
  for (long i = 0; i < c->wN; i++) if (c->wW[i] > 1e-8) c->w[i] = (oc_real)((double)c->w[i] / c->wW[i]);
}

/* ------------------------------------------------- boundary conditions */
/* _applyBCSlice, solver.js:1130-1171: domain planes 0, a face next to a
 * SOLID cell takes that cell's velocity (the left/lower one first; 0 for
 * rock, the ball's velocity for a ball cell) */
void oc_apply_bc(oc_ctx *c) {
  int nx = c->nx, ny = c->ny, nz = c->nz;
#define SV(a, i) (c->a ? c->a[i] : (oc_real)0)
  const uint8_t *t = c->cellType;
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < ny; j++) {
      c->u[((long)k * ny + j) * (nx + 1)] = 0;
      c->u[((long)k * ny + j) * (nx + 1) + nx] = 0;
    }
  for (int k = 0; k < nz; k++)
    for (int i = 0; i < nx; i++) {
      c->v[((long)k * (ny + 1)) * nx + i] = 0;
      c->v[((long)k * (ny + 1) + ny) * nx + i] = 0;
    }
  for (int j = 0; j < ny; j++)
    for (int i = 0; i < nx; i++) {
      c->w[(long)j * nx + i] = 0;
      c->w[((long)nz * ny + j) * nx + i] = 0;
    }
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < ny; j++)
      for (int i = 1; i < nx; i++) {
        long cl = ((long)k * ny + j) * nx + i - 1;
        if (t[cl] == SOLID) c->u[((long)k * ny + j) * (nx + 1) + i] = SV(svx, cl);
        else if (t[cl + 1] == SOLID) c->u[((long)k * ny + j) * (nx + 1) + i] = SV(svx, cl + 1);
      }
  for (int k = 0; k < nz; k++)
    for (int j = 1; j < ny; j++)
      for (int i = 0; i < nx; i++) {
        long cr = ((long)k * ny + j) * nx + i;
        if (t[cr - nx] == SOLID) c->v[((long)k * (ny + 1) + j) * nx + i] = SV(svy, cr - nx);
        else if (t[cr] == SOLID) c->v[((long)k * (ny + 1) + j) * nx + i] = SV(svy, cr);
      }
  for (int k = 1; k < nz; k++)
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++) {
        long cr = ((long)k * ny + j) * nx + i;
        if (t[cr - (long)nx * ny] == SOLID) c->w[cr] = SV(svz, cr - (long)nx * ny);
        else if (t[cr] == SOLID) c->w[cr] = SV(svz, cr);
      }
#undef SV
}

/* FLIP snapshot, solver.js:2579 */
void oc_snapshot(oc_ctx *c) {
  memcpy(c->uO, c->u, sizeof(oc_real) * c->uN);
  memcpy(c->vO, c->v, sizeof(oc_real) * c->vN);
  memcpy(c->wO, c->w, sizeof(oc_real) * c->wN);
}

/* ------------------------------------------------------------ viscosity */
/* _viscosity, solver.js:1259-1304: in place, lexicographic, fluid-fluid
 * faces only, u += coef (6u - sum of the six same-component neighbours),
 * a missing neighbour reads the face itself. */
static void visc_comp(oc_real *A, int d0, int d1, int d2, int comp, const oc_ctx *c, double coef) {
  int nx = c->nx, ny = c->ny;
  const uint8_t *t = c->cellType;
  long s1 = d0, s2 = (long)d0 * d1;
  for (int k = (comp == 2 ? 1 : 0); k < (comp == 2 ? d2 - 1 : d2); k++)
    for (int j = (comp == 1 ? 1 : 0); j < (comp == 1 ? d1 - 1 : d1); j++)
      for (int i = (comp == 0 ? 1 : 0); i < (comp == 0 ? d0 - 1 : d0); i++) {
        /* the face sits between cell cr (its + side) and cl (its - side) */
        long cr = ((long)k * ny + j) * nx + i;
        long cl = comp == 0 ? cr - 1 : comp == 1 ? cr - nx : cr - (long)nx * ny;
        if (t[cl] != FLUID || t[cr] != FLUID) continue;
        long idx = ((long)k * d1 + j) * d0 + i;
        double val = A[idx] * 6.0;
        val -= (i - 1 >= 0) ? A[idx - 1] : A[idx];
        val -= (i + 1 < d0) ? A[idx + 1] : A[idx];
        val -= (j - 1 >= 0) ? A[idx - s1] : A[idx];
        val -= (j + 1 < d1) ? A[idx + s1] : A[idx];
        val -= (k - 1 >= 0) ? A[idx - s2] : A[idx];
        val -= (k + 1 < d2) ? A[idx + s2] : A[idx];
        A[idx] = (oc_real)(A[idx] + coef * val);
      }
}

void oc_viscosity(oc_ctx *c, double dt) {
  double nu = c->viscosity;
  if (nu <= 1e-6) return;
  double coef = nu * dt / (c->dx * c->dx);
  if (coef > 0.24) coef = 0.24;
  int nx = c->nx, ny = c->ny, nz = c->nz;
  visc_comp(c->u, nx + 1, ny, nz, 0, c, coef);
  visc_comp(c->v, nx, ny + 1, nz, 1, c, coef);
  visc_comp(c->w, nx, ny, nz + 1, 2, c, coef);
}

/* -------------------------------------------------------- planet gravity */
/* _applyPlanetGravity (the definition that wins, solver.js:2656-2662) and
 * _applyPlanetGravitySlice, solver.js:1229-1256 */
void oc_apply_gravity(oc_ctx *c, double dt) {
  if (!c->gu) return;
  double gs = c->gRamp / 1.5;
  if (gs > 1) gs = 1;
  c->gRamp += dt;
  if (gs <= 0) return;
  gs *= c->gravity;
  int nx = c->nx, ny = c->ny, nz = c->nz;
  const uint8_t *t = c->cellType;
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < ny; j++)
      for (int i = 1; i < nx; i++) {
        long cr = ((long)k * ny + j) * nx + i;
        if (t[cr - 1] == FLUID || t[cr] == FLUID) {
          long idx = ((long)k * ny + j) * (nx + 1) + i;
          c->u[idx] = (oc_real)(c->u[idx] + c->gu[idx] * dt * gs);
        }
      }
  for (int k = 0; k < nz; k++)
    for (int j = 1; j < ny; j++)
      for (int i = 0; i < nx; i++) {
        long cr = ((long)k * ny + j) * nx + i;
        if (t[cr - nx] == FLUID || t[cr] == FLUID) {
          long idx = ((long)k * (ny + 1) + j) * nx + i;
          c->v[idx] = (oc_real)(c->v[idx] + c->gv[idx] * dt * gs);
        }
      }
  for (int k = 1; k < nz; k++)
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++) {
        long cr = ((long)k * ny + j) * nx + i;
        if (t[cr - (long)nx * ny] == FLUID || t[cr] == FLUID)
          c->w[cr] = (oc_real)(c->w[cr] + c->gw[cr] * dt * gs);
      }
}

/* -------------------------------------------------------- pressure solve */
/* _pressureSolve passes 1-2 and the CSR build, solver.js:1312-1395.
 * Returns the number of fluid cells. */
int oc_pressure_assemble(oc_ctx *c) {
  int nx = c->nx, ny = c->ny, nz = c->nz;
  long nxy = (long)nx * ny;
  const uint8_t *t = c->cellType;
  int nF = 0;
  for (long cc = 0; cc < c->nC; cc++) {
    if (t[cc] != FLUID) { c->q[cc] = 0; continue; }
    c->fluidList[nF++] = (int)cc;
  }
  c->nFluid = nF;
  int na = 0;
  for (int fi = 0; fi < nF; fi++) {
    long cc = c->fluidList[fi];
    int i = (int)(cc % nx), j = (int)((cc / nx) % ny), k = (int)(cc / nxy);
    long nb[6] = { i > 0 ? cc - 1 : -1, i < nx - 1 ? cc + 1 : -1, j > 0 ? cc - nx : -1,
                   j < ny - 1 ? cc + nx : -1, k > 0 ? cc - nxy : -1, k < nz - 1 ? cc + nxy : -1 };
    int Nf = 0, Na = 0;
    c->pStart[fi] = na;
    for (int d = 0; d < 6; d++) {
      if (nb[d] < 0) continue;
      if (t[nb[d]] == FLUID) { c->pAdj[na++] = (int)nb[d]; Nf++; }
      else if (t[nb[d]] == AIR) Na++;
    }
    c->diag[cc] = (oc_real)(Nf + Na);
    long iu = ((long)k * ny + j) * (nx + 1) + i;
    long iv = ((long)k * (ny + 1) + j) * nx + i;
    long iw = ((long)k * ny + j) * nx + i;
    double div = ((double)c->u[iu + 1] - c->u[iu] + c->v[iv + nx] - c->v[iv] + c->w[iw + nxy] - c->w[iw]) / c->dx;
    c->rhs[cc] = (oc_real)(-c->dx * c->dx * div);
    double dg = c->diag[cc];
    c->pW[fi] = (float)(dg > 0 ? c->omega / dg : 0);
  }
  c->pStart[nF] = na;
  return nF;
}

/* SOR Gauss-Seidel sweeps in the reference's lexicographic order,
 * solver.js:1397-1404 */
void oc_sor_lex(oc_ctx *c, int iters) {
  double om1 = 1 - c->omega;
  for (int it = 0; it < iters; it++)
    for (int fi = 0; fi < c->nFluid; fi++) {
      double sum = 0;
      for (int a = c->pStart[fi]; a < c->pStart[fi + 1]; a++) sum += c->q[c->pAdj[a]];
      long cc = c->fluidList[fi];
      c->q[cc] = (oc_real)(c->q[cc] * om1 + (double)c->pW[fi] * (c->rhs[cc] + sum));
    }
}

/* The same relaxation (solver.js:1402) in red-black order: cells with
 * (i+j+k) even first, then odd. One "iteration" relaxes both colours. */
void oc_sor_redblack(oc_ctx *c, int iters) {
  double om1 = 1 - c->omega;
  int nx = c->nx, ny = c->ny;
  long nxy = (long)nx * ny;
  for (int it = 0; it < iters; it++)
    for (int color = 0; color < 2; color++)
      for (int fi = 0; fi < c->nFluid; fi++) {
        long cc = c->fluidList[fi];
        int par = (int)((cc % nx) + ((cc / nx) % ny) + (cc / nxy)) & 1;
        if (par != color) continue;
        double sum = 0;
        for (int a = c->pStart[fi]; a < c->pStart[fi + 1]; a++) sum += c->q[c->pAdj[a]];
        c->q[cc] = (oc_real)(c->q[cc] * om1 + (double)c->pW[fi] * (c->rhs[cc] + sum));
      }
}

/* Relative residual ||rhs - A q||_2 / ||rhs||_2 of the assembled system
 * (A q)_c = diag_c q_c - sum of fluid neighbours, over fluid cells, in FP64. */
double oc_residual(oc_ctx *c, double *rhs_norm) {
  double rr = 0, bb = 0;
  for (int fi = 0; fi < c->nFluid; fi++) {
    long cc = c->fluidList[fi];
    double sum = 0;
    for (int a = c->pStart[fi]; a < c->pStart[fi + 1]; a++) sum += c->q[c->pAdj[a]];
    double r = (double)c->rhs[cc] - ((double)c->diag[cc] * c->q[cc] - sum);
    rr += r * r;
    bb += (double)c->rhs[cc] * c->rhs[cc];
  }
  if (rhs_norm) *rhs_norm = sqrt(bb);
  return bb > 0 ? sqrt(rr) / sqrt(bb) : 0;
}

void oc_pressure_finish(oc_ctx *c) {
  memcpy(c->cellPhi, c->q, sizeof(oc_real) * c->nC); /* solver.js:1409 */
}

void oc_pressure_solve(oc_ctx *c) {
  int nF = oc_pressure_assemble(c);
  if (nF > 0) {
    if (c->pressureTol > 0) {
      int done = 0, every = c->residualEvery > 0 ? c->residualEvery : 8;
      int mx = c->maxPressureIters > 0 ? c->maxPressureIters : 4000;
      double res = oc_residual(c, 0);
      while (res > c->pressureTol && done < mx) {
        int k = done + every > mx ? mx - done : every;
        if (c->sorRedBlack) oc_sor_redblack(c, k); else oc_sor_lex(c, k);
        done += k;
        res = oc_residual(c, 0);
      }
      c->itersDone = done;
      c->residual = res;
    } else {
      if (c->sorRedBlack) oc_sor_redblack(c, c->pressureIters); else oc_sor_lex(c, c->pressureIters);
      c->itersDone = c->pressureIters;
    }
  }
  oc_pressure_finish(c);
}

/* ------------------------------------------------------------- project */
/* _projectSlice, solver.js:1417-1443 */
void oc_project(oc_ctx *c) {
  int nx = c->nx, ny = c->ny, nz = c->nz;
  double inv = 1 / c->dx;
  const uint8_t *t = c->cellType;
  const oc_real *phi = c->cellPhi;
  long nxy = (long)nx * ny;
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < ny; j++)
      for (int i = 1; i < nx; i++) {
        long cr = ((long)k * ny + j) * nx + i, cl = cr - 1;
        if (t[cl] == SOLID || t[cr] == SOLID) continue;
        double pL = t[cl] == FLUID ? phi[cl] : 0, pR = t[cr] == FLUID ? phi[cr] : 0;
        long idx = ((long)k * ny + j) * (nx + 1) + i;
        c->u[idx] = (oc_real)(c->u[idx] - (pR - pL) * inv);
      }
  for (int k = 0; k < nz; k++)
    for (int j = 1; j < ny; j++)
      for (int i = 0; i < nx; i++) {
        long cr = ((long)k * ny + j) * nx + i, cl = cr - nx;
        if (t[cl] == SOLID || t[cr] == SOLID) continue;
        double pL = t[cl] == FLUID ? phi[cl] : 0, pR = t[cr] == FLUID ? phi[cr] : 0;
        long idx = ((long)k * (ny + 1) + j) * nx + i;
        c->v[idx] = (oc_real)(c->v[idx] - (pR - pL) * inv);
      }
  for (int k = 1; k < nz; k++)
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++) {
        long cr = ((long)k * ny + j) * nx + i, cl = cr - nxy;
        if (t[cl] == SOLID || t[cr] == SOLID) continue;
        double pL = t[cl] == FLUID ? phi[cl] : 0, pR = t[cr] == FLUID ? phi[cr] : 0;
        c->w[cr] = (oc_real)(c->w[cr] - (pR - pL) * inv);
      }
}

/* ------------------------------------------------- vorticity confinement */
/* _vorticityConfinement, solver.js:1454-1559. Scratch is float, as the
 * reference's Float32Array scratch is. */
void oc_vorticity(oc_ctx *c, double dt) {
  double eps = c->vorticity;
  if (!(eps > 0)) return;
  int nx = c->nx, ny = c->ny, nz = c->nz;
  long nxy = (long)nx * ny;
  const uint8_t *t = c->cellType;
  float *vx = c->vcx, *vy = c->vcy, *vz = c->vcz, *om = c->omg, *ox = c->omx, *oy = c->omy, *oz = c->omz;
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++) {
        long cc = ((long)k * ny + j) * nx + i;
        long ru = ((long)k * ny + j) * (nx + 1) + i, rv = ((long)k * (ny + 1) + j) * nx + i;
        vx[cc] = (float)(0.5 * ((double)c->u[ru] + c->u[ru + 1]));
        vy[cc] = (float)(0.5 * ((double)c->v[rv] + c->v[rv + nx]));
        vz[cc] = (float)(0.5 * ((double)c->w[cc] + c->w[cc + nxy]));
      }
  double inv2 = 0.5 / c->dx;
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++) {
        long cc = ((long)k * ny + j) * nx + i;
        long im = i > 0 ? cc - 1 : cc, ip = i < nx - 1 ? cc + 1 : cc;
        long jm = j > 0 ? cc - nx : cc, jp = j < ny - 1 ? cc + nx : cc;
        long km = k > 0 ? cc - nxy : cc, kp = k < nz - 1 ? cc + nxy : cc;
        double wx = ((double)vz[jp] - vz[jm] - ((double)vy[kp] - vy[km])) * inv2;
        double wy = ((double)vx[kp] - vx[km] - ((double)vz[ip] - vz[im])) * inv2;
        double wz = ((double)vy[ip] - vy[im] - ((double)vx[jp] - vx[jm])) * inv2;
        ox[cc] = (float)wx; oy[cc] = (float)wy; oz[cc] = (float)wz;
        om[cc] = (float)sqrt(wx * wx + wy * wy + wz * wz);
      }
  const double fCap = 6.0, omFloor = 0.35;
  for (int fi = 0; fi < c->nFluid; fi++) {
    long cc = c->fluidList[fi];
    int i = (int)(cc % nx), j = (int)((cc / nx) % ny), k = (int)(cc / nxy);
    if (om[cc] < omFloor || (i > 0 && t[cc - 1] != FLUID) || (i < nx - 1 && t[cc + 1] != FLUID) ||
        (j > 0 && t[cc - nx] != FLUID) || (j < ny - 1 && t[cc + nx] != FLUID) ||
        (k > 0 && t[cc - nxy] != FLUID) || (k < nz - 1 && t[cc + nxy] != FLUID)) {
      vx[cc] = 0; vy[cc] = 0; vz[cc] = 0;
      continue;
    }
    double gx = (double)om[i < nx - 1 ? cc + 1 : cc] - om[i > 0 ? cc - 1 : cc];
    double gy = (double)om[j < ny - 1 ? cc + nx : cc] - om[j > 0 ? cc - nx : cc];
    double gz = (double)om[k < nz - 1 ? cc + nxy : cc] - om[k > 0 ? cc - nxy : cc];
    double g2 = gx * gx + gy * gy + gz * gz;
    if (g2 < 1e-12) { vx[cc] = 0; vy[cc] = 0; vz[cc] = 0; continue; }
    double inv = 1 / sqrt(g2);
    double Nx = gx * inv, Ny = gy * inv, Nz = gz * inv;
    double fx = eps * (Ny * oz[cc] - Nz * oy[cc]);
    double fy = eps * (Nz * ox[cc] - Nx * oz[cc]);
    double fz = eps * (Nx * oy[cc] - Ny * ox[cc]);
    double f2 = fx * fx + fy * fy + fz * fz;
    if (f2 > fCap * fCap) { double fs = fCap / sqrt(f2); fx *= fs; fy *= fs; fz *= fs; }
    vx[cc] = (float)(fx * dt); vy[cc] = (float)(fy * dt); vz[cc] = (float)(fz * dt);
  }
  for (int fi = 0; fi < c->nFluid; fi++) {
    long cc = c->fluidList[fi];
    int i = (int)(cc % nx), j = (int)((cc / nx) % ny), k = (int)(cc / nxy);
    if (i < nx - 1 && t[cc + 1] == FLUID) {
      long f = ((long)k * ny + j) * (nx + 1) + i + 1;
      c->u[f] = (oc_real)(c->u[f] + 0.5 * ((double)vx[cc] + vx[cc + 1]));
    }
    if (j < ny - 1 && t[cc + nx] == FLUID) {
      long f = ((long)k * (ny + 1) + j + 1) * nx + i;
      c->v[f] = (oc_real)(c->v[f] + 0.5 * ((double)vy[cc] + vy[cc + nx]));
    }
    if (k < nz - 1 && t[cc + nxy] == FLUID) {
      long f = ((long)(k + 1) * ny + j) * nx + i;
      c->w[f] = (oc_real)(c->w[f] + 0.5 * ((double)vz[cc] + vz[cc + nxy]));
    }
  }
}

/* ---------------------------------------------------------- grid clamp */
/* _clampUmSlice, solver.js:2629-2653 */
static double clamp_arr(oc_real *A, long n, double vcap, double um) {
  for (long i = 0; i < n; i++) {
    if (A[i] > vcap) A[i] = (oc_real)vcap; else if (A[i] < -vcap) A[i] = (oc_real)(-vcap);
    double av = A[i];
    if (av > um) um = av; else if (-av > um) um = -av;
  }
  return um;
}

double oc_clamp_umax(oc_ctx *c) {
  double vcap = c->maxSpeed * 1.1, um = 0;
  um = clamp_arr(c->u, c->uN, vcap, um);
  um = clamp_arr(c->v, c->vN, vcap, um);
  um = clamp_arr(c->w, c->wN, vcap, um);
  c->umax = um;
  return um;
}

/* -------------------------------------------------------- extrapolation */
/* _extrapolate, solver.js:1562-1636: valid = weighted or solid-adjacent or
 * a domain plane; one in-place lexicographic pass where faces filled earlier
 * in the pass (mark 2) already count as valid sources. */
static void extrap_comp(oc_ctx *c, int comp, oc_real *A, const oc_real *Wt, uint8_t *valid, int d0, int d1, int d2) {
  int nx = c->nx, ny = c->ny, nz = c->nz;
  long nxy = (long)nx * ny;
  const uint8_t *t = c->cellType;
  for (int k = 0; k < d2; k++)
    for (int j = 0; j < d1; j++)
      for (int i = 0; i < d0; i++) {
        long idx = ((long)k * d1 + j) * d0 + i;
        long cr = ((long)k * ny + j) * nx + i;
        int sa;
        if (comp == 0) sa = (i == 0 || i == nx) ? 1 : (t[cr - 1] == SOLID || t[cr] == SOLID);
        else if (comp == 1) sa = (j == 0 || j == ny) ? 1 : (t[cr - nx] == SOLID || t[cr] == SOLID);
        else sa = (k == 0 || k == nz) ? 1 : (t[cr - nxy] == SOLID || t[cr] == SOLID);
        valid[idx] = (Wt[idx] > 1e-8 || sa) ? 1 : 0;
      }
  long sK = (long)d0 * d1;
  for (int k = 0; k < d2; k++)
    for (int j = 0; j < d1; j++)
      for (int i = 0; i < d0; i++) {
        long idx = ((long)k * d1 + j) * d0 + i;
        if (valid[idx]) continue;
        double sum = 0;
        int cnt = 0;
        if (i > 0 && valid[idx - 1]) { sum += A[idx - 1]; cnt++; }
        if (i < d0 - 1 && valid[idx + 1]) { sum += A[idx + 1]; cnt++; }
        if (j > 0 && valid[idx - d0]) { sum += A[idx - d0]; cnt++; }
        if (j < d1 - 1 && valid[idx + d0]) { sum += A[idx + d0]; cnt++; }
        if (k > 0 && valid[idx - sK]) { sum += A[idx - sK]; cnt++; }
        if (k < d2 - 1 && valid[idx + sK]) { sum += A[idx + sK]; cnt++; }
        if (cnt > 0) { A[idx] = (oc_real)(sum / cnt); valid[idx] = 2; }
      }
  for (long i = 0; i < sK * d2; i++) if (valid[i] == 2) valid[i] = 1;
}

void oc_extrapolate(oc_ctx *c) {
  int nx = c->nx, ny = c->ny, nz = c->nz;
  extrap_comp(c, 0, c->u, c->uW, c->validU, nx + 1, ny, nz);
  extrap_comp(c, 1, c->v, c->vW, c->validV, nx, ny + 1, nz);
  extrap_comp(c, 2, c->w, c->wW, c->validW, nx, ny, nz + 1);
}

/* ------------------------------------------------------------------ G2P */
/* _g2pChunk, solver.js:1646-1756. Returns the airborne count. */
int oc_g2p(oc_ctx *c, double dt) {
  int nx = c->nx, ny = c->ny, nz = c->nz, air = 0;
  double dx = c->dx, pic = c->pic;
  const uint8_t *t = c->cellType;
  for (int p = 0; p < c->nP; p++) {
    if (c->pflag[p] != 0) { air++; continue; }
    if (c->pcool[p] > 0) {
      double tmp[3];
      c->pcool[p]--;
      double dtt = dt != 0 ? dt : (c->dtLast != 0 ? c->dtLast : 0.01);
      c->pvy[p] = (oc_real)(c->pvy[p] - c->gravity * dtt);
      oc_sample_vel3(c, c->px[p], c->py[p], c->pz[p], tmp);
      c->advU[p] = tmp[0]; c->advV[p] = tmp[1]; c->advW[p] = tmp[2];
      continue;
    }
    double x = c->px[p], y = c->py[p], z = c->pz[p];
    oc_stencil su, sv, sw;
    face_stencil(c, 0, x, y, z, &su);
    face_stencil(c, 1, x, y, z, &sv);
    face_stencil(c, 2, x, y, z, &sw);
    double ou = stencil_eval(&su, c->uO), nu = stencil_eval(&su, c->u);
    double ov = stencil_eval(&sv, c->vO), nv = stencil_eval(&sv, c->v);
    double ow = stencil_eval(&sw, c->wO), nw = stencil_eval(&sw, c->w);
    c->advU[p] = nu; c->advV[p] = nv; c->advW[p] = nw;
    c->pvx[p] = (oc_real)(((double)c->pvx[p] + nu - ou) * (1 - pic) + nu * pic);
    c->pvy[p] = (oc_real)(((double)c->pvy[p] + nv - ov) * (1 - pic) + nv * pic);
    c->pvz[p] = (oc_real)(((double)c->pvz[p] + nw - ow) * (1 - pic) + nw * pic);
    double pv2 = (double)c->pvx[p] * c->pvx[p] + (double)c->pvy[p] * c->pvy[p] + (double)c->pvz[p] * c->pvz[p];
    if (pv2 > c->maxSpeed * c->maxSpeed) {
      double psc = c->maxSpeed / sqrt(pv2);
      c->pvx[p] = (oc_real)(c->pvx[p] * psc);
      c->pvy[p] = (oc_real)(c->pvy[p] * psc);
      c->pvz[p] = (oc_real)(c->pvz[p] * psc);
    }
    int ci = clampi((int)floor(x / dx), 0, nx - 1);
    int cj = clampi((int)floor(y / dx), 0, ny - 1);
    int ck = clampi((int)floor(z / dx), 0, nz - 1);
    long cc = ((long)ck * ny + cj) * nx + ci;
    if (t[cc] == AIR) {
      long nxy = (long)nx * ny;
      int fluidNbr = (ci > 0 && t[cc - 1] == FLUID) || (ci < nx - 1 && t[cc + 1] == FLUID) ||
                     (cj > 0 && t[cc - nx] == FLUID) || (cj < ny - 1 && t[cc + nx] == FLUID) ||
                     (ck > 0 && t[cc - nxy] == FLUID) || (ck < nz - 1 && t[cc + nxy] == FLUID);
      if (!fluidNbr && !on_rock_below(c, x, y, z)) { c->pflag[p] = 1; air++; }
    }
  }
  c->airborne = air;
  return air;
}

/* ------------------------------------------------------------ advection */
static uint32_t lcg(uint32_t s) { return s * 1664525u + 1013904223u; }

static double rlen(double x, double y, double z) {
  double r = sqrt(x * x + y * y + z * z);
  return (r != 0) ? r : 1e-9; /* "|| 1e-9" in the reference */
}

/* _advectChunk, solver.js:1769-2292, sphere mode. Positions are carried in
 * double within a particle's update and stored to OC_REAL at the end. */
void oc_advect(oc_ctx *c, double dt) {
  int nx = c->nx, ny = c->ny, nz = c->nz;
  double dx = c->dx, g = c->gravity;
  const uint8_t *t = c->cellType;
  double xLo = dx * 1.02, xHi = c->W - dx * 1.02, yLo = dx * 1.02, yHi = c->H - dx * 1.02;
  double zLo = dx * 1.02, zHi = c->D - dx * 1.02;
  double cx = c->cx, cy = c->cy, cz = c->cz;
  double domHi = c->domainR - dx * 0.02;
  double coreLo = c->coreR + dx * 1.02;
  int hasTerr = c->tsolid != 0;
  double wSX = 1, wSY = 0, wSZ = 0;
  if (c->hasSun) {
    wSX = c->sunX - cx; wSY = c->sunY - cy; wSZ = c->sunZ - cz;
    double wl = sqrt(wSX * wSX + wSY * wSY + wSZ * wSZ);
    if (!(wl != 0)) wl = 1;
    wSX /= wl; wSY /= wl; wSZ /= wl;
  }
  double rainTe = c->rainT < c->cloudT - 0.02 ? c->rainT : c->cloudT - 0.02;
  if (rainTe < 0) rainTe = 0;
  double ms2 = c->maxSpeed * c->maxSpeed;

  for (int p = 0; p < c->nP; p++) {
    double x = c->px[p], y = c->py[p], z = c->pz[p];
    double vx = c->pvx[p], vy = c->pvy[p], vz = c->pvz[p];
    int fl = c->pflag[p];

    if (fl == 5) { /* ice, solver.js:1797-1864 */
      if (c->pAir[p] == 1 || !hasTerr) { c->pvx[p] = 0; c->pvy[p] = 0; c->pvz[p] = 0; continue; }
      double dragK = 12, gI = c->gravity != 0 ? c->gravity : 9.81;
      double exI = x - cx, eyI = y - cy, ezI = z - cz;
      double elI = rlen(exI, eyI, ezI);
      double axI = -gI * 0.6 * exI / elI, ayI = -gI * 0.6 * eyI / elI, azI = -gI * 0.6 * ezI / elI;
      c->pvx[p] = (oc_real)(c->pvx[p] + (axI - dragK * c->pvx[p]) * dt);
      c->pvy[p] = (oc_real)(c->pvy[p] + (ayI - dragK * c->pvy[p]) * dt);
      c->pvz[p] = (oc_real)(c->pvz[p] + (azI - dragK * c->pvz[p]) * dt);
      x += c->pvx[p] * dt; y += c->pvy[p] * dt; z += c->pvz[p] * dt;
      double rrI = sqrt((x - cx) * (x - cx) + (y - cy) * (y - cy) + (z - cz) * (z - cz));
      if (rrI > c->oceanR + c->atmosphereH) {
        double scI = (c->oceanR + c->atmosphereH) / rrI;
        x = cx + (x - cx) * scI; y = cy + (y - cy) * scI; z = cz + (z - cz) * scI;
        c->pvx[p] = (oc_real)(c->pvx[p] * 0.5); c->pvy[p] = (oc_real)(c->pvy[p] * 0.5); c->pvz[p] = (oc_real)(c->pvz[p] * 0.5);
      }
      if (oc_rock_cell_at(c, x, y, z)) {
        c->pAir[p] = 1; c->pvx[p] = 0; c->pvy[p] = 0; c->pvz[p] = 0;
        c->px[p] = (oc_real)x; c->py[p] = (oc_real)y; c->pz[p] = (oc_real)z;
        continue;
      }
      int ciI = (int)floor(x / dx), cjI = (int)floor(y / dx), ckI = (int)floor(z / dx);
      if (ciI >= 0 && cjI >= 0 && ckI >= 0 && ciI < nx && cjI < ny && ckI < nz &&
          t[((long)ckI * ny + cjI) * nx + ciI] == FLUID) {
        double exF = x - cx, eyF = y - cy, ezF = z - cz;
        double elF = rlen(exF, eyF, ezF);
        double stepF = dx * 0.5;
        for (int mf = 0; mf < 12; mf++) {
          double qx = x + exF / elF * stepF, qy = y + eyF / elF * stepF, qz = z + ezF / elF * stepF;
          int qi = (int)floor(qx / dx), qj = (int)floor(qy / dx), qk = (int)floor(qz / dx);
          if (qi < 0 || qj < 0 || qk < 0 || qi >= nx || qj >= ny || qk >= nz) break;
          if (t[((long)qk * ny + qj) * nx + qi] != FLUID) break;
          x = qx; y = qy; z = qz;
        }
        c->pAir[p] = 1; c->pvx[p] = 0; c->pvy[p] = 0; c->pvz[p] = 0;
        if (c->pWx) { c->pWx[p] = 0; c->pWy[p] = 0; c->pWz[p] = 0; }
      }
      c->px[p] = (oc_real)x; c->py[p] = (oc_real)y; c->pz[p] = (oc_real)z;
      continue;
    }

    if (fl == 1 || fl == 4) { /* ballistic spray / rain, solver.js:1865-1976 */
      double gx = x - cx, gy = y - cy, gz = z - cz;
      double gr2 = gx * gx + gy * gy + gz * gz;
      double ga = g * c->oceanR * c->oceanR * dt / (gr2 * sqrt(gr2));
      c->pvx[p] = (oc_real)(vx - gx * ga); c->pvy[p] = (oc_real)(vy - gy * ga); c->pvz[p] = (oc_real)(vz - gz * ga);
      double dr = 1 - 2.5 * dt;
      if (dr < 0) dr = 0;
      c->pvx[p] = (oc_real)(c->pvx[p] * dr); c->pvy[p] = (oc_real)(c->pvy[p] * dr); c->pvz[p] = (oc_real)(c->pvz[p] * dr);
      double dsp = sqrt((double)c->pvx[p] * c->pvx[p] + (double)c->pvy[p] * c->pvy[p] + (double)c->pvz[p] * c->pvz[p]);
      if (dsp > c->maxSpeed) {
        double dsc = c->maxSpeed / dsp;
        c->pvx[p] = (oc_real)(c->pvx[p] * dsc); c->pvy[p] = (oc_real)(c->pvy[p] * dsc); c->pvz[p] = (oc_real)(c->pvz[p] * dsc);
      }
      x += c->pvx[p] * dt; y += c->pvy[p] * dt; z += c->pvz[p] * dt;
      double bx = x - cx, by = y - cy, bz = z - cz;
      double br = rlen(bx, by, bz);
      double skyHi = c->oceanR + c->atmosphereH;
      if (br > skyHi) {
        double sc0 = skyHi / br;
        x = cx + bx * sc0; y = cy + by * sc0; z = cz + bz * sc0;
        double vr0 = ((double)c->pvx[p] * bx + (double)c->pvy[p] * by + (double)c->pvz[p] * bz) / br;
        if (vr0 > 0) {
          c->pvx[p] = (oc_real)(c->pvx[p] - bx / br * vr0 * 1.15);
          c->pvy[p] = (oc_real)(c->pvy[p] - by / br * vr0 * 1.15);
          c->pvz[p] = (oc_real)(c->pvz[p] - bz / br * vr0 * 1.15);
        }
      } else if (br <= domHi) {
        int tries0 = 0;
        while (hasTerr && oc_rock_cell_at(c, x, y, z) && tries0 < 8) {
          double rW0 = sqrt((x - cx) * (x - cx) + (y - cy) * (y - cy) + (z - cz) * (z - cz));
          double stW0 = dx * 0.5 < domHi - rW0 ? dx * 0.5 : domHi - rW0;
          if (stW0 <= 1e-5) { tries0 = 8; break; }
          x += bx / br * stW0; y += by / br * stW0; z += bz / br * stW0;
          tries0++;
        }
        if (!hasTerr && br < coreLo) {
          double sc1 = coreLo / br;
          x = cx + bx * sc1; y = cy + by * sc1; z = cz + bz * sc1;
          tries0 = 1;
        }
        if (tries0 > 0) {
          if (hasTerr) {
            double exS = x - cx, eyS = y - cy, ezS = z - cz;
            double erS = rlen(exS, eyS, ezS);
            double wantS = oc_terrain_radius_at(c, x, y, z) + dx * 0.06;
            if (erS > wantS) { x = cx + exS / erS * wantS; y = cy + eyS / erS * wantS; z = cz + ezS / erS * wantS; }
            int kS = 0;
            while (oc_rock_cell_at(c, x, y, z) && kS < 4) {
              x += exS / erS * dx * 0.5; y += eyS / erS * dx * 0.5; z += ezS / erS * dx * 0.5;
              kS++;
            }
          }
          double vr1 = ((double)c->pvx[p] * bx + (double)c->pvy[p] * by + (double)c->pvz[p] * bz) / br;
          c->pvx[p] = (oc_real)(c->pvx[p] - bx / br * vr1);
          c->pvy[p] = (oc_real)(c->pvy[p] - by / br * vr1);
          c->pvz[p] = (oc_real)(c->pvz[p] - bz / br * vr1);
          c->pvx[p] = (oc_real)(c->pvx[p] * 0.3); c->pvy[p] = (oc_real)(c->pvy[p] * 0.3); c->pvz[p] = (oc_real)(c->pvz[p] * 0.3);
          c->pflag[p] = 0;
        }
      }
      c->px[p] = (oc_real)x; c->py[p] = (oc_real)y; c->pz[p] = (oc_real)z;
      int ci = clampi((int)floor(x / dx), 0, nx - 1), cj = clampi((int)floor(y / dx), 0, ny - 1);
      int ck = clampi((int)floor(z / dx), 0, nz - 1);
      if (t[((long)ck * ny + cj) * nx + ci] == FLUID) {
        c->pvx[p] = (oc_real)(c->pvx[p] * 0.45); c->pvy[p] = (oc_real)(c->pvy[p] * 0.45); c->pvz[p] = (oc_real)(c->pvz[p] * 0.45);
        if (c->pvy[p] > 0.5) c->pvy[p] = (oc_real)0.5;
        c->pcool[p] = 2;
        c->pflag[p] = 0;
      }
      continue;
    }

    if (fl == 2 || fl == 3) { /* vapor / cloud, solver.js:1978-2174 */
      double Tv = c->pT[p];
      c->pAir[p] = (oc_real)(c->pAir[p] + dt);
      double vEx = x - cx, vEy = y - cy, vEz = z - cz;
      double vEr = rlen(vEx, vEy, vEz);
      double ux = vEx / vEr, uy = vEy / vEr, uz = vEz / vEr;
      double exposure = ux * wSX + uy * wSY + uz * wSZ;
      double jet = c->oceanR * 0.58 < c->maxSpeed * 0.45 ? c->oceanR * 0.58 : c->maxSpeed * 0.45;
      double ret = c->oceanR * 0.32 < c->maxSpeed * 0.25 ? c->oceanR * 0.32 : c->maxSpeed * 0.25;
      double vr = vx * ux + vy * uy + vz * uz;
      double radial = -c->gravity * 0.06;
      double wx = -uz * jet - (wSX - exposure * ux) * ret;
      double wy = -(wSY - exposure * uy) * ret;
      double wz = ux * jet - (wSZ - exposure * uz) * ret;
      double relax = dt * 1.3 < 1 ? dt * 1.3 : 1;
      c->pvx[p] = (oc_real)(vx + (wx - (vx - vr * ux)) * relax + ux * radial * dt);
      c->pvy[p] = (oc_real)(vy + (wy - (vy - vr * uy)) * relax + uy * radial * dt);
      c->pvz[p] = (oc_real)(vz + (wz - (vz - vr * uz)) * relax + uz * radial * dt);
      double svv = (double)c->pvx[p] * c->pvx[p] + (double)c->pvy[p] * c->pvy[p] + (double)c->pvz[p] * c->pvz[p];
      if (svv > ms2) {
        double svs = c->maxSpeed / sqrt(svv);
        c->pvx[p] = (oc_real)(c->pvx[p] * svs); c->pvy[p] = (oc_real)(c->pvy[p] * svs); c->pvz[p] = (oc_real)(c->pvz[p] * svs);
      }
      x += c->pvx[p] * dt; y += c->pvy[p] * dt; z += c->pvz[p] * dt;
      double vCx = x - cx, vCy = y - cy, vCz = z - cz;
      double vCr = rlen(vCx, vCy, vCz);
      double ceilR = c->oceanR + c->atmosphereH - 1e-3 * dx;
      if (hasTerr) {
        double RtE = oc_terrain_radius_at(c, x, y, z);
        if (RtE + 0.02 > ceilR) ceilR = RtE + 0.02;
      }
      if (vCr > ceilR) {
        double scE = ceilR / vCr;
        x = cx + vCx * scE; y = cy + vCy * scE; z = cz + vCz * scE;
        double vrE = ((double)c->pvx[p] * vCx + (double)c->pvy[p] * vCy + (double)c->pvz[p] * vCz) / vCr;
        if (vrE > 0) {
          c->pvx[p] = (oc_real)(c->pvx[p] - vCx / vCr * vrE * 2);
          c->pvy[p] = (oc_real)(c->pvy[p] - vCy / vCr * vrE * 2);
          c->pvz[p] = (oc_real)(c->pvz[p] - vCz / vCr * vrE * 2);
        }
      }
      double vrC2 = ((double)c->pvx[p] * vCx + (double)c->pvy[p] * vCy + (double)c->pvz[p] * vCz) / vCr;
      if (vrC2 > 0 && vCr > c->oceanR) {
        double hh = (vCr - c->oceanR) / (c->atmosphereH > 1e-4 ? c->atmosphereH : 1e-4);
        if (hh > 1) hh = 1;
        double prf = c->ceilReflect == 1 ? hh * hh : c->ceilReflect == 2 ? (exp(3 * hh) - 1) * 0.0523957 : hh;
        if (prf > 0) {
          c->cfS = lcg(c->cfS);
          if (c->cfS / 4294967296.0 < prf * dt * 4) {
            c->pvx[p] = (oc_real)(c->pvx[p] - vCx / vCr * vrC2 * 2);
            c->pvy[p] = (oc_real)(c->pvy[p] - vCy / vCr * vrC2 * 2);
            c->pvz[p] = (oc_real)(c->pvz[p] - vCz / vCr * vrC2 * 2);
          }
        }
      }
      double fCx = x - cx, fCy = y - cy, fCz = z - cz;
      double fCr = rlen(fCx, fCy, fCz);
      double floorR = c->oceanR + 0.15 * dx;
      if (fCr < floorR) {
        double vrF = ((double)c->pvx[p] * fCx + (double)c->pvy[p] * fCy + (double)c->pvz[p] * fCz) / fCr;
        if (vrF < 0) {
          c->bfS = lcg(c->bfS);
          double sc = (c->bfS / 4294967296.0 < 0.07) ? floorR / fCr : (floorR + 0.02 * dx) / fCr;
          x = cx + fCx * sc; y = cy + fCy * sc; z = cz + fCz * sc;
          c->pvx[p] = (oc_real)(c->pvx[p] - fCx / fCr * vrF * 2);
          c->pvy[p] = (oc_real)(c->pvy[p] - fCy / fCr * vrF * 2);
          c->pvz[p] = (oc_real)(c->pvz[p] - fCz / fCr * vrF * 2);
        }
      }
      int triesE = 0;
      double rNow3 = sqrt((x - cx) * (x - cx) + (y - cy) * (y - cy) + (z - cz) * (z - cz));
      while (hasTerr && rNow3 < c->domainR && oc_rock_cell_at(c, x, y, z) && triesE < 8) {
        double rW3 = sqrt((x - cx) * (x - cx) + (y - cy) * (y - cy) + (z - cz) * (z - cz));
        double stW3 = dx * 0.5 < domHi - rW3 ? dx * 0.5 : domHi - rW3;
        if (stW3 <= 1e-5) break;
        x += vCx / vCr * stW3; y += vCy / vCr * stW3; z += vCz / vCr * stW3;
        triesE++;
      }
      if (triesE > 0) {
        double vrT = ((double)c->pvx[p] * vCx + (double)c->pvy[p] * vCy + (double)c->pvz[p] * vCz) / vCr;
        if (vrT < 0) {
          double bT = 1.5 * vrT / vCr;
          c->pvx[p] = (oc_real)(c->pvx[p] - vCx * bT); c->pvy[p] = (oc_real)(c->pvy[p] - vCy * bT);
          c->pvz[p] = (oc_real)(c->pvz[p] - vCz * bT);
        }
        c->pvx[p] = (oc_real)(c->pvx[p] * 0.92); c->pvy[p] = (oc_real)(c->pvy[p] * 0.92); c->pvz[p] = (oc_real)(c->pvz[p] * 0.92);
      }
      int ciV = clampi((int)floor(x / dx), 0, nx - 1), cjV = clampi((int)floor(y / dx), 0, ny - 1);
      int ckV = clampi((int)floor(z / dx), 0, nz - 1);
      int backToSea = t[((long)ckV * ny + cjV) * nx + ciV] == FLUID;
      if (backToSea) {
        double bx2 = x - cx, by2 = y - cy, bz2 = z - cz;
        double br2 = rlen(bx2, by2, bz2);
        if (br2 > c->oceanR + 0.15 * dx) backToSea = 0;
        else if (((double)c->pvx[p] * bx2 + (double)c->pvy[p] * by2 + (double)c->pvz[p] * bz2) / br2 > 0.02) backToSea = 0;
      }
      c->px[p] = (oc_real)x; c->py[p] = (oc_real)y; c->pz[p] = (oc_real)z;
      if (backToSea) {
        c->pvx[p] = 0; c->pvy[p] = 0; c->pvz[p] = 0; c->pAir[p] = 0; c->pflag[p] = 0;
      } else if (fl == 2 && ((Tv <= rainTe && c->pAir[p] >= c->rainLiftAge) || c->pAir[p] > 40)) {
        c->pvx[p] = 0; c->pvy[p] = 0; c->pvz[p] = 0; c->pAir[p] = 0; c->pflag[p] = 4;
      }
      continue;
    }

    /* grid-coupled fluid: RK2/RK1 through the grid, solver.js:2176-2290 */
    double t0x = c->advU[p], t0y = c->advV[p], t0z = c->advW[p];
    double sp2 = t0x * t0x + t0y * t0y + t0z * t0z;
    double v2x = t0x, v2y = t0y, v2z = t0z;
    if (sp2 > 0.25) {
      double mx = clampd(x + 0.5 * dt * t0x, xLo, xHi);
      double my = clampd(y + 0.5 * dt * t0y, yLo, yHi);
      double mz = clampd(z + 0.5 * dt * t0z, zLo, zHi);
      double t2[3];
      oc_sample_vel3(c, mx, my, mz, t2);
      v2x = t2[0]; v2y = t2[1]; v2z = t2[2];
    }
    sp2 = v2x * v2x + v2y * v2y + v2z * v2z;
    if (sp2 > ms2) {
      double sc = c->maxSpeed / sqrt(sp2);
      v2x *= sc; v2y *= sc; v2z *= sc;
      c->pvx[p] = (oc_real)(c->pvx[p] * sc); c->pvy[p] = (oc_real)(c->pvy[p] * sc); c->pvz[p] = (oc_real)(c->pvz[p] * sc);
    }
    x += v2x * dt; y += v2y * dt; z += v2z * dt;
    double nx2 = x - cx, ny2 = y - cy, nz2 = z - cz;
    double nr = rlen(nx2, ny2, nz2);
    if (nr > domHi) {
      double sc2 = domHi / nr;
      x = cx + nx2 * sc2; y = cy + ny2 * sc2; z = cz + nz2 * sc2;
      c->pvy[p] = (oc_real)(-fabs((double)c->pvy[p]) * 0.1 - 0.3);
      c->pflag[p] = 1;
    }
    if (nr < domHi) {
      int tries = 0;
      while (hasTerr && oc_rock_cell_at(c, x, y, z) && tries < 8) {
        double rW2 = sqrt((x - cx) * (x - cx) + (y - cy) * (y - cy) + (z - cz) * (z - cz));
        double stW2 = dx * 0.5 < domHi - rW2 ? dx * 0.5 : domHi - rW2;
        if (stW2 <= 1e-5) { tries = 8; break; }
        x += nx2 / nr * stW2; y += ny2 / nr * stW2; z += nz2 / nr * stW2;
        tries++;
      }
      if (tries > 0) {
        if (hasTerr) {
          double exS2 = x - cx, eyS2 = y - cy, ezS2 = z - cz;
          double erS2 = rlen(exS2, eyS2, ezS2);
          double wantS2 = oc_terrain_radius_at(c, x, y, z) + dx * 0.06;
          if (erS2 > wantS2) { x = cx + exS2 / erS2 * wantS2; y = cy + eyS2 / erS2 * wantS2; z = cz + ezS2 / erS2 * wantS2; }
          int kS2 = 0;
          while (oc_rock_cell_at(c, x, y, z) && kS2 < 4) {
            x += exS2 / erS2 * dx * 0.5; y += eyS2 / erS2 * dx * 0.5; z += ezS2 / erS2 * dx * 0.5;
            kS2++;
          }
        }
        double vr2 = ((double)c->pvx[p] * nx2 + (double)c->pvy[p] * ny2 + (double)c->pvz[p] * nz2) / nr;
        if (vr2 < 0) {
          c->pvx[p] = (oc_real)(c->pvx[p] - nx2 / nr * vr2);
          c->pvy[p] = (oc_real)(c->pvy[p] - ny2 / nr * vr2);
          c->pvz[p] = (oc_real)(c->pvz[p] - nz2 / nr * vr2);
        }
      }
      if (hasTerr && (tries > 0 || oc_rock_cell_at(c, x - nx2 / nr * dx, y - ny2 / nr * dx, z - nz2 / nr * dx))) {
        double fr0 = 1 - 0.6 * dt;
        if (fr0 < 0) fr0 = 0;
        c->pvx[p] = (oc_real)(c->pvx[p] * fr0); c->pvy[p] = (oc_real)(c->pvy[p] * fr0); c->pvz[p] = (oc_real)(c->pvz[p] * fr0);
      }
    }
    double sx = x - cx, sy = y - cy, sz = z - cz;
    double sr = rlen(sx, sy, sz);
    double vrad = ((double)c->pvx[p] * sx + (double)c->pvy[p] * sy + (double)c->pvz[p] * sz) / sr;
    double da = 1.0 * dt;
    if (da > 0.5) da = 0.5;
    c->pvx[p] = (oc_real)(c->pvx[p] - sx / sr * vrad * da);
    c->pvy[p] = (oc_real)(c->pvy[p] - sy / sr * vrad * da);
    c->pvz[p] = (oc_real)(c->pvz[p] - sz / sr * vrad * da);
    double dg = 1 - 0.42 * dt;
    if (dg < 0) dg = 0;
    c->pvx[p] = (oc_real)(c->pvx[p] * dg); c->pvy[p] = (oc_real)(c->pvy[p] * dg); c->pvz[p] = (oc_real)(c->pvz[p] * dg);
    c->px[p] = (oc_real)x; c->py[p] = (oc_real)y; c->pz[p] = (oc_real)z;
  }
  /* ball collisions after advection, solver.js:1762-1763 */
  oc_push_balls(c);
}

/* --------------------------------------------------------------- substep */
/* _substep, solver.js:2568-2624 (sphere mode, no balls) */
/* ------------------------------------------------------------- balls */
int oc_set_balls(oc_ctx *c, int n, const double *s9) {
  if (n < 0 || n > OC_MAX_BALLS) return -1;
  c->nBalls = n;
  for (int i = 0; i < n; i++) {
    oc_ball *b = &c->balls[i];
    const double *q = s9 + 9 * i;
    b->x = q[0]; b->y = q[1]; b->z = q[2]; b->vx = q[3]; b->vy = q[4]; b->vz = q[5];
    b->r = q[6]; b->rho = q[7]; b->submerged = q[8];
  }
  return 0;
}

int oc_get_balls(oc_ctx *c, int n, double *s9) {
  if (n < 0 || n > c->nBalls) return -1;
  for (int i = 0; i < n; i++) {
    const oc_ball *b = &c->balls[i];
    double *q = s9 + 9 * i;
    q[0] = b->x; q[1] = b->y; q[2] = b->z; q[3] = b->vx; q[4] = b->vy; q[5] = b->vz;
    q[6] = b->r; q[7] = b->rho; q[8] = b->submerged;
  }
  return 0;
}

const float *oc_ball_hist(oc_ctx *c) { return c->hist; }
const float *oc_ball_surf(oc_ctx *c) { return c->surfY; }

/* _ballVsTerrain, solver.js:2299-2342: the deepest penetration of the ball
 * into a solid voxel of the terrain lattice (first voxel in k, j, i order
 * among equals) pushes the ball out along the voxel's closest-point normal
 * (radially when the centre is inside the voxel) */
static void ball_vs_terrain(oc_ctx *c, oc_ball *b) {
  int n = c->tn;
  double dv = c->dv;
  const uint8_t *solid = c->tsolid;
  int i0 = (int)((b->x - b->r) / dv), i1 = (int)((b->x + b->r) / dv);
  int j0 = (int)((b->y - b->r) / dv), j1 = (int)((b->y + b->r) / dv);
  int k0 = (int)((b->z - b->r) / dv), k1 = (int)((b->z + b->r) / dv);
  if (i0 < 0) i0 = 0;
  if (j0 < 0) j0 = 0;
  if (k0 < 0) k0 = 0;
  if (i1 > n - 1) i1 = n - 1;
  if (j1 > n - 1) j1 = n - 1;
  if (k1 > n - 1) k1 = n - 1;
  double bestPen = 0, bpx = 0, bpy = 0, bpz = 0;
  for (int k = k0; k <= k1; k++)
    for (int j = j0; j <= j1; j++) {
      long rowB = ((long)k * n + j) * n;
      for (int i = i0; i <= i1; i++) {
        if (!solid[rowB + i]) continue;
        double qx = b->x, qy = b->y, qz = b->z;
        double vx0 = i * dv, vy0 = j * dv, vz0 = k * dv;
        if (qx < vx0) qx = vx0; else if (qx > vx0 + dv) qx = vx0 + dv;
        if (qy < vy0) qy = vy0; else if (qy > vy0 + dv) qy = vy0 + dv;
        if (qz < vz0) qz = vz0; else if (qz > vz0 + dv) qz = vz0 + dv;
        double ddx = b->x - qx, ddy = b->y - qy, ddz = b->z - qz;
        double dd = sqrt(ddx * ddx + ddy * ddy + ddz * ddz);
        double pen = b->r - dd;
        if (pen > bestPen) {
          if (dd > 1e-6) {
            bestPen = pen; bpx = ddx / dd; bpy = ddy / dd; bpz = ddz / dd;
          } else {
            double rx = b->x - c->cx, ry = b->y - c->cy, rz = b->z - c->cz;
            double rl = sqrt(rx * rx + ry * ry + rz * rz);
            if (rl == 0) rl = 1e-9;
            bestPen = pen; bpx = rx / rl; bpy = ry / rl; bpz = rz / rl;
          }
        }
      }
    }
  if (bestPen > 0) {
    b->x += bpx * (bestPen + 0.002); b->y += bpy * (bestPen + 0.002); b->z += bpz * (bestPen + 0.002);
    double vn = b->vx * bpx + b->vy * bpy + b->vz * bpz;
    if (vn < 0) {
      b->vx -= bpx * vn; b->vy -= bpy * vn; b->vz -= bpz * vn;
      b->vx *= 0.985; b->vy *= 0.985; b->vz *= 0.985;
    }
  }
}

/* _updateBalls, solver.js:2344-2564 (sphere mode): the local sea level of
 * each ball from a 24-bucket radial histogram of the liquid particles in a
 * shell 2.2r..3.2r around it (90th percentile, fallbacks to oceanR); then
 * per ball Archimedes buoyancy of the submerged cap, weight under the
 * inverse-square law, quadratic drag against the grid velocity at the centre
 * (clamped to 0.8 m g), added mass, damping, a 4 m/s clamp, explicit
 * integration, terrain (or core) contact and the domain sphere; finally the
 * ball-ball contacts in pair order */
void oc_update_balls(oc_ctx *c, double dt) {
  const double rhoF = 1000, Cd = 0.5;
  const int HIST = 24;
  int nb = c->nBalls;
  if (nb > 0) {
    double bcx = c->cx, bcy = c->cy, bcz = c->cz;
    double hLo = c->coreR, hSpan = (c->oceanR - c->coreR) * 1.5;
    for (int i = 0; i < nb * HIST; i++) c->hist[i] = 0;
    for (int i = 0; i < nb; i++) c->histC[i] = 0;
    for (int p = 0; p < c->nP; p++) {
      if (c->pflag[p] != 0) continue;
      double ppx = c->px[p], ppy = c->py[p], ppz = c->pz[p];
      double ldx = ppx - bcx, ldy = ppy - bcy, ldz = ppz - bcz;
      double level = sqrt(ldx * ldx + ldy * ldy + ldz * ldz);
      for (int sb = 0; sb < nb; sb++) {
        const oc_ball *bb = &c->balls[sb];
        double bdx = ppx - bb->x, bdy = ppy - bb->y, bdz = ppz - bb->z;
        double dh2 = bdx * bdx + bdy * bdy + bdz * bdz;
        double r1 = 2.2 * bb->r, r2 = 3.2 * bb->r;
        if (dh2 < r1 * r1 || dh2 > r2 * r2) continue;
        int bucket = (int)floor(((level - hLo) / hSpan) * HIST);
        if (bucket < 0) bucket = 0; else if (bucket >= HIST) bucket = HIST - 1;
        c->hist[sb * HIST + bucket] += 1.0f;
        c->histC[sb]++;
      }
    }
    for (int sb = 0; sb < nb; sb++) {
      int total = c->histC[sb];
      double wl = -INFINITY;
      if (total > 20) {
        double acc = 0, target = total * 0.10;
        for (int bkt = HIST - 1; bkt >= 0; bkt--) {
          acc += c->hist[sb * HIST + bkt];
          if (acc >= target) { wl = hLo + (bkt + 0.5) / HIST * hSpan; break; }
        }
        if (wl == -INFINITY || total < 140) wl = c->oceanR;
        double wlMax = c->oceanR + (c->oceanR - c->coreR) * 0.3;
        if (wl > wlMax) wl = wlMax;
        if (wl < c->coreR) wl = c->coreR;
      } else {
        wl = c->oceanR;
      }
      c->surfY[sb] = (float)wl;
    }
  }
  for (int n = 0; n < nb; n++) {
    oc_ball *b = &c->balls[n];
    double m = b->rho * 4 / 3 * JS_PI * b->r * b->r * b->r;
    double area = JS_PI * b->r * b->r;
    double surf = c->surfY[n];
    double bR = sqrt((b->x - c->cx) * (b->x - c->cx) + (b->y - c->cy) * (b->y - c->cy) + (b->z - c->cz) * (b->z - c->cz));
    double inner = bR - b->r;
    double h = surf - inner;
    if (!(h > 0)) h = 0;
    if (h > 2 * b->r) h = 2 * b->r;
    double Vs = JS_PI * h * h * (3 * b->r - h) / 3;
    b->submerged = Vs / (4.0 / 3 * JS_PI * b->r * b->r * b->r);
    double upx = 0, upy = 1, upz = 0, gEff = c->gravity;
    if (bR > 1e-6) {
      upx = (b->x - c->cx) / bR; upy = (b->y - c->cy) / bR; upz = (b->z - c->cz) / bR;
      gEff = c->gravity * (c->oceanR * c->oceanR) / (bR * bR);
    }
    double Fb = (rhoF * Vs - m) * gEff;
    double tmp[3];
    oc_sample_vel3(c, b->x, b->y, b->z, tmp);
    double rvx = b->vx - tmp[0], rvy = b->vy - tmp[1], rvz = b->vz - tmp[2];
    double vr = sqrt(rvx * rvx + rvy * rvy + rvz * rvz);
    double kd = 0.5 * rhoF * Cd * area * vr;
    double Fx = -kd * rvx + upx * Fb, Fz = -kd * rvz + upz * Fb, Fyd = -kd * rvy + upy * Fb;
    {
      double fdx = Fx - upx * Fb, fdy = Fyd - upy * Fb, fdz = Fz - upz * Fb;
      double fdm = sqrt(fdx * fdx + fdy * fdy + fdz * fdz);
      double fdmMax = 0.8 * m * c->gravity;
      if (fdm > fdmMax) {
        double fdc = fdmMax / fdm;
        Fx = upx * Fb + fdx * fdc; Fyd = upy * Fb + fdy * fdc; Fz = upz * Fb + fdz * fdc;
      }
    }
    double mEff = m + 0.5 * rhoF * Vs;
    b->vx += (Fx / mEff) * dt;
    b->vz += (Fz / mEff) * dt;
    b->vy += (Fyd / mEff) * dt;
    b->vx *= 0.999; b->vy *= 0.999; b->vz *= 0.999;
    const double bvmax = 4;
    if (b->vx > bvmax) b->vx = bvmax; else if (b->vx < -bvmax) b->vx = -bvmax;
    if (b->vz > bvmax) b->vz = bvmax; else if (b->vz < -bvmax) b->vz = -bvmax;
    if (b->vy > bvmax) b->vy = bvmax; else if (b->vy < -bvmax) b->vy = -bvmax;
    b->x += b->vx * dt; b->y += b->vy * dt; b->z += b->vz * dt;
    if (c->tsolid) {
      ball_vs_terrain(c, b);
    } else {
      double cbx = b->x - c->cx, cby = b->y - c->cy, cbz = b->z - c->cz;
      double cbr = sqrt(cbx * cbx + cby * cby + cbz * cbz);
      if (cbr == 0) cbr = 1e-9;
      if (cbr < c->coreR + b->r) {
        double rnx0 = cbx / cbr, rny0 = cby / cbr, rnz0 = cbz / cbr;
        double rr0 = c->coreR + b->r;
        b->x = c->cx + rnx0 * rr0; b->y = c->cy + rny0 * rr0; b->z = c->cz + rnz0 * rr0;
        double vn0 = b->vx * rnx0 + b->vy * rny0 + b->vz * rnz0;
        if (vn0 < 0) {
          b->vx -= rnx0 * vn0 * 1.3; b->vy -= rny0 * vn0 * 1.3; b->vz -= rnz0 * vn0 * 1.3;
          b->vx *= 0.98; b->vy *= 0.98; b->vz *= 0.98;
        }
      }
    }
    double sbx = b->x - c->cx, sby = b->y - c->cy, sbz = b->z - c->cz;
    double sbr = sqrt(sbx * sbx + sby * sby + sbz * sbz);
    if (sbr == 0) sbr = 1e-9;
    double rnx = sbx / sbr, rny = sby / sbr, rnz = sbz / sbr;
    if (sbr > c->domainR - b->r) {
      double rr1 = c->domainR - b->r;
      b->x = c->cx + rnx * rr1; b->y = c->cy + rny * rr1; b->z = c->cz + rnz * rr1;
      double vn1 = b->vx * rnx + b->vy * rny + b->vz * rnz;
      if (vn1 > 0) { b->vx -= rnx * vn1 * 1.3; b->vy -= rny * vn1 * 1.3; b->vz -= rnz * vn1 * 1.3; }
    }
  }
  for (int a = 0; a < nb; a++)
    for (int b2 = a + 1; b2 < nb; b2++) {
      oc_ball *A = &c->balls[a], *B = &c->balls[b2];
      double ex = B->x - A->x, ey = B->y - A->y, ez = B->z - A->z;
      double d2 = ex * ex + ey * ey + ez * ez, rSum = A->r + B->r;
      if (d2 > rSum * rSum || d2 < 1e-12) continue;
      double d = sqrt(d2), nxn = ex / d, nyn = ey / d, nzn = ez / d;
      double overlap = rSum - d;
      double mA = A->rho * A->r * A->r * A->r, mB = B->rho * B->r * B->r * B->r;
      double tA = mB / (mA + mB), tB = mA / (mA + mB);
      A->x -= nxn * overlap * tA; A->y -= nyn * overlap * tA; A->z -= nzn * overlap * tA;
      B->x += nxn * overlap * tB; B->y += nyn * overlap * tB; B->z += nzn * overlap * tB;
      double rvn = (B->vx - A->vx) * nxn + (B->vy - A->vy) * nyn + (B->vz - A->vz) * nzn;
      if (rvn < 0) {
        double jimp = -(1 + 0.3) * rvn / (1 / mA + 1 / mB);
        A->vx -= nxn * jimp / mA; A->vy -= nyn * jimp / mA; A->vz -= nzn * jimp / mA;
        B->vx += nxn * jimp / mB; B->vy += nyn * jimp / mB; B->vz += nzn * jimp / mB;
      }
    }
}

void oc_substep(oc_ctx *c, double dt) {
  oc_clear_faces(c);
  oc_rasterize_static(c);
  oc_mark_fluid(c, 0, c->nP);
  oc_push_balls(c);
  oc_p2g(c);
  oc_apply_bc(c);
  oc_snapshot(c);
  oc_viscosity(c, dt);
  oc_apply_gravity(c, dt);
  oc_apply_bc(c);
  oc_pressure_solve(c);
  oc_project(c);
  if (c->vorticity > 0) oc_vorticity(c, dt);
  oc_clamp_umax(c);
  oc_extrapolate(c);
  oc_g2p(c, dt);
  oc_advect(c, dt);
  oc_update_balls(c, dt);
  c->dtLast = dt;
}

/* CFL law of _stepSync for one substep, solver.js:2689-2690 */
double oc_dt_law(oc_ctx *c) {
  double um = c->umax > 0.08 ? c->umax : 0.08;
  double dt = c->cfl * c->dx / um;
  if (dt > 1.0 / 100) dt = 1.0 / 100;
  if (dt < 1e-5) dt = 1e-5;
  return dt;
}

/* ------------------------------------------------------- frame tick */
static int trunc_clamp(double v, int n) {
  /* (v | 0) then clamped to [0, n-1] */
  int i = (int)v;
  return i < 0 ? 0 : (i >= n ? n - 1 : i);
}

/* _updateHeat, solver.js:2715-2994: exposure + relaxation per particle,
 * conservative cell conduction on cell means (f64), radial convection */
void oc_update_heat(oc_ctx *c, double dt) {
  int nP = c->nP, nx = c->nx, ny = c->ny, nz = c->nz;
  if (nP < 1) return;
  double dx = c->dx, alpha = c->heatK * dt * 0.15;
  if (alpha > 0.14) alpha = 0.14;
  long nC = c->nC;
  int *cnt = (int *)calloc(nC, sizeof(int));
  int *cell = (int *)malloc(sizeof(int) * nP);
  for (int p = 0; p < nP; p++) {
    int gx = trunc_clamp(c->px[p] / dx, nx), gy = trunc_clamp(c->py[p] / dx, ny), gz = trunc_clamp(c->pz[p] / dx, nz);
    cell[p] = (gz * ny + gy) * nx + gx;
    cnt[cell[p]]++;
  }
  /* _heatShadowChunk, :2773-2918 */
  double cx = c->cx, cy = c->cy, cz = c->cz, coreR2 = c->coreR * c->coreR;
  long sj = nx + 1, sk = (long)(ny + 1) * sj;
  double sunPow = c->sunActivity / 3, dissip = 0.045, half = dx, marchMax = 0.55, absK = 1.1;
  const uint8_t *type = c->cellType;
  for (int p = 0; p < nP; p++) {
    double X = c->px[p], Y = c->py[p], Z = c->pz[p];
    double lit = c->hasSun ? 1 : 0;
    double dwx = X - cx, dwy = Y - cy, dwz = Z - cz;
    double dwl = sqrt(dwx * dwx + dwy * dwy + dwz * dwz);
    if (!(dwl != 0)) dwl = 1e-9;
    double stpU = dx / dwl;
    int above = 0;
    for (int st = 1; st <= 4; st++) {
      double qx = X + dwx * stpU * st, qy = Y + dwy * stpU * st, qz = Z + dwz * stpU * st;
      int qi = (int)(qx / dx), qj = (int)(qy / dx), qk = (int)(qz / dx);
      if (qi < 0 || qj < 0 || qk < 0 || qi >= nx || qj >= ny || qk >= nz) break;
      if (type[((long)qk * ny + qj) * nx + qi] != FLUID) break;
      above++;
    }
    if (c->pDepth) c->pDepth[p] = (uint8_t)above;
    if (c->hasSun) {
      if (above >= 4) lit = 0;
      else {
        double lx = c->sunX - X, ly = c->sunY - Y, lz = c->sunZ - Z;
        double ld = sqrt(lx * lx + ly * ly + lz * lz);
        if (!(ld != 0)) ld = 1e-9;
        double dxn = lx / ld, dyn = ly / ld, dzn = lz / ld;
        double ex = cx - X, ey = cy - Y, ez = cz - Z;
        double b = ex * dxn + ey * dyn + ez * dzn;
        if (c->tsolid) {
          double d2c = ex * ex + ey * ey + ez * ez - b * b;
          if (b > 0 && d2c < c->Rlo * c->Rlo * 0.96) lit = 0;
          else if (b > 0 || d2c < c->Rhi * c->Rhi * 1.1) {
            double stp = dx * 1.5, mx = X, my = Y, mz = Z, esc2 = (c->Rhi + 0.1) * (c->Rhi + 0.1);
            int tvn = c->tn;
            for (int st2 = 0; st2 < 26; st2++) {
              mx += dxn * stp; my += dyn * stp; mz += dzn * stp;
              double rex = mx - cx, rey = my - cy, rez = mz - cz;
              if (rex * rex + rey * rey + rez * rez > esc2) break;
              int ivx = (int)(mx / c->dv); if (ivx < 0 || ivx >= tvn) break;
              int ivy = (int)(my / c->dv); if (ivy < 0 || ivy >= tvn) break;
              int ivz = (int)(mz / c->dv); if (ivz < 0 || ivz >= tvn) break;
              if (c->tsolid[((long)ivz * tvn + ivy) * tvn + ivx]) { lit = 0; break; }
            }
          }
        } else if (b > 0) {
          double d2c2 = ex * ex + ey * ey + ez * ez - b * b;
          if (d2c2 < coreR2) lit = 0;
        }
        if (lit != 0) {
          double od = 0, sx = X, sy = Y, sz = Z;
          for (int st = 0; st < 6 && st * half < marchMax; st++) {
            sx += dxn * half; sy += dyn * half; sz += dzn * half;
            int i0 = (int)(sx / dx), j0 = (int)(sy / dx), k0 = (int)(sz / dx);
            if (i0 < 0 || j0 < 0 || k0 < 0 || i0 >= nx || j0 >= ny || k0 >= nz) break;
            od += c->dens[k0 * sk + j0 * sj + i0] * half * absK;
            if (od > 3.5) break;
          }
          lit = od < 3.5 ? exp(-od) : 0;
        }
      }
    }
    if (c->pLight) c->pLight[p] = (oc_real)lit;
    double T, Tp = c->pT[p];
    int fl = c->pflag[p];
    if (fl == 2) {
      double amb = lit > 0.05 ? c->Tamb + 0.18 : 0.06, k = c->hasSun ? 0.16 + (1 - lit) * 0.20 : 1.2;
      T = Tp + (amb - Tp) * k * dt;
      if (lit > 0) T += sunPow * 0.6 * lit * dt;
    } else if (fl == 3) {
      double amb = lit > 0.05 ? c->Tamb + 0.18 : 0.06, k = c->hasSun ? 0.10 + (1 - lit) * 0.12 : 0.9;
      T = Tp + (amb - Tp) * k * dt;
      if (lit > 0) T += sunPow * 0.35 * lit * dt;
    } else if (fl == 4) {
      double amb = lit > 0.05 ? c->Tamb + 0.18 : 0.06, k = c->hasSun ? 0.14 + (1 - lit) * 0.16 : 1.0;
      T = Tp + (amb - Tp) * k * dt;
      if (lit > 0) T += sunPow * 0.45 * lit * dt;
    } else {
      double amb = lit > 0.05 ? c->Tamb : 0;
      T = Tp + (amb - Tp) * (dissip + (1 - lit) * c->shadeCool) * dt;
      if (lit > 0) T += sunPow * lit * dt;
    }
    if (T > 1.15) T = 1.15; else if (T < 0) T = 0;
    c->pT[p] = (oc_real)T;
  }
  if (c->heatK > 0) {
    /* _heatMeansSlice, _heatConductFlux, _heatApplySlice, :2921-2979 */
    double *means = (double *)calloc(nC, sizeof(double)), *changes = (double *)calloc(nC, sizeof(double));
    for (int p = 0; p < nP; p++) means[cell[p]] += c->pT[p];
    for (long cc = 0; cc < nC; cc++) means[cc] = cnt[cc] ? means[cc] / cnt[cc] : 0;
    long nxy = (long)nx * ny;
    for (int k = 0; k < nz; k++)
      for (int j = 0; j < ny; j++)
        for (int i = 0; i < nx; i++) {
          long cc = ((long)k * ny + j) * nx + i;
          if (!cnt[cc]) continue;
          for (int axis = 0; axis < 3; axis++) {
            if ((axis == 0 && i == nx - 1) || (axis == 1 && j == ny - 1) || (axis == 2 && k == nz - 1)) continue;
            long nb = cc + (axis == 0 ? 1 : axis == 1 ? nx : nxy);
            if (!cnt[nb]) continue;
            double e = alpha * (cnt[cc] < cnt[nb] ? cnt[cc] : cnt[nb]) * (means[nb] - means[cc]);
            changes[cc] += e;
            changes[nb] -= e;
          }
        }
    for (int p = 0; p < nP; p++) {
      long cc = cell[p];
      double deltaT = changes[cc] / cnt[cc];
      c->pT[p] = (oc_real)(c->pT[p] + (alpha * (means[cc] - c->pT[p]) + deltaT));
    }
    free(means);
    free(changes);
  }
  /* _heatConvectChunk, :2982-2994 */
  for (int p = 0; p < nP; p++) {
    if (c->pflag[p] != 0) continue;
    double rx = c->px[p] - cx, ry = c->py[p] - cy, rz = c->pz[p] - cz;
    double rl = sqrt(rx * rx + ry * ry + rz * rz);
    if (!(rl != 0)) rl = 1e-9;
    double ab = 3.0 * (c->pT[p] - c->Tamb) * dt / rl;
    c->pvx[p] = (oc_real)(c->pvx[p] + rx * ab);
    c->pvy[p] = (oc_real)(c->pvy[p] + ry * ab);
    c->pvz[p] = (oc_real)(c->pvz[p] + rz * ab);
  }
  free(cnt);
  free(cell);
}

static double ev_gauss(oc_ctx *c) {
  /* _evGauss, solver.js:3120-3125 */
  double u1;
  do { c->evS = lcg(c->evS); u1 = c->evS / 4294967296.0; } while (u1 < 1e-9);
  c->evS = lcg(c->evS);
  return sqrt(-2 * log(u1)) * cos(6.283185306 * (c->evS / 4294967296.0));
}

/* _updateEvaporation / _evapChunk, solver.js:3008-3116 (sphere world) */
void oc_update_evaporation(oc_ctx *c, double dt) {
  double rate = c->sunActivity * c->evapIntensity;
  c->evapCount = 0;
  if (c->sunActivity <= 0 || c->evapIntensity <= 0 || rate <= 0) return;
  double freezeT = c->meltT > c->snowT * 1.05 ? c->meltT : c->snowT * 1.05;
  int nx = c->nx, ny = c->ny, nz = c->nz, n = 0;
  double dx = c->dx, cx = c->cx, cy = c->cy, cz = c->cz;
  for (int p = 0; p < c->nP; p++) {
    if (c->pflag[p] != 0) continue;
    double T = c->pT[p];
    if (T <= freezeT) continue;
    int i = (int)(c->px[p] / dx), j = (int)(c->py[p] / dx), k = (int)(c->pz[p] / dx);
    i = clampi(i, 1, nx - 2); j = clampi(j, 1, ny - 2); k = clampi(k, 1, nz - 2);
    long cc = ((long)k * ny + j) * nx + i;
    if (c->cellType[cc] != FLUID) continue;
    double ex = c->px[p] - cx, ey = c->py[p] - cy, ez = c->pz[p] - cz;
    double el = sqrt(ex * ex + ey * ey + ez * ez);
    if (!(el != 0)) el = 1e-9;
    int ui = (int)((c->px[p] + ex / el * dx) / dx), uj = (int)((c->py[p] + ey / el * dx) / dx);
    int uk = (int)((c->pz[p] + ez / el * dx) / dx);
    if (ui < 0 || uj < 0 || uk < 0 || ui >= nx || uj >= ny || uk >= nz) continue;
    if (c->cellType[((long)uk * ny + uj) * nx + ui] == FLUID) continue;
    double tau = (T - freezeT) / (1.15 - freezeT);
    if (tau < 0) tau = 0; else if (tau > 1) tau = 1;
    double prob = rate * dt * exp(3.5 * (T - freezeT));
    if (prob > 1) prob = 1;
    c->evS = lcg(c->evS);
    double rnd = c->evS / 4294967296.0;
    if (rnd > prob) continue;
    c->pflag[p] = 2;
    c->pAir[p] = 0;
    double sig = 0.30 + 0.32 * sqrt(tau);
    double kx = sig * ev_gauss(c), ky = sig * ev_gauss(c), kz = sig * ev_gauss(c);
    double ex2 = c->px[p] - cx, ey2 = c->py[p] - cy, ez2 = c->pz[p] - cz;
    double el2 = ex2 * ex2 + ey2 * ey2 + ez2 * ez2;
    double elL = sqrt(el2);
    if (!(elL != 0)) elL = 1e-9;
    double kUp = (kx * ex2 + ky * ey2 + kz * ez2) / el2;
    if (kUp < 0) { kx -= 2 * kUp * ex2 / elL; ky -= 2 * kUp * ey2 / elL; kz -= 2 * kUp * ez2 / elL; }
    double guS = c->u[((long)k * ny + j) * (nx + 1) + i];
    double gvS = c->v[((long)k * (ny + 1) + j) * nx + i];
    double gwS = c->w[((long)k * ny + j) * nx + i];
    double nl2 = 1 / el2, vN = ((double)c->pvx[p] * ex2 + (double)c->pvy[p] * ey2 + (double)c->pvz[p] * ez2) * nl2;
    double vNx = vN * ex2, vNy = vN * ey2, vNz = vN * ez2;
    double gN = (guS * ex2 + gvS * ey2 + gwS * ez2) * nl2;
    double vtx = guS - gN * ex2, vty = gvS - gN * ey2, vtz = gwS - gN * ez2;
    c->pvx[p] = (oc_real)(vNx + vtx + kx);
    c->pvy[p] = (oc_real)(vNy + vty + ky);
    c->pvz[p] = (oc_real)(vNz + vtz + kz);
    n++;
  }
  c->evapCount = n;
}

/* _updateCurrents, solver.js:3634-3782 */
void oc_update_currents(oc_ctx *c, double dt) {
  double strength = c->currents, shell = c->oceanR - c->coreR;
  if (!(strength > 0) || !(shell > 0)) return;
  double seed = 1234567;
  double K[3] = {1.5, 4.0, 9.0}, amp3[3] = {0.10, 0.085, 0.05}, drift[3] = {0.021, 0.05, 0.11};
  double wave[3][3][5];
  for (int o = 0; o < 3; o++)
    for (int cmp = 0; cmp < 3; cmp++) {
      double r3[3];
      for (int q = 0; q < 3; q++) { seed = fmod(seed * 1664525 + 1013904223, 4294967296.0); r3[q] = seed / 4294967296.0 * 2 - 1; }
      double L = sqrt(r3[0] * r3[0] + r3[1] * r3[1] + r3[2] * r3[2]);
      if (!(L != 0)) L = 1;
      wave[o][cmp][0] = r3[0] / L * K[o]; wave[o][cmp][1] = r3[1] / L * K[o]; wave[o][cmp][2] = r3[2] / L * K[o];
      wave[o][cmp][3] = amp3[o] / K[o];
      seed = fmod(seed * 1664525 + 1013904223, 4294967296.0);
      wave[o][cmp][4] = seed / 4294967296.0 * 6.283;
    }
  int stride = 4, gnx = c->nx / stride + 2, gny = c->ny / stride + 2, gnz = c->nz / stride + 2;
  long gN = (long)gnx * gny * gnz;
  float *cf = (float *)malloc(sizeof(float) * 3 * gN);
  double sx = c->dx * stride, t = c->simTime, amp = strength * 2;
  for (int kk = 0; kk < gnz; kk++)
    for (int jj = 0; jj < gny; jj++)
      for (int ii = 0; ii < gnx; ii++) {
        double wx2 = ii * sx, wy2 = jj * sx, wz2 = kk * sx, ux = 0, uy = 0, uz = 0;
        for (int o = 0; o < 3; o++) {
          double dph = drift[o] * t, th[3];
          for (int q = 0; q < 3; q++) th[q] = wave[o][q][0] * wx2 + wave[o][q][1] * wy2 + wave[o][q][2] * wz2 + dph + wave[o][q][4];
          ux += wave[o][2][3] * wave[o][2][1] * cos(th[2]) - wave[o][1][3] * wave[o][1][2] * cos(th[1]);
          uy += wave[o][0][3] * wave[o][0][2] * cos(th[0]) - wave[o][2][3] * wave[o][2][0] * cos(th[2]);
          uz += wave[o][1][3] * wave[o][1][0] * cos(th[1]) - wave[o][0][3] * wave[o][0][1] * cos(th[0]);
        }
        long gi = ((long)kk * gny + jj) * gnx + ii;
        cf[gi] = (float)(ux * amp); cf[gN + gi] = (float)(uy * amp); cf[2 * gN + gi] = (float)(uz * amp);
      }
  double relax = 2.2 * dt;
  if (relax > 0.35) relax = 0.35;
  double invS = 1 / (c->dx * stride);
  long gxy = (long)gnx * gny;
  int kMax = c->nz / stride;
  for (int p = 0; p < c->nP; p++) {
    if (c->pflag[p] != 0) continue;
    double rx = c->px[p] - c->cx, ry = c->py[p] - c->cy, rz = c->pz[p] - c->cz;
    double rl = sqrt(rx * rx + ry * ry + rz * rz);
    if (!(rl != 0)) rl = 1e-9;
    double dpt = (c->oceanR - rl) / shell, wgt = (dpt - 0.06) * 4.0;
    if (wgt <= 0) continue;
    if (wgt > 1) wgt = 1;
    double qa = c->px[p] * invS, qb = c->py[p] * invS, qc = c->pz[p] * invS;
    int i0 = (int)qa, j0 = (int)qb, k0 = (int)qc;
    if (i0 > gnx - 2) i0 = gnx - 2; else if (i0 < 0) i0 = 0;
    if (j0 > gny - 2) j0 = gny - 2; else if (j0 < 0) j0 = 0;
    if (k0 > kMax) k0 = kMax; else if (k0 < 0) k0 = 0;
    double ta = clampd(qa - i0, 0, 1), tb = clampd(qb - j0, 0, 1), tc = clampd(qc - k0, 0, 1);
    long g0 = ((long)k0 * gny + j0) * gnx + i0;
    double a0 = 1 - ta, a1 = ta, b0 = 1 - tb, b1 = tb, c0 = 1 - tc, c1 = tc, U[3];
    for (int d = 0; d < 3; d++) {
      const float *f = cf + d * gN;
      U[d] = (f[g0] * a0 + f[g0 + 1] * a1) * b0 * c0 + (f[g0 + gnx] * a0 + f[g0 + gnx + 1] * a1) * b1 * c0 +
             (f[g0 + gxy] * a0 + f[g0 + gxy + 1] * a1) * b0 * c1 +
             (f[g0 + gxy + gnx] * a0 + f[g0 + gxy + gnx + 1] * a1) * b1 * c1;
    }
    double ir = 1 / rl, Ur = (U[0] * rx + U[1] * ry + U[2] * rz) * ir;
    double Tx = (U[0] - rx * ir * Ur) * wgt, Ty = (U[1] - ry * ir * Ur) * wgt, Tz = (U[2] - rz * ir * Ur) * wgt;
    double vr = ((double)c->pvx[p] * rx + (double)c->pvy[p] * ry + (double)c->pvz[p] * rz) * ir;
    double vx = c->pvx[p] - rx * ir * vr, vy = c->pvy[p] - ry * ir * vr, vz = c->pvz[p] - rz * ir * vr;
    double ex = relax * (Tx - vx), ey = relax * (Ty - vy), ez = relax * (Tz - vz);
    double e2 = ex * ex + ey * ey + ez * ez;
    if (e2 > 0.15 * 0.15) { double es = 0.15 / sqrt(e2); ex *= es; ey *= es; ez *= es; }
    c->pvx[p] = (oc_real)(c->pvx[p] + ex);
    c->pvy[p] = (oc_real)(c->pvy[p] + ey);
    c->pvz[p] = (oc_real)(c->pvz[p] + ez);
  }
  free(cf);
}

/* (a * b) >>> 0 with JavaScript's double product */
static uint32_t js_mul(uint32_t a, uint32_t b) {
  double pr = (double)a * (double)b;
  return (uint32_t)fmod(pr, 4294967296.0);
}

static uint32_t spin_hash(uint32_t a, uint32_t b, uint32_t cc) {
  /* _spinHash, solver.js:3145-3150 */
  uint32_t s = a ^ js_mul(b, 0x9e3779b9u) ^ js_mul(cc, 0x85ebca6bu) ^ js_mul(0x51ab3c77u, 0xc2b2ae35u);
  s = (uint32_t)fmod((double)s * 1664525 + 1013904223, 4294967296.0);
  s ^= s >> 13;
  s = js_mul(s, 0x5bd1e995u);
  s ^= s >> 15;
  return s;
}

/* _updatePhaseChanges, solver.js:3152-3281 (sphere world) */
void oc_update_phase(oc_ctx *c, double dt) {
  int nP = c->nP, nx = c->nx, ny = c->ny, nz = c->nz;
  double dx = c->dx;
  int *cnt = (int *)calloc(c->nC, sizeof(int));
  for (int p = 0; p < nP; p++) {
    if (c->pflag[p] != 2) continue;
    int i = trunc_clamp(c->px[p] / dx, nx), j = trunc_clamp(c->py[p] / dx, ny), k = trunc_clamp(c->pz[p] / dx, nz);
    cnt[((long)k * ny + j) * nx + i]++;
  }
  double cloudT = c->cloudT, cloudP = c->cloudP;
  double rainT = c->rainT < cloudT - 0.02 ? c->rainT : cloudT - 0.02;
  if (rainT < 0) rainT = 0;
  double snowT = c->snowT < rainT - 0.02 ? c->snowT : rainT - 0.02;
  if (snowT < 0) snowT = 0;
  double meltT = c->meltT > snowT * 1.05 ? c->meltT : snowT * 1.05;
  const uint8_t *type = c->cellType;
  double cloudBase = c->oceanR + 0.3 * c->atmosphereH;
  long nxy = (long)nx * ny;
  for (int p = 0; p < nP; p++) {
    int f = c->pflag[p];
    int i = trunc_clamp(c->px[p] / dx, nx), j = trunc_clamp(c->py[p] / dx, ny), k = trunc_clamp(c->pz[p] / dx, nz);
    long cB = ((long)k * ny + j) * nx + i;
    if (f == 0) {
      if (c->pT[p] <= snowT && c->pT[p] > 1e-3) {
        int atSurf = (k > 0 && type[cB - nxy] != FLUID) || (k < nz - 1 && type[cB + nxy] != FLUID) ||
                     (j > 0 && type[cB - nx] != FLUID) || (j < ny - 1 && type[cB + nx] != FLUID) ||
                     (i > 0 && type[cB - 1] != FLUID) || (i < nx - 1 && type[cB + 1] != FLUID);
        if (atSurf) { c->pflag[p] = 5; c->pvx[p] = c->pvy[p] = c->pvz[p] = 0; c->pAir[p] = 1; }
      }
      continue;
    }
    if (f == 5) {
      if (c->pT[p] >= meltT) {
        c->pflag[p] = type[cB] == FLUID ? 0 : 4;
        c->pAir[p] = 0;
        if (c->pWx) c->pWx[p] = c->pWy[p] = c->pWz[p] = 0;
      }
      continue;
    }
    double T = c->pT[p];
    if (T <= 1e-3) continue;
    if (f == 2) {
      if (T <= rainT && c->pAir[p] >= c->rainLiftAge) {
        c->pflag[p] = 4; c->pAir[p] = 0;
        c->pvx[p] = c->pvy[p] = c->pvz[p] = 0;
        if (c->pWx) c->pWx[p] = c->pWy[p] = c->pWz[p] = 0;
        continue;
      }
      if (T > cloudT) continue;
      double rx = c->px[p] - c->cx, ry = c->py[p] - c->cy, rz = c->pz[p] - c->cz;
      if (sqrt(rx * rx + ry * ry + rz * rz) < cloudBase) continue;
      int ii0 = i > 0 ? i - 1 : 0, ii1 = i < nx - 1 ? i + 1 : nx - 1;
      int jj0 = j > 0 ? j - 1 : 0, jj1 = j < ny - 1 ? j + 1 : ny - 1;
      int kk0 = k > 0 ? k - 1 : 0, kk1 = k < nz - 1 ? k + 1 : nz - 1;
      double sum = 0;
      for (int kk = kk0; kk <= kk1; kk++)
        for (int jj = jj0; jj <= jj1; jj++)
          for (int ii = ii0; ii <= ii1; ii++) sum += cnt[((long)kk * ny + jj) * nx + ii];
      if (sum / 27 >= cloudP) { c->pflag[p] = 3; c->pAir[p] = 0; }
      continue;
    }
    if (f == 3) {
      if (T >= cloudT + 0.05) c->pflag[p] = 2;
      continue;
    }
    if (T <= snowT) {
      c->pflag[p] = 5; c->pvx[p] = c->pvy[p] = c->pvz[p] = 0; c->pAir[p] = 0;
      if (c->pWx) c->pWx[p] = c->pWy[p] = c->pWz[p] = 0;
    }
  }
  if (c->spinOn && c->pWx && dt > 0) {
    uint32_t tick = (uint32_t)(int)(c->simTime * 60), sd = 0x51ab3c77u;
    double damp = 1 - 0.12 * dt;
    if (damp < 0) damp = 0;
    for (int p = 0; p < nP; p++) {
      int f = c->pflag[p];
      if (f != 2 && f != 3) continue;
      if (f == 2 && c->pAir[p] <= 1e-6 && c->pWx[p] == 0 && c->pWy[p] == 0 && c->pWz[p] == 0) {
        double h1 = spin_hash(sd, (uint32_t)p, tick) / 4294967296.0;
        double h2 = spin_hash(sd + 1, (uint32_t)p, tick) / 4294967296.0;
        double h3 = spin_hash(sd + 2, (uint32_t)p, tick) / 4294967296.0;
        double zz = 2 * h2 - 1, rr = sqrt(1 - zz * zz > 0 ? 1 - zz * zz : 0), th = 6.283185306 * h1;
        double Tp = c->pT[p];
        if (!(Tp > 0.05)) Tp = 0.05; else if (Tp > 1.15) Tp = 1.15;
        double wSpd = 3.0 * sqrt(Tp) * (0.5 + h3);
        c->pWx[p] = (oc_real)(rr * cos(th) * wSpd); c->pWy[p] = (oc_real)(zz * wSpd); c->pWz[p] = (oc_real)(rr * sin(th) * wSpd);
      }
      double wm = sqrt((double)c->pWx[p] * c->pWx[p] + (double)c->pWy[p] * c->pWy[p] + (double)c->pWz[p] * c->pWz[p]);
      if (wm > 1e-6 && c->pPh) c->pPh[p] = (oc_real)(c->pPh[p] + wm * dt);
      c->pWx[p] = (oc_real)(c->pWx[p] * damp); c->pWy[p] = (oc_real)(c->pWy[p] * damp); c->pWz[p] = (oc_real)(c->pWz[p] * damp);
    }
  }
  free(cnt);
}

/* ------------------------------------------------------------- surface */
/* splatDensity / _splatInto, solver.js:3785-3820 */
void oc_splat_density(oc_ctx *c) {
  int nx = c->nx, ny = c->ny, nz = c->nz;
  double dx = c->dx, r = 1.6 * c->spacing, r2 = r * r;
  long sj = nx + 1, sk = (long)(ny + 1) * sj, nD = sk * (nz + 1);
  for (long i = 0; i < nD; i++) c->dens[i] = 0;
  for (int p = 0; p < c->nP; p++) {
    if (c->pflag[p] != 0) continue;
    double X = c->px[p], Y = c->py[p], Z = c->pz[p];
    int i0 = clampi((int)floor(X / dx), 0, nx - 1);
    int j0 = clampi((int)floor(Y / dx), 0, ny - 1);
    int k0 = clampi((int)floor(Z / dx), 0, nz - 1);
    double fx = X - i0 * dx, fy = Y - j0 * dx, fz = Z - k0 * dx;
    for (int kk = 0; kk < 2; kk++) {
      double qz = kk == 0 ? fz : fz - dx;
      for (int jj = 0; jj < 2; jj++) {
        double qy = jj == 0 ? fy : fy - dx;
        for (int ii = 0; ii < 2; ii++) {
          double qx = ii == 0 ? fx : fx - dx;
          double d2 = qx * qx + qy * qy + qz * qz;
          if (d2 >= r2) continue;
          double ww = 1 - d2 / r2;
          ww = ww * ww;
          long o = (k0 + kk) * sk + (j0 + jj) * sj + (i0 + ii);
          c->dens[o] = (oc_real)(c->dens[o] + ww);
        }
      }
    }
  }
}

/* Marching tetrahedra, surface.js:17-46 (tables) and :66-177 (build). */
static const int MT_CORNER[8][3] = { {0,0,0},{1,0,0},{1,1,0},{0,1,0},{0,0,1},{1,0,1},{1,1,1},{0,1,1} };
static const int MT_TET[6][4] = { {0,5,1,6},{0,1,2,6},{0,2,3,6},{0,3,7,6},{0,7,4,6},{0,4,5,6} };
static const int MT_EDGE[6][2] = { {0,1},{1,2},{2,0},{0,3},{1,3},{2,3} };
/* per case (bit c = tet corner c above iso): triangle count and edge triples */
static const int MT_NTRI[16] = { 0,1,1,2,1,2,2,1,1,2,2,1,2,1,1,0 };
static const int MT_TRI[16][2][3] = {
  {{0,0,0},{0,0,0}}, {{0,2,3},{0,0,0}}, {{0,1,4},{0,0,0}}, {{2,3,4},{2,4,1}},
  {{2,1,5},{0,0,0}}, {{0,3,5},{0,5,1}}, {{0,4,5},{0,5,2}}, {{3,5,4},{0,0,0}},
  {{3,4,5},{0,0,0}}, {{0,2,5},{0,5,4}}, {{0,1,5},{0,5,3}}, {{2,1,5},{0,0,0}},
  {{2,1,4},{2,4,3}}, {{0,1,4},{0,0,0}}, {{0,2,3},{0,0,0}}, {{0,0,0},{0,0,0}} };

long oc_mt_build(const oc_real *dens, int nx, int ny, int nz, double dx, double iso,
                 float *pos, float *nrm, long cap, double *volume) {
  long sj = nx + 1, sk = (long)(ny + 1) * sj;
  long vcount = 0;
  double vol = 0, cv[8], gx[8], gy[8], gz[8];
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++) {
        double mn = INFINITY, mxv = -INFINITY;
        for (int cc = 0; cc < 8; cc++) {
          double val = dens[(k + MT_CORNER[cc][2]) * sk + (j + MT_CORNER[cc][1]) * sj + i + MT_CORNER[cc][0]];
          cv[cc] = val;
          if (val < mn) mn = val;
          if (val > mxv) mxv = val;
        }
        if (mxv < iso || mn > iso) continue;
        for (int cc = 0; cc < 8; cc++) {
          int ci = i + MT_CORNER[cc][0], cj = j + MT_CORNER[cc][1], ck = k + MT_CORNER[cc][2];
          int ip = ci + 1 < nx ? ci + 1 : nx, im = ci - 1 > 0 ? ci - 1 : 0;
          int jp = cj + 1 < ny ? cj + 1 : ny, jm = cj - 1 > 0 ? cj - 1 : 0;
          int kp = ck + 1 < nz ? ck + 1 : nz, km = ck - 1 > 0 ? ck - 1 : 0;
          gx[cc] = ((double)dens[ck * sk + cj * sj + ip] - dens[ck * sk + cj * sj + im]) / (2 * dx);
          gy[cc] = ((double)dens[ck * sk + jp * sj + ci] - dens[ck * sk + jm * sj + ci]) / (2 * dx);
          gz[cc] = ((double)dens[kp * sk + cj * sj + ci] - dens[km * sk + cj * sj + ci]) / (2 * dx);
        }
        for (int tt = 0; tt < 6; tt++) {
          const int *tet = MT_TET[tt];
          int code = 0;
          for (int cc = 0; cc < 4; cc++) if (cv[tet[cc]] > iso) code |= 1 << cc;
          for (int e = 0; e < MT_NTRI[code]; e++) {
            double P[3][3], Nn[3][3];
            for (int vi = 0; vi < 3; vi++) {
              const int *ed = MT_EDGE[MT_TRI[code][e][vi]];
              int a = tet[ed[0]], b = tet[ed[1]];
              double va = cv[a], vb = cv[b], den = vb - va;
              double tl = den != 0 ? (iso - va) / den : 0.5;
              if (tl < 0) tl = 0; else if (tl > 1) tl = 1;
              const int *pa = MT_CORNER[a], *pb = MT_CORNER[b];
              P[vi][0] = ((i + pa[0]) + (pb[0] - pa[0]) * tl) * dx;
              P[vi][1] = ((j + pa[1]) + (pb[1] - pa[1]) * tl) * dx;
              P[vi][2] = ((k + pa[2]) + (pb[2] - pa[2]) * tl) * dx;
              double nxx = -(gx[a] + (gx[b] - gx[a]) * tl);
              double nyy = -(gy[a] + (gy[b] - gy[a]) * tl);
              double nzz = -(gz[a] + (gz[b] - gz[a]) * tl);
              double nl = sqrt(nxx * nxx + nyy * nyy + nzz * nzz);
              if (!(nl != 0)) nl = 1;
              Nn[vi][0] = nxx / nl; Nn[vi][1] = nyy / nl; Nn[vi][2] = nzz / nl;
            }
            double ux = P[1][0] - P[0][0], uy = P[1][1] - P[0][1], uz = P[1][2] - P[0][2];
            double wx = P[2][0] - P[0][0], wy = P[2][1] - P[0][1], wz = P[2][2] - P[0][2];
            double fnx = uy * wz - uz * wy, fny = uz * wx - ux * wz, fnz = ux * wy - uy * wx;
            double cgx = 0, cgy = 0, cgz = 0;
            for (int vi = 0; vi < 3; vi++) { cgx += Nn[vi][0]; cgy += Nn[vi][1]; cgz += Nn[vi][2]; }
            if (fnx * cgx + fny * cgy + fnz * cgz < 0) {
              for (int d = 0; d < 3; d++) {
                double tp = P[2][d]; P[2][d] = P[1][d]; P[1][d] = tp;
                tp = Nn[2][d]; Nn[2][d] = Nn[1][d]; Nn[1][d] = tp;
              }
            }
            vol += (P[0][0] * (P[1][1] * P[2][2] - P[2][1] * P[1][2]) -
                    P[0][1] * (P[1][0] * P[2][2] - P[2][0] * P[1][2]) +
                    P[0][2] * (P[1][0] * P[2][1] - P[2][0] * P[1][1])) / 6;
            if (pos && vcount + 3 <= cap) {
              for (int vi = 0; vi < 3; vi++)
                for (int d = 0; d < 3; d++) {
                  pos[(vcount + vi) * 3 + d] = (float)P[vi][d];
                  nrm[(vcount + vi) * 3 + d] = (float)Nn[vi][d];
                }
            }
            vcount += 3;
          }
        }
      }
  if (volume) *volume = fabs(vol);
  return vcount;
}

/* Math.round: nearest integer, halves toward +infinity (x - floor(x) is
 * exact, unlike floor(x + 0.5)) */
static double js_round(double x) {
  double r = floor(x);
  return x - r >= 0.5 ? r + 1 : r;
}

/* ---------------------------------------------------------- metaballs
 * Metaballs.build, surface.js:178-284: the kernel radius in lattice cells,
 * the contour level and the field lattice of a tension / detail setting,
 * then the summed compact kernels of the liquid particles. */
void oc_metaball_params(int nx, double dx, double spacing, double tension, double detail, int *m_out,
                        double *mdx_out, double *R_out, int *rad_out, double *iso_out) {
  const double R_MIN = 1.6, R_SPAN = 2.2, ISO0 = 1.5, ALPHA = 1.5, D_MIN = 0.5, D_MAX = 1.6;
  double f = 1;
  if (isfinite(detail) && detail > 0) f = detail < D_MIN ? D_MIN : (detail > D_MAX ? D_MAX : detail);
  /* Math.max(6, Math.round(nx * f)); Math.round rounds halves up */
  int m = (int)js_round(nx * f);
  if (m < 6) m = 6;
  double W = nx * dx;
  double mdx = m == nx ? dx : W / m;
  double t = isfinite(tension) ? tension : 0.5;
  t = t < 0 ? 0 : (t > 1 ? 1 : t);
  double cell = mdx > 0 ? mdx : dx;
  double R = (R_MIN + R_SPAN * t) * spacing / cell;
  double rc = ceil(R);
  *m_out = m;
  *mdx_out = mdx;
  *R_out = R;
  *rad_out = (int)(3 < rc ? 3 : rc);
  *iso_out = ISO0 * pow((R_MIN + R_SPAN * t) / R_MIN, ALPHA);
}

/* the field, surface.js:256-281: every liquid particle adds (1 - d^2/R^2)^2
 * to the lattice corners within `rad` of its nearest corner and within R,
 * in particle order, each sum rounded to the field type */
void oc_metaball_field(const oc_real *px, const oc_real *py, const oc_real *pz, const uint8_t *pflag, int nP, int m,
                       double mdx, double R, int rad, oc_real *field) {
  long sj = m + 1, sk = (long)(m + 1) * (m + 1), len = sk * (m + 1);
  for (long i = 0; i < len; i++) field[i] = 0;
  double R2 = R * R;
  for (int p = 0; p < nP; p++) {
    if (pflag[p] != 0) continue;
    double gx = px[p] / mdx, gy = py[p] / mdx, gz = pz[p] / mdx;
    int i0 = (int)js_round(gx), j0 = (int)js_round(gy), k0 = (int)js_round(gz);
    int ia = i0 - rad > 0 ? i0 - rad : 0, ib = i0 + rad < m ? i0 + rad : m;
    int ja = j0 - rad > 0 ? j0 - rad : 0, jb = j0 + rad < m ? j0 + rad : m;
    int ka = k0 - rad > 0 ? k0 - rad : 0, kb = k0 + rad < m ? k0 + rad : m;
    for (int k = ka; k <= kb; k++) {
      double ez = k - gz, dz2 = ez * ez;
      if (dz2 >= R2) continue;
      long row = k * sk;
      for (int j = ja; j <= jb; j++) {
        double ey = j - gy, dy2 = dz2 + ey * ey;
        if (dy2 >= R2) continue;
        long base = row + j * sj;
        for (int i = ia; i <= ib; i++) {
          double ex = i - gx;
          double d2 = dy2 + ex * ex;
          if (d2 >= R2) continue;
          double u = 1 - d2 / R2;
          field[base + i] = (oc_real)(field[base + i] + u * u);
        }
      }
    }
  }
}

/* autotuneIso, solver.js:3838-3849 */
double oc_autotune_iso(oc_ctx *c) {
  double Vt = c->nP * c->particleVolume, lo = 0.3, hi = 3.0;
  for (int it = 0; it < 8; it++) {
    double mid = 0.5 * (lo + hi), vol;
    long cnt = oc_mt_build(c->dens, c->nx, c->ny, c->nz, c->dx, mid, 0, 0, 0, &vol);
    if (cnt > 900000 || vol > Vt) lo = mid; else hi = mid;
  }
  return 0.5 * (lo + hi);
}

/* kineticEnergy, solver.js:3914-3921 */
double oc_kinetic_energy(oc_ctx *c) {
  double e = 0, mp = 1000 * c->particleVolume;
  for (int p = 0; p < c->nP; p++) {
    double v2 = (double)c->pvx[p] * c->pvx[p] + (double)c->pvy[p] * c->pvy[p] + (double)c->pvz[p] * c->pvz[p];
    e += 0.5 * mp * v2;
  }
  return e;
}

/* _pushSurfaceChunk, solver.js:3496-3587 (sphere world): the frame-end
 * guarantee that no particle sits in rasterized rock or under the terrain
 * surface, that airborne water stays under the atmosphere cap and fluid
 * inside the domain. Per particle; walks in double like the reference and
 * stores float. */
void oc_push_surface(oc_ctx *c) {
  if (!c->tsolid) return;
  double cx = c->cx, cy = c->cy, cz = c->cz, dx = c->dx;
  for (int p = 0; p < c->nP; p++) {
    double x = c->px[p], y = c->py[p], z = c->pz[p];
    double ex = x - cx, ey = y - cy, ez = z - cz;
    double er = sqrt(ex * ex + ey * ey + ez * ez);
    if (!(er != 0)) er = 1e-9;
    int fixed = 0, rockFixed = 0;
    double outer = c->domainR - dx * 0.02;
    double skyHi = c->oceanR + c->atmosphereH - 1e-3 * dx;
    if (er > skyHi) {
      x = cx + ex / er * skyHi; y = cy + ey / er * skyHi; z = cz + ez / er * skyHi;
      double outward = ((double)c->pvx[p] * ex + (double)c->pvy[p] * ey + (double)c->pvz[p] * ez) / er;
      if (outward > 0) {
        c->pvx[p] = (oc_real)(c->pvx[p] - ex / er * outward);
        c->pvy[p] = (oc_real)(c->pvy[p] - ey / er * outward);
        c->pvz[p] = (oc_real)(c->pvz[p] - ez / er * outward);
      }
      c->px[p] = (oc_real)x; c->py[p] = (oc_real)y; c->pz[p] = (oc_real)z;
      ex = x - cx; ey = y - cy; ez = z - cz; er = skyHi;
      fixed = 1;
    }
    if (er <= c->domainR && oc_rock_cell_at(c, x, y, z)) {
      rockFixed = 1;
      int tries = 0;
      while (oc_rock_cell_at(c, x, y, z) && tries < c->nx * 2) {
        x += ex / er * dx * 0.5; y += ey / er * dx * 0.5; z += ez / er * dx * 0.5;
        tries++;
      }
      double exR = x - cx, eyR = y - cy, ezR = z - cz;
      double erR = sqrt(exR * exR + eyR * eyR + ezR * ezR);
      if (!(erR != 0)) erR = 1e-9;
      double wantR = oc_terrain_radius_at(c, x, y, z) + dx * 0.06;
      if (erR > wantR) {
        x = cx + exR / erR * wantR; y = cy + eyR / erR * wantR; z = cz + ezR / erR * wantR;
      }
      int kR = 0;
      while (oc_rock_cell_at(c, x, y, z) && kR < 4) {
        x += exR / erR * dx * 0.5; y += eyR / erR * dx * 0.5; z += ezR / erR * dx * 0.5;
        kR++;
      }
      fixed = 1;
    } else if (c->pflag[p] == 0 && er <= c->domainR) {
      double Rt = oc_terrain_radius_at(c, x, y, z);
      if (er < Rt - dx * 0.5) {
        rockFixed = 1;
        int t2 = 0;
        x = cx + ex / er * (Rt + dx * 0.05);
        y = cy + ey / er * (Rt + dx * 0.05);
        z = cz + ez / er * (Rt + dx * 0.05);
        while (oc_rock_cell_at(c, x, y, z) && t2 < 8) {
          x += ex / er * dx * 0.5; y += ey / er * dx * 0.5; z += ez / er * dx * 0.5;
          t2++;
        }
        fixed = 1;
      }
    }
    double finalR = sqrt((x - cx) * (x - cx) + (y - cy) * (y - cy) + (z - cz) * (z - cz));
    double capR = c->pflag[p] == 0 ? outer : skyHi;
    if (finalR > capR) {
      x = cx + (x - cx) * capR / finalR; y = cy + (y - cy) * capR / finalR; z = cz + (z - cz) * capR / finalR;
      fixed = 1;
    }
    if (fixed) {
      c->px[p] = (oc_real)x; c->py[p] = (oc_real)y; c->pz[p] = (oc_real)z;
      double vr = ((double)c->pvx[p] * ex + (double)c->pvy[p] * ey + (double)c->pvz[p] * ez) / er;
      if (rockFixed && vr < 0) {
        c->pvx[p] = (oc_real)(c->pvx[p] - ex / er * vr);
        c->pvy[p] = (oc_real)(c->pvy[p] - ey / er * vr);
        c->pvz[p] = (oc_real)(c->pvz[p] - ez / er * vr);
      }
    }
  }
}

/* ordering of the x-sorted sweeps: px, then particle index */
static const oc_real *g_sort_px;
static int by_px(const void *a, const void *b) {
  int i = *(const int *)a, j = *(const int *)b;
  double d = (double)g_sort_px[i] - (double)g_sort_px[j];
  if (d < 0) return -1;
  if (d > 0) return 1;
  return i < j ? -1 : (i > j);
}

/* _vaporCollisions, solver.js:3345-3479: steam (2) and cloud (3) parcels
 * sorted by x; every overlapping pair (in sorted order, with the early
 * break on the x gap) exchanges the normal velocity if approaching, is
 * separated and pushed apart, and with spin on trades linear momentum for
 * rotation; then every liquid particle (surface band, spray, rain) pushes
 * the parcels it touches out of the water. Serial, in place, as written. */
void oc_vapor_collisions(oc_ctx *c) {
  int nP = c->nP, nV = 0, nSteam = 0, night = 0;
  int *v = (int *)malloc(sizeof(int) * (nP > 0 ? nP : 1));
  double sx = c->hasSun ? c->sunX - c->cx : 0, sy = c->hasSun ? c->sunY - c->cy : 0;
  double sz = c->hasSun ? c->sunZ - c->cz : 0;
  for (int p = 0; p < nP; p++) {
    int f = c->pflag[p];
    if (f != 2 && f != 3) continue;
    v[nV++] = p;
    if (f == 2) {
      nSteam++;
      if (!c->hasSun || ((double)c->px[p] - c->cx) * sx + ((double)c->py[p] - c->cy) * sy +
                                ((double)c->pz[p] - c->cz) * sz < 0)
        night++;
    }
  }
  c->vaporCount = nSteam;
  c->nightVaporCount = night;
  if (nV < 2) { free(v); return; }
  oc_real *px = c->px, *py = c->py, *pz = c->pz, *vx = c->pvx, *vy = c->pvy, *vz = c->pvz;
  oc_real *wx = c->pWx, *wy = c->pWy, *wz = c->pWz;
  int spin = c->spinOn && wx != 0;
  double rad = c->spacing * 0.6, rad2 = rad * rad;
  g_sort_px = px;
  qsort(v, nV, sizeof(int), by_px);
  for (int a = 0; a < nV; a++) {
    int i = v[a];
    for (int b = a + 1; b < nV; b++) {
      int j = v[b];
      if ((double)px[j] - px[i] >= rad) break;
      double ddx = (double)px[j] - px[i], ddy = (double)py[j] - py[i], ddz = (double)pz[j] - pz[i];
      double d2 = ddx * ddx + ddy * ddy + ddz * ddz;
      if (d2 >= rad2) continue;
      double d = sqrt(d2);
      if (!(d != 0)) d = 1e-9;
      double nX = ddx / d, nY = ddy / d, nZ = ddz / d;
      double vRel = ((double)vx[i] - vx[j]) * nX + ((double)vy[i] - vy[j]) * nY + ((double)vz[i] - vz[j]) * nZ;
      if (vRel > 0) {
        vx[i] = (oc_real)(vx[i] - vRel * nX); vy[i] = (oc_real)(vy[i] - vRel * nY); vz[i] = (oc_real)(vz[i] - vRel * nZ);
        vx[j] = (oc_real)(vx[j] + vRel * nX); vy[j] = (oc_real)(vy[j] + vRel * nY); vz[j] = (oc_real)(vz[j] + vRel * nZ);
      }
      double half = (rad - d) * 0.5;
      px[i] = (oc_real)(px[i] - nX * half); py[i] = (oc_real)(py[i] - nY * half); pz[i] = (oc_real)(pz[i] - nZ * half);
      px[j] = (oc_real)(px[j] + nX * half); py[j] = (oc_real)(py[j] + nY * half); pz[j] = (oc_real)(pz[j] + nZ * half);
      double kick = (rad - d) * 2;
      vx[i] = (oc_real)(vx[i] - nX * kick); vy[i] = (oc_real)(vy[i] - nY * kick); vz[i] = (oc_real)(vz[i] - nZ * kick);
      vx[j] = (oc_real)(vx[j] + nX * kick); vy[j] = (oc_real)(vy[j] + nY * kick); vz[j] = (oc_real)(vz[j] + nZ * kick);
      if (!spin) continue;
      double R = rad * 0.5;
      double wSx = (double)wx[i] + wx[j], wSy = (double)wy[i] + wy[j], wSz = (double)wz[i] + wz[j];
      double usx = ((double)vx[i] - vx[j]) + R * (wSy * nZ - wSz * nY);
      double usy = ((double)vy[i] - vy[j]) + R * (wSz * nX - wSx * nZ);
      double usz = ((double)vz[i] - vz[j]) + R * (wSx * nY - wSy * nX);
      double uN = usx * nX + usy * nY + usz * nZ;
      double utx = usx - uN * nX, uty = usy - uN * nY, utz = usz - uN * nZ;
      double utM = sqrt(utx * utx + uty * uty + utz * utz);
      if (utM < 1e-9) continue;
      double jt = utM / 7;
      if (vRel > 0) { double mu = 0.25 * vRel; if (jt > mu) jt = mu; }
      double jx = -utx / utM * jt, jy = -uty / utM * jt, jz = -utz / utM * jt;
      vx[i] = (oc_real)(vx[i] + jx); vy[i] = (oc_real)(vy[i] + jy); vz[i] = (oc_real)(vz[i] + jz);
      vx[j] = (oc_real)(vx[j] - jx); vy[j] = (oc_real)(vy[j] - jy); vz[j] = (oc_real)(vz[j] - jz);
      double invI = 1 / (0.4 * R * R);
      double dwx = R * (nY * jz - nZ * jy) * invI, dwy = R * (nZ * jx - nX * jz) * invI;
      double dwz = R * (nX * jy - nY * jx) * invI;
      wx[i] = (oc_real)(wx[i] + dwx); wy[i] = (oc_real)(wy[i] + dwy); wz[i] = (oc_real)(wz[i] + dwz);
      wx[j] = (oc_real)(wx[j] + dwx); wy[j] = (oc_real)(wy[j] + dwy); wz[j] = (oc_real)(wz[j] + dwz);
    }
  }
  /* airborne x liquid contacts */
  double band = c->oceanR - c->dx * 2, bandLo2 = band * band;
  for (int p = 0; p < nP; p++) {
    int fw = c->pflag[p];
    if (fw != 0 && fw != 1 && fw != 4) continue;
    double wxp = px[p], wyp = py[p], wzp = pz[p];
    if (fw == 0) {
      double ex = wxp - c->cx, ey = wyp - c->cy, ez = wzp - c->cz;
      if (ex * ex + ey * ey + ez * ez < bandLo2) continue;
    }
    int lo = 0, hi = nV;
    while (lo < hi) {
      int mid = (lo + hi) >> 1;
      if ((double)px[v[mid]] < wxp - rad) lo = mid + 1; else hi = mid;
    }
    for (int b = lo; b < nV; b++) {
      int j = v[b];
      double ddx = (double)px[j] - wxp;
      if (ddx >= rad) break;
      double ddy = (double)py[j] - wyp, ddz = (double)pz[j] - wzp;
      double dd2 = ddx * ddx + ddy * ddy + ddz * ddz;
      if (dd2 >= rad2 || dd2 < 1e-12) continue;
      double dd = sqrt(dd2);
      double nX = ddx / dd, nY = ddy / dd, nZ = ddz / dd;
      double vR = ((double)vx[j] - vx[p]) * nX + ((double)vy[j] - vy[p]) * nY + ((double)vz[j] - vz[p]) * nZ;
      if (vR < 0) {
        double vB = vR * (1 + 0.4);
        vx[j] = (oc_real)(vx[j] - vB * nX); vy[j] = (oc_real)(vy[j] - vB * nY); vz[j] = (oc_real)(vz[j] - vB * nZ);
      }
      double ov = rad - dd;
      px[j] = (oc_real)(px[j] + nX * ov); py[j] = (oc_real)(py[j] + nY * ov); pz[j] = (oc_real)(pz[j] + nZ * ov);
      double kk = ov * 2;
      vx[j] = (oc_real)(vx[j] + nX * kk); vy[j] = (oc_real)(vy[j] + nY * kk); vz[j] = (oc_real)(vz[j] + nZ * kk);
    }
  }
  free(v);
}

/* _iceContacts, solver.js:3290-3324: snow grains (5) sorted by x; a free
 * grain touching a stuck one (pAir == 1) sticks, two free grains pair up
 * at their mean velocity. Serial, in place. */
void oc_ice_contacts(oc_ctx *c) {
  int nP = c->nP, nI = 0;
  int *v = (int *)malloc(sizeof(int) * (nP > 0 ? nP : 1));
  for (int p = 0; p < nP; p++)
    if (c->pflag[p] == 5) v[nI++] = p;
  if (nI < 2) { free(v); return; }
  oc_real *px = c->px, *py = c->py, *pz = c->pz, *vx = c->pvx, *vy = c->pvy, *vz = c->pvz, *air = c->pAir;
  double rad = c->spacing * 0.7, rad2 = rad * rad;
  g_sort_px = px;
  qsort(v, nI, sizeof(int), by_px);
  for (int a = 0; a < nI; a++) {
    int i = v[a];
    int stuckI = air[i] == 1;
    for (int b = a + 1; b < nI; b++) {
      int j = v[b];
      if ((double)px[j] - px[i] >= rad) break;
      double ddx = (double)px[j] - px[i], ddy = (double)py[j] - py[i], ddz = (double)pz[j] - pz[i];
      if (ddx * ddx + ddy * ddy + ddz * ddz >= rad2) continue;
      int stuckJ = air[j] == 1;
      if (stuckI && stuckJ) continue;
      if (stuckJ) {
        air[i] = 1; vx[i] = vy[i] = vz[i] = 0; stuckI = 1;
      } else if (stuckI) {
        air[j] = 1; vx[j] = vy[j] = vz[j] = 0;
      } else {
        double mx = ((double)vx[i] + vx[j]) * 0.5, my = ((double)vy[i] + vy[j]) * 0.5, mz = ((double)vz[i] + vz[j]) * 0.5;
        vx[i] = vx[j] = (oc_real)mx; vy[i] = vy[j] = (oc_real)my; vz[i] = vz[j] = (oc_real)mz;
      }
    }
  }
  free(v);
}
