/*
 * watersim C ABI: the FLIP substep pipeline and marching-tetrahedra surface of
 * johnnybuggy/watersim-v1 (js/solver.js, js/surface.js) as a HIP library for
 * MI355X (gfx950).
 *
 * The reference has no FFI: its boundary is the FluidSolver JavaScript object
 * (js/solver.js:45-3924). Each entry point below replaces a method or a piece
 * of state of that object; the citation on each names it. A host binding
 * (N-API addon, ctypes, cgo ...) maps the object's constructor, typed-array
 * fields and methods onto these calls.
 *
 * Conventions
 *  - Plain C: opaque context, plain pointers and byte sizes, int status
 *    (0 = ok, negative = error; ws_last_error() gives the message).
 *  - Host buffers belong to the caller; device buffers belong to the library.
 *  - One HIP stream per context. Calls are synchronous unless named *_async;
 *    a context must not be used from two threads at once.
 *  - Arrays keep the reference's layouts (solver.js:93-131): cells
 *    (k*ny+j)*nx+i, u faces (k*ny+j)*(nx+1)+i, v faces (k*(ny+1)+j)*nx+i,
 *    w faces (k*ny+j)*nx+i, corners (k*(ny+1)+j)*(nx+1)+i, particle SoA.
 */
#ifndef WATERSIM_H
#define WATERSIM_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ws_ctx ws_ctx;

/* Error codes */
#define WS_OK 0
#define WS_EINVAL -1   /* bad argument, unknown name or size mismatch */
#define WS_EHIP -2     /* HIP runtime error */
#define WS_ENODEV -3   /* no usable GPU */
#define WS_ESTATE -4   /* call out of order (e.g. no world uploaded) */
#define WS_ENUMERIC -5 /* non-finite values detected */

/* Stage identifiers for ws_stage(): the reference's underscore methods. */
enum ws_stage_id {
  WS_STAGE_CLEAR = 0,      /* u..wW.fill(0), solver.js:2569-2570 */
  WS_STAGE_RASTERIZE = 1,  /* _rasterizeStatic + _markFluidCells, :860-1000 */
  WS_STAGE_P2G = 2,        /* _p2g, :1030-1120 */
  WS_STAGE_BC = 3,         /* _applyBC, :1123-1171 */
  WS_STAGE_SNAPSHOT = 4,   /* uO/vO/wO.set(u/v/w), :2579 */
  WS_STAGE_VISCOSITY = 5,  /* _viscosity, :1259-1304 */
  WS_STAGE_GRAVITY = 6,    /* _applyPlanetGravity + _applyBC, :2656-2662, :1229-1256 */
  WS_STAGE_PRESSURE = 7,   /* _pressureSolve, :1312-1410 */
  WS_STAGE_PROJECT = 8,    /* _project, :1412-1443 */
  WS_STAGE_VORTICITY = 9,  /* _vorticityConfinement, :1454-1559 */
  WS_STAGE_CLAMP = 10,     /* umax = _clampUmSlice, :2629-2653 */
  WS_STAGE_EXTRAPOLATE = 11, /* _extrapolate, :1562-1636 */
  WS_STAGE_G2P = 12,       /* _g2p, :1639-1756 */
  WS_STAGE_ADVECT = 13,    /* _advect, :1759-2292 */
  WS_STAGE_HEAT = 14,      /* _updateHeat(dt), :2715-2994 (dt = tick length) */
  WS_STAGE_EVAPORATION = 15, /* _updateEvaporation(dt), :3008-3125 */
  WS_STAGE_CURRENTS = 16,  /* _updateCurrents(dt), :3658-3782 */
  WS_STAGE_PHASE = 17,     /* _updatePhaseChanges(dtFrame), :3152-3281 */
  WS_STAGE_PUSH_SURFACE = 18, /* _pushSurfaceParticles, :3486-3587 */
  WS_STAGE_VAPOR_COLLISIONS = 19, /* _vaporCollisions, :3345-3479 */
  WS_STAGE_ICE_CONTACTS = 20,  /* _iceContacts, :3290-3324 */
  WS_STAGE_PUSH_BALLS = 21,  /* _pushOutOfBalls, :1005-1026 (then the particle bins are rebuilt) */
  WS_STAGE_UPDATE_BALLS = 22, /* _updateBalls(dt), :2344-2564 */
  WS_STAGE_COUNT = 23
};

/* Most balls a context holds. */
#define WS_MAX_BALLS 64

/* new FluidSolver({nx,ny,nz,dx}) plus _allocParticles(capacity),
 * solver.js:45-231. Device memory for every field is allocated here. */
int ws_create(ws_ctx **out, int nx, int ny, int nz, double dx, int capacity);
/* Releases the context and its device memory. */
void ws_destroy(ws_ctx *c);
/* Message of the last failed call on this context (never NULL). */
const char *ws_last_error(const ws_ctx *c);

/* Scalar state by name: the tunables and world constants of FluidSolver
 * (solver.js:56-75, 81-87, 137-193, 203-210), e.g. "pic", "omega",
 * "pressureIters", "gravity", "maxSpeed", "viscosity", "vorticity", "cfl",
 * "coreR", "oceanR", "spacing", "particleVolume", "atmosphereH", "gRamp",
 * "nP", "dv", "tn", "pressureTol", "maxPressureIters", "iso"; and the
 * kernel choices of this library (no reference counterpart; each gives the
 * same results, except that the network's forms round its prediction
 * differently): "sorMode", "sorThreads", "netTiled" (the warm start's
 * convolutions: 0 untiled, 1 LDS-tiled where the grid has enough tiles to
 * fill the GPU and four waves per 16 voxels elsewhere, 2 LDS-tiled on every
 * grid, for a batch whose planets fill the GPU together), "p2gBin" (0..3: P2G
 * bins of 8x8x8, 4x4x4, 16x8x8, 8x8x4 cells), "xcdGroups" (workgroups of
 * sorMode 6, 0 = one per compute unit of an XCD), "xcdUnits" (its cells per
 * thread and colour, 1..3, 0 = chosen from the list lengths); the counter
 * "xcdFallbacks" (sweep batches that sorMode 6 handed to the list kernel);
 * the warm-start switches
 * "warmStart", "recordSamples", "netCoarse" (include/watersim_learn.h).
 * ws_get_f also reads results: "umax", "dtLast", "residual",
 * "pressureIters_done", "airborneCount", "nFluid", "dfRefined". */
int ws_set_f(ws_ctx *c, const char *name, double v);
int ws_get_f(ws_ctx *c, const char *name, double *v);

/* Typed-array fields by name (solver.js:93-131, 218-231, 499-504, 1181-1183):
 * "px","py","pz","pvx","pvy","pvz","pT","pAir" (f32, nP), "pflag","pcool"
 * (u8, nP), "cellType" (u8, C), "u","v","w","uO","vO","wO","uW","vW","wW"
 * (f32 faces), "validU","validV","validW" (u8 faces), "q","rhs" (f32, C),
 * "dens" (f32, corners), "advU","advV","advW" (f32, nP; the G2P samples),
 * and the static world: "staticTypes" (u8, C), "gu","gv","gw" (f32 faces),
 * "terrainR" (f32, tn^3); with balls also "svx","svy","svz" (f32, C, the
 * solid cell velocities), "ballHist" (u32, 25 per ball: the waterline
 * histogram and its count) and "ballSurf" (f32 per ball, the waterline);
 * after ws_metaballs_build "metaballField" (f32, (m+1)^3). Read-only:
 * "pcell" (i32, nP), the cell each particle was last binned to,
 * (k*ny+j)*nx+i with i = clamp(floor(x/dx), 0, nx-1) evaluated as the
 * reference does in double (solver.js:976-978), by the marking pass (-1 for
 * particles that mark nothing: flags other than 0 and the ones it demotes)
 * or by splatDensity's binning.
 * nbytes must equal the field's byte size (for particle fields: nP
 * elements). */
int ws_upload(ws_ctx *c, const char *name, const void *host, size_t nbytes);
int ws_download(ws_ctx *c, const char *name, void *host, size_t nbytes);

/* One operator-split substep of length dt, FluidSolver.prototype._substep,
 * solver.js:2568-2624 (sphere world). */
int ws_substep(ws_ctx *c, double dt);
/* One stage of the substep, for the reference's stage methods; dt is used by
 * the stages that take it. */
int ws_stage(ws_ctx *c, int stage, double dt);
/* step(dtFrame): _stepSync, solver.js:2663-2700: the 25 Hz thermal,
 * evaporation and currents tick, the CFL substep loop
 * (dt = min(dtFrame/substeps, cfl*dx/max(umax,0.08), 1/100), >= 1e-5,
 * at most 10 substeps), then the frame-end passes: vapor collisions, phase
 * changes, snow contacts, surface rescue. Returns the substep count through
 * *n_sub. */
int ws_step_frame(ws_ctx *c, double dt_frame, int *n_sub);
/* n substeps with the CFL law evaluated on the device (no host round trip
 * between substeps): the measurement path for substeps/s. */
int ws_run_substeps(ws_ctx *c, int n);

/* _markFluidCells(p0, p1), solver.js:945-1000: the marking pass over
 * particles [p0, p1) on the current cell types (no re-rasterization): sky and
 * ceiling demotions, AIR -> FLUID, and each marked particle's cell index
 * ("pcell"). */
int ws_mark_fluid(ws_ctx *c, int p0, int p1);
/* sampleVel(x, y, z), solver.js:811-854: the trilinear MAC sample of the
 * current face velocities at n points (xyz: 3n floats, host), written to out
 * (3n floats, host). The sampler is the one G2P uses, so a sample at a
 * particle's position equals its "advU/advV/advW" entry after G2P. */
int ws_sample_vel(ws_ctx *c, const float *xyz, long n, float *out);
/* splatDensity(), solver.js:3785-3820: corner density of fluid particles. */
int ws_splat_density(ws_ctx *c);
/* MarchingTetrahedra.build(dens,nx,ny,nz,dx,iso), surface.js:66-177, on the
 * context's density field (dens == NULL) or on a caller field of
 * (nx+1)(ny+1)(nz+1) floats. Writes up to cap vertices into pos/nrm
 * (3 floats each, may be NULL), the full vertex count into *count and the
 * enclosed volume (FP64) into *volume. */
int ws_mt_build(ws_ctx *c, const float *dens, int nx, int ny, int nz, double dx, double iso,
                float *pos, float *nrm, long cap, long *count, double *volume);
/* No counterpart in the reference: ws_mt_build for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). The field (dens_dev, (nx+1)(ny+1)(nz+1)
 * floats, or NULL for the context's own density) and the mesh (pos_dev /
 * nrm_dev, 3 floats per vertex, up to cap vertices; NULL for count and volume
 * only) are device buffers of the caller; *count and *volume are host
 * values, as ws_mt_build returns them, and every output equals
 * ws_mt_build's on the same field. All scratch sits in a static device array
 * of the library (about 8.5 MB per GPU, held until the process ends): at most
 * WS_MT_NOALLOC_MAX_CELLS cells, larger fields are refused with WS_EINVAL.
 * The call allocates nothing. Every call on one GPU shares that GPU's array:
 * calls that overlap on one GPU (two threads, two contexts) race. */
#define WS_MT_NOALLOC_MAX_CELLS (1L << 21)
int ws_mt_build_noalloc(ws_ctx *c, const float *dens_dev, int nx, int ny, int nz, double dx, double iso,
                        float *pos_dev, float *nrm_dev, long cap, long *count, double *volume);
/* autotuneIso(), solver.js:3838-3849: 8 bisections of iso on the device. */
int ws_autotune_iso(ws_ctx *c, double *iso);
/* Metaballs.build(solver, tension, detail), surface.js:178-284: the skin of
 * the liquid particles (flag 0) on its own cubic corner lattice of
 * m = max(6, round(nx * clamp(detail, 0.5, 1.6))) cells a side (detail NaN
 * or <= 0: the solver lattice), kernel radius (1.6 + 2.2 t) spacings,
 * contoured at Metaballs.iso(t) = 1.5 ((1.6 + 2.2 t) / 1.6)^1.5 with
 * t = clamp(tension, 0, 1) (NaN: 0.5). Mesh output as ws_mt_build; the
 * contour level and lattice size go to *iso and *lattice (either may be
 * NULL). The field stays readable as "metaballField". */
int ws_metaballs_build(ws_ctx *c, double tension, double detail, float *pos, float *nrm, long cap, long *count,
                       double *volume, double *iso, int *lattice);
/* kineticEnergy(), solver.js:3914-3921 (FP64 on the device). */
int ws_kinetic_energy(ws_ctx *c, double *ke);

/* Balls (moving solids): the `balls` array of FluidSolver (solver.js:115)
 * and addBall (solver.js:3910-3912). State is 9 doubles per ball, in order
 * x y z vx vy vz r rho submerged, for n <= WS_MAX_BALLS balls. ws_set_balls
 * replaces the context's balls (n = 0 removes them); the substep rasterizes
 * them as moving solids (:918-938), pushes particles out of them after the
 * marking pass and after advection (:1005-1026) and integrates them at its
 * end (:2344-2564). ws_get_balls reads the current state back. */
int ws_set_balls(ws_ctx *c, int n, const double *state);
int ws_get_balls(ws_ctx *c, int n, double *state);

/* Render readback: the particle fields a renderer reads every frame
 * (scene.updateParticles, js/scene.js:1595-1700: px, py, pz, pflag, pT,
 * pLight, pDepth, pPh and, for the spin colours, pWx, pWy, pWz). The fields
 * picked by `mask` (WS_RF_* bits) are copied with back-to-back DMA
 * transfers into pinned host memory owned by the context, with one wait;
 * `view` receives pointers to them (NULL for fields not picked) and the
 * particle count. The memory stays valid until the next ws_read_render or
 * ws_destroy on the context; callers read it in place (no further copy). */
enum {
  WS_RF_PX = 1 << 0, WS_RF_PY = 1 << 1, WS_RF_PZ = 1 << 2, WS_RF_PFLAG = 1 << 3, WS_RF_PT = 1 << 4,
  WS_RF_PLIGHT = 1 << 5, WS_RF_PDEPTH = 1 << 6, WS_RF_PPH = 1 << 7, WS_RF_PWX = 1 << 8, WS_RF_PWY = 1 << 9,
  WS_RF_PWZ = 1 << 10, WS_RF_ALL = (1 << 11) - 1
};
typedef struct {
  int nP;
  const float *px, *py, *pz, *pT, *pLight, *pPh, *pWx, *pWy, *pWz;
  const uint8_t *pflag, *pDepth;
} ws_render_view;
int ws_read_render(ws_ctx *c, unsigned mask, ws_render_view *view);

/* Autotuner (no counterpart in the reference, which has one SOR loop): on
 * the context's current state it times `sweeps` SOR sweeps of each SOR
 * kernel (sorMode 0 per-colour grid, 1 LDS tile, 2 fluid list, 3 two sweeps
 * per LDS launch, 4 the whole solve in one workgroup of 256, 512 or 1024
 * threads, 5 the whole solve on "flowGroups" workgroups that wait only for
 * their neighbours between half-sweeps, 6 the sweeps between two residual
 * checks in one launch on the workgroups of one XCD, each holding its range
 * of cells in LDS; all give the same bits for the same sweeps; a sorMode 6
 * launch that finds its workgroups off one XCD or its ranges too large
 * returns before writing, and the same sweeps run on the list kernel) and,
 * with the warm
 * start on, the three forms of the convolutions ("netTiled" 0, 1, 2), and
 * the P2G bins
 * (8x8x8, 4x4x4, 16x8x8, 8x8x4 cells per tile; the sums are the same for
 * every bin), and keeps the fastest ("sorMode", "sorThreads", "netTiled",
 * "p2gBin"). The pressure, the faces and all counters are left as they
 * were. `report` (may be NULL) receives one line per candidate
 * and the choice. */
int ws_tune(ws_ctx *c, int sweeps, char *report, size_t len);

/* Batched planets (no counterpart in the reference, which runs one solver
 * per page). A batch advances n contexts together, each by whole substeps
 * with the device dt law, as ws_run_substeps does for one context. Every
 * stage is one launch over all planets on the batch's stream: the planets'
 * own stages (each planet's substep is recorded once per ws_batch_run call,
 * with the settings it has then, and the recordings merged stage by stage),
 * the learned warm start's forward pass, the pressure solves (workgroups
 * per planet by size, residual checks on the device, sorMode 4 arithmetic)
 * and the extrapolation; with use_graph the batched substep is captured
 * once per call into a HIP graph that is replayed for each substep. (With
 * "stageBatch" 0, or when a planet has vapor parcels or balls, the planets'
 * own stages are queued on their own streams instead.) Each planet's results equal those of ws_run_substeps
 * with sorMode 4 and the same sorThreads (the batch's); a planet's solve may
 * span several workgroups, which changes no bits. Between calls the
 * contexts may be read, written and reconfigured as usual; while
 * ws_batch_run runs they must not be used from another thread. Contexts in
 * a batch may not have "timing", "recordSamples" or "pressureCompare" on.
 * `threads` is the workgroup size of the solves (256, 512 or 1024). The
 * contexts stay owned by the caller and must outlive the batch. */
typedef struct ws_batch ws_batch;
int ws_batch_create(ws_batch **out, ws_ctx *const *ctx, int n, int use_graph, int threads);
int ws_batch_run(ws_batch *b, int substeps);
/* Settings: "graph" (0/1), "threads" (256/512/1024), "cellsPerGroup" (a
 * planet's solve gets one workgroup per that many cells, 1 to 32),
 * "maxGroups" (1..256: workgroups of the joint solve in all; they must all be
 * resident at once, so the launch runs alone on the GPU between the planets'
 * stages), "timing" (0/1: the following runs go eagerly, timing the joint
 * solve launch with HIP events; setting it resets the totals), "xcdLocal"
 * (0/1: lay each planet's solve workgroups out on one XCD where they fit,
 * so that its barriers stay in that XCD's L2), "netBatch" (0/1: the warm
 * start's forward passes as one launch per network step over all planets;
 * implied by stageBatch), "stageBatch" (0/1: the planets' own stages merged
 * into one launch per stage). Values: those, "groups" (workgroups of the
 * last joint solve), "xcdLocalPlanets" (planets whose last solve ran on one
 * XCD), "stageLaunches" (merged stage launches per batched substep, 0 when
 * the stages run per planet), and the totals of the timed
 * runs: "solveMs", "solveLaunches", "sweeps" (summed over planets),
 * "cellSweeps" (fluid cells relaxed, summed over planets and sweeps). */
int ws_batch_set(ws_batch *b, const char *name, double v);
int ws_batch_get(ws_batch *b, const char *name, double *v);
const char *ws_batch_last_error(const ws_batch *b);
void ws_batch_destroy(ws_batch *b);

/* Waits for all work queued on the context's stream. */
int ws_sync(ws_ctx *c);

#ifdef __cplusplus
}
#endif
#endif
