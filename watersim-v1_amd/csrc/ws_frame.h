// The context type shared by the ABI translation units, and the frame-level
// passes (thermal tick, evaporation, currents, phase) they call.
#pragma once
#include "ws_kernels.h"

struct ws_ctx : public ws::Ctx {};

namespace ws {
int frame_heat(ws_ctx *c, double dt);
int frame_evaporation(ws_ctx *c, double dt);
int frame_currents(ws_ctx *c, double dt);
int frame_phase(ws_ctx *c, double dt);
int frame_push_surface(ws_ctx *c);
int frame_vapor_collisions(ws_ctx *c);
int frame_ice_contacts(ws_ctx *c);
void contacts_release(ws_ctx *c, bool all);
// balls (ws_balls.hip)
int balls_raster(ws_ctx *c);
int balls_push(ws_ctx *c, bool rebin);
int balls_update(ws_ctx *c);
void balls_release(ws_ctx *c);
// marching tetrahedra over a device corner field (ws_api_surface.hip)
int mt_build_dev(ws_ctx *c, const float *dd, int nx, int ny, int nz, double dx, double iso, float *pos, float *nrm,
                 long cap, long *count, double *volume);
void metaballs_release(ws_ctx *c);
World make_world(ws_ctx *c);
// the device address of a GPU's static marching-tets scratch, looked up once
// per device (ws_noalloc.hip)
int mt_scratch_setup(int device);
// a device-dt-law substep split around its pressure sweeps (ws_api.hip)
int batch_substep_pre(ws_ctx *c);
int batch_substep_post1(ws_ctx *c);
int batch_substep_post2(ws_ctx *c);
ExPlanet batch_ex_planet(ws_ctx *c);
SorJob batch_sor_job(ws_ctx *c);
void batch_solved(ws_ctx *c);
// exclusive prefix sums on the context's stream, one launch each (ws_scan.hip;
// 64-bit entries below 2^62); in may equal out
int scan_excl(ws_ctx *c, const unsigned *in, unsigned *out, long n);
int scan_excl(ws_ctx *c, const unsigned long long *in, unsigned long long *out, long n);
long scan_tiles_for(long n);
}  // namespace ws
