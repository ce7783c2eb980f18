/*
 * The learned pressure warm start of the watersim library (MI355X, gfx950).
 *
 * None of these entry points has a counterpart in the reference: the
 * reference solves the pressure by lexicographic SOR from the previous
 * substep's pressure (js/solver.js:1312-1410). The benchmark configuration
 * adds a small 3-D convolutional network that predicts the correction of
 * that previous pressure; plain SOR then iterates from the prediction to
 * the same residual, so SOR stays the fallback and the judge of the guess.
 *
 * Network (a 3-level U-Net, channels 16/32/64, 3x3x3 kernels, ReLU,
 * average pooling, nearest upsampling with skip concatenation, a 1x1x1
 * head): inputs per cell are the initial pressure residual r0 = rhs - A q
 * scaled by its RMS over fluid cells, the cell type one-hot (air, fluid,
 * solid) and the previous pressure scaled by its RMS; the output is the
 * correction of q in units of the residual scale. Weights and activations
 * are bfloat16, products accumulate in FP32 on the matrix cores. The flat
 * parameter vector holds, per layer, W[cout][tap][cin] then b[cout]
 * (ws_net_param_count() floats, 355009).
 *
 * Switches through ws_set_f (include/watersim.h):
 *   "warmStart"     1: every pressure solve starts from the prediction
 *   "recordSamples" 1: every pressure solve stores (inputs, target) in the
 *                   sample bank, the target being the converged correction
 *                   (q_solved - q_previous) / scale
 *   "netCoarse"     1: the network sees the solver's grid coarsened 2x per
 *                   axis (even sides): each input channel averaged over a
 *                   2x2x2 block (the one-hots become type fractions), and
 *                   its output is added to q constant over each block's
 *                   fluid cells; recorded targets are the block means of
 *                   the correction over fluid cells. A sample bank holds
 *                   samples of the netCoarse it was made with. The shipped
 *                   weights record theirs in their .json ("net_coarse").
 */
#ifndef WATERSIM_LEARN_H
#define WATERSIM_LEARN_H
#include <stdint.h>

#include "watersim.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Number of FP32 parameters of the network. */
long ws_net_param_count(void);
/* Copies the flat parameter vector in / out (n = ws_net_param_count()). */
int ws_net_set_params(ws_ctx *c, const float *w, long n);
int ws_net_get_params(ws_ctx *c, float *w, long n);
/* Runs the network on host inputs [nz][ny][nx][5] (sides multiples of 4)
 * and writes the output [nz][ny][nx]. */
int ws_net_forward(ws_ctx *c, int nz, int ny, int nx, const float *in5, float *out);
/* Loss sum(mask (out - target)^2) / sum(mask target^2) on host inputs and
 * its gradient with respect to every parameter (flat order). */
int ws_net_grad(ws_ctx *c, int nz, int ny, int nx, const float *in5, const float *target, const uint8_t *mask,
                float *grad, double *loss);
/* Allocates a device bank for `capacity` recorded samples of the context's
 * grid, or of its coarsened grid with netCoarse 1 (recording overwrites the
 * oldest once full). */
int ws_net_bank(ws_ctx *c, int capacity);
int ws_net_bank_count(ws_ctx *c, int *count);
/* Adam training on random crops (crop^3 cells of the bank's grid, crop a
 * multiple of 4) of recorded samples: `steps` updates of `batch` crops each; *loss is the
 * mean loss of the last update's batch. */
int ws_net_train(ws_ctx *c, int steps, int batch, int crop, double lr, unsigned seed, double *loss);

#ifdef __cplusplus
}
#endif
#endif
