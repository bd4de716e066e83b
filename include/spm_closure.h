/*
 * spm_closure.h -- learned closure of the coarse-grid SPM step (no
 * counterpart in the reference).
 *
 * A small MLP corrects the hydrodynamic force and torque that
 * Calc_f_hydro_correct_precision (md_force.cxx:263-389) computes on a coarse
 * grid, so that fewer grid points per radius are needed; the plain method on
 * a fine grid is the fallback and the judge.  Per particle it reads 126
 * inputs and returns 6 outputs:
 *
 *   inputs   [0, 3)     relative velocity: the phi-weighted mean of
 *                       v + w x r - u over the particle, -F_h dt / (rho V_p)
 *                       with the force F_h of Calc_f_hydro_correct_precision
 *                       and V_p = 4 pi a^3 / 3
 *            [3, 84)    u(x_p + a o) - v_p at the 27 offsets o in {-1,0,1}^3
 *                       (o lexicographic, x slowest; components x, y, z),
 *                       trilinear in the grid
 *            [84, 111)  phi(x_p + a o): the particles' profiles summed at
 *                       the same points, clamped at 1
 *            [111, 114) gap vector to the nearest other particle j: the
 *                       centre separation r times (1 - 2a / |r|), minimum
 *                       image
 *            [114, 117) its relative velocity v_j - v_p
 *            [117, 120) its relative angular velocity w_j - w_p
 *            [120, 123) gap vector to the second nearest particle k
 *            [123, 126) its relative velocity v_k - v_p
 *                       (zeros where there is no such particle)
 *   outputs  [0, 3) force correction, [3, 6) torque correction
 *
 * with x_p the particle's centre and a its radius.  The network is
 * 126 (padded to 128) -> 128 -> 128 -> 128 -> 6 with exact GELU between the
 * layers, BF16 weights and activations into the matrix cores and FP32
 * accumulation; inputs are normalised as (x - shift) * scale and outputs
 * multiplied by out_scale.  FP32 master weights, trained with Adam on the
 * GPU by the library's own forward and backward kernels.
 *
 * Parameter layout of spm_closure_get/put_params (SPM_CLOSURE_NPARAM
 * floats): W1 [128 in][128 out], b1[128], W2, b2, W3, b3 (the same),
 * W4 [128 in][16 out], b4[16]; input rows 126..127 of W1 and output
 * columns 6..15 of W4 / b4 are padding and stay zero.
 */
#ifndef SPM_CLOSURE_H
#define SPM_CLOSURE_H

#include <stdint.h>

#include "spm.h"

#ifdef __cplusplus
extern "C" {
#endif

#define SPM_CLOSURE_IN 126
#define SPM_CLOSURE_OUT 6
#define SPM_CLOSURE_NPARAM 51600

typedef struct spm_closure spm_closure;

/* A network on `device` for batches of up to max_batch particles; weights
 * start from spm_closure_init(seed 0). */
int spm_closure_create(int device, int max_batch, spm_closure **out);
int spm_closure_destroy(spm_closure *cl);
/* Uniform Glorot initialisation from a seeded generator, biases zero,
 * Adam moments and step count reset. */
int spm_closure_init(spm_closure *cl, uint64_t seed);
int spm_closure_get_params(spm_closure *cl, float *host);
int spm_closure_put_params(spm_closure *cl, const float *host);
/* Input normalisation and output scale (host arrays of 126, 126, 6). */
int spm_closure_set_scaling(spm_closure *cl, const float *in_shift, const float *in_scale, const float *out_scale);
int spm_closure_get_scaling(spm_closure *cl, float *in_shift, float *in_scale, float *out_scale);
/* y[n][6] = net(x[n][126]) for host arrays, n <= max_batch. */
int spm_closure_forward(spm_closure *cl, const float *x, float *y, int n);
/* Gradient of the mean squared error (in normalised outputs, averaged over
 * the n samples and 6 outputs) with respect to every parameter, for host
 * arrays (test hook for the backward kernels); grad holds
 * SPM_CLOSURE_NPARAM floats, loss one. */
int spm_closure_gradient(spm_closure *cl, const float *x, const float *target, int n, float *grad, float *loss);
/* Training set held on the device: x[n][126], target[n][6] (physical
 * units; divided by out_scale inside). */
int spm_closure_set_data(spm_closure *cl, const float *x, const float *target, int n);
/* `steps` Adam steps (lr, beta1 0.9, beta2 0.999, eps 1e-8) over
 * minibatches of `batch` consecutive samples of the training set, cycling;
 * loss[steps] receives each step's minibatch loss (may be NULL). */
int spm_closure_train(spm_closure *cl, int steps, int batch, float lr, float *loss);

/* Attach a network to an SPM context (NULL detaches): every step then adds
 * its correction to f_hydro and torque_hydro right after
 * Calc_f_hydro_correct_precision, before the velocity update.  The context
 * keeps the pointer; the network must outlive the attachment.  n_particles
 * must not exceed the network's max_batch. */
int spm_set_closure(spm_ctx *ctx, spm_closure *cl);
/* The 126 inputs of every particle for the context's current real-space u,
 * particles and f_hydro (as the step leaves them after
 * spm_calc_f_hydro / spm_zeta_k2u; host array [n_particles][126]). */
int spm_closure_features(spm_ctx *ctx, float *x);

/* Particle motion of the step (no counterpart in the reference):
 * SPM_MOTION_FREE (default) integrates positions and velocities as the
 * reference does; SPM_MOTION_PRESCRIBED holds every particle's position,
 * velocity and angular velocity fixed (the resistance problem: the step
 * still computes f_hydro / torque_hydro, the fluid forcing and the
 * closure's correction). */
enum spm_motion { SPM_MOTION_FREE = 0, SPM_MOTION_PRESCRIBED = 1 };
int spm_set_particle_motion(spm_ctx *ctx, int motion);

#ifdef __cplusplus
}
#endif

#endif /* SPM_CLOSURE_H */
