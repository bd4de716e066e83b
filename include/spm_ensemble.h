/*
 * spm_ensemble.h -- ensembles of independent SPM boxes on one GPU (no
 * counterpart in the reference, which runs one box per process).
 *
 * An ensemble owns nbox contexts, each with its own spm_params (grid,
 * particle count, radius, precision, thermostat seed ...), all on one
 * device.  Each box is an ordinary spm_ctx: spm_ensemble_box() returns it
 * for spm_put_particles, spm_init_fluid, spm_get_field, spm_get_particles
 * and the rest of include/spm.h.  spm_ensemble_step advances every box by
 * the same number of steps, each box exactly as spm_step would (the same
 * kernels in the same order, so every box's state is bit-identical to a
 * lone context stepped with spm_step from the same state).
 *
 * How it runs: one step of every box is captured once into a HIP graph.
 * The boxes are split over `lanes` streams (graph branches) by longest
 * processing time first on their measured step times, so the branches run
 * side by side on the device's hardware queues; one graph launch then
 * advances the whole ensemble by one step with no host work per box.
 * The thermostat's normals are drawn on the host ahead of time,
 * SPM_ENSEMBLE_CHUNK steps at a time, into a device ring per box that the
 * graph reads through a device-side counter (the host draws the next chunk
 * while the device runs the current one); the draw order is the one
 * spm_step uses, so the MT19937 stream is unchanged.  A box's first step (ts == 0,
 * the Euler start of the AB2 integrators) runs outside the graph.
 *
 * Boxes must use the plain Navier-Stokes path: no electrolyte, no
 * Lees-Edwards shear and timing off; a box may have a closure network
 * attached (spm_set_closure), one network per box (those steps are
 * rejected with SPM_ERR_UNSUPPORTED).  A one-box ensemble is the way to
 * replay a single context's step from a graph.
 */
#ifndef SPM_ENSEMBLE_H
#define SPM_ENSEMBLE_H

#include "spm.h"

#ifdef __cplusplus
extern "C" {
#endif

#define SPM_ENSEMBLE_CHUNK 32
#define SPM_ENSEMBLE_MAX_LANES 8

typedef struct spm_ensemble spm_ensemble;

/* nbox contexts from params[0 .. nbox-1] (all on params[0].device, which
 * every entry must name).  Grids with a mixed length (48, 80, 96, 160,
 * 192) have their transform launch shapes autotuned here (once per shape
 * and process). */
int spm_ensemble_create(int nbox, const spm_params *params, spm_ensemble **out);
int spm_ensemble_destroy(spm_ensemble *e);
/* Box b's context (owned by the ensemble; do not spm_destroy it), or NULL. */
spm_ctx *spm_ensemble_box(spm_ensemble *e, int b);
int spm_ensemble_size(const spm_ensemble *e);
/* Every box advances nsteps steps; waits for the device before returning. */
int spm_ensemble_step(spm_ensemble *e, int nsteps);
/* graph = 1 (default): replay the captured step; 0: launch every kernel
 * from the host each step (same results; for comparison). */
int spm_ensemble_set_graph(spm_ensemble *e, int graph);
/* Chooses the lane count: times steps_per_candidate steps (after one
 * untimed step) for each of 1, 2, 4 and 8 lanes (at most nbox) and keeps
 * the fastest; these are real steps of the ensemble (each box advances
 * 4 * (steps_per_candidate + 1) + 1 steps at most, the first one also timing
 * every box alone for the split).  Until called, lanes = min(4, nbox)
 * split on estimated costs. */
int spm_ensemble_autotune(spm_ensemble *e, int steps_per_candidate);
/* Lanes in use; ms[k] (k < 4) the measured milliseconds per ensemble step
 * with 1, 2, 4, 8 lanes (0 where not measured), box_ms[b] (nbox entries,
 * may be NULL) each box's measured step time (0 before autotuning). */
int spm_ensemble_tuning(const spm_ensemble *e, int *lanes, double ms[4], double *box_ms);
/* Text of the transform autotuner's choices (one line per tuned pass):
 * writes at most len bytes including the terminating NUL, returns the full
 * length. */
int spm_autotune_report(char *buf, int len);
/* Last error message of an ensemble call (never NULL). */
const char *spm_ensemble_last_error(const spm_ensemble *e);

#ifdef __cplusplus
}
#endif

#endif /* SPM_ENSEMBLE_H */
