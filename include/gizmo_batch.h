/*
 * gizmo_batch.h -- C ABI of a batch of independent periodic boxes on one GPU.
 *
 * A batch holds many boxes of uneven particle count in one device state and advances all of
 * them with shared launches: every box keeps its own timeline (the run() loop of
 * gizmo_hydro.h: power-of-two time bins, wake-ups, kick-drift-kick), its own search grid and
 * its own Morton order, and the active particles of all boxes go through each stage of the hot
 * path (density() with the radius iteration, hydro_gradient_calc(), hydro_force()) together.
 * Boxes may join and leave between steps (continuous batching).  A box's results do not depend
 * on the other boxes of the batch: it computes the same sums in the same order alone or in any
 * mix.
 *
 * This interface has no counterpart in the reference, which runs one simulation per process;
 * it is for ensembles and for serving many small simulations on one GPU.  The physics of each
 * box is that of the gizmo_hydro.h entry points (reference citations there).
 */
#ifndef GIZMO_BATCH_H
#define GIZMO_BATCH_H

#include "gizmo_hydro.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Per-box run parameters: the box's own All fields. */
typedef struct {
    double box_size;           /* All.BoxSize */
    double min_kernel_radius;  /* All.MinKernelRadius */
    double max_kernel_radius;  /* All.MaxKernelRadius */
    gz_integ_params integ;     /* the box's timestep-loop parameters */
} gz_box_params;

typedef struct gz_batch gz_batch;

/* A batch on HIP device `device` for up to `capacity` particles in all and `max_boxes` boxes at
   once.  shared: kernel, gamma, des_num_ngb, max_num_ngb_dev and precision apply to every box
   (its box_size, radius limits and times are ignored).  NULL on failure. */
gz_batch *gz_batch_create(const gz_params *shared, long capacity, int max_boxes, int device);
void gz_batch_destroy(gz_batch *b);
const char *gz_batch_last_error(const gz_batch *b);

/* Load a box (pos, vel, mass, u, h required; the other fields optional, as gz_load).  The box
   waits until the next gz_batch_begin; *box_id receives its handle.  GZ_E_MEMORY when the
   particles or the box do not fit. */
int gz_batch_add(gz_batch *b, long n, const gz_fields *host, const gz_box_params *bp, int *box_id);
/* Start every box added since the last call: init() and run()'s first hot-path pass
   (gz_run_begin), all such boxes in shared launches.  They then step with the others. */
int gz_batch_begin(gz_batch *b);
/* One sync point of every started box (each at its own time): gz_run_step for all of them. */
int gz_batch_step(gz_batch *b);
/* Leave the batch; the box's particles are retired and the other boxes packed together (the box
   id may then be given to a new box).  A box whose timeline reached its end (gz_timeline's
   stop_code set) no longer steps and waits for this call. */
int gz_batch_remove(gz_batch *b, int box_id);
/* Device -> host for one box, in the order its particles were given to gz_batch_add. */
int gz_batch_store(gz_batch *b, int box_id, gz_fields *host);
int gz_batch_timeline(gz_batch *b, int box_id, gz_timeline *out);
/* Particles resident (all live boxes) and live boxes. */
long gz_batch_particles(const gz_batch *b);
int gz_batch_boxes(const gz_batch *b);
int gz_batch_synchronize(gz_batch *b);
/* Totals since the last gz_batch_timing_reset: "active_updates" (active particles through the
   hot path), "steps", "begins", "resorts", "force_pairs_total"; and "device_bytes", the device
   memory the batch holds now. */
int gz_batch_stat(gz_batch *b, const char *name, double *value);
int gz_batch_timing_reset(gz_batch *b);
/* Options: "few_rows" (default 8192) -- passes over at most this many active rows (all boxes
   together) take one wavefront per row; more rows are walked in groups that share a cell of
   their box's grid (the same lists and sums either way).  No counterpart in the reference (a
   tuning knob of this library, as gz_set_option's "few_rows").  "fd_correction" (default 0):
   the first-derivative volume correction of the density for every box, as gz_set_option's. */
int gz_batch_set_option(gz_batch *b, const char *name, double value);

#ifdef __cplusplus
}
#endif
#endif /* GIZMO_BATCH_H */
