/*
 * gamera_op.h -- C ABI of the MI355X (gfx950) device solver for the
 * GAMERA-OP per-step ideal-MHD update.
 *
 * This is the thin layer the C host facade (initialize_solver / solve /
 * finalize_solver, src/solver.h:17,26,65 in the reference) calls to run the
 * time step on the GPU.  It replaces, for a single-GPU domain:
 *   - the solve() loop body           src/solver_solve.c:102-508
 *   - the numerical kernels            src/solver_func.c:215-1529
 *   - the halo exchange                src/boundary_exchange.c:7-201
 *   - the boundary-condition library   src/boundary_conditions.c:9-593
 * All entry points take plain pointers and sizes and return 0 on success or a
 * negative gk_status code; nothing throws across the ABI.  Host arrays use
 * the reference layout: one contiguous block per field, [i][j][k] with k
 * fastest, NI = ni + 9 points per axis (4 ghost layers each side plus the
 * upper-face slot), i.e. the layout of src/array.c:87-114.
 */
#ifndef GAMERA_OP_H
#define GAMERA_OP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  GK_OK = 0,
  GK_ERR_ARG = -1,      /* invalid argument or size */
  GK_ERR_DEVICE = -2,   /* HIP runtime error */
  GK_ERR_NOMEM = -3,    /* device allocation failed */
  GK_ERR_STATE = -4,    /* call out of order (e.g. step before prologue) */
  GK_ERR_NAN = -5,      /* non-finite state detected by the fused reduction */
} gk_status;

/* boundary kinds per array, dimension and side.  GK_BC_POLE (j sides of a
   spherical grid whose theta range ends at a pole; replaces gas_bc_pole_j_* /
   gem_bc_pole_j_*, src/spherical/spherical_boundary.c:23-138): each ghost row
   takes the row mirrored across the pole, rotated by pi in phi
   (k -> 4 + (k - 4 + nk/2) mod nk), for every i and k of the box, with v_theta,
   v_phi and B_theta, B_phi negated.  It needs ring averaging (below).
   GK_BC_REFLECTIVE (gas only; gas_bc_reflective_*,
   src/boundary_conditions.c:114-245): the symmetric mirror with the normal
   velocity negated.  GK_BC_EXTRAPOLATED (gas_bc_extrapolated_*,
   gem_bc_extrapolated_*, src/boundary_conditions.c:404-496,687-770): every
   ghost layer copies the first active layer (low side) or the last active
   layer or boundary face (high side).  GK_BC_FIXED (gas_bc_fixed_*,
   gem_bc_fixed_*, src/boundary_conditions.c:247-403,602-686): the ghost
   layers hold rho = 1, v = 0, p = 1 and face B = 0; p_S is left to the other
   sides' maps, as when the plugin calls the library i, j, k (low, high) for
   gas, then gem.  GK_BC_NONE leaves the ghosts at the values uploaded with
   the state. */
enum {
  GK_BC_PERIODIC = 0, GK_BC_SYMMETRIC = 1, GK_BC_NONE = 2, GK_BC_POLE = 3,
  GK_BC_REFLECTIVE = 4, GK_BC_EXTRAPOLATED = 5, GK_BC_FIXED = 6
};

/* separable geometry quantities: q(i,j,k) = f[q][0][i] * f[q][1][j] * f[q][2][k] */
enum {
  GK_G_EDGE_I = 0, GK_G_EDGE_J, GK_G_EDGE_K,
  GK_G_FACE_I, GK_G_FACE_J, GK_G_FACE_K,
  GK_G_VOL, GK_G_DXI, GK_G_DXJ, GK_G_DXK,
  GK_G_COUNT
};

/* per-face reconstruction table entry (per axis, indexed by face index) */
#define GK_RECW_STRIDE 18 /* wL[7] wR[7] c1L c2L c1R c2R */

typedef struct {
  int ni, nj, nk;          /* active cells per axis */
  int coord;               /* 0 Cartesian, 2 spherical (geometric sources) */
  int bc[2][6];            /* [gas|gem][i_lo,i_hi,j_lo,j_hi,k_lo,k_hi] */
  int precision;           /* 0 FP64 (reference arithmetic), 1 FP32-bulk */
  double CA;               /* Boris speed of light (src/config.c:85) */
  /* per-axis tables; axis a has N_a = n_a + 9 entries, indexed like the
     reference arrays */
  const double *geom[GK_G_COUNT][3];
  /* cell-centre B from the two faces (src/solver_func.c:310-322):
     b[n] = (B[n] * cc[0][n]) / cc[2][n] + (B[n+1] * cc[1][n]) / cc[2][n],
     cc = {x[n+1] - xctr[n], xctr[n] - x[n], dx[n]} */
  const double *cell_coef[3][3];
  /* reconstruction tables of sweep axis a, indexed
     [((face * ncls[a][0] + class_b) * ncls[a][1] + class_c) * GK_RECW_STRIDE],
     b < c the transverse axes; cls[a][0|1][t] gives the class of transverse
     coordinate t (on Cartesian grids the weights depend on the transverse
     cell widths through roundoff only; one class per distinct width keeps
     them exact) */
  const double *recw[3];
  const int *cls[3][2];
  int ncls[3][2];
  /* background field B0 (src/solver_func.c:1828-2077): mode 0 none,
     1 the same at every point (constants below), 2 given per point
     (b0_field) */
  int b0_mode;
  double b0_cell[3];
  double b0_face[3][8];    /* per face direction: b1 b2 b3 bn bsq bn1 bn2 bn3 */
  double b0_edge[3][3];    /* per edge direction: b1 b2 b3 */
  double b0_lorentz[3];    /* dmagsf_G, constant */
  /* pole ring averaging (doRingAverage; src/spherical/spherical_RingAverage.c,
     called from src/solver_solve.c:258-260,308-310,486-491): ring = 1 turns it
     on (spherical, k periodic; either precision form).  ring_pole[0|1]: theta = 0 / theta = pi
     is a pole of the grid; the ring_rows cell rows next to each pole are
     averaged, row r (0 at the pole) over ring_nchunk[r] azimuthal blocks, and
     the CFL time step uses the phi width times nk / ring_nchunk[r] on them.
     Zero-initialise the descriptor: ring = 0 leaves all of this unused. */
  int ring;
  int ring_pole[2];
  int ring_rows;
  int ring_nchunk[16];
  /* positivity floors and uniform gravity: the reference's mutable globals
     rho_floor, p_floor (src/config.c:89-90) and G01..G03 (:95-97), which a
     problem file may set; applied in check_positivity
     (src/solver_func.c:2117-2135) and the momentum update
     (src/solver_solve.c:469-471).  A floor of 0 selects the reference's
     initial value (0.00316227, 0.000316227), so a zero-initialised
     descriptor is the reference's default set-up; negative or non-finite
     values are rejected (GK_ERR_ARG). */
  double rho_floor, p_floor;
  double gravity[3];
  /* b0_mode 2: the reference's background-field arrays, GK_B0_FIELDS host
     arrays in the reference layout (NI*NJ*NK doubles each), in this order:
     face averages B0AfaceI_b1 b2 b3 bn, B0AfaceJ_b1 b2 b3 bn, B0AfaceK_b1 b2
     b3 bn (12); edge averages B0AedgeI_b2 b3, B0AedgeJ_b3 b1, B0AedgeK_b1 b2
     (6); cell-centre B0_b1 b2 b3 (3); Lorentz force dmagsf_G_x1dir x2dir
     x3dir (3).  Copied to the device by gk_create. */
  const double *b0_field[24];
  /* HIP stream (a hipStream_t) for the solver's own work: uploads, steps,
     read-backs.  NULL: the solver takes a stream of its own.  A given stream
     must outlive the solver; gk_batch_side_stream provides one whose work
     never waits behind a batch's steps. */
  void *stream;
} gk_grid_desc;

#define GK_B0_FIELDS 24

typedef struct gk_solver gk_solver;

/* Create a solver on the current HIP device. */
int gk_create(const gk_grid_desc *desc, gk_solver **out);
int gk_destroy(gk_solver *s);

/* Upload the primitive state (6 cell fields rho v1 v2 v3 p p_S) and the
   face-normal magnetic field (3 face fields), reference layout, NI*NJ*NK
   doubles each.  prev_* may be NULL (then X_p := X, as the solve() prologue
   does); otherwise they carry the AB2 history of a restart. */
int gk_set_state(gk_solver *s, const double *const prim[6], const double *const bface[3],
                 const double *const prim_prev[6], const double *const bface_prev[3]);

/* solve() prologue (src/solver_solve.c:25-99): ghost fill + boundary
   conditions, then dt0 := CFL time step.  With history uploaded, dt0 is the
   value passed in (restart). */
int gk_prologue(gk_solver *s, double dt0_restart);

/* Advance n AB2 steps; dt, dt0 and time stay on the device.  Returns
   GK_ERR_NAN when the state holds a non-finite value after the steps (the
   cell update's fused check; the steps are taken regardless). */
int gk_step(gk_solver *s, int n);

/* Up to n steps of the reference's solve() loop (src/solver_solve.c:102-529)
   without a host round trip per step: the device ends the run after the
   first step whose time exceeds t_stop (the loop's time_stop break,
   :523-527) or reaches t_event (time >= t_event: the caller's next dump or
   status line is due, :510-521).  *done (may be NULL) receives the steps
   taken.  Returns GK_ERR_NAN as gk_step does. */
int gk_step_until(gk_solver *s, int n, double t_stop, double t_event, int *done);

/* Download state and history (any pointer may be NULL). */
int gk_get_state(gk_solver *s, double *prim[6], double *bface[3], double *prim_prev[6],
                 double *bface_prev[3]);

/* gk_get_state without any allocation (no counterpart in the reference).
   For callers that cannot allocate device or pinned memory: code under HIP
   graph capture, tight serving loops, harnesses that forbid it.  gk_get_state
   allocates a device staging field and a pinned host buffer on its first
   call; this variant stages each field through a fixed 32 MiB device array
   that the library holds on every GPU it runs on until the process ends,
   and copies into the caller's arrays directly.  Same arguments and results
   as gk_get_state; GK_ERR_ARG for a grid whose reference-layout field
   exceeds the array ((ni+9)(nj+9)(nk+9) > 4194304 points, e.g. above
   152^3 cells).  Calls that overlap on one GPU (several threads or streams)
   share the array and race: serialise them. */
int gk_get_state_noalloc(gk_solver *s, double *prim[6], double *bface[3], double *prim_prev[6],
                         double *bface_prev[3]);

/* scalars: [0] dt of the last step, [1] dt0, [2] time_sim, [3] max|div B|,
   [4] next CFL time step, [5] non-finite flag */
int gk_get_scalars(gk_solver *s, double out[6]);
int gk_set_time(gk_solver *s, double time_sim);

/* Average kernel durations (ms) over the last gk_step call when timing is
   enabled: names/ms arrays of length *n (in/out).  gk_time_stage turns timing
   on for the launches of one stage only (name as gk_kernel_times reports it,
   e.g. "k_cell"; NULL: every stage), so that a timed run carries two events
   per step instead of two per launch; gk_enable_timing(s, 1) times every
   stage again.  (No counterpart in the reference.) */
int gk_enable_timing(gk_solver *s, int on);
int gk_time_stage(gk_solver *s, const char *name);
int gk_kernel_times(gk_solver *s, const char **names, double *ms, int *n);

/* Copy one intermediate field of the last step to the host (reference
   layout): which 0..2 face fluxes of sweep i/j/k (field 0..8: rho, rho v1..3,
   energy, entropy, magnetic stress 1..3), 3..5 face-averaged velocities of
   sweep i/j/k (field 0..1), 6 edge electric fields (field 0..2; on a
   ring-averaged grid the rows next to the poles then hold the ring average's
   edge fields), 7 the cell-centred magnetic field of the current state
   (field 0..2). */
int gk_debug_fetch(gk_solver *s, int which, int field, double *host);

/* device bytes held by the solver */
size_t gk_device_bytes(const gk_solver *s);

/* Launch-shape parameters (no counterpart in the reference; they change
   speed, never results).  GK_TUNE_CHUNK: about how many faces each thread
   walks in the i and j sweeps (1..4096; they walk twice that on grids
   large enough to fill the GPU with it); GK_TUNE_EMF_CHUNK: about how
   many edges each thread walks in the edge-field kernels (1..4096).  Their
   defaults depend on the grid's cell count: 2 and 2 below 1e5 cells, 4 and
   2 below 5e5, 16 (FP32-bulk: 8) and 4 below 4e6, 32 and 16 below 1.2e7,
   64 and 17 above.  A column is split into near-equal chunks of
   about that length (within 1.5x of it).  GK_TUNE_EMF_STAGE: bit D (D = 0
   for E_i, 1 for E_j) runs that edge field over whole 64-edge tiles along
   k with its k-direction reconstructions staged in LDS (0..3; default 3 in
   FP64, 2 in FP32-bulk).  GK_ERR_STATE while the solver is in a batch (the
   batch's launch tables hold its shapes). */
enum { GK_TUNE_CHUNK = 0, GK_TUNE_EMF_CHUNK = 1, GK_TUNE_EMF_STAGE = 2 };
int gk_set_tuning(gk_solver *s, int param, int value);
int gk_get_tuning(const gk_solver *s, int param);

/* Launch-shape autotuner (no counterpart in the reference): times
   trial_steps steps for each candidate march length of the sweeps and of the
   edge-field kernels and, on grids with whole 64-edge tiles along k, each
   GK_TUNE_EMF_STAGE setting; keeps the fastest of each (the sweeps' march
   length is written to *best) and restores the state it started from, so
   results do not depend on tuning.  Picks are cached per grid shape and
   precision for the life of the process.  trial_steps = 0 only applies a
   cached pick (*best = 0 when there is none) and times nothing. */
int gk_autotune(gk_solver *s, int trial_steps, int *best);

/* Batched stepping (no counterpart in the reference; for ensembles and the
   serving loop).  A batch steps many prepared solvers (after gk_prologue)
   together: per stage one launch covers every grid of a precision, and the
   launches of a step are captured in HIP graphs.  Each grid keeps its own dt,
   time and history and gets exactly the results gk_step would give it.
   While a solver is in a batch, gk_step on it returns GK_ERR_STATE;
   gk_get_state / gk_get_scalars work after gk_batch_step (or gk_batch_sync). */
typedef struct gk_batch gk_batch;
int gk_batch_create(gk_batch **out);
int gk_batch_destroy(gk_batch *b);
int gk_batch_add(gk_batch *b, gk_solver *s);
int gk_batch_remove(gk_batch *b, gk_solver *s);
int gk_batch_size(const gk_batch *b);
int gk_batch_step(gk_batch *b, int n);        /* n steps, waits for them */
int gk_batch_step_async(gk_batch *b, int n);  /* n steps, returns at once */
/* waits for the launched steps; GK_ERR_NAN if a member's state holds a
   non-finite value (gk_batch_status names which) */
int gk_batch_sync(gk_batch *b);
/* Step budget of a member (serving): it takes `steps` more steps, counted
   from the steps already launched, and then stops on the device while the
   batch goes on (its later launches do nothing), so one batch step count
   need not stop at the member that finishes first.  steps < 0: no limit
   (the default on gk_batch_add).  gk_batch_budget returns what is left of
   it after the launched steps (-1: no limit). */
int gk_batch_set_budget(gk_batch *b, gk_solver *s, long long steps);
long long gk_batch_budget(gk_batch *b, gk_solver *s);
/* after gk_batch_sync: GK_ERR_NAN if this member's state holds a
   non-finite value, GK_OK if not, GK_ERR_ARG if it is not a member */
int gk_batch_status(gk_batch *b, gk_solver *s);
int gk_batch_set_graphs(gk_batch *b, int on); /* default on */
/* A second stream of the batch, created with its own so that the two map to
   different hardware queues: solvers created on it (gk_grid_desc.stream) can
   be set up, read back and destroyed while the batch steps, without waiting
   for its launched steps.  It stays valid after gk_batch_destroy. */
int gk_batch_side_stream(gk_batch *b, void **stream);

/* Device time of the batch's runs (no counterpart in the reference; for the
   serving loop's GPU-busy share).  With run timing on, each gk_batch_step /
   gk_batch_step_async call is timed on the batch's stream from its first
   launch to the read-back of its non-finite flags; gk_batch_run_ms gives the
   milliseconds of the runs completed (synchronised) since it was turned on
   (turning it on resets the total).  Work on the side stream is not
   counted. */
int gk_batch_run_timing(gk_batch *b, int on);
int gk_batch_run_ms(gk_batch *b, double *ms);
int gk_batch_enable_timing(gk_batch *b, int on);
int gk_batch_kernel_times(gk_batch *b, const char **names, double *ms, int *n);

/* Measurement helper (no counterpart in the reference): a plain device copy
   of `bytes` bytes (rounded down to 16-byte units), `reps` times after one
   warm-up; *ms receives the average duration of one copy kernel.  Used to
   measure the achievable HBM bandwidth and to scale hardware byte counters. */
int gk_copy_calibrate(size_t bytes, int reps, double *ms);

#ifdef __cplusplus
}
#endif
#endif
