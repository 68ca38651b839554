/*
 * scx.h -- C ABI of the MI355X (gfx950) contact pipeline for circle (2D) and
 * sphere (3D) scenes: uniform-grid broad phase, narrow phase, Delassus assembly,
 * graph-coloured projected Gauss-Seidel Coulomb solve and the symplectic
 * Euler / Verlet update.
 *
 * Plain C: pointers, sizes and int status codes. All array arguments are host
 * pointers in the reference's layouts unless stated otherwise; the library keeps
 * its own device-resident copies (one context = one device + one HIP stream).
 * Every entry point returns SCX_OK (0) or a negative SCX_E* code and sets a
 * message readable with scx_last_error(). No exception crosses the boundary.
 * A context is not re-entrant: calls on one context must be serialised by the
 * caller (as the reference's single-threaded step loop does).
 *
 * Reference interfaces replaced (paths relative to the breannansmith/scisim tree):
 *   scx_potential_overlaps  <- SpatialGridDetector::getPotentialOverlaps
 *                              (ball2d/SpatialGridDetector.h:39, rigidbody3d/SpatialGridDetector.h:37)
 *   scx_step                <- ImpactFrictionMap::flow for GeometricImpactFrictionMap
 *                              (scisim/ConstrainedMaps/ImpactFrictionMap.h:36,
 *                               GeometricImpactFrictionMap.cpp:267-442), called per step by
 *                              Ball2DSim::flow (ball2d/Ball2DSim.cpp:305-326) and
 *                              RigidBody3DSim::flow (rigidbody3d/RigidBody3DSim.cpp:557-630)
 *   scx_friction_solve      <- FrictionSolver::solve / Sobogus::solve
 *                              (scisim/ConstrainedMaps/FrictionSolver.h:28, Sobogus.cpp:880-1088)
 *   scx_unconstrained_flow  <- UnconstrainedMap::flow (scisim/UnconstrainedMaps/UnconstrainedMap.h:33)
 *                              for SymplecticEulerMap (ball2d/SymplecticEulerMap.cpp:21-32), VerletMap
 *                              (ball2d/VerletMap.cpp:15-41) and the translational part of SplitHamMap
 *                              (rigidbody3d/UnconstrainedMaps/SplitHamMap.cpp:17-182)
 */
#ifndef SCX_H
#define SCX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SCX_OK 0
#define SCX_E_ARG (-1)      /* invalid argument */
#define SCX_E_HIP (-2)      /* HIP runtime error */
#define SCX_E_CAPACITY (-3) /* a capacity of the context was exceeded */
#define SCX_E_STATE (-4)    /* call out of order (e.g. step before scx_set_bodies) */
#define SCX_E_COLOURS (-5)  /* contact graph needs more than 64 colours */

#define SCX_INTEGRATOR_SYMPLECTIC_EULER 0
#define SCX_INTEGRATOR_VERLET 1

/* Staggering of the impact-friction map: the "staggering" attribute of the reference's
 * <sobogus_friction_solver> (rigidbody3dutils/RigidBody3DSceneParser.cpp:1865-1962). */
#define SCX_STAGGER_GEOMETRIC 0        /* GeometricImpactFrictionMap::flow (GeometricImpactFrictionMap.cpp:267-442):
                                          solve on v0, then the unconstrained map from q0 with v2 */
#define SCX_STAGGER_SYMPLECTIC_EULER 1 /* SymplecticEulerImpactFrictionMap::flow
                                          (SymplecticEulerImpactFrictionMap.cpp:201-407): the force
                                          increment dt M^-1 F(q0) enters the solve, q1 = q0 + dt v1;
                                          needs CoR = 0 (RigidBody3DSceneParser.cpp:1881-1888) */
#define SCX_STAGGER_STABILIZED 2       /* StabilizedImpactFrictionMap::flow (StabilizedImpactFrictionMap.cpp:48-183):
                                          cold solve on the predictor velocity, q1 = q0 + dt v1; no
                                          impulse cache (warm_start must be SCX_WARM_NONE) */

#define SCX_PRECISION_FP64 0       /* every sweep in FP64 */
#define SCX_PRECISION_FP32_POLISH 1 /* FP32 sweeps, then FP64 polishing sweeps to tolerance */

/* Impulse cache between steps (ImpulsesToCache, GeometricImpactFrictionMap.cpp:174-264):
 * the previous step's impulses of the same constraint, (i, j) or (plane, i), warm start
 * the solve; constraints new this step start from zero. */
#define SCX_WARM_NONE 0      /* ImpulsesToCache::NONE */
#define SCX_WARM_CACHE 1     /* ImpulsesToCache::NORMAL_AND_FRICTION: alpha and beta (2D only, as the reference) */
#define SCX_WARM_NORMAL 2    /* ImpulsesToCache::NORMAL: alpha only (2D and 3D) */
#define SCX_WARM_TRANSPORT 3 /* world impulse rotated with the contact normal, then projected on the
                                new basis (the 3D cache of SymplecticEulerImpactFrictionMap.cpp:87-127) */
#define SCX_WARM_LEARNED 4   /* impulses predicted per contact by a 16-128-128-3 BF16 perceptron from
                                the contact's features and its transported cached impulse; no
                                counterpart in the reference (the solve still runs to tolerance) */

/* Parameters of the warm-start perceptron (FP32): W1 [16][128], b1 [128], W2 [128][128],
 * b2 [128], W3 [128][3], b3 [3], row-major, y = W3^T relu(W2^T relu(W1^T x + b1) + b2) + b3.
 * The prediction is y + x[8:11]: the network learns a correction to the transported cached
 * impulse (features 8-10), so a zero network reproduces SCX_WARM_TRANSPORT. */
#define SCX_MLP_PARAMS 19075

typedef struct scx_ctx scx_ctx;

typedef struct scx_config {
  int32_t dim;          /* 2 (ball2d) or 3 (rigidbody3d spheres) */
  int32_t device;       /* HIP device ordinal */
  uint32_t max_bodies;  /* body capacity */
  uint32_t max_planes;  /* static plane capacity (grows on demand; at most 64 per scene) */
  uint32_t max_contacts; /* contact capacity; 0 = 8 * max_bodies */
  uint32_t max_pairs;    /* broad-phase pair capacity for scx_potential_overlaps; 0 = 16 * max_bodies */
} scx_config;

/* Step parameters: the <integrator> and <sobogus_friction_solver> options of the
 * reference's scene files (ball2dutils/Ball2DSceneParser.cpp:587-640, 1193-1424). */
typedef struct scx_step_params {
  double dt;
  double g[3];          /* gravity; force = m g (ball2d/Forces/Ball2DGravityForce.cpp:36-46) */
  double mu;            /* friction coefficient, all contacts */
  double CoR;           /* coefficient of restitution, all contacts */
  double tol;           /* residual tolerance: max over contacts of ||FB||^2 < tol */
  uint32_t max_iters;   /* Gauss-Seidel sweep limit */
  uint32_t eval_every;  /* residual evaluation period in sweeps */
  int32_t integrator;   /* SCX_INTEGRATOR_* */
  int32_t precision;    /* SCX_PRECISION_* */
  int32_t warm_start;   /* SCX_WARM_* */
  uint32_t iteration;   /* step index (informational) */
  int32_t staggering;   /* SCX_STAGGER_* (0 = geometric) */
  int32_t stabilize;    /* symplectic Euler: add min(0, gap(q0)) / dt to each normal forcing term
                           (SymplecticEulerImpactFrictionMap.cpp:309-327) */
  double penetration_threshold; /* symplectic Euler with stabilize: gaps of smaller magnitude count as 0 */
} scx_step_params;

typedef struct scx_step_stats {
  uint64_t nentries;    /* grid entries (cells covered by all AABBs) */
  uint32_t npairs;      /* body-body contacts after the exact narrow phase */
  uint32_t nplane_contacts;
  uint32_t ncontacts;   /* active set size */
  uint32_t ncolours;
  uint32_t sweeps;      /* total GS sweeps (FP32 + FP64 phases) */
  uint32_t sweeps_fp32;
  double error;         /* best residual, max ||FB||^2 */
  int32_t succeeded;    /* error < tol */
  uint32_t ncells;      /* cells of the broad-phase grid */
} scx_step_stats;

const char* scx_version(void);
const char* scx_last_error(const scx_ctx* ctx);

int scx_create(const scx_config* cfg, scx_ctx** out);
void scx_destroy(scx_ctx* ctx);

/* SpatialGridDetector::getPotentialOverlaps: all pairs (i<j) of axis-aligned boxes that
 * overlap (a shared face counts), lexicographically ordered, as pairs_out[2k], pairs_out[2k+1].
 * aabb_min/aabb_max: n * dim doubles each. Returns SCX_E_CAPACITY with *npairs set to the
 * required count if cap is too small. */
int scx_potential_overlaps(scx_ctx* ctx, uint32_t n, const double* aabb_min, const double* aabb_max,
                           uint64_t cap, uint32_t* pairs_out, uint64_t* npairs);

/* scx_potential_overlaps without any allocation. No counterpart in the reference: this is for
 * callers that cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc). Same arguments and results as
 * scx_potential_overlaps, bit for bit, but the broad phase's working memory is the library's
 * file-scope scratch array (40 MiB of device memory on each device the library runs on, held
 * until the process ends) instead of the context's buffers, and nothing is allocated: at most
 * 65,536 boxes, 16 (3D) or 8 (2D) cell entries per box and 1,048,576 pairs; beyond these it
 * returns SCX_E_CAPACITY (with *npairs set when the pair count is the limit). Every call on a
 * device shares that device's array: calls that overlap in time on one device (from several
 * threads or contexts) race and must be serialised by the caller. */
int scx_potential_overlaps_noalloc(scx_ctx* ctx, uint32_t n, const double* aabb_min, const double* aabb_max,
                                   uint64_t cap, uint32_t* pairs_out, uint64_t* npairs);

/* Scene state, in the reference's layouts: q, v = dim*n interleaved (ball2d q, v;
 * rigidbody3d first 3n entries of q and v); omega = 3n angular velocities (3D; NULL = 0);
 * r radii; m masses; inertia = isotropic sphere inertia (3D; NULL = 2/5 m r^2). */
int scx_set_bodies(scx_ctx* ctx, uint32_t n, const double* q, const double* v, const double* omega,
                   const double* r, const double* m, const double* inertia);
/* Static planes: point x and normal n (normalised here as StaticPlane does), dim each.
 * At most 64 for one scene; an ensemble (scx_set_scenes) has at most 64 per scene. */
int scx_set_planes(scx_ctx* ctx, uint32_t nplanes, const double* x, const double* n);

/* Ensemble of independent scenes stepped together (no counterpart in the reference,
 * which runs one simulation per process; this is for callers that batch many small
 * scenes onto one GPU). Scene k owns bodies [body_off[k], body_off[k+1]) of the arrays
 * given to scx_set_bodies and static planes [plane_off[k], plane_off[k+1]) of those given
 * to scx_set_planes (at most 64), and has friction coefficient mu[k] (scx_step_params.mu
 * is then unused). Bodies of different scenes never interact; within a scene the active
 * set, its order and the contact data are those of the scene alone. The Gauss-Seidel
 * solve runs until every scene meets the tolerance; a scene that meets it is frozen (its
 * contacts are no longer updated) for the rest of the solve. Call after scx_set_bodies and scx_set_planes; a new body or plane count
 * clears the ensemble; nscenes = 0 returns to one scene. */
int scx_set_scenes(scx_ctx* ctx, uint32_t nscenes, const uint32_t* body_off, const uint32_t* plane_off,
                   const double* mu);
/* Per-scene results of the last step of an ensemble: contacts, residual (max over the
 * scene's contacts of ||FB||^2, in FP64) and the sweep at which the scene met the tolerance
 * (its contacts are not updated after that; a scene that never met it reports the step's
 * total sweeps). Any array may be NULL. */
int scx_get_scene_stats(scx_ctx* ctx, uint32_t nscenes, uint32_t* ncontacts, double* error, uint32_t* sweeps);
int scx_get_state(scx_ctx* ctx, double* q, double* v, double* omega);

/* One impact-friction map step on the resident state (GeometricImpactFrictionMap, or the
 * staggering named by p->staggering). stats may be NULL. */
int scx_step(scx_ctx* ctx, const scx_step_params* p, scx_step_stats* stats);
/* Several steps without host synchronisation between them except where the
 * pipeline needs counts; stats (may be NULL) receives the last step's stats. */
int scx_run(scx_ctx* ctx, const scx_step_params* p, uint32_t nsteps, scx_step_stats* stats);

/* Last step's active set in the reference's order (body-body pairs lexicographic, then
 * plane contacts plane-major): body0, body1 (-1 for planes), plane (-1 for body pairs),
 * bases (dim*dim per contact, column-major [n s t]), alpha (normal impulse), beta
 * ((dim-1) per contact), colour. Any output pointer may be NULL. */
int scx_get_contacts(scx_ctx* ctx, uint32_t cap, int32_t* body0, int32_t* body1, int32_t* plane, double* bases,
                     double* alpha, double* beta, int32_t* colour, uint32_t* ncontacts);
/* Last step's contacts in the layout of the reference's force output
 * (ImpactFrictionMap::exportConstraintForcesToBinaryFile, scisim/ConstrainedMaps/ImpactFrictionMap.cpp:58-163,
 * datasets collision_indices, collision_points, collision_normals, collision_forces):
 * indices 2 per contact (body0, then body1 or -plane - 2 for a static plane), points dim per
 * contact (the world-space contact point at q0), normals dim, forces dim (the impulse
 * alpha n + sum_k beta_k t_k, as the reference writes it). All arrays must be given. */
int scx_get_contact_forces(scx_ctx* ctx, uint32_t cap, int32_t* indices, double* points, double* normals,
                           double* forces, uint32_t* ncontacts);
/* Last step's post-impact velocity v2 (dim*n; 3D: plus omega2 3n if not NULL). Under the
 * symplectic Euler and stabilized staggerings this is the step's final velocity v1. */
int scx_get_v2(scx_ctx* ctx, double* v2, double* omega2);

/* An explicit active set for scx_friction_solve, in the reference's conventions
 * (Sobogus::solve inputs, scisim/ConstrainedMaps/Sobogus.cpp:880-1088): body1 = -1 for
 * static geometry; bases column-major [n s t] per contact (right-handed, orthonormal,
 * as Constraint::computeBasis asserts, scisim/Constraints/Constraint.cpp:10-16);
 * 3D arms = contact point - body centre at q0 for each body (sphere arms are parallel to n);
 * nrel_extra is added to the normal forcing term; drel_extra must be NULL. */
typedef struct scx_contact_set {
  uint32_t ncontacts;
  const int32_t* body0;
  const int32_t* body1;
  const double* bases;
  const double* arm0;      /* 3D: 3 per contact; NULL in 2D */
  const double* arm1;      /* 3D: 3 per contact; NULL in 2D */
  const double* mu;        /* per contact */
  const double* CoR;       /* per contact, NULL = 0 */
  const double* nrel_extra;
  const double* drel_extra;
} scx_contact_set;

/* FrictionSolver::solve (scisim/ConstrainedMaps/FrictionSolver.h:28; Sobogus::solve,
 * scisim/ConstrainedMaps/Sobogus.cpp:880-1088) on an explicit active set: masses m (and
 * 3D isotropic inertia), start velocities v0 (dim per body; 3D omega0 3 per body). alpha
 * (ncontacts) and beta ((dim-1) per contact) carry the warm start in and the impulses out.
 * Outputs: v2 (and omega2) = v0 + M^-1 H^T r; f = friction-only generalized impulse
 * ([lin | ang] per body in 3D); succeeded = error < tol; error = max ||FB||^2; sweeps.
 * Replaces any resident scene (call scx_set_bodies again before stepping). */
int scx_friction_solve(scx_ctx* ctx, uint32_t nbodies, const double* m, const double* inertia, const double* v0,
                       const double* omega0, const scx_contact_set* contacts, double tol, uint32_t max_iters,
                       uint32_t eval_every, int32_t precision, double* alpha, double* beta, double* v2,
                       double* omega2, double* f, int32_t* succeeded, double* error, uint32_t* sweeps);

/* Learned warm start (SCX_WARM_LEARNED); no counterpart in the reference.
 * A sample is one contact: 16 features (normal; pre-solve relative velocity in the contact
 * frame over |g| dt; w_t / w_n; mu; transported previous impulse over s and its presence
 * flag; overlap over the radius sum; plane flag; 0; 1) and 3 targets, the converged local
 * impulse over s = |g| dt / w_n. */
int scx_set_warm_model(scx_ctx* ctx, const float* params, uint32_t n);  /* n = SCX_MLP_PARAMS */
int scx_get_warm_model(scx_ctx* ctx, float* params, uint32_t n);
/* The last step's samples (that step must have used SCX_WARM_TRANSPORT or SCX_WARM_LEARNED):
 * features 16 and targets 3 per contact, in active-set order. */
int scx_get_samples(scx_ctx* ctx, uint32_t cap, float* features, float* targets, uint32_t* count);
/* Mean-squared-error loss of the prediction, mean of (y + x[8:11] - t)^2, and its gradient
 * (SCX_MLP_PARAMS, the layout above) over n samples with the current model, from the
 * device's BF16 forward and backward kernels. */
int scx_warm_model_gradients(scx_ctx* ctx, uint32_t n, const float* features, const float* targets, float* grads,
                             double* loss);
/* Adam training of the current model on the device: `epochs` passes over the n samples in
 * shuffled batches of `batch` (seeded); loss = the final MSE on the first batch of samples. */
int scx_train_warm_model(scx_ctx* ctx, uint32_t n, const float* features, const float* targets, uint32_t epochs,
                         uint32_t batch, float lr, uint64_t seed, double* loss);

/* Checkpoint and resume (the serializeSystem / deserializeSystem pair of
 * ball2dcli/ball2d_cli.cpp:199-316 for this path, with ConstraintCache::serialize,
 * ball2d/ConstraintCache.cpp:125-175, and the maps' serialize,
 * GeometricImpactFrictionMap.cpp:33-44): the resident bodies (q, v, omega, r, m, inertia),
 * the static planes, the ensemble layout, the impulse cache and the warm-start model with
 * its Adam moments, as one byte image (a 64-byte header with magic "SCXCKPT1", version,
 * dimension and counts, then the arrays, guarded by a hash). A context of the same
 * dimension that loads it steps bit-identically to the one that saved it (the step counter
 * and the step parameters stay with the caller, as the reference's CLI serializes its own).
 * scx_save_checkpoint: buf == NULL sets *len to the size needed and returns SCX_OK; a
 * smaller cap returns SCX_E_CAPACITY with *len set. scx_load_checkpoint replaces the
 * context's scene (growing its buffers as needed); a malformed or corrupt image returns
 * SCX_E_ARG and leaves the context unchanged. */
int scx_save_checkpoint(scx_ctx* ctx, void* buf, uint64_t cap, uint64_t* len);
int scx_load_checkpoint(scx_ctx* ctx, const void* buf, uint64_t len);
/* Validates a checkpoint image on the host (no device, no context) and reports its
 * contents; any output may be NULL. */
int scx_checkpoint_info(const void* buf, uint64_t len, int32_t* dim, uint32_t* nbodies, uint32_t* nplanes,
                        uint32_t* nscenes, uint32_t* cache_entries, int32_t* has_model);

/* UnconstrainedMap::flow on host arrays (q0, v0 -> q1, v1), gravity only. */
int scx_unconstrained_flow(scx_ctx* ctx, int32_t integrator, uint32_t n, const double* m, const double* g, double dt,
                           const double* q0, const double* v0, double* q1, double* v1);

/* Per-phase device time of the last step, from HIP events on the context stream
 * (enabled with scx_set_timing). names/ms: up to cap entries; returns count. */
int scx_set_timing(scx_ctx* ctx, int32_t enable);
int scx_get_timing(scx_ctx* ctx, uint32_t cap, const char** names, double* ms, uint32_t* count);

/* Counters of the sweeps of the last step that returned stats. No counterpart in the
 * reference (which evaluates after every eval_every-th sweep and sweeps colour by colour);
 * they give the sweep kernel's byte count and show which paths a step took. */
typedef struct scx_sweep_counters {
  uint32_t full_evals;          /* residual evaluations that were full passes over the contacts */
  uint32_t full_evals_fp32;     /* of which in the FP32 phase */
  uint32_t skipped_evals;       /* FP32 evaluations replaced by the check of the contacts the last
                                   full one listed as near its maximum */
  uint32_t component_launches;  /* sweep launches that swept their late colours as components */
  uint32_t component_overflows; /* launches whose components overflowed a workgroup's list */
  uint32_t component_width;     /* most contacts of one colour in a workgroup's list */
  double component_share;       /* percent of the contacts the next step's component phase may hold */
} scx_sweep_counters;
int scx_get_sweep_counters(const scx_ctx* ctx, scx_sweep_counters* out);

/* Online autotuning of the persistent Gauss-Seidel launches (no counterpart in the
 * reference, whose So-bogus solve has no launch shape; this is for callers that step
 * one kind of scene many times). Enabled, each sweep launch runs on the cost model's
 * workgroup count times 2^e, e in [-3, 3]; per class of launch (dimension, precision
 * phase, ensemble or not, colours, contacts in half-octaves) the tuner times launches
 * at neighbouring e with HIP events, climbs while the device time per sweep falls by
 * more than 3 %, then keeps the fastest. Launches with fewer than 16 sweeps are not
 * timed. Results do not depend on it: every width gives the same iterates. Off by
 * default; the table lives as long as the context. */
int scx_set_autotune(scx_ctx* ctx, int32_t enable);
/* The tuner's table, one entry per class: key (bits 0-7 contacts in half-octaves, 8 3D,
 * 9 FP64 phase, 10 ensemble, 11-15 barrier intervals per sweep / 2), the exponent in use,
 * its workgroups at the last launch, its least time per sweep in microseconds (< 0: not
 * timed yet) and whether the class has settled. Up to cap entries; count = all. */
int scx_get_autotune(scx_ctx* ctx, uint32_t cap, uint32_t* keys, int32_t* exponent, int32_t* blocks,
                     float* us_per_sweep, int32_t* settled, uint32_t* count);

#ifdef __cplusplus
}
#endif

#endif /* SCX_H */
