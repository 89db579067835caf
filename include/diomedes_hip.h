/*
 * diomedes_hip.h — C ABI of the MI355X-native diomedes physics step.
 *
 * The step is the reference engine's broad phase, narrow phase, contact
 * build (raw contacts, container walls, warm start, CSR) and AVBD solve
 * (init, I x (primal, dual), recover).  Two families of entry points:
 *
 *  - granular calls that replace the reference's Vulkan compute objects
 *    one for one (NarrowPhaseCompute and AvbdCompute behind the six
 *    Renderer physics methods, engine/src/render/renderer.rs:75-175);
 *  - a fused, device-resident world (dio_world_upload / dio_step /
 *    dio_world_download) that runs the demo step protocol
 *    (demos/avbd_box/src/main.rs:189-272) entirely on the GPU, replacing
 *    AvbdSolver::prepare_contacts and the host decode/serialise loops.
 *
 * Records are byte-identical to the reference's #[repr(C)] structs so a Rust
 * binding can pass its existing slices unchanged.  Every function returns a
 * dio_status; no function aborts or throws across the ABI.  A dio_ctx is not
 * thread-safe (it mirrors the reference's `&mut self` objects); separate
 * contexts may be used from separate threads, each owns one HIP stream.
 * Submit calls stage host inputs before returning, so callers may free them
 * immediately; read calls block until the GPU work they read is complete.
 */
#ifndef DIOMEDES_HIP_H
#define DIOMEDES_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DIO_ABI_VERSION 2u

/* `b` of a contact against the container box (physics/gpu.rs:15, avbd/src/lib.rs:75) */
#define DIO_CONTAINER 0xFFFFFFFFu

typedef enum dio_status {
    DIO_OK = 0,
    DIO_CAPACITY = 1,  /* a buffer limit was exceeded; dio_required_capacity names the size */
    DIO_INVALID = 2,   /* bad argument (null pointer, out-of-range index, zero iterations, ...) */
    DIO_HIP = 3,       /* a HIP runtime call failed; dio_last_error has the message */
    DIO_NOT_READY = 4  /* read before the matching submit */
} dio_status;

/* shape tags (compute.rs:37-39; BOX and HULL extend the reference) */
enum {
    DIO_SHAPE_SPHERE = 0,      /* corners[0][0] = radius */
    DIO_SHAPE_CUBE = 1,        /* corners[0][0..3] = [h, h, h]; SAT and walls read only [0][0] */
    DIO_SHAPE_TETRAHEDRON = 2, /* corners[0..4] = the four local corners */
    DIO_SHAPE_BOX = 3,         /* corners[0][0..3] = per-axis half extents (extension) */
    DIO_SHAPE_HULL = 4         /* corners[0][0] = first vertex, [0][1] = vertex count, as
                                  floats holding exact integers, into the hull vertex pool
                                  given to dio_world_upload_hulls (extension).  A range
                                  that is not whole, has a count below 1, or reaches
                                  past the uploaded pool is DIO_INVALID (at upload for a
                                  malformed range; at dio_step and the granular broad and
                                  narrow submits for one past the pool), before any launch */
};

/* constraint kinds (avbd/src/lib.rs:76-79) */
enum {
    DIO_CONSTRAINT_SPRING = 0,
    DIO_CONSTRAINT_ROD = 1,
    DIO_CONSTRAINT_BALL_SOCKET = 2,
    DIO_CONSTRAINT_ROPE = 3,
    /* Hard joints: no counterpart in the reference, whose joints are the soft
       kinds above (README.md:11 lists hard joints as missing).  The rod,
       socket and rope geometry held to the constraint by a multiplier and a
       penalty per row, with the contacts' augmented-Lagrangian updates and
       warm start; the stiffness field is unused.  World path only (not
       dio_avbd_submit); between two dynamic bodies they need dio_step
       mode 1. */
    DIO_CONSTRAINT_HARD_ROD = 4,
    DIO_CONSTRAINT_HARD_BALL_SOCKET = 5,
    DIO_CONSTRAINT_HARD_ROPE = 6,
    /* Bending joints (no counterpart in the reference; README.md:11 lists
       bending as missing): hold b's orientation at a's times a rest rotation,
       whose quaternion (x, y, z, w) is anchor_a; three angular rows, with the
       stiffness anchor_b[3] (BEND) or multipliers (HARD_BEND, which also needs
       mode 1 between two dynamic bodies).  World path only. */
    DIO_CONSTRAINT_BEND = 7,
    DIO_CONSTRAINT_HARD_BEND = 8
};

/* GpuShape, 80 bytes (compute.rs:28-34) */
typedef struct dio_gpu_shape {
    uint32_t tag;
    uint32_t pad[3];
    float corners[4][4];
} dio_gpu_shape;

/* GpuBody, 112 bytes (compute.rs:78-82) */
typedef struct dio_gpu_body {
    float position[4];
    float orientation[4]; /* quaternion x, y, z, w */
    dio_gpu_shape shape;
} dio_gpu_body;

/* GpuPair, 8 bytes (compute.rs:97-100) */
typedef struct dio_gpu_pair {
    uint32_t a;
    uint32_t b;
} dio_gpu_pair;

/* GpuContactOut, 80 bytes (compute.rs:123-133) */
typedef struct dio_gpu_contact_out {
    uint32_t valid;
    uint32_t a;
    uint32_t b;
    uint32_t pad;
    float normal[4];
    float depth;
    float pad2[3];
    float point_a[4];
    float point_b[4];
} dio_gpu_contact_out;

/* GpuAvbdBodyState, 80 bytes (gpu_physics.rs:18-26) */
typedef struct dio_avbd_body_state {
    float vel[4];
    float ang[4];
    float prev_vel[4];
    float inv_moment[4];
    float inv_mass;
    float friction;
    float pad[2];
} dio_avbd_body_state;

/* GpuAvbdContact, 144 bytes (gpu_physics.rs:31-44) */
typedef struct dio_avbd_contact {
    uint32_t a;
    uint32_t b;
    float friction;
    float pad;
    float normal[4];
    float tangent1[4];
    float tangent2[4];
    float r_a[4];
    float r_b[4];
    float c0[4];
    float penalty[4];
    float lambda[4];
} dio_avbd_contact;

/* GpuAvbdConstraint, 48 bytes (gpu_physics.rs:49-56): anchor_a[3] is the
   rest / max length (0 for sockets), anchor_b[3] the stiffness */
typedef struct dio_avbd_constraint {
    uint32_t a;
    uint32_t b;
    uint32_t kind;
    uint32_t pad;
    float anchor_a[4];
    float anchor_b[4];
} dio_avbd_constraint;

/* AvbdRunOptions (gpu_physics.rs:76-83) */
typedef struct dio_run_options {
    float dt;
    float gravity[3];
    float alpha;
    float beta_lin;
    float penalty_max;
    uint32_t iterations;
} dio_run_options;

/* AvbdOptions (physics/gpu.rs:127-138); defaults in dio_default_options */
typedef struct dio_avbd_options {
    float dt;
    float gravity[3];
    uint32_t iterations;
    float alpha;
    float beta_lin;
    float margin;
    float penalty_min;
    float penalty_max;
    float gamma;
    float friction;
} dio_avbd_options;

/* AvbdContainer (physics/gpu.rs:117-123) plus a presence flag (the
   reference's Option<AvbdContainer>) */
typedef struct dio_container {
    float center[3];
    float rotation[4]; /* x, y, z, w */
    float half_extent;
    uint32_t enabled;
} dio_container;

/* Buffer capacities; zero fields take defaults derived from max_bodies. */
typedef struct dio_limits {
    uint32_t max_bodies;
    uint32_t max_pairs;
    uint32_t max_contacts;
    uint32_t max_constraints;
    uint32_t max_hull_vertices;
} dio_limits;

/* Step protocol options for the fused path. */
typedef struct dio_step_opts {
    uint32_t contact_lag; /* 1: the reference demos' protocol (contacts found on the
                             previous state, lever arms on the current one); 0: synchronous.
                             A lag-1 step after a lag-0 step builds its contacts from the
                             narrow output that step found on the poses at its start, as
                             if it had run with lag 1 */
    uint32_t mode;        /* 0: Jacobi sweep with ping-pong pose buffers (reference);
                             1: coloured Gauss-Seidel (no counterpart in the reference): the
                             step's body graph (contacts and joints between dynamic bodies) is
                             coloured, and each iteration runs the primal colour by colour on
                             one pose buffer updated in place, then the dual update.  The step
                             reads the colour counts back once (not graph-captured). */
    uint32_t profile;     /* 1: launch eagerly and time every AVBD kernel with HIP events */
} dio_step_opts;

/* Counters and times of the last dio_step call. */
typedef struct dio_step_stats {
    uint32_t steps;
    uint32_t pairs;            /* candidate pairs of the last step's broad phase */
    uint32_t contacts;         /* contacts solved in the last step (narrow + container) */
    uint32_t csr_nnz;          /* CSR entries of the last step */
    uint32_t narrow_contacts;  /* narrow-phase part of `contacts` */
    uint32_t dynamic_bodies;
    uint64_t sum_pairs;        /* sums over the steps of the call */
    uint64_t sum_contacts;
    uint64_t sum_csr_nnz;
    float ms_total;            /* GPU time of the whole call (events around all steps) */
    float ms_avbd;             /* summed AVBD time: init, I primal launches (dual updates fused into
                                  iterations 1 .. I-1), final dual + recover */
    /* profile mode only: summed kernel times and launch counts.  ms_dual / n_dual
       count standalone dual launches (none: the dual updates are fused), ms_recover /
       n_recover the final dual + recover launch. */
    float ms_init, ms_primal, ms_dual, ms_recover;
    uint32_t n_init, n_primal, n_dual, n_recover;
    uint32_t colours;          /* mode 1: colours and colouring rounds of the last step */
    uint32_t colour_rounds;
} dio_step_stats;

/* Kernel launch shapes (the autotuner's search space). */
#define DIO_TUNE_PRIMAL_SHAPES 6
#define DIO_TUNE_BROAD_SHAPES 6
#define DIO_TUNE_COLOUR_SHAPES 6
#define DIO_TUNE_SLOTS (DIO_TUNE_PRIMAL_SHAPES + DIO_TUNE_BROAD_SHAPES + 4 + DIO_TUNE_COLOUR_SHAPES + 2)
typedef struct dio_tuning {
    uint32_t primal_lanes;     /* lanes per body in the primal: 1, 4, 8, 16, 32 or 64 (0 = default 32) */
    uint32_t broad_lanes;      /* lanes per query in the BVH traversal: 1, 2, 4, 8, 16 or 32; with 16
                                  or 32 a step pops 2 or 4 nodes.  0 (default): by world size, 16 up
                                  to 65,536 bodies (a short chain of dependent steps), 8 above (the
                                  GPU is full and extra lanes only add work) */
    uint32_t resort_interval;  /* fused step: Morton re-sort every this many steps; the tree is refit
                                  over the previous order in between (0 = default 16) */
    uint32_t fuse_dual;        /* 1: each primal iteration k >= 1 applies dual update k - 1 itself
                                  (fewer launches, both bodies of a contact evaluate it; default);
                                  0: a dual launch after each primal */
    uint32_t adjoint_lanes;    /* lanes per body in the adjoint's reverse primal: 1, 2, 4, 8 or 16
                                  (0 = default 8) */
    uint32_t overlap_broad;    /* 1: with contact lag 1, the next step's broad and narrow phases run
                                  on a second stream beside this step's contact build and CSR (a
                                  shorter step; the second hardware queue slows the dispatch of the
                                  AVBD launches that follow); 0 (default): after them, on one stream */
    uint32_t primal_order;     /* the order the Jacobi primal takes the bodies in: 0, body index; 1
                                  (default), the broad phase's Morton order with each XCD given one
                                  contiguous range, so bodies in contact share an L2 (every body's
                                  result is the same either way; on the settled 8192-box field the
                                  primal's L2 misses halve).  The coloured mode takes each colour's
                                  list */
    uint32_t scene_solve;      /* a batched world (scene table) without joints: N > 0 runs every primal,
                                  dual and the final dual + recover of a step's solve in one launch, each
                                  workgroup taking bundles of whole consecutive scenes (at least N
                                  bodies each, or one larger scene) from a queue, most bodies first (the
                                  result is the same either way); 0 (default): one launch per primal
                                  and dual over the whole world.  At most 65536.  Ignored for other
                                  worlds */
    uint32_t colour_lanes;     /* lanes per body in the coloured mode's primal (one launch per colour):
                                  1, 4, 8, 16, 32 or 64; 0 (default): primal_lanes */
    uint32_t primal_packed;    /* the Jacobi primal of dio_step: 1 (default), packed -- a workgroup of 256
                                  threads takes 8 bodies, every thread builds the rows of one item of their
                                  joined item list (so a body with many contacts takes no second round
                                  of loads), and 32 lanes per body add its block; primal_lanes is then
                                  unused by the step.  2: span -- a workgroup takes 256 consecutive
                                  bodies, one thread each, builds the rows of their contiguous range of
                                  the contact CSR across all its threads, and each thread adds its own
                                  body's rows (bodies with few contacts; a world with constraints takes
                                  1 instead).  0: primal_lanes lanes per body.  The result is the same
                                  every way.  The granular solve, the coloured mode and the adjoint's
                                  replay take primal_lanes */
} dio_tuning;

typedef struct dio_ctx dio_ctx;

/* ---- library / context ------------------------------------------------ */
uint32_t dio_abi_version(void);
void dio_default_options(dio_avbd_options* out); /* AvbdOptions::default(), physics/gpu.rs:140-155 */
dio_status dio_create(int device, const dio_limits* limits, dio_ctx** out);
void dio_destroy(dio_ctx* ctx);
const char* dio_last_error(const dio_ctx* ctx);
uint64_t dio_required_capacity(const dio_ctx* ctx); /* size asked for by the last DIO_CAPACITY */
dio_status dio_synchronize(dio_ctx* ctx);

/* ---- broad + narrow phase (NarrowPhaseCompute, render/compute.rs:490-750) ---- */
/* Replaces submit_broad (compute.rs:490-567): upload bodies, compute AABBs and the
   candidate pairs.  The pair list is the set {(i, j): i < j, closed AABB overlap}
   sorted lexicographically, without the reference's 16384-pair truncation. */
dio_status dio_broad_submit(dio_ctx* ctx, const dio_gpu_body* bodies, uint32_t n);
/* Replaces read_pair_count (compute.rs:570-579). */
dio_status dio_broad_count(dio_ctx* ctx, uint32_t* count);
/* Replaces read_pairs (compute.rs:583-598); DIO_CAPACITY when cap is too small. */
dio_status dio_broad_read_pairs(dio_ctx* ctx, dio_gpu_pair* pairs, uint32_t cap, uint32_t* n);
/* Replaces read_aabbs-style parity reads: 2 float4 (min, max) per body. */
dio_status dio_broad_read_aabbs(dio_ctx* ctx, float* aabbs, uint32_t cap_bodies);
/* Replaces submit_narrow (compute.rs:601-671): narrow phase over the first
   pair_count pairs of the last broad phase. */
dio_status dio_narrow_submit(dio_ctx* ctx, uint32_t pair_count);
/* Replaces submit (compute.rs:404-477): narrow phase over caller-given pairs. */
dio_status dio_narrow_submit_pairs(dio_ctx* ctx, const dio_gpu_body* bodies, uint32_t n,
                                   const dio_gpu_pair* pairs, uint32_t pair_count);
/* Replaces read (compute.rs:675-693): 4 slots per pair, slot pair*4+k. */
dio_status dio_narrow_read(dio_ctx* ctx, dio_gpu_contact_out* slots, uint32_t cap_slots,
                           uint32_t* n_slots);

/* No counterpart in the reference.  For callers that cannot allocate (tight
   serving loops, code that must not call hipMalloc): dio_narrow_submit_pairs
   followed by dio_narrow_read in one synchronous call that needs no context
   and allocates nothing.  The same kernels give the same slots, bit for bit.
   The caller passes host arrays: n bodies, pair_count pairs (a, b < n), the
   hull vertex pool (x, y, z per vertex) and slots (cap_slots >= 4 * pair_count).
   Every other buffer is one fixed device array per GPU, held from the
   library's load until the process ends (DIO_NOALLOC_SCRATCH_BYTES, about
   45 MiB).  Limits: DIO_CAPACITY beyond DIO_NOALLOC_MAX_BODIES bodies,
   DIO_NOALLOC_MAX_PAIRS pairs or DIO_NOALLOC_MAX_HULL_VERTICES vertices, or
   when cap_slots is short; DIO_INVALID for a null array, a device outside
   0 .. 63, a pair index >= n, or a hull range that is malformed or reaches
   past hull_count.  Calls on one device share its array: a mutex per device
   serialises them, so calls from several threads do not overlap. */
#define DIO_NOALLOC_MAX_BODIES 16384u
#define DIO_NOALLOC_MAX_PAIRS 65536u
#define DIO_NOALLOC_MAX_HULL_VERTICES 16384u
#define DIO_NOALLOC_SCRATCH_BYTES 47453184u
dio_status dio_narrow_pairs_noalloc(int device, const dio_gpu_body* bodies, uint32_t n,
                                    const dio_gpu_pair* pairs, uint32_t pair_count,
                                    const float* hull_vertices, uint32_t hull_count,
                                    dio_gpu_contact_out* slots, uint32_t cap_slots);

/* ---- AVBD solve (AvbdCompute, render/gpu_physics.rs:439-622) ---------- */
/* Replaces AvbdCompute::submit.  positions: n x 3 floats, orientations: n x 4
   (x, y, z, w).  offsets: n + 1 CSR offsets, indices: nidx contact indices. */
dio_status dio_avbd_submit(dio_ctx* ctx, const dio_avbd_body_state* state,
                           const float* positions, const float* orientations, uint32_t n,
                           const dio_avbd_contact* contacts, uint32_t m,
                           const dio_avbd_constraint* constraints, uint32_t c,
                           const uint32_t* offsets, const uint32_t* indices, uint32_t nidx,
                           const dio_run_options* options);
/* Replaces AvbdCompute::read (AvbdGpuResult, gpu_physics.rs:86-94).  Any
   output pointer may be null.  positions n x 3, orientations n x 4, velocity
   arrays n x 3, lambda / penalty m x 3. */
dio_status dio_avbd_read(dio_ctx* ctx, float* positions, float* orientations,
                         float* velocities, float* angular_velocities, float* prev_velocities,
                         float* lambda, float* penalty);

/* ---- fused device-resident world (AvbdSolver + Renderer step) --------- */
/* Upload the world: bodies (pose + shape), their dynamics state, joints and
   the optional container.  Resets contacts and the lagged narrow output. */
dio_status dio_world_upload(dio_ctx* ctx, const dio_gpu_body* bodies,
                            const dio_avbd_body_state* state, uint32_t n,
                            const dio_avbd_constraint* constraints, uint32_t c,
                            const dio_container* container);
/* Batched world: num_scenes independent scenes in one world.  Scene s owns
   bodies scene_offsets[s] .. scene_offsets[s+1] - 1 (offsets run from 0 to
   n) and has its own container, containers[s].  Bodies of different scenes
   never form pairs and constraints may not join two scenes (DIO_INVALID).
   dio_step then steps every scene in the same launches, and each scene's
   trajectory is bit-identical to stepping it alone; dio_world_download and
   dio_world_contacts return the concatenation.  num_scenes <= 65536.  No
   counterpart in the reference (its demos step one scene per Renderer);
   replaces one plain context per scene. */
dio_status dio_batch_upload(dio_ctx* ctx, const dio_gpu_body* bodies,
                            const dio_avbd_body_state* state, uint32_t n,
                            const dio_avbd_constraint* constraints, uint32_t c,
                            const uint32_t* scene_offsets, uint32_t num_scenes,
                            const dio_container* containers);
/* Vertex pool for DIO_SHAPE_HULL bodies (x, y, z per vertex). */
dio_status dio_world_upload_hulls(dio_ctx* ctx, const float* vertices, uint32_t count);
/* Run nsteps physics steps on the device without host round trips.
   DIO_CAPACITY when a step needed more candidate pairs than max_pairs or more
   contacts than max_contacts: dio_required_capacity gives the largest count
   needed (pairs are reported before contacts) and dio_last_error names the
   limit.  The step then ran on truncated lists and no longer follows the
   reference; every later dio_step reports the same until the next upload. */
dio_status dio_step(dio_ctx* ctx, const dio_avbd_options* options, const dio_step_opts* opts,
                    uint32_t nsteps);
dio_status dio_step_get_stats(dio_ctx* ctx, dio_step_stats* out);
/* Launch shapes; changing them drops the captured step graphs. */
dio_status dio_set_tuning(dio_ctx* ctx, const dio_tuning* tuning);
dio_status dio_get_tuning(dio_ctx* ctx, dio_tuning* out);
/* Autotuner: after at least one dio_step, time every primal shape on the
   current contact set and every broad-phase traversal shape on the current
   poses (reps launches each, HIP events), keep the fastest of each, and report
   the average microseconds per launch in us_per_launch[0 .. DIO_TUNE_SLOTS - 1]:
   primal lanes 1, 4, 8, 16, 32, 64; broad lanes 1, 2, 4, 8, 16, 32 (a refit + traversal
   pass); then, at the chosen primal shape, an unfused primal plus a dual
   launch and a fused primal (fuse_dual is set to the faster); then that
   primal in index order and in Morton order (primal_order is set to the
   faster; the shapes above are timed in the order current at the call); then
   the coloured mode's primal, one launch per colour of a colouring of the
   current contact graph, at lanes 1, 4, 8, 16, 32, 64 (microseconds per
   iteration over all colours; colour_lanes is set to the fastest); then the
   packed primal at the chosen dual schedule and order; then (a world without
   constraints; else 0 is reported) the span primal in both schedules,
   reported as microseconds per iteration at its faster one, a dual launch
   included when unfused.  primal_packed is set to 1 or 2 for the packed or
   the span primal if its cost per iteration beats the chosen primal_lanes,
   else 0; the span primal brings its schedule into fuse_dual.  The
   primal with a fixed parity is idempotent, the broad timing runs in the
   granular broad buffers and the colouring in the coloured mode's buffers,
   so the world state is unchanged (dio_world_colours then reports that
   colouring); a granular broad result submitted before the call must be
   submitted again. */
dio_status dio_autotune(dio_ctx* ctx, const dio_avbd_options* options, uint32_t reps, dio_tuning* chosen,
                        float* us_per_launch);
/* ---- adjoint of the AVBD step (no counterpart in the reference) -------
   The differentiated function is the frozen-contact replay: each recorded
   step's contact set (geometry and the warm-start lambda / penalty as the
   solve received them) and CSR are kept, and only the AVBD solve runs, with
   every contact's friction set to options->friction.  At the recorded
   parameters it reproduces the recorded run bit for bit.  Worlds with joints
   are rejected (DIO_INVALID).

   dio_adjoint_begin: record the following dio_step calls (at most max_steps
   steps and max_recorded_contacts contacts over all of them; 0 = max_steps x
   max_contacts), with a state checkpoint every checkpoint_every steps.
   dio_adjoint_end frees the recording. */
dio_status dio_adjoint_begin(dio_ctx* ctx, uint32_t max_steps, uint32_t checkpoint_every,
                             uint64_t max_recorded_contacts);
dio_status dio_adjoint_end(dio_ctx* ctx);
dio_status dio_adjoint_steps(dio_ctx* ctx, uint32_t* steps);
/* What dio_adjoint_backward differentiates (a setting of the context, kept
   across recordings; default DIO_ADJOINT_EXACT):
   - DIO_ADJOINT_EXACT: the replay itself.  Within a step, the dual update
     feeds each iteration's constraint error back into the multipliers and
     the penalty ramp; over many contact-rich steps this derivative grows
     without bound (on the 64 x 32 wall throw, |d x / d v0| passes 1e3 by
     step 40 and 1e26 by step 120).
   - DIO_ADJOINT_FROZEN_MULTIPLIERS: the replay with the multipliers and
     penalties that each dual update writes held as constants (their values
     are the replay's own).  Forces still respond to the poses through the
     penalty term, so contacts stay stiff in the derivative; on the same
     wall |d x / d v0| stays below 1e4 over 120 steps in float64.
   Other values: DIO_INVALID. */
#define DIO_ADJOINT_EXACT 0u
#define DIO_ADJOINT_FROZEN_MULTIPLIERS 1u
dio_status dio_adjoint_set_mode(dio_ctx* ctx, uint32_t mode);
/* Replay and backward need at least one recorded step (DIO_NOT_READY
   otherwise) and a recording that did not overflow (DIO_CAPACITY).
   Replay every recorded step from the first checkpoint.  state0 (n records
   or null) replaces the velocities and masses of the recorded start.  Final
   positions (n x 3) and orientations (n x 4); the world is not modified. */
dio_status dio_adjoint_replay(dio_ctx* ctx, const dio_avbd_options* options,
                              const dio_avbd_body_state* state0, float* positions,
                              float* orientations);
/* Reverse mode through every recorded step (checkpointed: each segment is
   replayed once to tape its iterations), at the replay from state0 (null:
   the recorded run itself, whose checkpoints the recording holds; otherwise
   a forward sweep from state0 saves the segment checkpoints first, unless the
   last dio_adjoint_replay ran from the same state0, options and steps).
   Inputs: the adjoint of the final state (any pointer may be null = 0; n x 3
   or n x 4 floats).  Outputs (any may be null): the adjoint of the initial
   positions, orientations, velocities, angular and previous velocities, of
   each body's inverse mass (n) and inverse moments (n x 3), and of the
   friction coefficient.  A new dio_world_upload / dio_batch_upload frees the
   recording. */
dio_status dio_adjoint_backward(dio_ctx* ctx, const dio_avbd_options* options,
                                const dio_avbd_body_state* state0,
                                const float* d_positions, const float* d_orientations,
                                const float* d_velocities, const float* d_angular,
                                float* g_positions, float* g_orientations, float* g_velocities,
                                float* g_angular, float* g_prev_velocities, float* g_inv_mass,
                                float* g_inv_moment, double* g_friction);

/* Download body state (any pointer may be null): n x 3 / n x 4 floats. */
dio_status dio_world_download(dio_ctx* ctx, float* positions, float* orientations,
                              float* velocities, float* angular_velocities,
                              float* prev_velocities);
/* Contacts of the last step with their solved lambda / penalty (the reference's
   AvbdSolver::contacts() after sync_bodies_from_gpu). */
dio_status dio_world_contacts(dio_ctx* ctx, dio_avbd_contact* out, uint32_t cap, uint32_t* m);
/* The last coloured step's colour per body (0xFFFFFFFE: static, not in the
   graph; cap >= n) and the number of colours.  No counterpart in the
   reference. */
dio_status dio_world_colours(dio_ctx* ctx, uint32_t* colour, uint32_t cap, uint32_t* num_colours);
/* Each constraint's multipliers after the last step: lambda and penalty per
   row, c x 3 floats each (x, y, z rows of a hard socket or hard bend; x alone
   of a hard rod or rope; zero for soft joints, which carry none).  cap >= the constraint
   count, *c set to it.  No counterpart in the reference. */
dio_status dio_world_joints(dio_ctx* ctx, float* lambda, float* penalty, uint32_t cap, uint32_t* c);
/* Candidate pairs of the last step's broad phase (sorted, a < b). */
dio_status dio_world_pairs(dio_ctx* ctx, dio_gpu_pair* out, uint32_t cap, uint32_t* n);

/* ---- instanced-renderer hand-off --------------------------------------
   The reference's frame copies each body's pose into its scene instance
   (demos/avbd_box/src/main.rs:274-280), builds one model matrix per
   instance, Mat4::from_scale_rotation_translation(scale, orientation,
   position), in the renderer's instance order (instances grouped by shape
   and style, a stable sort: engine/src/render/renderer.rs:222-255), and
   uploads (projection * view * model, model) per instance into a uniform
   buffer of at most 256 instances (engine/src/render/context.rs:171-199,
   :18-21; engine/shaders/mesh.vert).  Here the pairs are written on the
   device, from the world's device-resident poses, into device memory the
   caller names (an imported renderer buffer, below), with no instance limit.

   dio_world_set_instances: scales (3 floats per body, the world's n bodies)
   and order (count entries: order[s] is the body drawn in instance slot s,
   each < n, any count; null: slot s draws body s, count <= n).  Copied to
   the device; kept until the next call or upload.  count 0 clears them.
   dio_world_instances: for every slot s < count, writes the two column-major
   4 x 4 float matrices (view_proj * model, model) -- mesh.vert's
   InstanceTransform, 128 bytes -- at (char*)dst + 128 * s, where view_proj
   (16 floats) is the caller's projection * view (column-major, as glam stores a Mat4)
   and model is built from the body's current pose and scale.  dst_bytes <
   128 * count is DIO_CAPACITY (the required size from
   dio_required_capacity).  Enqueued on the context's stream after its last
   step; dio_synchronize waits for it. */
dio_status dio_world_set_instances(dio_ctx* ctx, const float* scales, const uint32_t* order, uint32_t count);
dio_status dio_world_instances(dio_ctx* ctx, const float* view_proj, void* dst, uint64_t dst_bytes);
/* Device memory another API exported as a file descriptor -- a Vulkan
   buffer's memory exported with VK_EXTERNAL_MEMORY_HANDLE_TYPE_OPAQUE_FD_BIT
   (a DMA-BUF on AMD's drivers), or a DMA-BUF of another allocation -- mapped
   into the context's device: *dev_ptr addresses bytes [offset, offset + size)
   of the object whose total size is object_size.  The descriptor names a
   whole memory object: for a Vulkan buffer, the VkDeviceMemory it is bound
   to (vkGetMemoryFdKHR), with offset the buffer's bind offset in it.  The
   library imports a
   duplicate of fd, so the caller keeps fd and may close it at once.  The
   mapping lives until dio_release_external_memory (or dio_destroy).  fd < 0,
   size 0 or offset + size > object_size: DIO_INVALID. */
typedef struct dio_external_memory dio_external_memory;
dio_status dio_import_external_memory(dio_ctx* ctx, int fd, uint64_t object_size, uint64_t offset,
                                      uint64_t size, void** dev_ptr, dio_external_memory** out);
dio_status dio_release_external_memory(dio_ctx* ctx, dio_external_memory* mem);

#ifdef __cplusplus
}
#endif

#endif /* DIOMEDES_HIP_H */
