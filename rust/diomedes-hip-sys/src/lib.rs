//! Raw bindings of the MI355X physics step's C ABI (`include/diomedes_hip.h`,
//! library `libdiomedes_hip.so`).
//!
//! The engine's `Renderer` physics methods (`engine/src/render/renderer.rs:75-175`)
//! delegate to these functions in place of `NarrowPhaseCompute` and
//! `AvbdCompute`; the records are the engine's own `#[repr(C)]` records of
//! `render/compute.rs` and `render/gpu_physics.rs`, byte for byte (the aliases
//! `GpuBody`, `GpuContactOut`, `GpuAvbdBodyState`, `GpuAvbdContact` and
//! `GpuAvbdConstraint` name them as the engine does), so a caller passes its
//! existing slices.  Every function returns a `dio_status` and never aborts; a
//! `dio_ctx` is not thread-safe (one per thread, as the reference's `&mut self`).
#![allow(non_camel_case_types)]
#![no_std]

use core::ffi::{c_char, c_void};
use core::mem::size_of;

/// The library context (opaque; created by `dio_create`, freed by `dio_destroy`).
pub enum dio_ctx {}

/// Device memory imported from another API (opaque; `dio_import_external_memory`).
pub enum dio_external_memory {}

/// The status every call returns.
pub type dio_status = i32;

pub const DIO_OK: dio_status = 0;
pub const DIO_CAPACITY: dio_status = 1;
pub const DIO_INVALID: dio_status = 2;
pub const DIO_HIP: dio_status = 3;
pub const DIO_NOT_READY: dio_status = 4;

/// Shape tags of `dio_gpu_shape::tag` (compute.rs:37-39; BOX and HULL extend the reference).
pub const DIO_SHAPE_SPHERE: u32 = 0;
pub const DIO_SHAPE_CUBE: u32 = 1;
pub const DIO_SHAPE_TETRAHEDRON: u32 = 2;
pub const DIO_SHAPE_BOX: u32 = 3;
pub const DIO_SHAPE_HULL: u32 = 4;

/// Constraint kinds of `dio_avbd_constraint::kind` (avbd/src/lib.rs:76-79; 4-8 extend the reference).
pub const DIO_CONSTRAINT_SPRING: u32 = 0;
pub const DIO_CONSTRAINT_ROD: u32 = 1;
pub const DIO_CONSTRAINT_BALL_SOCKET: u32 = 2;
pub const DIO_CONSTRAINT_ROPE: u32 = 3;
pub const DIO_CONSTRAINT_HARD_ROD: u32 = 4;
pub const DIO_CONSTRAINT_HARD_BALL_SOCKET: u32 = 5;
pub const DIO_CONSTRAINT_HARD_ROPE: u32 = 6;
pub const DIO_CONSTRAINT_BEND: u32 = 7;
pub const DIO_CONSTRAINT_HARD_BEND: u32 = 8;

/// ABI version, the container index of a wall contact's `b`, the autotuner's slot counts, the
/// allocation-free narrow phase's limits.
pub const DIO_ABI_VERSION: u32 = 2;
pub const DIO_CONTAINER: u32 = 0xFFFFFFFF;
pub const DIO_TUNE_PRIMAL_SHAPES: u32 = 6;
pub const DIO_TUNE_BROAD_SHAPES: u32 = 6;
pub const DIO_NOALLOC_MAX_BODIES: u32 = 16384;
pub const DIO_NOALLOC_MAX_PAIRS: u32 = 65536;
pub const DIO_NOALLOC_MAX_HULL_VERTICES: u32 = 16384;
pub const DIO_NOALLOC_SCRATCH_BYTES: u32 = 47453184;
pub const DIO_TUNE_COLOUR_SHAPES: u32 = 6;
pub const DIO_TUNE_SLOTS: usize = 24;
pub const DIO_ADJOINT_EXACT: u32 = 0;
pub const DIO_ADJOINT_FROZEN_MULTIPLIERS: u32 = 1;

/// GpuShape, 80 bytes (compute.rs:28-34)
#[repr(C)]
#[derive(Clone, Copy, Debug, Default, PartialEq)]
pub struct dio_gpu_shape {
    pub tag: u32,
    pub pad: [u32; 3],
    pub corners: [[f32; 4]; 4],
}

/// GpuBody, 112 bytes (compute.rs:78-82)
#[repr(C)]
#[derive(Clone, Copy, Debug, Default, PartialEq)]
pub struct dio_gpu_body {
    pub position: [f32; 4],
    pub orientation: [f32; 4],
    pub shape: dio_gpu_shape,
}

/// GpuPair, 8 bytes (compute.rs:97-100)
#[repr(C)]
#[derive(Clone, Copy, Debug, Default, PartialEq)]
pub struct dio_gpu_pair {
    pub a: u32,
    pub b: u32,
}

/// GpuContactOut, 80 bytes (compute.rs:123-133)
#[repr(C)]
#[derive(Clone, Copy, Debug, Default, PartialEq)]
pub struct dio_gpu_contact_out {
    pub valid: u32,
    pub a: u32,
    pub b: u32,
    pub pad: u32,
    pub normal: [f32; 4],
    pub depth: f32,
    pub pad2: [f32; 3],
    pub point_a: [f32; 4],
    pub point_b: [f32; 4],
}

/// GpuAvbdBodyState, 80 bytes (gpu_physics.rs:18-26)
#[repr(C)]
#[derive(Clone, Copy, Debug, Default, PartialEq)]
pub struct dio_avbd_body_state {
    pub vel: [f32; 4],
    pub ang: [f32; 4],
    pub prev_vel: [f32; 4],
    pub inv_moment: [f32; 4],
    pub inv_mass: f32,
    pub friction: f32,
    pub pad: [f32; 2],
}

/// GpuAvbdContact, 144 bytes (gpu_physics.rs:31-44)
#[repr(C)]
#[derive(Clone, Copy, Debug, Default, PartialEq)]
pub struct dio_avbd_contact {
    pub a: u32,
    pub b: u32,
    pub friction: f32,
    pub pad: f32,
    pub normal: [f32; 4],
    pub tangent1: [f32; 4],
    pub tangent2: [f32; 4],
    pub r_a: [f32; 4],
    pub r_b: [f32; 4],
    pub c0: [f32; 4],
    pub penalty: [f32; 4],
    pub lambda: [f32; 4],
}

/// GpuAvbdConstraint, 48 bytes (gpu_physics.rs:49-56): anchor_a[3] is the
/// rest / max length (0 for sockets), anchor_b[3] the stiffness
#[repr(C)]
#[derive(Clone, Copy, Debug, Default, PartialEq)]
pub struct dio_avbd_constraint {
    pub a: u32,
    pub b: u32,
    pub kind: u32,
    pub pad: u32,
    pub anchor_a: [f32; 4],
    pub anchor_b: [f32; 4],
}

/// AvbdRunOptions (gpu_physics.rs:76-83)
#[repr(C)]
#[derive(Clone, Copy, Debug, Default, PartialEq)]
pub struct dio_run_options {
    pub dt: f32,
    pub gravity: [f32; 3],
    pub alpha: f32,
    pub beta_lin: f32,
    pub penalty_max: f32,
    pub iterations: u32,
}

/// AvbdOptions (physics/gpu.rs:127-138); defaults in dio_default_options
#[repr(C)]
#[derive(Clone, Copy, Debug, Default, PartialEq)]
pub struct dio_avbd_options {
    pub dt: f32,
    pub gravity: [f32; 3],
    pub iterations: u32,
    pub alpha: f32,
    pub beta_lin: f32,
    pub margin: f32,
    pub penalty_min: f32,
    pub penalty_max: f32,
    pub gamma: f32,
    pub friction: f32,
}

/// AvbdContainer (physics/gpu.rs:117-123) plus a presence flag (the
/// reference's Option<AvbdContainer>)
#[repr(C)]
#[derive(Clone, Copy, Debug, Default, PartialEq)]
pub struct dio_container {
    pub center: [f32; 3],
    pub rotation: [f32; 4],
    pub half_extent: f32,
    pub enabled: u32,
}

/// Buffer capacities; zero fields take defaults derived from max_bodies.
#[repr(C)]
#[derive(Clone, Copy, Debug, Default, PartialEq)]
pub struct dio_limits {
    pub max_bodies: u32,
    pub max_pairs: u32,
    pub max_contacts: u32,
    pub max_constraints: u32,
    pub max_hull_vertices: u32,
}

/// Step protocol options for the fused path.
#[repr(C)]
#[derive(Clone, Copy, Debug, Default, PartialEq)]
pub struct dio_step_opts {
    pub contact_lag: u32,
    pub mode: u32,
    pub profile: u32,
}

/// Counters and times of the last dio_step call.
#[repr(C)]
#[derive(Clone, Copy, Debug, Default, PartialEq)]
pub struct dio_step_stats {
    pub steps: u32,
    pub pairs: u32,
    pub contacts: u32,
    pub csr_nnz: u32,
    pub narrow_contacts: u32,
    pub dynamic_bodies: u32,
    pub sum_pairs: u64,
    pub sum_contacts: u64,
    pub sum_csr_nnz: u64,
    pub ms_total: f32,
    pub ms_avbd: f32,
    pub ms_init: f32,
    pub ms_primal: f32,
    pub ms_dual: f32,
    pub ms_recover: f32,
    pub n_init: u32,
    pub n_primal: u32,
    pub n_dual: u32,
    pub n_recover: u32,
    pub colours: u32,
    pub colour_rounds: u32,
}

/// Kernel launch shapes (the autotuner's search space).
#[repr(C)]
#[derive(Clone, Copy, Debug, Default, PartialEq)]
pub struct dio_tuning {
    pub primal_lanes: u32,
    pub broad_lanes: u32,
    pub resort_interval: u32,
    pub fuse_dual: u32,
    pub adjoint_lanes: u32,
    pub overlap_broad: u32,
    pub primal_order: u32,
    pub scene_solve: u32,
    pub colour_lanes: u32,
    /// the Jacobi primal of dio_step: 1 (default) packed, 2 span (worlds without
    /// constraints), 0 primal_lanes lanes per body; the result is the same every way
    pub primal_packed: u32,
}

/// The engine's names of the records it already defines (render/compute.rs, render/gpu_physics.rs).
pub type GpuBody = dio_gpu_body;
pub type GpuPair = dio_gpu_pair;
pub type GpuContactOut = dio_gpu_contact_out;
pub type GpuAvbdBodyState = dio_avbd_body_state;
pub type GpuAvbdContact = dio_avbd_contact;
pub type GpuAvbdConstraint = dio_avbd_constraint;
pub type AvbdRunOptions = dio_run_options;

// the records' sizes, as the C header lays them out (a mismatch fails to compile)
const _: () = assert!(size_of::<dio_gpu_shape>() == 80);
const _: () = assert!(size_of::<dio_gpu_body>() == 112);
const _: () = assert!(size_of::<dio_gpu_pair>() == 8);
const _: () = assert!(size_of::<dio_gpu_contact_out>() == 80);
const _: () = assert!(size_of::<dio_avbd_body_state>() == 80);
const _: () = assert!(size_of::<dio_avbd_contact>() == 144);
const _: () = assert!(size_of::<dio_avbd_constraint>() == 48);
const _: () = assert!(size_of::<dio_run_options>() == 32);
const _: () = assert!(size_of::<dio_avbd_options>() == 48);
const _: () = assert!(size_of::<dio_container>() == 36);
const _: () = assert!(size_of::<dio_limits>() == 20);
const _: () = assert!(size_of::<dio_step_opts>() == 12);
const _: () = assert!(size_of::<dio_step_stats>() == 96);
const _: () = assert!(size_of::<dio_tuning>() == 40);

#[link(name = "diomedes_hip")]
extern "C" {
    // ---- library / context
    /// DIO_ABI_VERSION of the loaded library.
    pub fn dio_abi_version() -> u32;
    /// AvbdOptions::default(), physics/gpu.rs:140-155.
    pub fn dio_default_options(out: *mut dio_avbd_options);
    /// Replaces NarrowPhaseCompute::new + AvbdCompute::new (render/compute.rs:176-397,
    /// render/gpu_physics.rs:224-433): capacities from `limits` instead of the fixed 1024 / 16384.
    pub fn dio_create(device: i32, limits: *const dio_limits, out: *mut *mut dio_ctx) -> dio_status;
    pub fn dio_destroy(ctx: *mut dio_ctx);
    /// The message of the last failed call on `ctx`.
    pub fn dio_last_error(ctx: *const dio_ctx) -> *const c_char;
    /// The size asked for by the last DIO_CAPACITY.
    pub fn dio_required_capacity(ctx: *const dio_ctx) -> u64;
    pub fn dio_synchronize(ctx: *mut dio_ctx) -> dio_status;
    /// Replaces submit_broad (compute.rs:490-567): upload bodies, compute AABBs and the
    /// candidate pairs. The pair list is the set {(i, j): i < j, closed AABB overlap}
    /// sorted lexicographically, without the reference's 16384-pair truncation.
    pub fn dio_broad_submit(ctx: *mut dio_ctx, bodies: *const dio_gpu_body, n: u32) -> dio_status;
    /// Replaces read_pair_count (compute.rs:570-579).
    pub fn dio_broad_count(ctx: *mut dio_ctx, count: *mut u32) -> dio_status;
    /// Replaces read_pairs (compute.rs:583-598); DIO_CAPACITY when cap is too small.
    pub fn dio_broad_read_pairs(ctx: *mut dio_ctx, pairs: *mut dio_gpu_pair, cap: u32, n: *mut u32) -> dio_status;
    /// Replaces read_aabbs-style parity reads: 2 float4 (min, max) per body.
    pub fn dio_broad_read_aabbs(ctx: *mut dio_ctx, aabbs: *mut f32, cap_bodies: u32) -> dio_status;
    /// Replaces submit_narrow (compute.rs:601-671): narrow phase over the first
    /// pair_count pairs of the last broad phase.
    pub fn dio_narrow_submit(ctx: *mut dio_ctx, pair_count: u32) -> dio_status;
    /// Replaces submit (compute.rs:404-477): narrow phase over caller-given pairs.
    pub fn dio_narrow_submit_pairs(ctx: *mut dio_ctx, bodies: *const dio_gpu_body, n: u32, pairs: *const dio_gpu_pair, pair_count: u32) -> dio_status;
    /// Replaces read (compute.rs:675-693): 4 slots per pair, slot pair*4+k.
    pub fn dio_narrow_read(ctx: *mut dio_ctx, slots: *mut dio_gpu_contact_out, cap_slots: u32, n_slots: *mut u32) -> dio_status;
    /// No counterpart in the reference. For callers that cannot allocate (tight
    /// serving loops, code that must not call hipMalloc): dio_narrow_submit_pairs
    /// followed by dio_narrow_read in one synchronous call that needs no context
    /// and allocates nothing. The same kernels give the same slots, bit for bit.
    /// The caller passes host arrays: n bodies, pair_count pairs (a, b < n), the
    /// hull vertex pool (x, y, z per vertex) and slots (cap_slots >= 4 * pair_count).
    /// Every other buffer is one fixed device array per GPU, held from the
    /// library's load until the process ends (DIO_NOALLOC_SCRATCH_BYTES, about
    /// 45 MiB). Limits: DIO_CAPACITY beyond DIO_NOALLOC_MAX_BODIES bodies,
    /// DIO_NOALLOC_MAX_PAIRS pairs or DIO_NOALLOC_MAX_HULL_VERTICES vertices, or
    /// when cap_slots is short; DIO_INVALID for a null array, a device outside
    /// 0 .. 63, a pair index >= n, or a hull range that is malformed or reaches
    /// past hull_count. Calls on one device share its array: a mutex per device
    /// serialises them, so calls from several threads do not overlap.
    pub fn dio_narrow_pairs_noalloc(device: i32, bodies: *const dio_gpu_body, n: u32, pairs: *const dio_gpu_pair, pair_count: u32, hull_vertices: *const f32, hull_count: u32, slots: *mut dio_gpu_contact_out, cap_slots: u32) -> dio_status;
    /// Replaces AvbdCompute::submit. positions: n x 3 floats, orientations: n x 4
    /// (x, y, z, w). offsets: n + 1 CSR offsets, indices: nidx contact indices.
    pub fn dio_avbd_submit(ctx: *mut dio_ctx, state: *const dio_avbd_body_state, positions: *const f32, orientations: *const f32, n: u32, contacts: *const dio_avbd_contact, m: u32, constraints: *const dio_avbd_constraint, c: u32, offsets: *const u32, indices: *const u32, nidx: u32, options: *const dio_run_options) -> dio_status;
    /// Replaces AvbdCompute::read (AvbdGpuResult, gpu_physics.rs:86-94). Any
    /// output pointer may be null. positions n x 3, orientations n x 4, velocity
    /// arrays n x 3, lambda / penalty m x 3.
    pub fn dio_avbd_read(ctx: *mut dio_ctx, positions: *mut f32, orientations: *mut f32, velocities: *mut f32, angular_velocities: *mut f32, prev_velocities: *mut f32, lambda: *mut f32, penalty: *mut f32) -> dio_status;
    /// Upload the world: bodies (pose + shape), their dynamics state, joints and
    /// the optional container. Resets contacts and the lagged narrow output.
    pub fn dio_world_upload(ctx: *mut dio_ctx, bodies: *const dio_gpu_body, state: *const dio_avbd_body_state, n: u32, constraints: *const dio_avbd_constraint, c: u32, container: *const dio_container) -> dio_status;
    /// Batched world: num_scenes independent scenes in one world. Scene s owns
    /// bodies scene_offsets[s] .. scene_offsets[s+1] - 1 (offsets run from 0 to
    /// n) and has its own container, containers[s]. Bodies of different scenes
    /// never form pairs and constraints may not join two scenes (DIO_INVALID).
    /// dio_step then steps every scene in the same launches, and each scene's
    /// trajectory is bit-identical to stepping it alone; dio_world_download and
    /// dio_world_contacts return the concatenation. num_scenes <= 65536. No
    /// counterpart in the reference (its demos step one scene per Renderer);
    /// replaces one plain context per scene.
    pub fn dio_batch_upload(ctx: *mut dio_ctx, bodies: *const dio_gpu_body, state: *const dio_avbd_body_state, n: u32, constraints: *const dio_avbd_constraint, c: u32, scene_offsets: *const u32, num_scenes: u32, containers: *const dio_container) -> dio_status;
    /// Vertex pool for DIO_SHAPE_HULL bodies (x, y, z per vertex).
    pub fn dio_world_upload_hulls(ctx: *mut dio_ctx, vertices: *const f32, count: u32) -> dio_status;
    /// Run nsteps physics steps on the device without host round trips.
    /// DIO_CAPACITY when a step needed more candidate pairs than max_pairs or more
    /// contacts than max_contacts: dio_required_capacity gives the largest count
    /// needed (pairs are reported before contacts) and dio_last_error names the
    /// limit. The step then ran on truncated lists and no longer follows the
    /// reference; every later dio_step reports the same until the next upload.
    pub fn dio_step(ctx: *mut dio_ctx, options: *const dio_avbd_options, opts: *const dio_step_opts, nsteps: u32) -> dio_status;
    pub fn dio_step_get_stats(ctx: *mut dio_ctx, out: *mut dio_step_stats) -> dio_status;
    /// Launch shapes; changing them drops the captured step graphs.
    pub fn dio_set_tuning(ctx: *mut dio_ctx, tuning: *const dio_tuning) -> dio_status;
    pub fn dio_get_tuning(ctx: *mut dio_ctx, out: *mut dio_tuning) -> dio_status;
    /// Autotuner: after at least one dio_step, time every primal shape on the
    /// current contact set and every broad-phase traversal shape on the current
    /// poses (reps launches each, HIP events), keep the fastest of each, and report
    /// the average microseconds per launch in us_per_launch[0 .. DIO_TUNE_SLOTS - 1]:
    /// primal lanes 1, 4, 8, 16, 32, 64; broad lanes 1, 2, 4, 8, 16, 32 (a refit + traversal
    /// pass); then, at the chosen primal shape, an unfused primal plus a dual
    /// launch and a fused primal (fuse_dual is set to the faster); then that
    /// primal in index order and in Morton order (primal_order is set to the
    /// faster; the shapes above are timed in the order current at the call); then
    /// the coloured mode's primal per colour shape (colour_lanes); then the packed
    /// primal and, in a world without constraints, the span primal in both
    /// schedules, compared per iteration with a dual launch counted when unfused
    /// (primal_packed is set to 1 or 2 for the winner, else 0; the span primal
    /// brings its schedule into fuse_dual). The primal with a fixed parity is
    /// idempotent and the broad timing runs in the granular broad buffers, so the
    /// world state is unchanged; a granular broad result submitted before the call
    /// must be submitted again.
    pub fn dio_autotune(ctx: *mut dio_ctx, options: *const dio_avbd_options, reps: u32, chosen: *mut dio_tuning, us_per_launch: *mut f32) -> dio_status;
    // ---- adjoint of the AVBD step (no counterpart in the reference)
    /// The differentiated function is the frozen-contact replay: each recorded
    /// step's contact set (geometry and the warm-start lambda / penalty as the
    /// solve received them) and CSR are kept, and only the AVBD solve runs, with
    /// every contact's friction set to options->friction. At the recorded
    /// parameters it reproduces the recorded run bit for bit. Worlds with joints
    /// are rejected (DIO_INVALID).
    /// 
    /// dio_adjoint_begin: record the following dio_step calls (at most max_steps
    /// steps and max_recorded_contacts contacts over all of them; 0 = max_steps x
    /// max_contacts), with a state checkpoint every checkpoint_every steps.
    /// dio_adjoint_end frees the recording.
    pub fn dio_adjoint_begin(ctx: *mut dio_ctx, max_steps: u32, checkpoint_every: u32, max_recorded_contacts: u64) -> dio_status;
    pub fn dio_adjoint_end(ctx: *mut dio_ctx) -> dio_status;
    pub fn dio_adjoint_steps(ctx: *mut dio_ctx, steps: *mut u32) -> dio_status;
    /// What dio_adjoint_backward differentiates (a setting of the context,
    /// kept across recordings; default DIO_ADJOINT_EXACT): the replay itself,
    /// or (DIO_ADJOINT_FROZEN_MULTIPLIERS) the replay with the multipliers and
    /// penalties each dual update writes held as constants.  Other values:
    /// DIO_INVALID.
    pub fn dio_adjoint_set_mode(ctx: *mut dio_ctx, mode: u32) -> dio_status;
    /// Replay and backward need at least one recorded step (DIO_NOT_READY
    /// otherwise) and a recording that did not overflow (DIO_CAPACITY).
    /// Replay every recorded step from the first checkpoint. state0 (n records
    /// or null) replaces the velocities and masses of the recorded start. Final
    /// positions (n x 3) and orientations (n x 4); the world is not modified.
    pub fn dio_adjoint_replay(ctx: *mut dio_ctx, options: *const dio_avbd_options, state0: *const dio_avbd_body_state, positions: *mut f32, orientations: *mut f32) -> dio_status;
    /// Reverse mode through every recorded step (checkpointed: each segment is
    /// replayed once to tape its iterations), at the replay from state0 (null:
    /// the recorded run itself, whose checkpoints the recording holds; otherwise
    /// a forward sweep from state0 saves the segment checkpoints first, unless the
    /// last dio_adjoint_replay ran from the same state0, options and steps).
    /// Inputs: the adjoint of the final state (any pointer may be null = 0; n x 3
    /// or n x 4 floats). Outputs (any may be null): the adjoint of the initial
    /// positions, orientations, velocities, angular and previous velocities, of
    /// each body's inverse mass (n) and inverse moments (n x 3), and of the
    /// friction coefficient. A new dio_world_upload / dio_batch_upload frees the
    /// recording.
    pub fn dio_adjoint_backward(ctx: *mut dio_ctx, options: *const dio_avbd_options, state0: *const dio_avbd_body_state, d_positions: *const f32, d_orientations: *const f32, d_velocities: *const f32, d_angular: *const f32, g_positions: *mut f32, g_orientations: *mut f32, g_velocities: *mut f32, g_angular: *mut f32, g_prev_velocities: *mut f32, g_inv_mass: *mut f32, g_inv_moment: *mut f32, g_friction: *mut f64) -> dio_status;
    /// Download body state (any pointer may be null): n x 3 / n x 4 floats.
    pub fn dio_world_download(ctx: *mut dio_ctx, positions: *mut f32, orientations: *mut f32, velocities: *mut f32, angular_velocities: *mut f32, prev_velocities: *mut f32) -> dio_status;
    /// Contacts of the last step with their solved lambda / penalty (the reference's
    /// AvbdSolver::contacts() after sync_bodies_from_gpu).
    pub fn dio_world_contacts(ctx: *mut dio_ctx, out: *mut dio_avbd_contact, cap: u32, m: *mut u32) -> dio_status;
    /// The last coloured step's colour per body (0xFFFFFFFE: static, not in the
    /// graph; cap >= n) and the number of colours. No counterpart in the
    /// reference.
    pub fn dio_world_colours(ctx: *mut dio_ctx, colour: *mut u32, cap: u32, num_colours: *mut u32) -> dio_status;
    /// Each constraint's multipliers after the last step: lambda and penalty per
    /// row, c x 3 floats each (x, y, z rows of a hard socket or hard bend; x alone
    /// of a hard rod or rope; zero for soft joints, which carry none). cap >= the constraint
    /// count, *c set to it. No counterpart in the reference.
    pub fn dio_world_joints(ctx: *mut dio_ctx, lambda: *mut f32, penalty: *mut f32, cap: u32, c: *mut u32) -> dio_status;
    /// Candidate pairs of the last step's broad phase (sorted, a < b).
    pub fn dio_world_pairs(ctx: *mut dio_ctx, out: *mut dio_gpu_pair, cap: u32, n: *mut u32) -> dio_status;
    /// The instanced renderer's data: per-body render scales (3 floats per
    /// body) and the instance order (order[s] = the body drawn in slot s; null:
    /// slot s draws body s), as the renderer groups instances by shape and style
    /// (engine/src/render/renderer.rs:222-255).
    pub fn dio_world_set_instances(ctx: *mut dio_ctx, scales: *const f32, order: *const u32, count: u32) -> dio_status;
    /// For every instance slot s: (view_proj * model, model), two column-major
    /// 4 x 4 float matrices (mesh.vert's InstanceTransform, 128 bytes) at
    /// dst + 128 s, model = Mat4::from_scale_rotation_translation of the body's
    /// pose and scale; view_proj = projection * view (16 floats, column-major).
    /// dst: 16-byte aligned device memory, e.g. an imported renderer buffer.
    pub fn dio_world_instances(ctx: *mut dio_ctx, view_proj: *const f32, dst: *mut c_void, dst_bytes: u64) -> dio_status;
    /// Map device memory exported as a file descriptor (a Vulkan buffer's
    /// opaque-fd memory, a DMA-BUF) into the context's device; the library
    /// imports a duplicate, so the caller keeps fd.
    pub fn dio_import_external_memory(ctx: *mut dio_ctx, fd: i32, object_size: u64, offset: u64, size: u64, dev_ptr: *mut *mut c_void, out: *mut *mut dio_external_memory) -> dio_status;
    pub fn dio_release_external_memory(ctx: *mut dio_ctx, mem: *mut dio_external_memory) -> dio_status;
}
