/*
 * phieng_fluid3.h — C ABI of the MI355X-native PhiEng 3D SPH fluid step.
 *
 * Every function in the first block keeps the name, argument meaning and
 * error behaviour of the reference export it replaces (cited as
 * crates/engine/src/ffi.rs:<line>), so an existing caller (the C# P/Invoke
 * layer, a C host) links against libphieng_hip.so unchanged.  The second
 * block holds additions with no reference counterpart: capacity lift,
 * parity readback, batching, timing and the learned pressure surrogate.
 *
 * Conventions (ffi.rs:1-19): opaque world pointers, caller-owned buffers,
 * null tolerated everywhere (no-op / 0 / -1), no exception or HIP error
 * crosses the boundary.  One world must not be used from two threads at
 * once; different worlds are independent.
 */
#ifndef PHIENG_FLUID3_H
#define PHIENG_FLUID3_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct World3 World3;

/* Shape tags of phieng_world3_add_body (ffi.rs:3193-3195). */
#define PHIENG_SHAPE3_SPHERE 0u
#define PHIENG_SHAPE3_BOX 1u
#define PHIENG_SHAPE3_CAPSULE 2u

/* #[repr(C)] FluidConfig3Ffi (ffi.rs:743-767); 112 bytes, align 4.
 * Field semantics follow FluidConfig3 (crates/dynamics/src/fluid3.rs:51-135). */
typedef struct FluidConfig3Ffi {
    float particle_mass;          /*   0 */
    float rest_density;           /*   4 */
    float smoothing_length;       /*   8 */
    float stiffness;              /*  12 */
    float gamma;                  /*  16 */
    float artificial_viscosity;   /*  20 */
    float surface_tension;        /*  24 */
    float cfl;                    /*  28 */
    uint32_t max_substeps;        /*  32 */
    uint8_t adaptive_resolution;  /*  36  (Rust bool) */
    float split_density_ratio;    /*  40 */
    float merge_density_ratio;    /*  44 */
    uint8_t boundary_particles;   /*  48  (Rust bool) */
    float viscosity;              /*  52 */
    float drag;                   /*  56 */
    float wall_friction;          /*  60 */
    uint8_t phase_count;          /*  64 */
    float phase_rest_density[4];  /*  68 */
    float phase_mass[4];          /*  84 */
    float max_velocity;           /* 100 */
    float particle_radius;        /* 104 */
    uint8_t enabled;              /* 108  (Rust bool) */
} FluidConfig3Ffi;

/* ---------------------------------------------------------------------
 * Reference exports (drop-in).
 * ------------------------------------------------------------------- */

/* ffi.rs:3199-3201 — never null; release with phieng_world3_free. */
World3 *phieng_world3_new(void);
/* ffi.rs:3209-3213 — null is a no-op. */
void phieng_world3_free(World3 *world);
/* ffi.rs:3221-3226 — null, non-finite or dt <= 0 is a no-op; dt is
 * clamped to 1/15 s. Runs World3::step (crates/dynamics/src/world.rs:
 * 2670-2697, 2766-2812) as far as the fluid and its bodies go: the gravity
 * kick of dynamic bodies, the fluid phase on the GPU with its two-way
 * coupling (fluid3.rs:1141-1213), the bodies' integration.  The rigid
 * contact solver (body-body contacts, joints) is not part of this
 * library: dynamic bodies meet only the fluid. */
void phieng_world3_step(World3 *world, float dt);
/* ffi.rs:3234-3244 — non-finite components are ignored. */
void phieng_world3_set_gravity(World3 *world, float gx, float gy, float gz);
/* ffi.rs:3703-3760 — returns the body id or -1.  A dynamic body
 * (is_static == 0) gets the mass and inertia of RigidBody3::set_shape
 * (body.rs:1266-1300) and is coupled both ways to the fluid; it is refused
 * (-1, reason in phieng_last_error) while the world runs the learned
 * surrogate, collects samples or has adaptive_resolution on. */
int32_t phieng_world3_add_body(World3 *world, float pos_x, float pos_y, float pos_z,
                               uint32_t shape_tag, float a, float b, float c, float density,
                               uint8_t is_static, float friction, float restitution);
/* #[repr(C)] BodyDesc3D (ffi.rs:3765-3784), 44 bytes. */
typedef struct BodyDesc3D {
    float pos_x, pos_y, pos_z;
    uint32_t shape_tag;
    float a, b, c;
    float density;
    uint8_t is_static;
    uint8_t _pad[3];
    float friction, restitution;
} BodyDesc3D;
/* ffi.rs:3786-3860 — each descriptor validated and added as
 * phieng_world3_add_body; out_ids[i] = id or -1 (no rollback).  Returns
 * the number added; 0 for null world, descs or out_ids; n capped at
 * 1,000,000. */
uint32_t phieng_world3_add_bodies(World3 *world, const BodyDesc3D *descs, uint32_t n, int32_t *out_ids);
/* ffi.rs:3387-3393 — -1 for null. */
int32_t phieng_world3_body_count(World3 *world);
/* ffi.rs:3885-3910 — position (3) and rotation quaternion x, y, z, w (4);
 * 1 on success, 0 on null or an unknown id. */
int32_t phieng_world3_get_body_transform(World3 *world, uint32_t body_id, float *out_pos, float *out_quat);
/* ffi.rs:3917-3937 — 1 on success, 0 on null or an unknown id. */
int32_t phieng_world3_get_body_velocity(World3 *world, uint32_t body_id, float *out_vel);
/* ffi.rs:1639-1657 — radians per second; 1 on success, 0 on null or an
 * unknown id. */
int32_t phieng_world3_get_body_angular_velocity(World3 *world, uint32_t body_id, float *out_w);
/* ffi.rs:1032-1044 — 1 on success, 0 on null or invalid input (world
 * unchanged); validation rules ffi.rs:857-914.  Also returns 0 for
 * boundary_particles or adaptive_resolution set while a surrogate is
 * attached (or collecting).  With adaptive_resolution each substep merges
 * and splits particles as fluid3.rs:923-987 does, so the count and the
 * insertion order change; the split cap is the fluid capacity (4096 as
 * MAX_PARTICLES_3D unless reserved higher). */
int32_t phieng_world3_set_fluid_config(World3 *world, const FluidConfig3Ffi *cfg);
/* ffi.rs:1050-1059 — 1 on success, 0 on null. */
int32_t phieng_world3_get_fluid_config(const World3 *world, FluidConfig3Ffi *out);
/* ffi.rs:3579-3600 — interleaved xyz; count capped at 1,000,000 per call;
 * any non-finite value rejects the whole call (returns 0).  Returns the
 * number added (total capped at the fluid capacity, 4096 by default as
 * MAX_PARTICLES_3D, fluid3.rs:28). */
uint32_t phieng_world3_fluid_add(World3 *world, const float *positions, uint32_t count);
/* ffi.rs:3609-3635 — as fluid_add with a phase index per particle. */
uint32_t phieng_world3_fluid_add_phase(World3 *world, const float *positions,
                                       const uint8_t *phases, uint32_t count);
/* ffi.rs:3643-3648 */
uint32_t phieng_world3_fluid_count(const World3 *world);
/* ffi.rs:3656-3662 — 1 on success, 0 on null. */
int32_t phieng_world3_fluid_clear(World3 *world);
/* ffi.rs:3672-3697 — writes [x,y,z,vx,vy,vz] per particle in insertion
 * order, up to max; returns the number written. */
uint32_t phieng_world3_fluid_get(const World3 *world, float *out, uint32_t max);

/* Snapshots (ffi.rs:3097-3105, 4036-4117; formats snapshot.rs).  Format
 * tags and load modes as the reference's. */
#define PHIENG_SNAPSHOT_JSON 0u
#define PHIENG_SNAPSHOT_BINARY 1u
#define PHIENG_SNAPSHOT_EXACT 0u
#define PHIENG_SNAPSHOT_FRESH 1u
/* ffi.rs:4042-4082 — write the world to path_ptr[0 .. path_len) (UTF-8,
 * at most 4096 bytes) as JSON or as the PHE1 binary envelope, in the
 * reference's World3Dto layout at the lowest schema the world needs (3 to
 * 7).  Returns 1 on success, 0 for invalid input (null world or path, bad
 * format or mode), -1 on IO failure, -2 on encode failure. */
int32_t phieng_world3_save(const World3 *world, const uint8_t *path_ptr, size_t path_len, uint32_t format,
                           uint32_t mode);
/* ffi.rs:4084-4111 — a new world from a file (format detected from its
 * first non-blank byte), or null on any failure (phieng_last_error says
 * why).  Besides the reference's checks it refuses what the GPU fluid path
 * cannot hold: schema 8 or later, dynamic, rotated, moving or hull bodies,
 * joints, contacts, ambient turbulence, configs the C ABI rejects; it
 * accepts more than MAX_PARTICLES_3D particles.  Both modes load the same
 * state (a fluid world holds no warm-start cache or joint impulse). */
World3 *phieng_world3_load(const uint8_t *path_ptr, size_t path_len, uint32_t mode);

/* ---------------------------------------------------------------------
 * Extensions (no reference counterpart).
 * ------------------------------------------------------------------- */

/* Lift the per-fluid particle cap (MAX_PARTICLES_3D = 4096, fluid3.rs:28)
 * to `capacity` (at most 16,777,216).  Returns the capacity in force. */
uint32_t phieng_world3_fluid_reserve(World3 *world, uint32_t capacity);

/* Mass and phase per particle in insertion order (as fluid_get; either
 * pointer may be null, not both), up to max; returns the number written.
 * Masses change under adaptive_resolution. */
uint32_t phieng_world3_fluid_get_mass(const World3 *world, float *mass, uint8_t *phase, uint32_t max);

/* Overwrite velocities (insertion order, vx,vy,vz per particle) — used to
 * start parity checks from an arbitrary state.  Returns the number set,
 * 0 on null or any non-finite value. */
uint32_t phieng_world3_fluid_set_velocity(World3 *world, const float *vel, uint32_t count);

/* Record per-particle intermediates of the next steps (bit flags below). */
#define PHIENG_RECORD_CELLS 1u   /* cell key triples + sorted order   */
#define PHIENG_RECORD_DENSITY 2u /* density, pressure, neighbour count */
#define PHIENG_RECORD_ACCEL 4u   /* a = g + F/rho                      */
int32_t phieng_world3_fluid_set_record(World3 *world, uint32_t flags);
/* Read what the most recent substep recorded, in insertion order (cell3:
 * 3 x int32 per particle; sorted_ids: original index of the particle at
 * each sorted slot; accel3: 3 floats).  Null outputs are skipped.
 * Returns the particle count, 0 if nothing was recorded. */
uint32_t phieng_world3_fluid_get_record(const World3 *world, int32_t *cell3,
                                        uint32_t *sorted_ids, float *density,
                                        float *pressure, float *accel3,
                                        uint32_t *nbr_count, uint32_t max);

/* Kernel timing: when enabled, every launch of the step is bracketed by
 * HIP events on the world's stream.  get_timing fills per-kernel totals
 * (ms) and launch counts for up to `max` kernels and returns the number
 * of kernel kinds; names via phieng_kernel_name. */
int32_t phieng_world3_set_timing(World3 *world, int32_t enable);
int32_t phieng_world3_get_timing(const World3 *world, double *total_ms, uint64_t *launches,
                                 int32_t max);
const char *phieng_kernel_name(int32_t kind);

/* Boundary ghosts (config boundary_particles, fluid3.rs:716-920): the
 * number of ghosts and how many times they were generated (the reference's
 * boundary_regens: once per change of the static bodies or the kernel
 * config, fingerprinted as fluid3.rs:660-714).  1, or 0 on null. */
int32_t phieng_world3_boundary_info(const World3 *world, uint32_t *count, uint64_t *regens);
/* Extension (no reference counterpart): neighbour search of the plain
 * substep.  verlet = 1 keeps per particle the particles within (1 + skin) h
 * at the last rebuild and walks that list; verlet = 0 sweeps the 27 cells
 * every substep; verlet = 2 (the default) keeps the lists from 131,072
 * particles up and sweeps below (where the list form's extra launches cost
 * more than its walk saves).  All give the reference's results bit for
 * bit.  skin in [0, 0.5] (default 0.1).  1, or 0 on a bad argument. */
int32_t phieng_world3_set_neighbour_list(World3 *world, int32_t verlet, float skin);
/* Counters of the Verlet path: Verlet substeps, list rebuilds, rows put
 * back in order after a neighbour changed cell, and substeps that swept the
 * cells instead because such rows had become too many (more than
 * max(1024, n / 512) in a substep: the list form then runs for 16
 * substeps, twice as many on each return to a flow still that fast, before
 * the rows are rebuilt; every fourth Verlet substep decides from the one
 * two before, so the choice is the same on every run).  Synchronises. */
int32_t phieng_world3_neighbour_stats(World3 *world, uint64_t *substeps, uint64_t *rebuilds, uint64_t *repaired,
                                      uint64_t *list_substeps);
/* The ghosts in generation order: positions (3 floats), and the density
 * and pressure the last substep gave them.  Null outputs are skipped;
 * returns the number written. */
uint32_t phieng_world3_boundary_get(const World3 *world, float *xyz, float *density, float *pressure,
                                    uint32_t max);

/* Block until all queued GPU work of the world has finished.  Returns 1,
 * or 0 on null, a HIP error, or a step that failed since the last sync
 * (phieng_world3_step returns nothing, as in the reference); then
 * phieng_last_error names the cause. */
int32_t phieng_world3_sync(World3 *world);
/* Text of the last error seen by any entry point of this thread ("" if none). */
const char *phieng_last_error(void);

/* ---------------------------------------------------------------------
 * Batched stepping (no reference counterpart): many independent worlds
 * advance together.  phieng_batch3_new copies the worlds' particles into
 * one concatenated device state (each world keeps its config, gravity,
 * static bodies and grid); the worlds are not touched again until
 * phieng_batch3_sync_worlds writes the state back.  Every scene's
 * trajectory is bit-identical to its world stepped alone.
 * ------------------------------------------------------------------- */
typedef struct Batch3 Batch3;
#define PHIENG_BATCH_KERNELS 0           /* the eight launches of a single world   */
#define PHIENG_BATCH_PERSISTENT 1        /* one persistent launch per substep      */
#define PHIENG_BATCH_PERSISTENT_GRAPH 2  /* the persistent launch from a hipGraph  */
#define PHIENG_BATCH_KERNELS_GRAPH 3     /* the eight launches from a hipGraph     */
/* Null (with phieng_batch3_last_error) if a world is null, disabled, uses
 * CFL substepping or boundary ghosts, or the total exceeds 2^24 particles. */
Batch3 *phieng_batch3_new(World3 *const *worlds, uint32_t count);
void phieng_batch3_free(Batch3 *batch);
/* Launch configuration; blocks_per_cu = 0 uses the occupancy limit.  Also
 * clears the persistent launch's grid barrier after a time-out. */
int32_t phieng_batch3_set_mode(Batch3 *batch, int32_t mode, uint32_t blocks_per_cu);
/* Bounded wait of the persistent launch's grid barrier: a block that spins
 * more than `spins` times raises the abort flag and every block leaves at
 * its next barrier (default 2^22). */
int32_t phieng_batch3_set_spin_limit(Batch3 *batch, uint32_t spins);
/* Advance every world by dt (same rules as phieng_world3_step). 1 / 0;
 * 0 once a sync has seen a timed-out barrier, until phieng_batch3_set_mode. */
int32_t phieng_batch3_step(Batch3 *batch, float dt);
/* Wait for the queued steps; 0 if a grid barrier timed out (the batch's
 * state is then invalid). */
int32_t phieng_batch3_sync(Batch3 *batch);
/* Write the batch state back into its worlds. 1 / 0. */
int32_t phieng_batch3_sync_worlds(Batch3 *batch);
uint32_t phieng_batch3_count(const Batch3 *batch);
/* Pick the fastest launch configuration by timing `trial_steps` substeps
 * of each from the same saved state (restored afterwards); the choice is
 * appended to the text file cache_path (may be null) and reused from it.
 * Returns the chosen mode or -1. */
int32_t phieng_batch3_autotune(Batch3 *batch, float dt, uint32_t trial_steps, const char *cache_path,
                               uint32_t *out_blocks_per_cu, double *out_ms_per_step);
int32_t phieng_batch3_set_timing(Batch3 *batch, int32_t enable);
int32_t phieng_batch3_get_timing(Batch3 *batch, double *total_ms, uint64_t *launches, int32_t max);
const char *phieng_batch3_last_error(void);

/* ---------------------------------------------------------------------
 * Learned pressure surrogate (no reference counterpart).  A 2-layer MLP,
 * 16 inputs -> 128 (tanh-GELU) -> 16 outputs, BF16 weights on MFMA with
 * FP32 accumulation, trained on the GPU by the library's own forward and
 * backward kernels (Adam).  Parameters are an fp32 vector of
 * phieng_surrogate3_num_params() values: W1 [16][128], b1 [128],
 * W2 [128][16], b2 [16].  A training sample is 32 floats: MLP input x[16],
 * three basis vectors b_k (9), the network-free acceleration a0 (3), the
 * target acceleration a_t (3) and a scale s; its loss is
 * |a0 + sum_k y_k b_k - a_t|^2 / s^2 with y the first three outputs.
 * ------------------------------------------------------------------- */
typedef struct Surrogate3 Surrogate3;
Surrogate3 *phieng_surrogate3_new(uint32_t seed);
void phieng_surrogate3_free(Surrogate3 *s);
uint32_t phieng_surrogate3_num_params(void);
int32_t phieng_surrogate3_get_params(Surrogate3 *s, float *out, uint32_t n);
int32_t phieng_surrogate3_set_params(Surrogate3 *s, const float *in, uint32_t n);
/* y = MLP(x) for `rows` rows (x: rows x 16, y: rows x 16). */
int32_t phieng_surrogate3_mlp_forward(Surrogate3 *s, const float *x, uint32_t rows, float *y);
/* phieng_surrogate3_mlp_forward without any allocation: no counterpart in
 * the reference.  For callers that cannot allocate device memory (tight
 * serving loops, code that must not call hipMalloc).  Inputs and outputs are
 * staged in a static per-GPU scratch array of 8 MiB, so rows must be at most
 * 65,536 (larger calls return 0 with phieng_surrogate3_last_error).  Calls
 * that overlap on one GPU share that array and race.  Same results as
 * phieng_surrogate3_mlp_forward, bit for bit. */
int32_t phieng_surrogate3_mlp_forward_noalloc(Surrogate3 *s, const float *x, uint32_t rows, float *y);
/* grad = d(sum dy . MLP(x)) / d params. */
int32_t phieng_surrogate3_mlp_backward(Surrogate3 *s, const float *x, const float *dy, uint32_t rows,
                                       float *grad);
int32_t phieng_surrogate3_add_samples(Surrogate3 *s, const float *samples, uint32_t count);
uint32_t phieng_surrogate3_num_samples(const Surrogate3 *s);
int32_t phieng_surrogate3_clear_samples(Surrogate3 *s);
int32_t phieng_surrogate3_get_samples(Surrogate3 *s, float *out, uint32_t first, uint32_t count);
/* Adam over shuffled batches; mean sample loss of the first and last epoch
 * and of a0 alone (the method without the network). */
int32_t phieng_surrogate3_train(Surrogate3 *s, uint32_t epochs, uint32_t batch, float lr, uint32_t seed,
                                float *out_loss_first, float *out_loss_last, float *out_loss_base);
int32_t phieng_surrogate3_eval(Surrogate3 *s, float *out_loss, float *out_loss_base);
const char *phieng_surrogate3_last_error(void);

/* Step `world` with the surrogate in the loop (null disables it): every
 * check_every-th substep is a plain substep that also judges the state on
 * two measures — the mean |rho/rho0 - 1| over particles with
 * rho >= 0.9 rho0, and the surrogate's acceleration error: before the plain
 * substep the surrogate is evaluated on the same state for 1/16 of the
 * particles (every 16th run of 64 consecutive slots of the particle order,
 * the offset rotating from judge to judge), and the error is
 * sum |a_plain - a_sur| / sum |a_plain| over those particles (pressure and
 * viscous accelerations, gravity excluded; not evaluated, -1, when no
 * surrogate substep could have run from that state or the judge also
 * collects samples).  The verdict is read back without stalling the stream
 * and applied two substeps after the judging one: if the density error
 * exceeds max_density_error or the acceleration error exceeds the limit of
 * phieng_world3_set_surrogate_accel_limit, the check_every substeps from
 * there on stay plain (a fallback), so while judges keep failing only the
 * one substep after each judge can be a surrogate one.  use_network = 0
 * runs the same substep with the network's coefficients fixed at zero (the
 * method without the network).  The world must not be in a batch; 1 / 0. */
int32_t phieng_world3_set_surrogate(World3 *world, Surrogate3 *s, uint32_t check_every,
                                    float max_density_error, int32_t use_network);
/* While set (null clears), every `every`-th substep of `world` is plain and
 * appends training samples to s from the particles whose original index
 * is congruent to a rotating phase modulo `stride`.  1 / 0. */
int32_t phieng_world3_collect_samples(World3 *world, Surrogate3 *s, uint32_t every, uint32_t stride);
/* Counts of surrogate, plain and fallback decisions and the newest judged
 * density error (-1 before any), including a verdict not yet applied.
 * 1 / 0. */
int32_t phieng_world3_surrogate_stats(const World3 *world, uint64_t *sur_steps, uint64_t *plain_steps,
                                      uint64_t *fallbacks, float *last_density_error);
/* Limit of the judge's acceleration error (default 0.2; > 0, +inf turns
 * the measure off).  No counterpart in the reference.  1 / 0. */
int32_t phieng_world3_set_surrogate_accel_limit(World3 *world, float max_accel_error);
/* The newest judged density and acceleration errors (-1 before any / when
 * not evaluated), including a verdict not yet applied, and the count of
 * judges that kept the particle order and the neighbour list of the judge
 * before.  A judge that builds the list sizes it for two judge intervals
 * when that costs the surrogate substeps less than the next judge saves
 * (from the largest particle speed at the previous list's judge); the next
 * judge, if no particle fell outside that list, then computes the plain
 * substep over it instead of sorting and building a list (the same
 * neighbours, summed in list order instead of cell order).  No
 * counterpart in the reference.  1 / 0. */
int32_t phieng_world3_surrogate_judge(const World3 *world, float *density_error, float *accel_error,
                                      uint64_t *kept_judges);
/* Minimum skin of the surrogate's neighbour list: surrogate substeps
 * between two plain substeps see the candidates that were within
 * (1 + s) h at the plain one, with the Verlet skin
 * s = max(skin, 4 check_every dt vmax / h) from the largest particle speed
 * at that plain substep (default skin 0.01; 0 <= skin <= 1).  1 / 0. */
int32_t phieng_world3_set_surrogate_skin(World3 *world, float skin);

#ifdef __cplusplus
}
#endif

#endif /* PHIENG_FLUID3_H */
