/*
 * phieng_fluid2.h — C ABI of the 2D SPH fluid step (World2) on MI355X.
 *
 * The first block keeps the name, argument meaning and error behaviour of
 * the reference's 2D exports (cited as crates/engine/src/ffi.rs:<line>);
 * the second block holds the same additions as the 3D library (capacity
 * lift, velocity upload, parity readback, sync).  A World2 runs the 3D
 * engine's kernels on particles in the z = 0 plane with the 2D kernel
 * constants (crates/dynamics/src/fluid.rs:265-285) and the 2D box rule,
 * which reproduces the 2D step (fluid.rs:541-1136): every 3D term that
 * involves z is an exact zero there.
 *
 * Conventions as in phieng_fluid3.h: opaque pointers, caller-owned
 * buffers, null tolerated, nothing thrown across the boundary, one world
 * per thread at a time.  phieng_last_error() reports the cause of a 0/-1.
 */
#ifndef PHIENG_FLUID2_H
#define PHIENG_FLUID2_H

#include <stdint.h>

#include "phieng_fluid3.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct World2 World2;

/* Shape tags of phieng_world_add_body (ffi.rs:26-28). */
#define PHIENG_SHAPE2_CIRCLE 0u
#define PHIENG_SHAPE2_BOX 1u

/* #[repr(C)] FluidConfig2Ffi (ffi.rs:716-739): the same fields, order and
 * 112-byte layout as FluidConfig3Ffi; semantics FluidConfig2
 * (crates/dynamics/src/fluid.rs:55-140, defaults :142-169). */
typedef FluidConfig3Ffi FluidConfig2Ffi;

/* ---------------------------------------------------------------------
 * Reference exports (drop-in).
 * ------------------------------------------------------------------- */

/* ffi.rs:51-53 — never null; release with phieng_world_free. */
World2 *phieng_world_new(void);
/* ffi.rs:62-66 — null is a no-op. */
void phieng_world_free(World2 *world);
/* ffi.rs:76-81 — null, non-finite or dt <= 0 is a no-op; dt is clamped
 * to 1/15 s.  Runs the fluid phase of World2::step (phase 1b,
 * crates/dynamics/src/world.rs:843-849) on the GPU. */
void phieng_world_step(World2 *world, float dt);
/* ffi.rs:89-94 — non-finite components are ignored. */
void phieng_world_set_gravity(World2 *world, float gx, float gy);
/* ffi.rs:1159-1203 — circle (a = radius) or box (a, b = half extents) at
 * angle 0; returns the body id or -1.  Static bodies only: a dynamic body
 * returns -1 (the 2D fluid's two-way coupling, fluid.rs, is not on the
 * GPU path; 3D worlds couple). */
int32_t phieng_world_add_body(World2 *world, float pos_x, float pos_y, uint32_t shape_tag, float a, float b,
                              float density, uint8_t is_static, float friction, float restitution);
/* #[repr(C)] BodyDesc2D (ffi.rs:1205-1222), 36 bytes. */
typedef struct BodyDesc2D {
    float pos_x, pos_y;
    uint32_t shape_tag;
    float a, b;
    float density;
    uint8_t is_static;
    uint8_t _pad[3];
    float friction, restitution;
} BodyDesc2D;
/* ffi.rs:1224-1290 — each descriptor validated and added as
 * phieng_world_add_body; out_ids[i] = id or -1 (no rollback).  Returns the
 * number added; 0 for null world, descs or out_ids; n capped at
 * 1,000,000. */
uint32_t phieng_world_add_bodies(World2 *world, const BodyDesc2D *descs, uint32_t n, int32_t *out_ids);
/* ffi.rs:999-1011 — 1 on success, 0 on null or invalid input (world
 * unchanged); validation ffi.rs:773-840.  adaptive_resolution merges and
 * splits particles each substep as fluid.rs:826-891 does (split cap: the
 * capacity, MAX_PARTICLES_2D = 4096 unless reserved). */
int32_t phieng_world_set_fluid_config(World2 *world, const FluidConfig2Ffi *cfg);
/* ffi.rs:1017-1026 — 1 on success, 0 on null. */
int32_t phieng_world_get_fluid_config(const World2 *world, FluidConfig2Ffi *out);
/* ffi.rs:2486-2504 — interleaved xy; count capped at 1,000,000 per call;
 * any non-finite value rejects the whole call (returns 0).  Returns the
 * number added (total capped at MAX_PARTICLES_2D = 4096, fluid.rs:26,
 * unless reserved). */
uint32_t phieng_world_fluid_add(World2 *world, const float *positions, uint32_t count);
/* ffi.rs:2515-2538 — as fluid_add with a phase index per particle. */
uint32_t phieng_world_fluid_add_phase(World2 *world, const float *positions, const uint8_t *phases,
                                      uint32_t count);
/* ffi.rs:2546-2551 */
uint32_t phieng_world_fluid_count(const World2 *world);
/* ffi.rs:2559-2565 — 1 on success, 0 on null. */
int32_t phieng_world_fluid_clear(World2 *world);
/* ffi.rs:2575-2594 — writes [x,y,vx,vy] per particle in insertion order,
 * up to max; returns the number written. */
uint32_t phieng_world_fluid_get(const World2 *world, float *out, uint32_t max);

/* ffi.rs:3107-3160 — as phieng_world3_save, for a World2 (World2Dto:
 * 2-vectors, the broad-phase cell size, circle and box shapes, kind 2). */
int32_t phieng_world_save(const World2 *world, const uint8_t *path_ptr, size_t path_len, uint32_t format,
                          uint32_t mode);
/* ffi.rs:3162-3190 — as phieng_world3_load, for a World2 (polygon bodies
 * are refused with the rest of phieng_world3_load's list). */
World2 *phieng_world_load(const uint8_t *path_ptr, size_t path_len, uint32_t mode);

/* ---------------------------------------------------------------------
 * Extensions (no reference counterpart), as their World3 namesakes.
 * ------------------------------------------------------------------- */

/* Lift the cap to `capacity` (at most 16,777,216); returns the cap in force. */
uint32_t phieng_world_fluid_reserve(World2 *world, uint32_t capacity);
/* Mass and phase per particle in insertion order, up to max. */
uint32_t phieng_world_fluid_get_mass(const World2 *world, float *mass, uint8_t *phase, uint32_t max);
/* Overwrite velocities (vx,vy per particle, insertion order); 0 on null
 * or any non-finite value, else the number set. */
uint32_t phieng_world_fluid_set_velocity(World2 *world, const float *vel, uint32_t count);
/* PHIENG_RECORD_* flags, as phieng_world3_fluid_set_record. */
int32_t phieng_world_fluid_set_record(World2 *world, uint32_t flags);
/* The last substep's intermediates in insertion order: cell2 (2 x int32),
 * sorted_ids, density, pressure, accel2 (2 floats), neighbour counts.
 * Null outputs are skipped; returns the particle count, 0 if none. */
uint32_t phieng_world_fluid_get_record(const World2 *world, int32_t *cell2, uint32_t *sorted_ids, float *density,
                                       float *pressure, float *accel2, uint32_t *nbr_count, uint32_t max);
/* Block until the world's GPU work is done; 1, or 0 with phieng_last_error set. */
int32_t phieng_world_sync(World2 *world);
/* Boundary ghosts (fluid.rs:661-819), as phieng_world3_boundary_info / _get
 * with xy pairs. */
int32_t phieng_world_boundary_info(const World2 *world, uint32_t *count, uint64_t *regens);
/* Extensions: the 2D forms of phieng_world3_set_neighbour_list and
 * phieng_world3_neighbour_stats. */
int32_t phieng_world_set_neighbour_list(World2 *world, int32_t verlet, float skin);
int32_t phieng_world_neighbour_stats(World2 *world, uint64_t *substeps, uint64_t *rebuilds, uint64_t *repaired,
                                     uint64_t *list_substeps);
uint32_t phieng_world_boundary_get(const World2 *world, float *xy, float *density, float *pressure, uint32_t max);

#ifdef __cplusplus
}
#endif

#endif /* PHIENG_FLUID2_H */
