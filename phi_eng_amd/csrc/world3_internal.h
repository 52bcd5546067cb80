// world3_internal.h — what the batch and the snapshot code need from a World3 (world3.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/phieng_fluid3.h"
#include "engine.h"

uint32_t world3_count(const World3 *w);
const FluidConfig3Ffi &world3_config(const World3 *w);
const std::vector<BodyDev> &world3_bodies(const World3 *w);
ph::SceneHost world3_scene_host(const World3 *w, float dt);
// Current device state of the world's particles (sorted order + tags).
void world3_device_state(const World3 *w, const float4 **posm, const float4 **vel, const uint32_t **tag,
                         hipStream_t *stream);
// Same, writable; marks the world's grid domain and records stale.
void world3_device_state_mut(World3 *w, float4 **posm, float4 **vel, uint32_t **tag, hipStream_t *stream);

// ---------------------------------------------------------------------------
// Snapshot view of a world (snapshot.hip).  A GPU world holds gravity,
// static bodies and the fluid; the other fields of the reference's World3Dto
// (crates/dynamics/src/snapshot.rs:3189-3250) are kept verbatim so that a
// loaded file saves back unchanged.  Defaults are World3::new's
// (world.rs:2456-2503) and RigidBody3::new's (body.rs:1225-1255).
struct SnapMeta {
    float dt = 1.0f / 60.0f;
    float linear_sleep = 0.05f, angular_sleep = 0.05f;  // math.rs:14-15
    uint32_t iterations = 8;                             // solver3d.rs:95
    float slop = 0.005f, baumgarte = 0.2f, early_exit = 1e-6f;
    bool friction = true, restitution = true;
    // ValidationConfig::default (validation.rs:26-39, release build)
    bool check_finite = false;
    float max_penetration = 1.0f, max_mass_ratio = 1000.0f, max_coordinate = 1e4f, max_velocity = 1e3f,
          max_angular_velocity = 1e3f;
    bool log_warnings = false, auto_correct = true;
    // GravityConfig3::default (gravity.rs:154-166)
    uint32_t g_mode = 0, g_integrator = 0;
    float big_g = 1.0f, softening = 0.0f, attractor[3] = {0.0f, 0.0f, 0.0f}, attractor_mu = 0.0f, theta = 0.5f;
    // AmbientConfig3::default (ambient.rs: spring, drag, turbulence)
    bool spring_on = false;
    float spring_anchor[3] = {0.0f, 0.0f, 0.0f}, spring_k = 0.0f, spring_c = 0.0f, spring_max = 0.0f;
    bool drag_on = false;
    float drag_linear = 0.0f, drag_quadratic = 0.0f, drag_air = 1.0f, drag_angular = 0.0f,
          drag_wind[3] = {0.0f, 0.0f, 0.0f};
    bool turb_on = false;
    float turb_amplitude = 0.0f, turb_frequency = 1.0f, turb_time_scale = 1.0f;
    uint32_t turb_seed = 0;
    uint8_t turb_octaves = 3;
    bool turb_div_free = false;
    // CcdConfig::default (ccd.rs:55-64)
    bool ccd_on = false;
    uint32_t ccd_tier = 0;
    float ccd_max_movement = 0.1f;
    uint32_t ccd_max_substeps = 8;
};

struct SnapBody {
    uint32_t shape = 0;  // PHIENG_SHAPE3_SPHERE / BOX / CAPSULE; 2D: circle 0, box 3 (the BodyDev shapes)
    float pos[3] = {0.0f, 0.0f, 0.0f};
    float a = 0.0f, b = 0.0f, c = 0.0f;  // radius | half extents | radius, half length
    float friction = 0.4f, restitution = 0.1f, density = 1.0f;  // material.rs:31-36
    float specific_heat = 900.0f, thermal_conductivity = 1.0f, emissivity = 0.9f;
    float linear_damping = 0.01f, angular_damping = 0.01f;
    bool sleeping = false;
    float gravity_scale = 1.0f, ambient_scale = 1.0f;
    // a dynamic body's motion state (3D worlds); a static body keeps the
    // identity rotation and zero velocities
    bool is_static = true;
    float rot[4] = {0.0f, 0.0f, 0.0f, 1.0f};
    float vel[3] = {0.0f, 0.0f, 0.0f}, angvel[3] = {0.0f, 0.0f, 0.0f};
};

struct SnapWorld {
    int dims = 3;          // 2: a World2 (World2Dto, snapshot.rs:3048-3113; z components unused)
    float cell_size = 1.0f;  // World2's broad-phase cell (world.rs:334), kept verbatim
    float gravity[3] = {0.0f, -9.81f, 0.0f};
    float time = 0.0f;
    SnapMeta meta;
    std::vector<SnapBody> bodies;
    FluidConfig3Ffi cfg;
    std::vector<float> pos, vel, mass;  // 3, 3, 1 per particle, insertion order
    std::vector<uint8_t> phase;
};

struct World2;
World3 *world2_base(World2 *w);
World2 *world2_of(World3 *w);  // w must have been made as a World2

// The world as a snapshot (reads the particles back from the device).
void world3_snapshot(World3 *w, SnapWorld &out);
// A new world from a snapshot (a World2 when s.dims == 2), or null with the
// reason in err.
World3 *world3_from_snapshot(const SnapWorld &s, std::string &err);
