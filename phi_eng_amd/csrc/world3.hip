// world3.hip — the PhiEng World3 fluid C ABI over the GPU engine.
//
// Each entry point keeps the reference's argument meaning and error
// behaviour (crates/engine/src/ffi.rs, cited per function); nothing throws
// across the boundary: HIP and allocation failures become the same return
// codes the reference uses, with the message kept for phieng_last_error().
#include <hip/hip_runtime.h>

#include <algorithm>
#include <memory>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/phieng_fluid2.h"
#include "../../include/phieng_fluid3.h"
#include "coupling.h"
#include "engine.h"
#include "resample.h"
#include "rigid3.h"
#include "sph_kernels.h"
#include "surrogate.h"
#include "surrogate_step.h"
#include "world3_internal.h"

static_assert(sizeof(FluidConfig3Ffi) == 112, "FluidConfig3Ffi must match ffi.rs:743-767");
static_assert(sizeof(BodyDesc3D) == 44, "BodyDesc3D must match ffi.rs:3765-3784");
static_assert(sizeof(BodyDesc2D) == 36, "BodyDesc2D must match ffi.rs:1205-1222");
static_assert(offsetof(FluidConfig3Ffi, max_substeps) == 32, "layout");
static_assert(offsetof(FluidConfig3Ffi, adaptive_resolution) == 36, "layout");
static_assert(offsetof(FluidConfig3Ffi, boundary_particles) == 48, "layout");
static_assert(offsetof(FluidConfig3Ffi, phase_count) == 64, "layout");
static_assert(offsetof(FluidConfig3Ffi, phase_rest_density) == 68, "layout");
static_assert(offsetof(FluidConfig3Ffi, phase_mass) == 84, "layout");
static_assert(offsetof(FluidConfig3Ffi, max_velocity) == 100, "layout");
static_assert(offsetof(FluidConfig3Ffi, enabled) == 108, "layout");

namespace {

thread_local std::string g_last_error;

void set_error(const char *where, const char *what) {
    g_last_error = std::string(where) + ": " + what;
}

constexpr uint32_t kDefaultCapacity = 4096;        // MAX_PARTICLES_3D, fluid3.rs:28
constexpr uint32_t kMaxPerCall = 1000000;          // ffi.rs:3587
constexpr uint32_t kMaxCapacity = PH_ID_MASK + 1;  // 24-bit original index

// FluidConfig3::default (fluid3.rs:105-134).
FluidConfig3Ffi default_config() {
    FluidConfig3Ffi c;
    std::memset(&c, 0, sizeof(c));
    c.particle_mass = 0.0625f;
    c.rest_density = 1.0f;
    c.smoothing_length = 0.5f;
    c.stiffness = 100.0f;
    c.gamma = 1.0f;
    c.artificial_viscosity = 0.0f;
    c.surface_tension = 0.0f;
    c.cfl = 0.0f;
    c.max_substeps = 1;
    c.adaptive_resolution = 0;
    c.split_density_ratio = 0.5f;
    c.merge_density_ratio = 1.6f;
    c.boundary_particles = 0;
    c.viscosity = 0.3f;
    c.drag = 4.0f;
    c.wall_friction = 0.0f;
    c.phase_count = 1;
    for (int k = 0; k < 4; ++k) {
        c.phase_rest_density[k] = 1.0f;
        c.phase_mass[k] = 0.0625f;
    }
    c.max_velocity = 30.0f;
    c.particle_radius = 0.125f;
    c.enabled = 1;
    return c;
}

// fluid3_from_ffi (ffi.rs:857-914).
bool config_valid(const FluidConfig3Ffi &c) {
    const float f[] = {c.particle_mass, c.rest_density, c.smoothing_length, c.stiffness, c.gamma,
                       c.artificial_viscosity, c.surface_tension, c.cfl, c.split_density_ratio,
                       c.merge_density_ratio, c.viscosity, c.drag, c.wall_friction, c.max_velocity,
                       c.particle_radius, c.phase_rest_density[0], c.phase_rest_density[1],
                       c.phase_rest_density[2], c.phase_rest_density[3], c.phase_mass[0], c.phase_mass[1],
                       c.phase_mass[2], c.phase_mass[3]};
    for (float v : f)
        if (!std::isfinite(v)) return false;
    if (c.particle_mass <= 0.0f || c.rest_density <= 0.0f || c.smoothing_length <= 0.0f || c.stiffness < 0.0f ||
        c.gamma < 1.0f || c.artificial_viscosity < 0.0f || c.artificial_viscosity > 2.0f ||
        c.surface_tension < 0.0f || c.cfl < 0.0f || c.cfl > 1.0f || c.max_substeps < 1 ||
        c.split_density_ratio <= 0.0f || c.merge_density_ratio <= 1.0f || c.viscosity < 0.0f ||
        c.viscosity > 1.0f || c.drag < 0.0f || c.wall_friction < 0.0f || c.wall_friction > 1.0f ||
        c.phase_count < 1 || c.phase_count > 4 || c.max_velocity < 0.0f || c.particle_radius < 0.0f)
        return false;
    for (int k = 0; k < 4; ++k)
        if (c.phase_rest_density[k] <= 0.0f || c.phase_mass[k] <= 0.0f) return false;
    return true;
}

// The snapshot record of a static body made by add_body: the material is
// sanitised as PhysicsMaterial::new and RigidBody3::set_material do
// (material.rs:40-60, body.rs:1306-1317).
SnapBody snap_body(const BodyDev &bd, float density, float friction, float restitution) {
    SnapBody b;
    b.shape = bd.shape;
    b.pos[0] = bd.px;
    b.pos[1] = bd.py;
    b.pos[2] = bd.pz;
    b.a = bd.a;
    b.b = bd.b;
    b.c = bd.c;
    b.friction = std::isfinite(friction) && friction >= 0.0f ? friction : 0.4f;
    b.restitution = std::isfinite(restitution) ? std::min(std::max(restitution, 0.0f), 1.0f) : 0.1f;
    b.density = density;
    return b;
}

}  // namespace

struct World3 {
    int dims = 3;  // 2 for a World2: particles, gravity and bodies in the z = 0 plane
    FluidConfig3Ffi cfg = default_config();
    float g[3] = {0.0f, -9.81f, 0.0f};  // World3::new (world.rs)
    std::vector<BodyDev> bodies;
    std::vector<SnapBody> body_meta;  // material, damping, ... of each body, for snapshots
    // Rigid state of each body (a static body's never changes); the
    // dynamic ones are advanced by the world step and coupled both ways to
    // the fluid (fluid3.rs:1141-1213).
    std::vector<ph::Rigid3> rigid;
    uint32_t n_dynamic = 0;
    // coupling buffers: deferred slots, their staging, sample points and
    // pressure terms
    ph::DBuf<uint32_t> cp_defer, cp_count;
    ph::DBuf<float4> cp_stage, cp_bound, cp_spos;
    ph::DBuf<uint2> cp_range;
    ph::DBuf<int4> cp_scell;
    ph::DBuf<uint4> cp_rec;
    uint32_t cp_rec_cap = 4096;
    uint64_t n_deferred = 0, n_terms = 0;
    SnapMeta meta;                    // world fields the GPU step does not read, for snapshots
    float cell_size = 1.0f;           // a World2's broad-phase cell, kept for its snapshots
    int32_t next_body_id = 0;
    float time = 0.0f;
    uint32_t capacity = kDefaultCapacity;
    uint32_t n = 0;
    bool domain_dirty = true;
    ph::Engine eng;
    // learned pressure term in the loop
    Surrogate3 *sur = nullptr;
    uint32_t check_every = 10;
    float max_err = 0.01f;
    int use_net = 1;
    uint32_t fallback_left = 0;
    uint64_t step_index = 0;
    uint64_t n_sur = 0, n_plain = 0, n_fallback = 0;
    float last_err = -1.0f;
    // The judge's second measure: before its plain substep the surrogate is
    // evaluated on the judged state for every kJudgeStride-th run of 64
    // slots (the offset rotating per judge); the plain force pass then sums
    // |a_plain - a_sur| and |a_plain| over the same particles.  The verdict
    // fails above max_accel_err of their ratio as well as above max_err.
    static constexpr uint32_t kJudgeStride = 16;
    float max_accel_err = 0.2f;
    float last_accel_err = -1.0f;
    uint32_t n_judges = 0;
    ph::DBuf<float4> judge_acc;
    // A list built at a judge may be sized for two judge intervals; the
    // next judge then keeps the particle order and the list (no sort, no
    // list build: substep_kept_judge).  The list's judge, the intervals it
    // covers, and what its verdict read back: the largest speed and the
    // count of particles whose list did not fit.
    int64_t list_judge = -1;
    uint32_t list_k = 1;
    bool list_checked = false;
    uint32_t list_slow = 0;
    float list_vmax = -1.0f;
    ph::DBuf<uint32_t> slow_cnt;
    uint64_t n_kept_judges = 0;
    // training-sample collection from plain substeps
    Surrogate3 *col = nullptr;
    uint32_t col_every = 10, col_stride = 64, col_phase = 0;
    // The judge's verdict is read back asynchronously (pinned memory +
    // event) and applied exactly two substeps after its plain substep, so
    // the host never drains the stream for it.
    float *judge_host = nullptr;
    hipEvent_t judge_ev = nullptr;
    bool judge_pending = false;
    uint64_t judge_at = 0;
    // skin list of the last sort (the engine's neighbour list, written with
    // skin > 0 while a surrogate is attached; surrogate substeps read it)
    // list radius (1 + skin') h with skin' = max(skin, 4 check_every dt vmax / h):
    // a Verlet skin over the substeps one list serves, at least `skin`
    float skin = 0.01f;
    // neighbour records of generation k: rec_m positions (float4), then
    // rec_m packed velocity words (uint2) in the same allocation
    ph::DBuf<float4> sl_rec[2];
    size_t rec_m = 0;
    SurRec recs(int k) const {
        SurRec r;
        r.pos = sl_rec[k].p;
        r.vd = reinterpret_cast<uint2 *>(sl_rec[k].p + rec_m);
        return r;
    }
    uint64_t list_sorts = ~0ull, list_epoch = ~0ull;
    uint32_t list_stride = 0;
    uint64_t n_col_skipped = 0;
    // first error of a step since the last sync (steps return nothing)
    std::string step_error;
    // boundary ghosts (fluid3.rs:1018-1027): generation-order copy, the
    // fingerprint they were made for, and how often they were made
    std::vector<float4> ghost_gen;
    uint64_t ghost_fp = 0, ghost_regens = 0;

    // boundary_fingerprint (fluid3.rs:660-714, fluid.rs:617-659): FNV-1a
    // over the kernel config and every static body's transform and shape
    // (identity rotation: quaternion (0, 0, 0, 1), or the 2D matrix columns).
    uint64_t boundary_fingerprint() const {
        uint64_t hsh = 1469598103934665603ull;
        auto mix = [&](const void *p, size_t k) {
            const uint8_t *b = static_cast<const uint8_t *>(p);
            for (size_t i = 0; i < k; ++i) {
                hsh ^= b[i];
                hsh *= 1099511628211ull;
            }
        };
        auto mixf = [&](float v) { mix(&v, 4); };  // to_le_bytes on a little-endian host
        auto mixb = [&](uint8_t v) { mix(&v, 1); };
        mixf(cfg.smoothing_length);
        mixf(cfg.particle_mass);
        mixf(cfg.rest_density);
        for (const BodyDev &b : bodies) {
            if (b.flags & BODY_DYNAMIC) continue;  // static bodies only (fluid3.rs:673)
            mixf(b.px);
            mixf(b.py);
            if (dims == 2) {
                mixf(1.0f), mixf(0.0f), mixf(0.0f), mixf(1.0f);
                if (b.shape == 0u) {
                    mixb(1);
                    mixf(b.a);
                } else {
                    mixb(2);
                    mixf(b.a);
                    mixf(b.b);
                }
                continue;
            }
            mixf(b.pz);
            mixf(0.0f), mixf(0.0f), mixf(0.0f), mixf(1.0f);
            if (b.shape == 0u) {
                mixb(1);
                mixf(b.a);
            } else if (b.shape == 1u) {
                mixb(2);
                mixf(b.a);
                mixf(b.b);
                mixf(b.c);
            } else {
                mixb(3);
                mixf(b.a);
                mixf(b.b);
            }
        }
        return hsh;
    }

    // Per substep, as the reference (fluid3.rs:1018-1025): regenerate the
    // ghosts when the fingerprint changed or there are none; with the option
    // off the ghosts are kept but unused.
    void update_ghosts() {
        SceneDev &s = eng.scenes[0];
        uint32_t want_count = 0, want_mask = 0;
        if (cfg.boundary_particles) {
            const uint64_t fp = boundary_fingerprint();
            if (fp != ghost_fp || ghost_gen.empty()) {
                ph::make_ghosts(scene_host(s.dt), ghost_gen);
                eng.set_ghosts(ghost_gen, std::max(cfg.smoothing_length, 1e-6f));
                ghost_fp = fp;
                ++ghost_regens;
            }
            want_count = eng.g_count;
            want_mask = eng.g_mask;
        }
        if (s.ghost_count != want_count || s.ghost_mask != want_mask) {
            s.ghost_count = want_count;
            s.ghost_mask = want_mask;
            eng.scenes_dirty = true;
            eng.upload_tables();
        }
    }

    ph::SceneHost scene_host(float dt) const {
        ph::SceneHost h;
        h.cfg = cfg;
        h.g[0] = g[0];
        h.g[1] = g[1];
        h.g[2] = g[2];
        h.dt = dt;
        h.bodies = bodies;
        h.dims = dims;
        return h;
    }

    // Bring the engine's single scene in line with the host state.
    void sync_scene(float dt) {
        ph::Engine &e = eng;
        if (e.scenes.size() != 1) {
            e.scenes.assign(1, SceneDev{});
            domain_dirty = true;
        }
        SceneDev &s = e.scenes[0];
        SceneDev before = s;
        ph::fill_scene_physics(s, scene_host(dt));
        s.part_begin = 0;
        s.part_end = n;
        s.body_begin = 0;
        s.body_count = (uint32_t)bodies.size();
        if (e.bodies.size() != bodies.size() ||
            (bodies.size() && std::memcmp(e.bodies.data(), bodies.data(), sizeof(BodyDev) * bodies.size()))) {
            // a dynamic body's move leaves the domain (it follows the particles)
            bool statics_same = e.bodies.size() == bodies.size();
            for (size_t k = 0; statics_same && k < bodies.size(); ++k)
                if (!(bodies[k].flags & BODY_DYNAMIC))
                    statics_same = std::memcmp(&e.bodies[k], &bodies[k], sizeof(BodyDev)) == 0;
            e.bodies = bodies;
            e.bodies_dirty = true;
            if (!statics_same) domain_dirty = true;
        }
        if (before.h_grid != s.h_grid) domain_dirty = true;
        // the tiles cover [part_begin, part_end): rebuilt when the count
        // changed since they were laid out (an add or a load lays them out
        // before the scene range is updated here)
        if (e.n_total != n || e.tiles.empty() || before.part_end != s.part_end) {
            e.n_total = n;
            e.build_tiles();
        }
        if (domain_dirty) {
            int32_t box[6];
            e.scenes_dirty = true;
            e.upload_tables();
            e.particle_cell_box(0, box);
            ph::choose_domain(s, box, n, bodies);
            e.assign_buckets();
            domain_dirty = false;
        } else if (std::memcmp(&before, &s, sizeof(SceneDev)) != 0) {
            e.scenes_dirty = true;
        }
        e.upload_tables();
    }

    uint32_t cfl_substeps(float dt) {
        // fluid3.rs:610-636
        const uint32_t mx = std::max<uint32_t>(cfg.max_substeps, 1u);
        if (cfg.cfl <= 0.0f || mx <= 1) return 1;
        auto rmax = [](float a, float b) { return a != a ? b : (b != b ? a : (a > b ? a : b)); };
        const float h = rmax(cfg.smoothing_length, 1e-6f);
        const float rho0 = rmax(cfg.rest_density, 1e-6f);
        const float c0 = cfg.gamma == 1.0f ? std::sqrt(rmax(cfg.stiffness, 0.0f))
                                           : std::sqrt(cfg.gamma * rmax(cfg.stiffness, 0.0f) * rho0);
        const float vmax = eng.max_speed(0, n);
        const float speed = c0 + vmax;
        if (!std::isfinite(speed) || speed <= 1e-6f) return 1;
        const float dt_max = cfg.cfl * h / speed;
        if (!std::isfinite(dt_max) || dt_max <= 0.0f) return 1;
        const float q = std::ceil(dt / dt_max);
        int64_t s;
        if (q != q) s = 0;
        else if (q >= 9.2233720368547758e18f) s = INT64_MAX;
        else if (q <= -9.2233720368547758e18f) s = INT64_MIN;
        else s = (int64_t)q;
        return (uint32_t)std::min<int64_t>(std::max<int64_t>(s, 1), (int64_t)mx);
    }

    // World3::step fluid phase (world.rs:2670-2697, 2809-2812) and
    // FluidState3::step_fluid (fluid3.rs:591-607).
    // phieng_world3_add_body (ffi.rs:3703-3760): the shape and density
    // checks, then a static body or a dynamic one with the mass properties
    // of RigidBody3::set_shape.  Dynamic bodies are refused while the fluid
    // runs the learned surrogate or adaptive resolution (their substeps do
    // not couple).
    int32_t add_body3(float px, float py, float pz, uint32_t shape_tag, float a, float b, float c, float density,
                      uint8_t is_static, float friction, float restitution, const char *where) {
        if (!std::isfinite(px) || !std::isfinite(py) || !std::isfinite(pz)) return -1;
        BodyDev bd{shape_tag, px, py, pz, a, b, c};
        switch (shape_tag) {
            case PHIENG_SHAPE3_SPHERE:
                if (!(a > 0.0f && std::isfinite(a))) return -1;
                bd.b = bd.c = 0.0f;
                break;
            case PHIENG_SHAPE3_BOX:
                if (!(a > 0.0f && b > 0.0f && c > 0.0f && std::isfinite(a) && std::isfinite(b) && std::isfinite(c)))
                    return -1;
                break;
            case PHIENG_SHAPE3_CAPSULE:
                if (!(a > 0.0f && b > 0.0f && std::isfinite(a) && std::isfinite(b))) return -1;
                bd.c = 0.0f;
                break;
            default:
                return -1;
        }
        if (!std::isfinite(density) || density <= 0.0f) return -1;
        if (!is_static && (sur || col || cfg.adaptive_resolution)) {
            set_error(where, "a dynamic body with the learned surrogate or adaptive resolution: those substeps do "
                             "not couple to bodies");
            return -1;
        }
        if (bodies.size() >= PH_MAX_BODIES) {
            set_error(where, "too many bodies");
            return -1;
        }
        ph::Rigid3 r;
        r.is_static = is_static != 0;
        r.pos = ph::f3(px, py, pz);
        r.set_shape(shape_tag, bd.a, bd.b, bd.c, density);
        if (!r.is_static) {
            bd = r.dev();
            ++n_dynamic;
        }
        bodies.push_back(bd);
        rigid.push_back(r);
        SnapBody sb = snap_body(bd, density, friction, restitution);
        sb.is_static = r.is_static;
        body_meta.push_back(sb);
        return next_body_id++;
    }

    // phieng_world_add_body (ffi.rs:1149-1203) for a World2: a circle (a
    // disc in the z = 0 plane) or a box.  Friction and restitution are
    // sanitised, never rejected, by set_material (body.rs:627-637), and
    // static walls take wall_friction from the config; they are kept for
    // snapshots.  A 2D world's bodies are static: the 2D fluid's coupling
    // (fluid.rs) is not built on the GPU path.
    int32_t add_body2(float px, float py, uint32_t shape_tag, float a, float b, float density, uint8_t is_static,
                      float friction, float restitution, const char *where) {
        if (!std::isfinite(px) || !std::isfinite(py)) return -1;
        BodyDev bd{0u, px, py, 0.0f, a, 0.0f, 0.0f};
        switch (shape_tag) {
            case PHIENG_SHAPE2_CIRCLE:
                if (!(a > 0.0f && std::isfinite(a))) return -1;
                bd.shape = 0u;
                break;
            case PHIENG_SHAPE2_BOX:
                if (!(a > 0.0f && b > 0.0f && std::isfinite(a) && std::isfinite(b))) return -1;
                bd.shape = 3u;
                bd.b = b;
                break;
            default:
                return -1;
        }
        if (!std::isfinite(density) || density <= 0.0f) return -1;
        if (!is_static) {
            set_error(where, "dynamic bodies in a 2D world: the GPU path couples 3D worlds only");
            return -1;
        }
        if (bodies.size() >= PH_MAX_BODIES) {
            set_error(where, "too many bodies");
            return -1;
        }
        ph::Rigid3 r;
        r.shape = bd.shape;
        r.pos = ph::f3(px, py, 0.0f);
        r.a = bd.a;
        r.b = bd.b;
        bodies.push_back(bd);
        rigid.push_back(r);
        body_meta.push_back(snap_body(bd, density, friction, restitution));
        return next_body_id++;
    }

    void step(float dt) {
        // World3::step_inner (world.rs:2766-2812) as far as bodies move
        // without the rigid contact solver: the gravity kick, the fluid
        // with its two-way coupling, the integration.
        if (n_dynamic) {
            const ph::F3 gv = ph::f3(g[0], g[1], g[2]);
            for (ph::Rigid3 &r : rigid) r.gravity_kick(gv, dt);
            for (size_t k = 0; k < rigid.size(); ++k)
                if (!rigid[k].is_static) bodies[k] = rigid[k].dev();
        }
        step_fluid(dt);
        if (n_dynamic)
            for (ph::Rigid3 &r : rigid) r.integrate(dt);
        time += dt;
    }

    void step_fluid(float dt) {
        if (cfg.enabled && n > 0 && dt > 0.0f) {
            eng.init_device();
            (void)hipGetLastError();  // a failed call before this step has already reported
            if (!domain_dirty) eng.poll_monitor(n);
            // plain substeps leave a skin list for surrogate substeps
            eng.list_skin = sur ? skin : 0.0f;
            eng.list_cells = sur != nullptr;
            sync_scene(dt);
            const uint32_t sub = cfl_substeps(dt);
            const float sdt = dt / (float)sub;
            eng.list_coef = sur ? 4.0f * (float)check_every * sdt / cfg.smoothing_length : 0.0f;
            if (sub != 1) sync_scene(sdt);
            for (uint32_t k = 0; k < sub; ++k) {
                if (cfg.boundary_particles || eng.scenes[0].ghost_count) update_ghosts();
                if (n_dynamic) {
                    substep_coupled(sdt);
                    continue;
                }
                if (cfg.adaptive_resolution) {
                    substep_adaptive(sdt);
                    continue;
                }
                if (!sur && !col) {
                    eng.substep();
                    continue;
                }
                ensure_aux();
                if (judge_pending && step_index >= judge_at + 2) apply_judge();
                const bool collect = col && (step_index % col_every) == 0;
                const bool judge = sur && (step_index % check_every) == 0;
                if (sur && !judge && !collect && fallback_left == 0 && list_usable()) {
                    substep_surrogate();
                    ++n_sur;
                } else if (judge && !collect && fallback_left == 0 && list_kept()) {
                    substep_kept_judge();
                    ++n_plain;
                    ++n_kept_judges;
                } else {
                    // the next substep may be a surrogate one: leave it a list
                    substep_plain_aux(judge, collect, sur && fallback_left <= 1, sdt);
                    ++n_plain;
                    if (fallback_left) --fallback_left;
                }
                ++step_index;
            }
        }
    }

    // One plain substep coupled to the dynamic bodies.  The force pass
    // leaves particles touching a dynamic body unclamped; the host replays
    // the reference's clamp loop for them in particle order (the bodies'
    // impulses land in that order and wall friction reads the running body
    // velocity), then the pressure is sampled on each dynamic body's
    // surface from the clamped state (the reference's second grid rebuild)
    // and the buoyancy and drag forces are applied in sample order.
    void substep_coupled(float sdt) {
        ph::Engine &e = eng;
        const size_t m = e.alloc_particles;
        cp_defer.ensure(m + 1);
        e.run(ph::KIND_FILL, [&] { return ph::launch_fill_u32(e.stream, cp_defer.p, 1u, 0u); });
        e.sort_phase(nullptr);
        AuxFuse af;
        af.defer = cp_defer.p;
        af.defer_cap = (uint32_t)m;
        e.plain_phase(nullptr, &af);
        e.finish_substep();
        uint32_t cnt = 0;
        PH_CK(hipMemcpyAsync(&cnt, cp_defer.p, sizeof(uint32_t), hipMemcpyDeviceToHost, e.stream));
        PH_CK(hipStreamSynchronize(e.stream));
        cnt = std::min<uint32_t>(cnt, n);
        if (cnt) replay_clamp(cnt);
        buoyancy(sdt);
    }

    // fluid3.rs:1141-1179 for the deferred particles, ascending index.
    void replay_clamp(uint32_t cnt) {
        ph::Engine &e = eng;
        cp_stage.ensure(2 * (size_t)cnt);
        e.run(ph::KIND_EXPORT, [&] {
            return ph::launch_defer_gather(e.stream, cp_defer.p, cnt, e.posm[e.cs].p, e.vel[e.cs].p, e.tag[e.ct].p,
                                           cp_stage.p);
        });
        std::vector<float4> st(2 * (size_t)cnt);
        PH_CK(hipMemcpyAsync(st.data(), cp_stage.p, sizeof(float4) * st.size(), hipMemcpyDeviceToHost, e.stream));
        PH_CK(hipStreamSynchronize(e.stream));
        std::vector<std::pair<uint32_t, uint32_t>> order(cnt);
        for (uint32_t k = 0; k < cnt; ++k) {
            uint32_t tg;
            std::memcpy(&tg, &st[2 * k + 1].w, 4);
            order[k] = {tg & PH_ID_MASK, k};
        }
        std::sort(order.begin(), order.end());
        const float pr = cfg.particle_radius > 0.0f ? cfg.particle_radius : 0.0f;
        const float mu = std::min(std::max(cfg.wall_friction, 0.0f), 1.0f);
        for (const auto &ok : order) {
            float4 &P = st[2 * ok.second];
            float4 &V = st[2 * ok.second + 1];
            float x = P.x, y = P.y, z = P.z, vx = V.x, vy = V.y, vz = V.z;
            const float mass = P.w;
            for (size_t k = 0; k < bodies.size(); ++k) {
                float nx, ny, nz, depth, cp[3];
                if (!collide(x, y, z, pr, bodies[k], nx, ny, nz, depth, cp)) continue;
                if (depth <= 0.0f) continue;
                x = x + nx * depth;
                y = y + ny * depth;
                z = z + nz * depth;
                const float vn = vx * nx + vy * ny + vz * nz;
                if (!(vn < 0.0f)) continue;
                vx = vx - nx * vn;
                vy = vy - ny * vn;
                vz = vz - nz * vn;
                ph::Rigid3 &r = rigid[k];
                const ph::F3 n3 = ph::f3(nx, ny, nz), point = ph::f3(cp[0], cp[1], cp[2]);
                if (!r.is_static) r.sleeping = false;
                r.apply_impulse(n3 * (vn * mass), point);
                if (cfg.wall_friction > 0.0f) {
                    const ph::F3 wv = r.vel;  // zero for a static body
                    const float wd = wv.x * nx + wv.y * ny + wv.z * nz;
                    const float wtx = wv.x - nx * wd, wty = wv.y - ny * wd, wtz = wv.z - nz * wd;
                    const float dvx = (wtx - vx) * mu, dvy = (wty - vy) * mu, dvz = (wtz - vz) * mu;
                    vx = vx + dvx;
                    vy = vy + dvy;
                    vz = vz + dvz;
                    r.apply_impulse(-ph::f3(dvx, dvy, dvz) * mass, point);
                }
            }
            P.x = x;
            P.y = y;
            P.z = z;
            V.x = vx;
            V.y = vy;
            V.z = vz;
        }
        PH_CK(hipMemcpyAsync(cp_stage.p, st.data(), sizeof(float4) * st.size(), hipMemcpyHostToDevice, e.stream));
        e.run(ph::KIND_EXPORT, [&] {
            return ph::launch_defer_scatter(e.stream, cp_defer.p, cnt, cp_stage.p, e.posm[e.cs].p, e.vel[e.cs].p);
        });
        n_deferred += cnt;
    }

    // fluid3.rs:1180-1213: pressure sampled on each dynamic body's surface,
    // then per sample (skipping dry ones) Archimedes' -n p A and the
    // Morison drag -v_rel drag A |v_rel|, applied in order.
    void buoyancy(float sdt) {
        ph::Engine &e = eng;
        std::vector<uint32_t> who;
        std::vector<std::vector<ph::Sample3>> smp;
        std::vector<float4> bound, spos;
        std::vector<uint2> range;
        std::vector<int4> scell;
        const float hk = std::max(cfg.smoothing_length, PH_EPS);
        auto cellf = [&](float v) {
            const float q = std::floor(v / hk);
            if (q != q) return 0;
            if (q >= 2147483648.0f) return INT32_MAX;
            if (q < -2147483648.0f) return INT32_MIN;
            return (int32_t)q;
        };
        for (size_t k = 0; k < rigid.size(); ++k) {
            if (rigid[k].is_static) continue;
            std::vector<ph::Sample3> ss;
            ph::surface_samples(rigid[k], ss);
            if (ss.empty()) continue;
            const float R = (rigid[k].bound() + hk) * 1.001f + 1e-6f;
            bound.push_back(make_float4(rigid[k].pos.x, rigid[k].pos.y, rigid[k].pos.z, R * R));
            range.push_back(make_uint2((uint32_t)spos.size(), (uint32_t)ss.size()));
            for (const ph::Sample3 &q : ss) {
                spos.push_back(make_float4(q.point.x, q.point.y, q.point.z, 0.0f));
                scell.push_back(make_int4(cellf(q.point.x), cellf(q.point.y), cellf(q.point.z), 0));
            }
            who.push_back((uint32_t)k);
            smp.push_back(std::move(ss));
        }
        if (who.empty()) return;
        cp_bound.ensure(bound.size());
        cp_range.ensure(range.size());
        cp_spos.ensure(spos.size());
        cp_scell.ensure(scell.size());
        cp_count.ensure(1);
        PH_CK(hipMemcpyAsync(cp_bound.p, bound.data(), sizeof(float4) * bound.size(), hipMemcpyHostToDevice, e.stream));
        PH_CK(hipMemcpyAsync(cp_range.p, range.data(), sizeof(uint2) * range.size(), hipMemcpyHostToDevice, e.stream));
        PH_CK(hipMemcpyAsync(cp_spos.p, spos.data(), sizeof(float4) * spos.size(), hipMemcpyHostToDevice, e.stream));
        PH_CK(hipMemcpyAsync(cp_scell.p, scell.data(), sizeof(int4) * scell.size(), hipMemcpyHostToDevice, e.stream));
        uint32_t want = 0;
        for (;;) {
            cp_rec.ensure(cp_rec_cap);
            e.run(ph::KIND_FILL, [&] { return ph::launch_fill_u32(e.stream, cp_count.p, 1u, 0u); });
            ph::SampleArgs a;
            a.scenes = e.d_scenes.p;
            a.n = n;
            a.posm = e.posm[e.cs].p;
            a.rq = e.rq.p;
            a.tag = e.tag[e.ct].p;
            a.bound = cp_bound.p;
            a.range = cp_range.p;
            a.nbodies = (uint32_t)bound.size();
            a.spos = cp_spos.p;
            a.scell = cp_scell.p;
            a.rec = cp_rec.p;
            a.count = cp_count.p;
            a.cap = cp_rec_cap;
            e.run(ph::KIND_EXPORT, [&] { return ph::launch_sample_terms(e.stream, a); });
            PH_CK(hipMemcpyAsync(&want, cp_count.p, sizeof(uint32_t), hipMemcpyDeviceToHost, e.stream));
            PH_CK(hipStreamSynchronize(e.stream));
            if (want <= cp_rec_cap) break;
            cp_rec_cap = want + want / 2;  // the record store was too small: grow, run again
        }
        std::vector<uint4> rec(want);
        if (want) {
            PH_CK(hipMemcpyAsync(rec.data(), cp_rec.p, sizeof(uint4) * want, hipMemcpyDeviceToHost, e.stream));
            PH_CK(hipStreamSynchronize(e.stream));
        }
        n_terms += want;
        // per sample, the terms in the reference's visiting order
        std::sort(rec.begin(), rec.end(), [](const uint4 &p, const uint4 &q) {
            return p.x != q.x ? p.x < q.x : p.y < q.y;
        });
        std::vector<float> pres(spos.size(), 0.0f);
        for (const uint4 &r : rec) {
            float t;
            std::memcpy(&t, &r.z, 4);
            pres[r.x] = pres[r.x] + t;
        }
        for (float &p : pres) p = p > 0.0f ? p : 0.0f;  // p.max(0.0): NaN gives 0
        const float drag = std::max(cfg.drag, 0.0f);
        for (size_t b = 0; b < who.size(); ++b) {
            ph::Rigid3 &r = rigid[who[b]];
            bool wet = false;
            for (size_t q = 0; q < smp[b].size(); ++q) {
                const float p = pres[range[b].x + q];
                if (p <= 0.0f) continue;
                wet = true;
                const ph::Sample3 &sm = smp[b][q];
                const ph::F3 rr = sm.point - r.pos;
                const ph::F3 wind = ph::f3(meta.drag_wind[0], meta.drag_wind[1], meta.drag_wind[2]);
                const ph::F3 v_rel = (r.vel + ph::cross(r.ang, rr)) - wind;
                const float speed = ph::length(v_rel);
                ph::F3 force = -sm.normal * (p * sm.area);
                if (speed > PH_EPS) force = force - v_rel * (drag * sm.area * speed);
                r.apply_force(force, sm.point, sdt);
            }
            if (wet) r.sleeping = false;
        }
    }

    ~World3() {
        if (judge_ev) (void)hipEventDestroy(judge_ev);
        if (judge_host) (void)hipHostFree(judge_host);
    }

    // Read the pending verdict (waiting for its plain substep if needed) and
    // start the fallback if it failed: the next check_every substeps are
    // plain, which covers the next judge and the substeps until its own
    // verdict arrives, so a run of failing judges stays plain throughout.
    // Called at fixed substep indices only, so the decisions never depend
    // on timing.
    void apply_judge() {
        if (!judge_pending) return;
        PH_CK(hipEventSynchronize(judge_ev));
        judge_pending = false;
        last_err = judge_host[0];
        last_accel_err = judge_host[2];
        if ((int64_t)judge_at == list_judge) {
            list_checked = true;
            list_slow = (uint32_t)judge_host[4];
            list_vmax = judge_host[5];
        }
        if (judge_fails(last_err, last_accel_err)) {
            fallback_left = check_every;
            ++n_fallback;
        }
    }
    bool judge_fails(float dens_err, float accel_err) const {
        return dens_err > max_err || accel_err > max_accel_err;  // accel_err < 0: not evaluated
    }

    float4 mass4() const {
        auto pm = [&](int p) { return std::max(cfg.phase_mass[p], 0.0f); };
        return make_float4(pm(0), pm(1), pm(2), pm(3));
    }

    float qs() const {
        const float g = cfg.gamma > 1.0f ? cfg.gamma : 1.0f;
        return std::max(g * cfg.stiffness * 1e-3f / std::max(cfg.rest_density, 1e-6f), 1e-12f);
    }

    // Per-slot history for the surrogate, reset to rest density.
    void ensure_aux() {
        ph::Engine &e = eng;
        if (e.aux_on && e.aux[0].n >= e.alloc_particles) return;
        const size_t m = e.alloc_particles;
        e.aux[0].ensure(m);
        e.aux[1].ensure(m);
        e.perm.ensure(m);
        e.acc_sorted.ensure(m);
        e.tile_err.ensure(std::max<size_t>(e.tiles.size(), 1));
        e.err_out.ensure(6);
        if (!slow_cnt.p) {
            slow_cnt.ensure(1);
            e.run(ph::KIND_FILL, [&] { return ph::launch_fill_u32(e.stream, slow_cnt.p, 1u, 0u); });
            // 0.0f is the all-zero word
            e.run(ph::KIND_FILL, [&] {
                return ph::launch_fill_u32(e.stream, reinterpret_cast<uint32_t *>(e.err_out.p), 6u, 0u);
            });
        }
        if (!judge_ev) {
            PH_CK(hipEventCreateWithFlags(&judge_ev, hipEventDisableTiming));
            PH_CK(hipHostMalloc((void **)&judge_host, 6 * sizeof(float), hipHostMallocDefault));
        }
        const float4 r0 = make_float4(cfg.phase_rest_density[0], cfg.phase_rest_density[1],
                                      cfg.phase_rest_density[2], cfg.phase_rest_density[3]);
        e.run(ph::KIND_SUR_AUX, [&] { return ph::launch_aux_reset(e.stream, n, e.tag[e.ct].p, r0, e.aux[e.cs].p); });
        e.aux_on = true;
    }

    bool list_usable() const {
        return eng.nl_cell.p && sl_rec[0].p && list_sorts == eng.sorts_done && list_epoch == eng.layout_epoch &&
               list_stride == eng.list_stride();
    }

    // A judge at this substep may keep the order and the list: the list
    // was built at a judge whose verdict has arrived with no particle
    // outside its list, and its radius covers the substeps up to the next
    // judge (whose evaluation reads it too).
    bool list_kept() const {
        return list_usable() && list_judge >= 0 && list_checked && list_slow == 0 &&
               (int64_t)(step_index + check_every) - list_judge <= (int64_t)(list_k * check_every);
    }

    // Judge intervals the list built at this judge is sized for.  Two when
    // the list for two, from the speed at the last list's judge, holds few
    // enough extra candidates: per interval, a full judge plus (c - 1)
    // surrogate substeps against half of a full and a kept judge plus
    // (c - 1) substeps over the larger list.  Costs in units of one
    // surrogate substep, measured on the 1,048,576-particle dam break
    // (full judge 3.8, kept judge 2.2); a surrogate substep's cost grows
    // with the candidates, (1 + skin)^3.
    uint32_t plan_k(float sdt) const {
        if (list_vmax < 0.0f || check_every < 2) return 1;
        const float c = 4.0f * (float)check_every * sdt * list_vmax / std::max(cfg.smoothing_length, 1e-6f);
        const float s1 = std::max(skin, c), s2 = std::max(skin, 2.0f * c);
        const float g = (1.0f + s2) / (1.0f + s1), infl = g * g * g;
        const float m = (float)(check_every - 1), full = 3.8f, kept = 2.2f;
        return 0.5f * (full + kept) + m * infl < full + m ? 2u : 1u;
    }

    // Plain substep that also refreshes the history; optionally judged
    // (bulk density deviation, read back), collecting samples, and/or
    // building the skin list for surrogate substeps.
    void substep_plain_aux(bool judge, bool collect, bool list, float sdt) {
        ph::Engine &e = eng;
        const uint32_t nt = (uint32_t)e.tiles.size();
        const int ns = 1 - e.cs, ntg = 1 - e.ct;
        // Room for this collection, made before any launch of the substep
        // so an allocation failure cannot cut the substep short; a full
        // store stops the collection.
        const uint32_t cnt = collect && n > col_phase ? (n - 1 - col_phase) / col_stride + 1 : 0;
        if (collect && (uint64_t)col->nsamples + cnt > SUR_MAX_SAMPLES) {
            collect = false;
            ++n_col_skipped;
        }
        if (collect) col->grow_samples((uint64_t)col->nsamples + cnt);
        // the surrogate on the judged state, from the list and records its
        // own substep would read (a judge with a collection skips this)
        const bool accel = judge && sur && !collect && list_usable();
        const uint32_t joff = accel ? judge_evaluate() : 0u;
        if (judge) ++n_judges;
        // a judge that builds the list sizes it for one or two intervals and
        // reads back the largest speed and the count of particles outside it
        const bool build = list && e.list_cells && e.list_skin == skin;
        const float coef1 = e.list_coef;
        const uint32_t k = judge && build && !collect ? plan_k(sdt) : 1u;
        e.list_coef = coef1 * (float)k;
        e.list_slow = sur ? slow_cnt.p : nullptr;
        e.list_vmax_out = judge ? e.err_out.p + 5 : nullptr;
        if (sur && rec_m != e.alloc_particles) {
            // both generations are rewritten before a surrogate substep reads them
            rec_m = e.alloc_particles;
            sl_rec[0].ensure(rec_m + (rec_m + 1) / 2);
            sl_rec[1].ensure(rec_m + (rec_m + 1) / 2);
        }
        e.sort_phase(e.perm.p);
        // Without a collection the force pass also writes the new history
        // (into the other aux buffer, then swapped in), the records and the
        // judge sums; a collection reads the gathered history, so it takes
        // the separate gather and update kernels.
        const bool fuse = !collect;
        if (fuse) {
            AuxFuse af;
            af.perm = e.perm.p;
            af.aux_in = e.aux[e.cs].p;
            af.aux_out = e.aux[ns].p;
            if (sur) af.rec_out = recs(e.cs);
            af.tile_err = e.tile_err.p;
            if (accel) {
                af.sur_acc = judge_acc.p;
                af.sur_stride = kJudgeStride;
                af.sur_off = joff;
            }

            e.plain_phase(nullptr, &af);
            e.aux[0].swap(e.aux[1]);  // the new history is generation cs
        } else {
            e.run(ph::KIND_SUR_AUX,
                  [&] { return ph::launch_gather_aux(e.stream, n, e.perm.p, e.aux[e.cs].p, e.aux[ns].p); });
            e.plain_phase(e.acc_sorted.p);
        }
        e.list_coef = coef1;
        if (build) {
            // the skin list of the order this substep sorted into
            list_sorts = e.sorts_done;
            list_epoch = e.layout_epoch;
            list_stride = e.list_stride();
            list_judge = judge && !collect ? (int64_t)step_index : -1;
            list_k = k;
            list_checked = false;
        }
        if (collect) {
            SurCollectArgs a{};
            a.tiles = e.d_tiles.p;
            a.scenes = e.d_scenes.p;
            a.start = e.start.p;
            a.posm = e.posm[ns].p;
            a.vel = e.vel[ns].p;
            a.aux = e.aux[ns].p;
            a.tag = e.tag[ntg].p;
            a.mass = mass4();
            a.qs = qs();
            a.acc_target = e.acc_sorted.p;
            a.samples = col->data.p;
            a.sample_base = col->nsamples;
            a.stride = col_stride;
            a.phase = col_phase;
            a.max_samples = col->nsamples + cnt;
            e.run(ph::KIND_SUR_COLLECT, [&] { return ph::launch_sur_collect(e.stream, nt, a); });
            PH_CK(hipStreamSynchronize(e.stream));
            col->nsamples += cnt;
            col_phase = (col_phase + 1) % col_stride;
        }
        if (!fuse)
            e.run(ph::KIND_SUR_AUX, [&] {
                return ph::launch_aux_update(e.stream, nt, e.d_tiles.p, e.d_scenes.p, e.rq.p, e.aux[ns].p,
                                             e.tag[ntg].p, e.aux[e.cs].p, e.tile_err.p, e.posm[e.cs].p, e.vel[e.cs].p,
                                             sur ? recs(e.cs) : SurRec());
            });
        if (judge) judge_readback();
        e.finish_substep();
    }

    // The judge's verdict from the tile sums, written by the reduction
    // kernel straight into the pinned host words and read after the
    // judge's event (applied two substeps later, so the stream is never
    // drained for it).  An older verdict still pending (check_every < 2) is
    // applied first: its words are reused.
    void judge_readback() {
        ph::Engine &e = eng;
        apply_judge();
        const uint32_t nt = (uint32_t)e.tiles.size();
        e.run(ph::KIND_SUR_AUX, [&] {
            return ph::launch_err_reduce(e.stream, e.tile_err.p, nt, judge_host, slow_cnt.p, e.err_out.p + 5);
        });
        PH_CK(hipEventRecord(judge_ev, e.stream));
        judge_pending = true;
        judge_at = step_index;
    }

    // Launch the surrogate's evaluation on the current generation for the
    // judge's sample (judge_slot); returns its offset.
    uint32_t judge_evaluate() {
        ph::Engine &e = eng;
        const uint32_t joff = n_judges % kJudgeStride;
        judge_acc.ensure((e.alloc_particles / (64 * kJudgeStride) + 1) * 64);
        SurStepArgs a = sur_args();
        a.judge_acc = judge_acc.p;
        a.judge_stride = kJudgeStride;
        a.judge_off = joff;
        // every run that starts below n (slots past n are idle lanes)
        const uint32_t runs = n / 64 + (n % 64 ? 1 : 0);
        const uint32_t cnt = runs > joff ? ((runs - 1 - joff) / kJudgeStride + 1) * 64 : 0;
        e.run(ph::KIND_SUR_STEP, [&] { return ph::launch_sur_judge(e.stream, cnt, a); });
        return joff;
    }

    // A judging plain substep that keeps the particle order and the skin
    // list of the judge before (list_kept): the surrogate's evaluation,
    // the density from the list, the force pass over it with the history,
    // the records and the judge's sums, then the verdict's readback.  Like
    // a surrogate substep it reads generation cs and writes 1 - cs, which
    // becomes current; tags keep their order.
    void substep_kept_judge() {
        ph::Engine &e = eng;
        const uint32_t nt = (uint32_t)e.tiles.size();
        const int ns = 1 - e.cs;
        const bool rd = e.rec_flags & PHIENG_RECORD_DENSITY, ra = e.rec_flags & PHIENG_RECORD_ACCEL;
        e.vl_valid = false;
        const uint32_t joff = judge_evaluate();
        ++n_judges;
        e.run(ph::KIND_DENSITY, [&] {
            return ph::launch_density_skin(e.stream, nt, e.d_tiles.p, e.d_scenes.p, e.start.p, e.posm[e.cs].p,
                                           e.tag[e.ct].p, e.nl_cnt.p, e.nl_idx.p, list_stride, e.rq.p,
                                           rd ? e.rec_rho.p : nullptr, rd ? e.rec_p.p : nullptr,
                                           rd ? e.rec_count.p : nullptr);
        });
        AuxFuse af;
        af.aux_in = e.aux[e.cs].p;
        af.aux_out = e.aux[ns].p;
        af.rec_out = recs(ns);
        af.tile_err = e.tile_err.p;
        af.sur_acc = judge_acc.p;
        af.sur_stride = kJudgeStride;
        af.sur_off = joff;
        e.run(ph::KIND_FORCE, [&] {
            return ph::launch_force_list(e.stream, nt, e.d_tiles.p, e.d_scenes.p, e.d_bodies.p, e.start.p,
                                         e.posm[e.cs].p, e.vel[e.cs].p, e.rq.p, e.tag[e.ct].p, e.nl_cnt.p,
                                         e.nl_idx.p, list_stride, e.posm[ns].p, e.vel[ns].p,
                                         ra ? e.rec_acc.p : nullptr, e.d_bbox.p, nullptr, GhostDev{}, af);
        });
        e.cs = ns;
        judge_readback();
        // the particle records by index are this substep's; the cell record
        // belongs to the sort this substep kept
        e.rec_valid = e.rec_flags != 0 && !(e.rec_flags & PHIENG_RECORD_CELLS);
        ++e.substeps_done;
    }

    // Arguments of a surrogate substep on the current generation (reading
    // generation cs, writing 1 - cs).
    SurStepArgs sur_args() {
        ph::Engine &e = eng;
        SurStepArgs a{};
        a.use_net = use_net;
        a.tiles = e.d_tiles.p;
        a.scenes = e.d_scenes.p;
        a.bodies = e.d_bodies.p;
        a.start = e.start.p;
        a.posm = e.posm[e.cs].p;
        a.vel = e.vel[e.cs].p;
        a.aux = e.aux[e.cs].p;
        a.tag = e.tag[e.ct].p;
        a.cell = e.nl_cell.p;
        a.cnt = e.nl_cnt.p;
        a.idx = e.nl_idx.p;
        a.mstride = list_stride;
        a.rec = recs(e.cs);
        a.rec_out = recs(1 - e.cs);
        a.mass = mass4();
        a.mlp = sur->dev();
        a.qs = qs();
        a.posm_out = e.posm[1 - e.cs].p;
        a.vel_out = e.vel[1 - e.cs].p;
        a.aux_out = e.aux[1 - e.cs].p;
        return a;
    }

    // Surrogate substep on the order of the last sort: generation cs is
    // read, 1 - cs written, then becomes current (tags keep their order).
    void substep_surrogate() {
        ph::Engine &e = eng;
        e.vl_valid = false;  // moves particles without the Verlet displacement bound
        const SurStepArgs a = sur_args();
        e.run(ph::KIND_SUR_STEP, [&] { return ph::launch_sur_step(e.stream, (uint32_t)e.tiles.size(), a); });
        e.cs = 1 - e.cs;
        e.rec_valid = false;  // the density and force passes did not run
        ++e.substeps_done;
    }

    // A substep with adaptive resolution (fluid3.rs:1001-1012): density,
    // resample (merge, split), then the cell sort and density of the new
    // particle set before the ghosts and the force pass.  The split cap is
    // the world's capacity (MAX_PARTICLES_3D, 4,096, unless reserved
    // higher).  One host synchronisation per substep reads the new count.
    void substep_adaptive(float sdt) {
        ph::Engine &e = eng;
        const uint64_t room = std::min<uint64_t>(n, capacity / 2 + 1);
        if (n + room > kMaxCapacity)
            throw std::runtime_error("adaptive resolution: the particle count plus its split room exceeds 2^24");
        e.ensure_particles((uint32_t)(n + room));
        e.sort_phase(nullptr);
        e.density_phase();
        const float rest = std::max(cfg.rest_density, 1e-6f);
        bool changed = false;
        const uint32_t n2 =
            e.resample_phase(cfg.merge_density_ratio * rest, cfg.split_density_ratio * rest,
                             std::max(cfg.particle_mass, 1e-6f) * 0.125f, std::max(cfg.smoothing_length, 1e-6f),
                             capacity, changed);
        if (changed) {
            n = n2;
            sync_scene(sdt);
            e.sort_phase(nullptr);
        }
        // nothing merged or split: the sorted state and its densities are
        // those of the second pass of fluid3.rs:1010-1011 already
        e.plain_phase(nullptr, nullptr, !changed);
        e.finish_substep();
    }

    // Mass and phase per particle, in index order (an extension: the
    // reference's C ABI does not read them back).
    uint32_t get_mass(float *mass, uint8_t *phase, uint32_t max) {
        const uint32_t m = std::min(max, n);
        if (m == 0) return 0;
        eng.init_device();
        eng.export_buf.ensure(n);
        eng.staging_u8.ensure(n);
        eng.run(ph::KIND_EXPORT, [&] {
            return ph::launch_export_mass(eng.stream, n, eng.posm[eng.cs].p, eng.tag[eng.ct].p, eng.export_buf.p,
                                          eng.staging_u8.p);
        });
        if (mass)
            PH_CK(hipMemcpyAsync(mass, eng.export_buf.p, sizeof(float) * m, hipMemcpyDeviceToHost, eng.stream));
        if (phase) PH_CK(hipMemcpyAsync(phase, eng.staging_u8.p, m, hipMemcpyDeviceToHost, eng.stream));
        PH_CK(hipStreamSynchronize(eng.stream));
        return m;
    }

    // Replace the fluid by n particles with the given state (insertion
    // order), as a snapshot load does (snapshot.rs:4587-4612).
    void load_particles(const float *xyz, const float *v3, const float *mass, const uint8_t *phases, uint32_t count) {
        eng.init_device();
        n = 0;
        eng.n_total = 0;
        eng.ensure_particles(count);
        std::vector<float4> pm(count), vv(count);
        std::vector<uint32_t> tg(count);
        for (uint32_t i = 0; i < count; ++i) {
            pm[i] = make_float4(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2], mass[i]);
            vv[i] = make_float4(v3[3 * i], v3[3 * i + 1], v3[3 * i + 2], 0.0f);
            tg[i] = i | ((uint32_t)phases[i] << 24);
        }
        PH_CK(hipMemcpyAsync(eng.posm[eng.cs].p, pm.data(), sizeof(float4) * count, hipMemcpyHostToDevice,
                             eng.stream));
        PH_CK(hipMemcpyAsync(eng.vel[eng.cs].p, vv.data(), sizeof(float4) * count, hipMemcpyHostToDevice,
                             eng.stream));
        PH_CK(hipMemcpyAsync(eng.tag[eng.ct].p, tg.data(), sizeof(uint32_t) * count, hipMemcpyHostToDevice,
                             eng.stream));
        PH_CK(hipStreamSynchronize(eng.stream));
        n = count;
        eng.n_total = n;
        eng.build_tiles();
        eng.rec_valid = false;
        eng.aux_on = false;
        domain_dirty = true;
    }

    uint32_t add(const float *xyz, const uint8_t *phases, uint32_t count) {
        const uint32_t m = std::min(count, kMaxPerCall);
        for (size_t i = 0; i < 3 * (size_t)m; ++i)
            if (!std::isfinite(xyz[i])) return 0;
        const uint32_t room = capacity > n ? capacity - n : 0;
        const uint32_t k = std::min(room, m);
        if (k == 0) return 0;
        eng.init_device();
        eng.n_total = n;
        eng.ensure_particles(n + k);
        eng.staging_f.ensure(3 * (size_t)k);
        PH_CK(hipMemcpyAsync(eng.staging_f.p, xyz, sizeof(float) * 3 * k, hipMemcpyHostToDevice, eng.stream));
        const uint8_t *dph = nullptr;
        if (phases) {
            eng.staging_u8.ensure(k);
            PH_CK(hipMemcpyAsync(eng.staging_u8.p, phases, k, hipMemcpyHostToDevice, eng.stream));
            dph = eng.staging_u8.p;
        }
        auto pm = [&](int p) { return std::max(cfg.phase_mass[p], 0.0f); };
        const float4 masses = make_float4(pm(0), pm(1), pm(2), pm(3));
        const uint32_t n0 = n;
        eng.run(ph::KIND_APPEND, [&] {
            return ph::launch_append(eng.stream, n0, k, eng.staging_f.p, dph, masses, eng.posm[eng.cs].p,
                                     eng.vel[eng.cs].p, eng.tag[eng.ct].p);
        });
        PH_CK(hipStreamSynchronize(eng.stream));
        n += k;
        eng.n_total = n;
        eng.build_tiles();
        eng.rec_valid = false;
        eng.aux_on = false;  // history restarts from rest
        domain_dirty = true;
        return k;
    }

    uint32_t get(float *out, uint32_t max) {
        const uint32_t m = std::min(max, n);
        if (m == 0) return 0;
        eng.init_device();
        eng.export_buf.ensure(6 * (size_t)n);
        eng.run(ph::KIND_EXPORT, [&] {
            return ph::launch_export(eng.stream, n, eng.posm[eng.cs].p, eng.vel[eng.cs].p, eng.tag[eng.ct].p, 0u,
                                     eng.export_buf.p);
        });
        PH_CK(hipMemcpyAsync(out, eng.export_buf.p, sizeof(float) * 6 * m, hipMemcpyDeviceToHost, eng.stream));
        PH_CK(hipStreamSynchronize(eng.stream));
        return m;
    }
};

// The 2D world: a World3 in the z = 0 plane (2D coefficients, ties of the
// 2D box rule); the 2D exports below take it.
struct World2 : World3 {
    World2() { dims = 2; }
};

World3 *world2_base(World2 *w) { return w; }
World2 *world2_of(World3 *w) { return static_cast<World2 *>(w); }

uint32_t world3_count(const World3 *w) { return w->n; }
const FluidConfig3Ffi &world3_config(const World3 *w) { return w->cfg; }
const std::vector<BodyDev> &world3_bodies(const World3 *w) { return w->bodies; }
ph::SceneHost world3_scene_host(const World3 *w, float dt) { return w->scene_host(dt); }
void world3_device_state(const World3 *w, const float4 **posm, const float4 **vel, const uint32_t **tag,
                         hipStream_t *stream) {
    const ph::Engine &e = w->eng;
    *posm = e.posm[e.cs].p;
    *vel = e.vel[e.cs].p;
    *tag = e.tag[e.ct].p;
    *stream = e.stream;
}
void world3_device_state_mut(World3 *w, float4 **posm, float4 **vel, uint32_t **tag, hipStream_t *stream) {
    ph::Engine &e = w->eng;
    *posm = e.posm[e.cs].p;
    *vel = e.vel[e.cs].p;
    *tag = e.tag[e.ct].p;
    *stream = e.stream;
    w->domain_dirty = true;
    e.rec_valid = false;
}

void snapshot_set_error(const char *where, const std::string &what) { set_error(where, what.c_str()); }

void world3_snapshot(World3 *w, SnapWorld &out) {
    out.dims = w->dims;
    out.cell_size = w->cell_size;
    out.gravity[0] = w->g[0];
    out.gravity[1] = w->g[1];
    out.gravity[2] = w->g[2];
    out.time = w->time;
    out.meta = w->meta;
    out.bodies = w->body_meta;
    for (size_t k = 0; k < out.bodies.size() && k < w->rigid.size(); ++k) {
        const ph::Rigid3 &r = w->rigid[k];
        SnapBody &b = out.bodies[k];
        b.is_static = r.is_static;
        if (r.is_static) continue;
        b.pos[0] = r.pos.x;
        b.pos[1] = r.pos.y;
        b.pos[2] = r.pos.z;
        b.rot[0] = r.rot.x;
        b.rot[1] = r.rot.y;
        b.rot[2] = r.rot.z;
        b.rot[3] = r.rot.w;
        b.vel[0] = r.vel.x;
        b.vel[1] = r.vel.y;
        b.vel[2] = r.vel.z;
        b.angvel[0] = r.ang.x;
        b.angvel[1] = r.ang.y;
        b.angvel[2] = r.ang.z;
        b.sleeping = r.sleeping;
    }
    out.cfg = w->cfg;
    const uint32_t n = w->n;
    out.pos.assign(3 * (size_t)n, 0.0f);
    out.vel.assign(3 * (size_t)n, 0.0f);
    out.mass.assign(n, 0.0f);
    out.phase.assign(n, 0);
    if (!n) return;
    std::vector<float> pv(6 * (size_t)n);
    if (w->get(pv.data(), n) != n || w->get_mass(out.mass.data(), out.phase.data(), n) != n)
        throw std::runtime_error("reading the particles back failed");
    for (size_t i = 0; i < n; ++i)
        for (int k = 0; k < 3; ++k) {
            out.pos[3 * i + k] = pv[6 * i + k];
            out.vel[3 * i + k] = pv[6 * i + 3 + k];
        }
}

World3 *world3_from_snapshot(const SnapWorld &s, std::string &err) {
    std::unique_ptr<World3> w(s.dims == 2 ? static_cast<World3 *>(new World2()) : new World3());
    w->cell_size = s.cell_size;
    if (!config_valid(s.cfg)) {
        err = "the fluid config fails the C ABI's validation (ffi.rs:857-914)";
        return nullptr;
    }
    if (s.meta.turb_on) {
        err = "ambient turbulence acts on the fluid and is not implemented on the GPU path";
        return nullptr;
    }
    if (s.bodies.size() > PH_MAX_BODIES) {
        err = "too many bodies";
        return nullptr;
    }
    w->cfg = s.cfg;
    w->cfg.adaptive_resolution = s.cfg.adaptive_resolution ? 1 : 0;
    w->cfg.boundary_particles = s.cfg.boundary_particles ? 1 : 0;
    w->cfg.enabled = s.cfg.enabled ? 1 : 0;
    for (int k = 0; k < 3; ++k) w->g[k] = s.gravity[k];
    w->time = s.time;
    w->meta = s.meta;
    for (const SnapBody &b : s.bodies) {
        ph::Rigid3 r;
        r.is_static = b.is_static;
        r.pos = ph::f3(b.pos[0], b.pos[1], b.pos[2]);
        r.rot.x = b.rot[0];
        r.rot.y = b.rot[1];
        r.rot.z = b.rot[2];
        r.rot.w = b.rot[3];
        r.vel = ph::f3(b.vel[0], b.vel[1], b.vel[2]);
        r.ang = ph::f3(b.angvel[0], b.angvel[1], b.angvel[2]);
        r.linear_damping = b.linear_damping;
        r.angular_damping = b.angular_damping;
        r.gravity_scale = b.gravity_scale;
        r.sleeping = b.sleeping;
        if (w->dims == 3) r.set_shape(b.shape, b.a, b.b, b.c, b.density);
        else {
            r.shape = b.shape;
            r.a = b.a;
            r.b = b.b;
        }
        BodyDev bd{b.shape, b.pos[0], b.pos[1], b.pos[2], b.a, b.b, b.c};
        if (!r.is_static) {
            bd = r.dev();
            ++w->n_dynamic;
        }
        w->bodies.push_back(bd);
        w->rigid.push_back(r);
        w->body_meta.push_back(b);
        ++w->next_body_id;
    }
    const uint32_t n = (uint32_t)s.mass.size();
    if (n > kMaxCapacity) {
        err = "more than 2^24 particles";
        return nullptr;
    }
    w->capacity = std::max(kDefaultCapacity, n);
    if (n) w->load_particles(s.pos.data(), s.vel.data(), s.mass.data(), s.phase.data(), n);
    return w.release();
}

extern "C" {

const char *phieng_last_error(void) { return g_last_error.c_str(); }

World3 *phieng_world3_new(void) {
    try {
        return new World3();
    } catch (...) {
        return nullptr;  // allocation of the host object itself failed
    }
}

void phieng_world3_free(World3 *world) {
    if (!world) return;
    try {
        delete world;
    } catch (...) {
    }
}

void phieng_world3_step(World3 *world, float dt) {
    if (!world || !std::isfinite(dt) || dt <= 0.0f) return;
    try {
        world->step(std::min(dt, 1.0f / 15.0f));
    } catch (const std::exception &e) {
        world->eng.recover();
        if (world->step_error.empty()) world->step_error = e.what();
        set_error("phieng_world3_step", e.what());
    }
}

void phieng_world3_set_gravity(World3 *world, float gx, float gy, float gz) {
    if (!world || !std::isfinite(gx) || !std::isfinite(gy) || !std::isfinite(gz)) return;
    world->g[0] = gx;
    world->g[1] = gy;
    world->g[2] = gz;
}

int32_t phieng_world3_add_body(World3 *world, float px, float py, float pz, uint32_t shape_tag, float a, float b,
                               float c, float density, uint8_t is_static, float friction, float restitution) {
    if (!world) return -1;
    return world->add_body3(px, py, pz, shape_tag, a, b, c, density, is_static, friction, restitution,
                            "phieng_world3_add_body");
}

uint32_t phieng_world3_add_bodies(World3 *world, const BodyDesc3D *descs, uint32_t n, int32_t *out_ids) {
    if (!world || !descs || !out_ids) return 0;
    n = std::min<uint32_t>(n, kMaxPerCall);
    uint32_t added = 0;
    for (uint32_t i = 0; i < n; ++i) {
        const BodyDesc3D &d = descs[i];
        out_ids[i] = world->add_body3(d.pos_x, d.pos_y, d.pos_z, d.shape_tag, d.a, d.b, d.c, d.density, d.is_static,
                                      d.friction, d.restitution, "phieng_world3_add_bodies");
        if (out_ids[i] >= 0) ++added;
    }
    return added;
}

int32_t phieng_world3_get_body_transform(World3 *world, uint32_t body_id, float *out_pos, float *out_quat) {
    if (!world || !out_pos || !out_quat || body_id >= world->rigid.size()) return 0;
    const ph::Rigid3 &r = world->rigid[body_id];
    out_pos[0] = r.pos.x;
    out_pos[1] = r.pos.y;
    out_pos[2] = r.pos.z;
    out_quat[0] = r.rot.x;
    out_quat[1] = r.rot.y;
    out_quat[2] = r.rot.z;
    out_quat[3] = r.rot.w;
    return 1;
}

int32_t phieng_world3_get_body_velocity(World3 *world, uint32_t body_id, float *out_vel) {
    if (!world || !out_vel || body_id >= world->rigid.size()) return 0;
    const ph::Rigid3 &r = world->rigid[body_id];
    out_vel[0] = r.vel.x;
    out_vel[1] = r.vel.y;
    out_vel[2] = r.vel.z;
    return 1;
}

int32_t phieng_world3_get_body_angular_velocity(World3 *world, uint32_t body_id, float *out_w) {
    if (!world || !out_w || body_id >= world->rigid.size()) return 0;
    const ph::Rigid3 &r = world->rigid[body_id];
    out_w[0] = r.ang.x;
    out_w[1] = r.ang.y;
    out_w[2] = r.ang.z;
    return 1;
}

int32_t phieng_world3_body_count(World3 *world) { return world ? (int32_t)world->rigid.size() : -1; }

int32_t phieng_world3_set_fluid_config(World3 *world, const FluidConfig3Ffi *cfg) {
    if (!world || !cfg) return 0;
    if (!config_valid(*cfg)) return 0;
    if (cfg->adaptive_resolution && (world->sur || world->col)) {
        set_error("phieng_world3_set_fluid_config",
                  "adaptive_resolution with the learned surrogate attached: the surrogate keeps a fixed particle set");
        return 0;
    }
    if (cfg->boundary_particles && world->sur) {
        set_error("phieng_world3_set_fluid_config",
                  "boundary_particles with the learned surrogate attached: the surrogate does not model ghosts");
        return 0;
    }
    if (cfg->adaptive_resolution && world->n_dynamic) {
        set_error("phieng_world3_set_fluid_config", "adaptive_resolution with dynamic bodies: its substeps do not couple");
        return 0;
    }
    FluidConfig3Ffi c = *cfg;
    // Rust bools read back as 0/1.
    c.adaptive_resolution = cfg->adaptive_resolution ? 1 : 0;
    c.boundary_particles = cfg->boundary_particles ? 1 : 0;
    c.enabled = cfg->enabled ? 1 : 0;
    world->cfg = c;
    return 1;
}

int32_t phieng_world3_get_fluid_config(const World3 *world, FluidConfig3Ffi *out) {
    if (!world || !out) return 0;
    *out = world->cfg;
    return 1;
}

uint32_t phieng_world3_fluid_add(World3 *world, const float *positions, uint32_t count) {
    if (!world || !positions) return 0;
    try {
        return world->add(positions, nullptr, count);
    } catch (const std::exception &e) {
        set_error("phieng_world3_fluid_add", e.what());
        return 0;
    }
}

uint32_t phieng_world3_fluid_add_phase(World3 *world, const float *positions, const uint8_t *phases,
                                       uint32_t count) {
    if (!world || !positions || !phases) return 0;
    try {
        return world->add(positions, phases, count);
    } catch (const std::exception &e) {
        set_error("phieng_world3_fluid_add_phase", e.what());
        return 0;
    }
}

uint32_t phieng_world3_fluid_count(const World3 *world) { return world ? world->n : 0; }

int32_t phieng_world3_fluid_clear(World3 *world) {
    if (!world) return 0;
    world->n = 0;
    world->eng.n_total = 0;
    world->eng.tiles.clear();
    world->eng.tiles_dirty = true;
    world->eng.rec_valid = false;
    world->domain_dirty = true;
    return 1;
}

uint32_t phieng_world3_fluid_get(const World3 *world, float *out, uint32_t max) {
    if (!world || !out) return 0;
    try {
        return const_cast<World3 *>(world)->get(out, max);
    } catch (const std::exception &e) {
        set_error("phieng_world3_fluid_get", e.what());
        return 0;
    }
}

uint32_t phieng_world3_fluid_get_mass(const World3 *world, float *mass, uint8_t *phase, uint32_t max) {
    if (!world || (!mass && !phase)) return 0;
    try {
        return const_cast<World3 *>(world)->get_mass(mass, phase, max);
    } catch (const std::exception &e) {
        set_error("phieng_world3_fluid_get_mass", e.what());
        return 0;
    }
}

uint32_t phieng_world3_fluid_reserve(World3 *world, uint32_t capacity) {
    if (!world) return 0;
    world->capacity = std::max(world->n, std::min(std::max(capacity, kDefaultCapacity), kMaxCapacity));
    return world->capacity;
}

uint32_t phieng_world3_fluid_set_velocity(World3 *world, const float *vel, uint32_t count) {
    if (!world || !vel) return 0;
    const uint32_t m = std::min(count, world->n);
    for (size_t i = 0; i < 3 * (size_t)m; ++i)
        if (!std::isfinite(vel[i])) return 0;
    if (m == 0) return 0;
    try {
        ph::Engine &e = world->eng;
        e.init_device();
        // Particles beyond `count` keep their velocity: start from the current one.
        std::vector<float> full(6 * (size_t)world->n);
        world->get(full.data(), world->n);
        std::vector<float> v3(3 * (size_t)world->n);
        for (uint32_t i = 0; i < world->n; ++i)
            for (int c = 0; c < 3; ++c) v3[3 * i + c] = i < m ? vel[3 * i + c] : full[6 * i + 3 + c];
        e.staging_f.ensure(v3.size());
        PH_CK(hipMemcpyAsync(e.staging_f.p, v3.data(), sizeof(float) * v3.size(), hipMemcpyHostToDevice, e.stream));
        e.run(ph::KIND_SETVEL, [&] {
            return ph::launch_set_velocity(e.stream, world->n, e.tag[e.ct].p, e.staging_f.p, e.vel[e.cs].p);
        });
        ++e.layout_epoch;  // surrogate records hold the old velocities: next substep plain
        PH_CK(hipStreamSynchronize(e.stream));
        return m;
    } catch (const std::exception &ex) {
        set_error("phieng_world3_fluid_set_velocity", ex.what());
        return 0;
    }
}

int32_t phieng_world3_fluid_set_record(World3 *world, uint32_t flags) {
    if (!world) return 0;
    try {
        world->eng.rec_flags = flags & (PHIENG_RECORD_CELLS | PHIENG_RECORD_DENSITY | PHIENG_RECORD_ACCEL);
        world->eng.rec_valid = false;
        if (world->eng.rec_flags && world->eng.alloc_particles) world->eng.ensure_record();
        return 1;
    } catch (const std::exception &e) {
        set_error("phieng_world3_fluid_set_record", e.what());
        return 0;
    }
}

uint32_t phieng_world3_fluid_get_record(const World3 *world, int32_t *cell3, uint32_t *sorted_ids, float *density,
                                        float *pressure, float *accel3, uint32_t *nbr_count, uint32_t max) {
    if (!world || !world->eng.rec_valid) return 0;
    try {
        const ph::Engine &e = world->eng;
        const uint32_t m = std::min(max, world->n);
        const uint32_t f = e.rec_flags;
        hipStream_t st = e.stream;
        if (cell3 && (f & PHIENG_RECORD_CELLS))
            PH_CK(hipMemcpyAsync(cell3, e.rec_cells.p, sizeof(int32_t) * 3 * m, hipMemcpyDeviceToHost, st));
        if (sorted_ids && (f & PHIENG_RECORD_CELLS))
            PH_CK(hipMemcpyAsync(sorted_ids, e.rec_sorted.p, sizeof(uint32_t) * m, hipMemcpyDeviceToHost, st));
        if (density && (f & PHIENG_RECORD_DENSITY))
            PH_CK(hipMemcpyAsync(density, e.rec_rho.p, sizeof(float) * m, hipMemcpyDeviceToHost, st));
        if (pressure && (f & PHIENG_RECORD_DENSITY))
            PH_CK(hipMemcpyAsync(pressure, e.rec_p.p, sizeof(float) * m, hipMemcpyDeviceToHost, st));
        if (nbr_count && (f & PHIENG_RECORD_DENSITY))
            PH_CK(hipMemcpyAsync(nbr_count, e.rec_count.p, sizeof(uint32_t) * m, hipMemcpyDeviceToHost, st));
        if (accel3 && (f & PHIENG_RECORD_ACCEL))
            PH_CK(hipMemcpyAsync(accel3, e.rec_acc.p, sizeof(float) * 3 * m, hipMemcpyDeviceToHost, st));
        PH_CK(hipStreamSynchronize(st));
        return m;
    } catch (const std::exception &ex) {
        set_error("phieng_world3_fluid_get_record", ex.what());
        return 0;
    }
}

int32_t phieng_world3_set_timing(World3 *world, int32_t enable) {
    if (!world) return 0;
    try {
        if (world->eng.device >= 0) world->eng.reset_timing();
        world->eng.timing = enable != 0;
        return 1;
    } catch (const std::exception &e) {
        set_error("phieng_world3_set_timing", e.what());
        return 0;
    }
}

int32_t phieng_world3_get_timing(const World3 *world, double *total_ms, uint64_t *launches, int32_t max) {
    if (!world) return 0;
    try {
        ph::Engine &e = const_cast<World3 *>(world)->eng;
        if (e.device >= 0) e.harvest_timing();
        const int32_t m = std::min<int32_t>(max, ph::KIND_COUNT);
        for (int32_t k = 0; k < m; ++k) {
            if (total_ms) total_ms[k] = e.total_ms[k];
            if (launches) launches[k] = e.launches[k];
        }
        return ph::KIND_COUNT;
    } catch (const std::exception &ex) {
        set_error("phieng_world3_get_timing", ex.what());
        return 0;
    }
}

const char *phieng_kernel_name(int32_t kind) { return ph::kind_name(kind); }

int32_t phieng_world3_set_surrogate(World3 *world, Surrogate3 *s, uint32_t check_every, float max_density_error,
                                    int32_t use_network) {
    if (!world) return 0;
    if (s && (check_every < 1 || !std::isfinite(max_density_error) || max_density_error <= 0.0f)) return 0;
    if (s && world->cfg.boundary_particles) {
        set_error("phieng_world3_set_surrogate", "the surrogate does not model boundary ghosts (boundary_particles)");
        return 0;
    }
    if (s && world->cfg.adaptive_resolution) {
        set_error("phieng_world3_set_surrogate", "the surrogate keeps a fixed particle set (adaptive_resolution)");
        return 0;
    }
    if (s && world->n_dynamic) {
        set_error("phieng_world3_set_surrogate", "the surrogate's substeps do not couple to dynamic bodies");
        return 0;
    }
    world->sur = s;
    world->check_every = check_every ? check_every : 1;
    world->max_err = max_density_error;
    world->use_net = use_network ? 1 : 0;
    world->fallback_left = 0;
    world->step_index = 0;
    world->judge_pending = false;
    world->list_judge = -1;
    world->list_checked = false;
    world->list_vmax = -1.0f;
    return 1;
}

int32_t phieng_world3_collect_samples(World3 *world, Surrogate3 *s, uint32_t every, uint32_t stride) {
    if (!world || (s && (every < 1 || stride < 1))) return 0;
    if (s && world->cfg.adaptive_resolution) {
        set_error("phieng_world3_collect_samples", "sample collection keeps a fixed particle set (adaptive_resolution)");
        return 0;
    }
    if (s && world->n_dynamic) {
        set_error("phieng_world3_collect_samples", "sample collection does not run coupled substeps (dynamic bodies)");
        return 0;
    }
    world->col = s;
    world->col_every = every ? every : 1;
    world->col_stride = stride ? stride : 1;
    world->col_phase = 0;
    world->step_index = 0;
    world->list_judge = -1;
    world->list_checked = false;
    return 1;
}

int32_t phieng_world3_set_surrogate_skin(World3 *world, float skin) {
    if (!world || !(skin >= 0.0f && skin <= 1.0f)) return 0;
    world->skin = skin;
    world->list_sorts = ~0ull;  // rebuild at the next plain substep
    return 1;
}

int32_t phieng_world3_surrogate_stats(const World3 *world, uint64_t *sur_steps, uint64_t *plain_steps,
                                      uint64_t *fallbacks, float *last_density_error) {
    if (!world) return 0;
    // Report a pending verdict without applying it: applying moves the
    // fallback window, which only the step loop may do.
    float err = world->last_err;
    uint64_t fb = world->n_fallback;
    if (world->judge_pending) {
        if (hipEventSynchronize(world->judge_ev) != hipSuccess) {
            set_error("phieng_world3_surrogate_stats", "waiting for the judge's readback failed");
            return 0;
        }
        err = world->judge_host[0];
        if (world->judge_fails(err, world->judge_host[2])) ++fb;
    }
    if (sur_steps) *sur_steps = world->n_sur;
    if (plain_steps) *plain_steps = world->n_plain;
    if (fallbacks) *fallbacks = fb;
    if (last_density_error) *last_density_error = err;
    return 1;
}

int32_t phieng_world3_set_surrogate_accel_limit(World3 *world, float max_accel_error) {
    if (!world || !(max_accel_error > 0.0f)) return 0;
    world->max_accel_err = max_accel_error;
    return 1;
}

int32_t phieng_world3_surrogate_judge(const World3 *world, float *density_error, float *accel_error,
                                      uint64_t *kept_judges) {
    if (!world) return 0;
    float de = world->last_err, ae = world->last_accel_err;
    if (world->judge_pending) {
        if (hipEventSynchronize(world->judge_ev) != hipSuccess) {
            set_error("phieng_world3_surrogate_judge", "waiting for the judge's readback failed");
            return 0;
        }
        de = world->judge_host[0];
        ae = world->judge_host[2];
    }
    if (density_error) *density_error = de;
    if (accel_error) *accel_error = ae;
    if (kept_judges) *kept_judges = world->n_kept_judges;
    return 1;
}

int32_t phieng_world3_set_neighbour_list(World3 *world, int32_t verlet, float skin) {
    if (!world || !(skin >= 0.0f && skin <= 0.5f) || verlet < 0 || verlet > 2) return 0;
    world->eng.vl_enabled = verlet != 0;
    world->eng.vl_min_particles = verlet == 2 ? ph::Engine::kVlAutoMin : 0u;
    world->eng.vl_skin = skin;
    world->eng.vl_valid = false;  // the next substep rebuilds with the new radius
    return 1;
}

int32_t phieng_world3_neighbour_stats(World3 *world, uint64_t *substeps, uint64_t *rebuilds, uint64_t *repaired,
                                      uint64_t *list_substeps) {
    if (!world) return 0;
    try {
        VerletDev v{};
        ph::Engine &e = world->eng;
        if (e.vl_state.p && e.vl_state_ready) {
            PH_CK(hipMemcpyAsync(&v, e.vl_state.p, sizeof(v), hipMemcpyDeviceToHost, e.stream));
            PH_CK(hipStreamSynchronize(e.stream));
        }
        if (substeps) *substeps = v.substeps;
        if (rebuilds) *rebuilds = v.rebuilds;
        if (repaired) *repaired = v.repaired;
        if (list_substeps) *list_substeps = e.vl_list_substeps;
        return 1;
    } catch (const std::exception &ex) {
        set_error("phieng_world3_neighbour_stats", ex.what());
        return 0;
    }
}

int32_t phieng_world3_boundary_info(const World3 *world, uint32_t *count, uint64_t *regens) {
    if (!world) return 0;
    if (count) *count = (uint32_t)world->ghost_gen.size();
    if (regens) *regens = world->ghost_regens;
    return 1;
}

uint32_t phieng_world3_boundary_get(const World3 *world, float *xyz, float *density, float *pressure,
                                    uint32_t max) {
    if (!world) return 0;
    try {
        const uint32_t g = (uint32_t)world->ghost_gen.size();
        const uint32_t m = std::min(max, g);
        if (m == 0) return 0;
        const ph::Engine &e = world->eng;
        std::vector<float4> st(e.g_count);
        if ((density || pressure) && e.g_count) {
            PH_CK(hipMemcpyAsync(st.data(), e.g_state.p, sizeof(float4) * e.g_count, hipMemcpyDeviceToHost,
                                 e.stream));
            PH_CK(hipStreamSynchronize(e.stream));
        }
        // cell-sorted slot -> generation index; a ghost in no cell (a
        // non-finite one) has no fluid neighbours: density 0, pressure 0
        std::vector<float2> rp(g, make_float2(0.0f, 0.0f));
        for (uint32_t k = 0; k < e.g_count; ++k) rp[e.g_order[k]] = make_float2(st[k].x, st[k].y);
        for (uint32_t k = 0; k < m; ++k) {
            if (xyz) {
                xyz[3 * k] = world->ghost_gen[k].x;
                xyz[3 * k + 1] = world->ghost_gen[k].y;
                xyz[3 * k + 2] = world->ghost_gen[k].z;
            }
            if (density) density[k] = rp[k].x;
            if (pressure) pressure[k] = rp[k].y;
        }
        return m;
    } catch (const std::exception &ex) {
        set_error("phieng_world3_boundary_get", ex.what());
        return 0;
    }
}

int32_t phieng_world3_sync(World3 *world) {
    if (!world) return 0;
    if (!world->step_error.empty()) {
        const std::string m = "an earlier step failed: " + world->step_error;
        world->step_error.clear();
        set_error("phieng_world3_step", m.c_str());
        return 0;
    }
    if (world->eng.device < 0) return 1;
    try {
        PH_CK(hipSetDevice(world->eng.device));
        PH_CK(hipStreamSynchronize(world->eng.stream));
        return 1;
    } catch (const std::exception &e) {
        set_error("phieng_world3_sync", e.what());
        return 0;
    }
}

// ---------------------------------------------------------------------
// World2: the reference's 2D fluid exports (ffi.rs:51-94, 999-1026,
// 1159-1203, 2486-2594) over the same engine, with the particles in the
// z = 0 plane.  With z, v_z and g_z exactly zero every z term of the 3D
// kernels adds an exact +0, the z row offsets hold no particles, and the
// remaining arithmetic is the 2D step's (fluid.rs:541-1136) operation for
// operation once the scene carries the 2D kernel constants and the 2D box.
// ---------------------------------------------------------------------

}  // extern "C"


namespace {
World3 *as3(World2 *w) { return static_cast<World3 *>(w); }
const World3 *as3(const World2 *w) { return static_cast<const World3 *>(w); }
}  // namespace

extern "C" {

World2 *phieng_world_new(void) {
    try {
        return new World2();
    } catch (...) {
        return nullptr;
    }
}

void phieng_world_free(World2 *world) {
    if (!world) return;
    try {
        delete world;
    } catch (...) {
    }
}

void phieng_world_step(World2 *world, float dt) {
    if (!world || !std::isfinite(dt) || dt <= 0.0f) return;
    try {
        world->step(std::min(dt, 1.0f / 15.0f));
    } catch (const std::exception &e) {
        world->eng.recover();
        if (world->step_error.empty()) world->step_error = e.what();
        set_error("phieng_world_step", e.what());
    }
}

void phieng_world_set_gravity(World2 *world, float gx, float gy) {
    if (!world || !std::isfinite(gx) || !std::isfinite(gy)) return;
    world->g[0] = gx;
    world->g[1] = gy;
    world->g[2] = 0.0f;
}

int32_t phieng_world_add_body(World2 *world, float px, float py, uint32_t shape_tag, float a, float b, float density,
                              uint8_t is_static, float friction, float restitution) {
    if (!world) return -1;
    return world2_base(world)->add_body2(px, py, shape_tag, a, b, density, is_static, friction, restitution,
                                         "phieng_world_add_body");
}

uint32_t phieng_world_add_bodies(World2 *world, const BodyDesc2D *descs, uint32_t n, int32_t *out_ids) {
    if (!world || !descs || !out_ids) return 0;
    n = std::min<uint32_t>(n, kMaxPerCall);
    uint32_t added = 0;
    for (uint32_t i = 0; i < n; ++i) {
        const BodyDesc2D &d = descs[i];
        out_ids[i] = world2_base(world)->add_body2(d.pos_x, d.pos_y, d.shape_tag, d.a, d.b, d.density, d.is_static,
                                                   d.friction, d.restitution, "phieng_world_add_bodies");
        if (out_ids[i] >= 0) ++added;
    }
    return added;
}

int32_t phieng_world_set_fluid_config(World2 *world, const FluidConfig2Ffi *cfg) {
    // fluid2_from_ffi (ffi.rs:773-840) applies the same rules as the 3D one
    return world ? phieng_world3_set_fluid_config(as3(world), cfg) : 0;
}

int32_t phieng_world_get_fluid_config(const World2 *world, FluidConfig2Ffi *out) {
    return world ? phieng_world3_get_fluid_config(as3(world), out) : 0;
}

}  // extern "C"

namespace {
// xy pairs -> xyz triples in the z = 0 plane (after the finite check of the
// caller's values, so a rejected call stays atomic)
std::vector<float> lift_xy(const float *xy, uint32_t m) {
    std::vector<float> xyz(3 * (size_t)m);
    for (uint32_t i = 0; i < m; ++i) {
        xyz[3 * i] = xy[2 * i];
        xyz[3 * i + 1] = xy[2 * i + 1];
        xyz[3 * i + 2] = 0.0f;
    }
    return xyz;
}
bool all_finite(const float *v, size_t k) {
    for (size_t i = 0; i < k; ++i)
        if (!std::isfinite(v[i])) return false;
    return true;
}
}  // namespace

extern "C" {

uint32_t phieng_world_fluid_add(World2 *world, const float *positions, uint32_t count) {
    if (!world || !positions) return 0;
    try {
        const uint32_t m = std::min(count, kMaxPerCall);
        if (!all_finite(positions, 2 * (size_t)m)) return 0;
        const std::vector<float> xyz = lift_xy(positions, m);
        return world->add(xyz.data(), nullptr, m);
    } catch (const std::exception &e) {
        set_error("phieng_world_fluid_add", e.what());
        return 0;
    }
}

uint32_t phieng_world_fluid_add_phase(World2 *world, const float *positions, const uint8_t *phases,
                                      uint32_t count) {
    if (!world || !positions || !phases) return 0;
    try {
        const uint32_t m = std::min(count, kMaxPerCall);
        if (!all_finite(positions, 2 * (size_t)m)) return 0;
        const std::vector<float> xyz = lift_xy(positions, m);
        return world->add(xyz.data(), phases, m);
    } catch (const std::exception &e) {
        set_error("phieng_world_fluid_add_phase", e.what());
        return 0;
    }
}

uint32_t phieng_world_fluid_count(const World2 *world) { return world ? world->n : 0; }

int32_t phieng_world_fluid_clear(World2 *world) { return world ? phieng_world3_fluid_clear(as3(world)) : 0; }
uint32_t phieng_world_fluid_get_mass(const World2 *world, float *mass, uint8_t *phase, uint32_t max) {
    return world ? phieng_world3_fluid_get_mass(as3(world), mass, phase, max) : 0;
}

uint32_t phieng_world_fluid_get(const World2 *world, float *out, uint32_t max) {
    if (!world || !out) return 0;
    try {
        const uint32_t m = std::min(max, world->n);
        if (m == 0) return 0;
        std::vector<float> six(6 * (size_t)m);
        const uint32_t got = const_cast<World2 *>(world)->get(six.data(), m);
        for (uint32_t i = 0; i < got; ++i) {
            out[4 * i] = six[6 * i];
            out[4 * i + 1] = six[6 * i + 1];
            out[4 * i + 2] = six[6 * i + 3];
            out[4 * i + 3] = six[6 * i + 4];
        }
        return got;
    } catch (const std::exception &e) {
        set_error("phieng_world_fluid_get", e.what());
        return 0;
    }
}

uint32_t phieng_world_fluid_reserve(World2 *world, uint32_t capacity) {
    return world ? phieng_world3_fluid_reserve(as3(world), capacity) : 0;
}

uint32_t phieng_world_fluid_set_velocity(World2 *world, const float *vel, uint32_t count) {
    if (!world || !vel) return 0;
    const uint32_t m = std::min(count, world->n);
    if (!all_finite(vel, 2 * (size_t)m)) return 0;
    std::vector<float> v3(3 * (size_t)m);
    for (uint32_t i = 0; i < m; ++i) {
        v3[3 * i] = vel[2 * i];
        v3[3 * i + 1] = vel[2 * i + 1];
        v3[3 * i + 2] = 0.0f;
    }
    return phieng_world3_fluid_set_velocity(as3(world), v3.data(), m);
}

int32_t phieng_world_fluid_set_record(World2 *world, uint32_t flags) {
    return world ? phieng_world3_fluid_set_record(as3(world), flags) : 0;
}

uint32_t phieng_world_fluid_get_record(const World2 *world, int32_t *cell2, uint32_t *sorted_ids, float *density,
                                       float *pressure, float *accel2, uint32_t *nbr_count, uint32_t max) {
    if (!world) return 0;
    try {
        const uint32_t m = std::min(max, world->n);
        std::vector<int32_t> c3(cell2 ? 3 * (size_t)m : 0);
        std::vector<float> a3(accel2 ? 3 * (size_t)m : 0);
        const uint32_t got = phieng_world3_fluid_get_record(as3(world), cell2 ? c3.data() : nullptr, sorted_ids,
                                                            density, pressure, accel2 ? a3.data() : nullptr,
                                                            nbr_count, m);
        for (uint32_t i = 0; i < got; ++i) {
            if (cell2) {
                cell2[2 * i] = c3[3 * i];
                cell2[2 * i + 1] = c3[3 * i + 1];
            }
            if (accel2) {
                accel2[2 * i] = a3[3 * i];
                accel2[2 * i + 1] = a3[3 * i + 1];
            }
        }
        return got;
    } catch (const std::exception &e) {
        set_error("phieng_world_fluid_get_record", e.what());
        return 0;
    }
}

int32_t phieng_world_sync(World2 *world) { return world ? phieng_world3_sync(as3(world)) : 0; }

int32_t phieng_world_set_neighbour_list(World2 *world, int32_t verlet, float skin) {
    return world ? phieng_world3_set_neighbour_list(world2_base(world), verlet, skin) : 0;
}

int32_t phieng_world_neighbour_stats(World2 *world, uint64_t *substeps, uint64_t *rebuilds, uint64_t *repaired,
                                     uint64_t *list_substeps) {
    return world ? phieng_world3_neighbour_stats(world2_base(world), substeps, rebuilds, repaired, list_substeps) : 0;
}

int32_t phieng_world_boundary_info(const World2 *world, uint32_t *count, uint64_t *regens) {
    return world ? phieng_world3_boundary_info(as3(world), count, regens) : 0;
}

uint32_t phieng_world_boundary_get(const World2 *world, float *xy, float *density, float *pressure, uint32_t max) {
    if (!world) return 0;
    try {
        const uint32_t m = std::min<uint32_t>(max, (uint32_t)world->ghost_gen.size());
        std::vector<float> xyz(xy ? 3 * (size_t)m : 0);
        const uint32_t got = phieng_world3_boundary_get(as3(world), xy ? xyz.data() : nullptr, density, pressure, m);
        for (uint32_t k = 0; xy && k < got; ++k) {
            xy[2 * k] = xyz[3 * k];
            xy[2 * k + 1] = xyz[3 * k + 1];
        }
        return got;
    } catch (const std::exception &e) {
        set_error("phieng_world_boundary_get", e.what());
        return 0;
    }
}

}  // extern "C"
