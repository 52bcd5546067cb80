// engine.hip — step orchestration of the device-resident SPH state.
#include "engine.h"

#include <algorithm>
#include <cmath>

#include "sph_kernels.h"

namespace ph {

namespace {
constexpr float kEps = 1e-6f;
constexpr float kPiF = 3.14159265358979323846f;  // std::f32::consts::PI

float rmax_h(float a, float b) {
    if (a != a) return b;
    if (b != b) return a;
    return a > b ? a : b;
}
float rclamp_h(float v, float lo, float hi) {
    if (v < lo) v = lo;
    if (v > hi) v = hi;
    return v;
}
int32_t sat_i32_h(float v) {
    if (v != v) return 0;
    if (v >= 2147483648.0f) return INT32_MAX;
    if (v < -2147483648.0f) return INT32_MIN;
    return (int32_t)v;
}
}  // namespace

const char *kind_name(int k) {
    static const char *names[KIND_COUNT] = {
        "cell_key", "scan_reduce", "scan_partials", "scan_apply", "scatter", "rank_gather",
        "density",  "force_integrate", "vmax", "export", "monitor", "append", "cell_bbox",
        "fill", "set_velocity", "batch_persistent", "surrogate_step", "surrogate_history",
        "surrogate_collect"};
    return (k >= 0 && k < KIND_COUNT) ? names[k] : "";
}

void fill_scene_physics(SceneDev &s, const SceneHost &h) {
    const FluidConfig3Ffi &c = h.cfg;
    s.h_grid = c.smoothing_length;
    s.h_grid2 = s.h_grid * s.h_grid;
    s.h = rmax_h(c.smoothing_length, kEps);
    const float hk = s.h;
    s.h2 = hk * hk;
    const float h2 = s.h2;
    if (h.dims == 2) {
        s.poly6_c = 4.0f / (kPiF * h2 * h2 * h2 * h2);              // fluid.rs:272
        s.spiky_c = -30.0f / (kPiF * hk * hk * hk * hk * hk);       // fluid.rs:283
    } else {
        s.poly6_c = 315.0f / (64.0f * kPiF * h2 * h2 * h2 * h2 * hk);  // fluid3.rs:227
        s.spiky_c = -45.0f / (kPiF * hk * hk * hk * hk * hk * hk);     // fluid3.rs:237
    }
    s.rho_max = 3.0f * rmax_h(c.rest_density, kEps);
    s.stiffness = c.stiffness;
    s.gamma = c.gamma;
    for (int k = 0; k < 4; ++k) s.phase_rho0[k] = c.phase_rest_density[k];
    s.av_alpha = c.artificial_viscosity;
    s.c_s = std::sqrt(rmax_h(c.stiffness, 0.0f));
    s.kappa = c.surface_tension;
    s.viscosity = c.viscosity;
    s.max_velocity = c.max_velocity;
    s.particle_radius = c.particle_radius;
    s.wall_friction = c.wall_friction;
    s.wall_mu = rclamp_h(c.wall_friction, 0.0f, 1.0f);
    s.gx = h.g[0];
    s.gy = h.g[1];
    s.gz = h.g[2];
    s.dt = h.dt;
}

void choose_domain(SceneDev &s, const int32_t *pb, uint32_t np, const std::vector<BodyDev> &bodies) {
    const int64_t limit = std::min<int64_t>(std::max<int64_t>(8 * (int64_t)np, 1 << 20), (int64_t)1 << 27);
    int64_t plo[3], phi[3];
    bool have_p = pb && pb[0] <= pb[3];
    for (int a = 0; a < 3; ++a) {
        plo[a] = have_p ? pb[a] : 0;
        phi[a] = have_p ? pb[3 + a] : 0;
    }
    auto vol = [](const int64_t *lo, const int64_t *hi) {
        int64_t v = 1;
        for (int a = 0; a < 3; ++a) {
            int64_t e = hi[a] - lo[a] + 1;
            if (e <= 0) return (int64_t)0;
            if (v > ((int64_t)1 << 40) / e) return ((int64_t)1 << 40);
            v *= e;
        }
        return v;
    };
    const int64_t margin = 2;
    int64_t dlo[3], dhi[3];
    for (int a = 0; a < 3; ++a) {
        dlo[a] = plo[a] - margin;
        dhi[a] = phi[a] + margin;
    }
    // Static bodies usually bound the fluid: cover them too if affordable.
    int64_t ulo[3] = {dlo[0], dlo[1], dlo[2]}, uhi[3] = {dhi[0], dhi[1], dhi[2]};
    const float h = s.h_grid;
    for (const BodyDev &b : bodies) {
        float ex = b.a, ey = b.b, ez = b.c;
        if (b.shape == 0u) ey = ez = b.a;
        if (b.shape == 3u) ez = 0.0f;
        if (b.shape == 2u) {
            ey = b.b + b.a;
            ez = b.a;
        }
        const float lo[3] = {b.px - ex, b.py - ey, b.pz - ez}, hi[3] = {b.px + ex, b.py + ey, b.pz + ez};
        for (int a = 0; a < 3; ++a) {
            const int64_t cl = sat_i32_h(std::floor(lo[a] / h)) - 1, ch = (int64_t)sat_i32_h(std::floor(hi[a] / h)) + 1;
            ulo[a] = std::min(ulo[a], cl);
            uhi[a] = std::max(uhi[a], ch);
        }
    }
    const int64_t *lo = ulo, *hi = uhi;
    if (vol(ulo, uhi) > limit) {
        lo = dlo;
        hi = dhi;
    }
    int64_t flo[3], fhi[3];
    for (int a = 0; a < 3; ++a) {
        flo[a] = std::max<int64_t>(lo[a], INT32_MIN / 2);
        fhi[a] = std::min<int64_t>(hi[a], INT32_MAX / 2);
    }
    // Still too large (a very sparse cloud): shrink around the box centre;
    // what falls outside becomes outliers.
    while (vol(flo, fhi) > limit) {
        int a = 0;
        for (int k = 1; k < 3; ++k)
            if (fhi[k] - flo[k] > fhi[a] - flo[a]) a = k;
        const int64_t c = (flo[a] + fhi[a]) / 2, e = (fhi[a] - flo[a]) / 4;
        flo[a] = c - e;
        fhi[a] = c + e;
    }
    s.x0 = (int32_t)flo[0];
    s.y0 = (int32_t)flo[1];
    s.z0 = (int32_t)flo[2];
    s.nx = (int32_t)(fhi[0] - flo[0] + 1);
    s.ny = (int32_t)(fhi[1] - flo[1] + 1);
    s.nz = (int32_t)(fhi[2] - flo[2] + 1);
    s.ncells = (uint32_t)((int64_t)s.nx * s.ny * s.nz);
}

Engine::~Engine() {
    if (device >= 0) {
        (void)hipSetDevice(device);
        if (stream) (void)hipStreamSynchronize(stream);
        for (auto &e : pending_ev) {
            (void)hipEventDestroy(e.a);
            (void)hipEventDestroy(e.b);
        }
        for (auto e : ev_pool) (void)hipEventDestroy(e);
        if (mon_event) (void)hipEventDestroy(mon_event);
        if (h_monitor) (void)hipHostFree(h_monitor);
        if (stream) (void)hipStreamDestroy(stream);
    }
}

void Engine::init_device() {
    if (device >= 0) {
        PH_CK(hipSetDevice(device));
        return;
    }
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count <= 0) throw std::runtime_error("no HIP device available");
    PH_CK(hipGetDevice(&device));
    PH_CK(hipSetDevice(device));
    PH_CK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    PH_CK(hipEventCreateWithFlags(&mon_event, hipEventDisableTiming));
    PH_CK(hipHostMalloc((void **)&h_monitor, 8 * sizeof(int32_t) * 4096, hipHostMallocDefault));
    d_vmax.ensure(1);
}

void Engine::ensure_particles(uint32_t cap) {
    if (cap <= alloc_particles && posm[0].p) return;
    ++layout_epoch;
    size_t want = std::max<size_t>(cap, std::min<size_t>(2 * alloc_particles, (size_t)1 << 24));
    want = std::max<size_t>(want, 1024);
    DBuf<float4> np0, nv0;
    DBuf<uint32_t> nt0;
    np0.ensure(want);
    nv0.ensure(want);
    nt0.ensure(want);
    if (n_total) {
        PH_CK(hipMemcpyAsync(np0.p, posm[cs].p, sizeof(float4) * n_total, hipMemcpyDeviceToDevice, stream));
        PH_CK(hipMemcpyAsync(nv0.p, vel[cs].p, sizeof(float4) * n_total, hipMemcpyDeviceToDevice, stream));
        PH_CK(hipMemcpyAsync(nt0.p, tag[ct].p, sizeof(uint32_t) * n_total, hipMemcpyDeviceToDevice, stream));
    }
    PH_CK(hipStreamSynchronize(stream));
    for (int k = 0; k < 2; ++k) {
        posm[k].release();
        vel[k].release();
        tag[k].release();
    }
    std::swap(posm[0].p, np0.p);
    std::swap(posm[0].n, np0.n);
    std::swap(vel[0].p, nv0.p);
    std::swap(vel[0].n, nv0.n);
    std::swap(tag[0].p, nt0.p);
    std::swap(tag[0].n, nt0.n);
    posm[1].ensure(want);
    vel[1].ensure(want);
    tag[1].ensure(want);
    cs = 0;
    ct = 0;
    keys.ensure(want);
    slots.ensure(want);
    tmp.ensure(want);
    rq.ensure(want);
    masks.ensure(9 * want);
    flags.ensure(want);
    alloc_particles = want;
    if (rec_flags) ensure_record();
}

void Engine::ensure_record() {
    const size_t m = std::max<size_t>(alloc_particles, 1);
    rec_cells.ensure(3 * m);
    rec_sorted.ensure(m);
    rec_count.ensure(m);
    rec_rho.ensure(m);
    rec_p.ensure(m);
    rec_acc.ensure(3 * m);
}

void Engine::build_tiles() {
    ++layout_epoch;
    tiles.clear();
    for (uint32_t k = 0; k < scenes.size(); ++k) {
        scenes[k].tile_begin = (uint32_t)tiles.size();
        for (uint32_t f = scenes[k].part_begin; f < scenes[k].part_end; f += PH_TILE) tiles.push_back({k, f});
        scenes[k].tile_count = (uint32_t)tiles.size() - scenes[k].tile_begin;
    }
    tiles_dirty = true;
    scenes_dirty = true;
}

void Engine::assign_buckets() {
    ++layout_epoch;
    uint32_t base = 0;
    for (auto &s : scenes) {
        s.bucket_base = base;
        base += s.ncells + 2;
    }
    nbuckets = base;
    if (nbuckets + 1 > alloc_buckets || !counts.p) {
        size_t want = (size_t)nbuckets + 1 + (nbuckets >> 3);
        counts.ensure(want);
        start.ensure(want + 1);
        partial.ensure(scan_partials_needed((uint32_t)want) + 1);
        alloc_buckets = want;
        run(KIND_FILL, [&] { return launch_fill_u32(stream, counts.p, (uint32_t)want, 0u); });
    }
    scenes_dirty = true;
}

void Engine::upload_tables() {
    const bool any = scenes_dirty || tiles_dirty || bodies_dirty;
    if (scenes_dirty) {
        d_scenes.ensure(scenes.size());
        PH_CK(hipMemcpyAsync(d_scenes.p, scenes.data(), sizeof(SceneDev) * scenes.size(), hipMemcpyHostToDevice,
                             stream));
        d_monitor.ensure(8 * scenes.size());
        scenes_dirty = false;
    }
    if (tiles_dirty) {
        d_tiles.ensure(tiles.size());
        d_bbox.ensure(6 * std::max<size_t>(tiles.size(), 1));
        if (!tiles.empty())
            PH_CK(hipMemcpyAsync(d_tiles.p, tiles.data(), sizeof(TileDev) * tiles.size(), hipMemcpyHostToDevice,
                                 stream));
        tiles_dirty = false;
    }
    if (bodies_dirty) {
        d_bodies.ensure(std::max<size_t>(bodies.size(), 1));
        if (!bodies.empty())
            PH_CK(hipMemcpyAsync(d_bodies.p, bodies.data(), sizeof(BodyDev) * bodies.size(), hipMemcpyHostToDevice,
                                 stream));
        bodies_dirty = false;
    }
    // Host vectors may be rewritten by the caller right after this call:
    // make the staged copies complete first.
    if (any) PH_CK(hipStreamSynchronize(stream));
}

void Engine::substep() {
    if (n_total == 0 || tiles.empty()) return;
    sort_phase(nullptr);
    plain_phase(nullptr);
    finish_substep();
}

// K1 .. K3: keys, bucket scan, scatter, rank + gather into generation 1-cs.
void Engine::sort_phase(uint32_t *perm_out) {
    const uint32_t nt = (uint32_t)tiles.size();
    const uint32_t n = n_total;
    const bool rc = rec_flags & PHIENG_RECORD_CELLS;
    const int ns = 1 - cs, nt_ = 1 - ct;
    // cell_key counts into zeroed buckets and scan_apply zeroes them again;
    // a sort cut short in between leaves them to be cleared here.
    if (counts_dirty)
        run(KIND_FILL, [&] { return launch_fill_u32(stream, counts.p, (uint32_t)alloc_buckets, 0u); });
    counts_dirty = true;
    ++sorts_done;
    run(KIND_CELL_KEY, [&] {
        return launch_cell_key(stream, nt, d_tiles.p, d_scenes.p, posm[cs].p, tag[ct].p, keys.p, slots.p, counts.p,
                               rc ? rec_cells.p : nullptr);
    });
    run(KIND_SCAN_REDUCE, [&] { return launch_scan_reduce(stream, counts.p, nbuckets, partial.p); });
    run(KIND_SCAN_PARTIALS, [&] {
        return launch_scan_partials(stream, partial.p, nbuckets, start.p, d_bbox.p, (uint32_t)scenes.size());
    });
    run(KIND_SCAN_APPLY, [&] { return launch_scan_apply(stream, counts.p, nbuckets, partial.p, start.p); });
    counts_dirty = false;
    run(KIND_SCATTER, [&] { return launch_scatter(stream, n, keys.p, slots.p, start.p, tmp.p); });
    run(KIND_RANK_GATHER, [&] {
        return launch_rank_gather(stream, nt, d_tiles.p, d_scenes.p, tmp.p, keys.p, start.p, posm[cs].p, vel[cs].p,
                                  tag[ct].p, posm[ns].p, vel[ns].p, tag[nt_].p, rc ? rec_sorted.p : nullptr, perm_out);
    });
}

// K4 + K5 of the plain step on generation 1-cs, writing generation cs.
void Engine::plain_phase(float4 *acc_sorted, const ListOut *lo) {
    const uint32_t nt = (uint32_t)tiles.size();
    const bool rd = rec_flags & PHIENG_RECORD_DENSITY, ra = rec_flags & PHIENG_RECORD_ACCEL;
    const int ns = 1 - cs, nt_ = 1 - ct;
    run(KIND_DENSITY, [&] {
        if (lo)
            return launch_density_list(stream, nt, d_tiles.p, d_scenes.p, start.p, posm[ns].p, tag[nt_].p, rq.p,
                                       masks.p, flags.p, (uint32_t)alloc_particles, rd ? rec_rho.p : nullptr,
                                       rd ? rec_p.p : nullptr, rd ? rec_count.p : nullptr, *lo);
        return launch_density(stream, nt, d_tiles.p, d_scenes.p, start.p, posm[ns].p, tag[nt_].p, rq.p, masks.p,
                              flags.p, (uint32_t)alloc_particles, rd ? rec_rho.p : nullptr, rd ? rec_p.p : nullptr,
                              rd ? rec_count.p : nullptr);
    });
    run(KIND_FORCE, [&] {
        return launch_force_integrate(stream, nt, d_tiles.p, d_scenes.p, d_bodies.p, start.p, posm[ns].p,
                                      vel[ns].p, rq.p, tag[nt_].p, masks.p, flags.p, (uint32_t)alloc_particles,
                                      posm[cs].p, vel[cs].p, ra ? rec_acc.p : nullptr, d_bbox.p, acc_sorted);
    });
    gen_open = true;
}

// Generation flip of the tags, records, domain monitor.
void Engine::finish_substep() {
    ct = 1 - ct;
    gen_open = false;
    rec_valid = rec_flags != 0;
    ++substeps_done;
    if (!mon_pending && (substeps_done % monitor_every) == 0 && scenes.size() <= 4096) {
        run(KIND_MONITOR, [&] {
            return launch_monitor(stream, d_scenes.p, (uint32_t)scenes.size(), start.p, d_bbox.p, d_monitor.p);
        });
        PH_CK(hipMemcpyAsync(h_monitor, d_monitor.p, 8 * sizeof(int32_t) * scenes.size(), hipMemcpyDeviceToHost,
                             stream));
        PH_CK(hipEventRecord(mon_event, stream));
        mon_pending = true;
    }
}

void Engine::recover() {
    (void)hipGetLastError();
    if (gen_open) {
        ct = 1 - ct;
        gen_open = false;
    }
    aux_on = false;
}

bool Engine::poll_monitor(uint32_t) {
    if (!mon_pending) return false;
    if (hipEventQuery(mon_event) != hipSuccess) return false;
    mon_pending = false;
    bool changed = false;
    for (uint32_t k = 0; k < scenes.size(); ++k) {
        const int32_t *m = h_monitor + 8 * k;
        if (m[0] <= 0) continue;
        // Outliers present: grow the domain to the current cell box.
        SceneDev &s = scenes[k];
        int32_t box[6] = {std::min(m[2], s.x0), std::min(m[3], s.y0), std::min(m[4], s.z0),
                          std::max(m[5], s.x0 + s.nx - 1), std::max(m[6], s.y0 + s.ny - 1),
                          std::max(m[7], s.z0 + s.nz - 1)};
        std::vector<BodyDev> none;
        choose_domain(s, box, s.part_end - s.part_begin, none);
        changed = true;
    }
    if (changed) assign_buckets();
    return changed;
}

void Engine::particle_cell_box(uint32_t k, int32_t out[6]) {
    const SceneDev &s = scenes[k];
    int32_t init[6] = {INT32_MAX, INT32_MAX, INT32_MAX, INT32_MIN, INT32_MIN, INT32_MIN};
    d_box6.ensure(6);
    PH_CK(hipMemcpyAsync(d_box6.p, init, sizeof(init), hipMemcpyHostToDevice, stream));
    const uint32_t n = s.part_end - s.part_begin;
    if (n)
        run(KIND_BBOX, [&] { return launch_cell_bbox(stream, n, posm[cs].p + s.part_begin, s.h_grid, d_box6.p); });
    PH_CK(hipMemcpyAsync(out, d_box6.p, sizeof(init), hipMemcpyDeviceToHost, stream));
    PH_CK(hipStreamSynchronize(stream));
}

float Engine::max_speed(uint32_t begin, uint32_t end) {
    uint32_t zero = 0, bits = 0;
    PH_CK(hipMemcpyAsync(d_vmax.p, &zero, sizeof(zero), hipMemcpyHostToDevice, stream));
    if (end > begin) run(KIND_VMAX, [&] { return launch_vmax(stream, end - begin, vel[cs].p + begin, d_vmax.p); });
    PH_CK(hipMemcpyAsync(&bits, d_vmax.p, sizeof(bits), hipMemcpyDeviceToHost, stream));
    PH_CK(hipStreamSynchronize(stream));
    float v;
    std::memcpy(&v, &bits, sizeof(v));
    return v;
}

hipEvent_t Engine::get_event() {
    if (!ev_pool.empty()) {
        hipEvent_t e = ev_pool.back();
        ev_pool.pop_back();
        return e;
    }
    hipEvent_t e;
    PH_CK(hipEventCreate(&e));
    return e;
}

void Engine::harvest_timing() {
    if (pending_ev.empty()) return;
    PH_CK(hipStreamSynchronize(stream));
    for (auto &e : pending_ev) {
        float ms = 0;
        PH_CK(hipEventElapsedTime(&ms, e.a, e.b));
        total_ms[e.kind] += ms;
        launches[e.kind] += 1;
        ev_pool.push_back(e.a);
        ev_pool.push_back(e.b);
    }
    pending_ev.clear();
}

void Engine::reset_timing() {
    harvest_timing();
    for (int k = 0; k < KIND_COUNT; ++k) {
        total_ms[k] = 0;
        launches[k] = 0;
    }
}

}  // namespace ph
