// engine.hip — step orchestration of the device-resident SPH state.
#include "engine.h"

#include <cstdlib>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <stdexcept>

#include "resample.h"
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
        "surrogate_collect", "boundary_ghosts", "resample"};
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
    s.rest_density = c.rest_density;
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

void make_ghosts(const SceneHost &h, std::vector<float4> &out) {
    constexpr size_t kMaxGhosts = size_t(1) << 22;
    out.clear();
    const FluidConfig3Ffi &c = h.cfg;
    const float hk = rmax_h(c.smoothing_length, kEps);
    const float spacing = rmax_h(hk * 0.5f, kEps);
    const float gm = rmax_h(c.particle_mass, kEps);
    const float sp2 = spacing * spacing;  // spacing.powi(2)
    // ceil(v).clamp(1, cap) as usize; cap 0: no upper clamp (n.max(1))
    auto runs = [](float v, float cap) -> uint32_t {
        float k = std::ceil(v);
        if (!(k >= 1.0f)) k = 1.0f;
        if (cap > 0.0f && k > cap) k = cap;
        return k >= 4294967295.0f ? 0xFFFFFFFFu : (uint32_t)k;
    };
    // transform_point with the identity rotation: local + position
    auto put = [&](const BodyDev &b, float x, float y, float z) {
        if (out.size() >= kMaxGhosts) throw std::runtime_error("boundary ghosts exceed 4,194,304");
        out.push_back(make_float4(x + b.px, y + b.py, z + b.pz, gm));
    };
    for (const BodyDev &b : h.bodies) {
        if (b.flags & BODY_DYNAMIC) continue;  // ghosts of static bodies only (fluid3.rs:724)
        if (h.dims == 2) {
            if (b.shape == 0u) {  // circle ring, fluid.rs:673-688
                const uint32_t m = runs((2.0f * kPiF * b.a) / spacing, 0.0f);
                for (uint32_t k = 0; k < m; ++k) {
                    const float a = 2.0f * kPiF * (float)k / (float)m;
                    put(b, std::cos(a) * b.a, std::sin(a) * b.a, 0.0f);
                }
            } else if (b.shape == 3u) {  // four edges, fluid.rs:689-714
                // origin and tangent per edge in the reference's order (the
                // first two run along y, the last two along x)
                const float ox[4] = {-b.a, -b.a, -b.a, b.a}, oy[4] = {-b.b, b.b, -b.b, -b.b};
                for (int e = 0; e < 4; ++e) {
                    const float tanx = (e >= 2) ? 1.0f : 0.0f, tany = (e >= 2) ? 0.0f : 1.0f;
                    const float extent = tanx == 1.0f ? b.a : b.b;
                    const uint32_t m = runs((2.0f * extent) / spacing, 0.0f);
                    for (uint32_t k = 0; k < m; ++k) {
                        const float t = ((float)k + 0.5f) / (float)m;
                        const float d = 2.0f * extent * t;
                        put(b, ox[e] + tanx * d, oy[e] + tany * d, 0.0f);
                    }
                }
            }
            continue;
        }
        if (b.shape == 0u) {  // Fibonacci sphere, fluid3.rs:725-748
            const uint32_t m = runs((4.0f * kPiF * b.a * b.a) / sp2, 512.0f);
            const float golden = kPiF * (3.0f - std::sqrt(5.0f));
            for (uint32_t k = 0; k < m; ++k) {
                const float z = 1.0f - 2.0f * ((float)k + 0.5f) / (float)m;
                const float r = std::sqrt(1.0f - z * z);
                const float th = golden * (float)k;
                put(b, std::cos(th) * r * b.a, z * b.a, std::sin(th) * r * b.a);
            }
        } else if (b.shape == 1u) {  // box, fluid3.rs:749-791: six sweeps, the
            // first coordinate of each axis pair left at 0 and z at the
            // "fixed" value, as the reference lays them out
            const float he[3] = {b.a, b.b, b.c};
            for (int f = 0; f < 6; ++f) {
                const int axis = f / 2;
                const float fixed = axis == 0 ? -he[1] : (axis == 1 ? -he[0] : -he[2]);
                const float area = axis == 0 ? 4.0f * he[1] * he[2]
                                             : (axis == 1 ? 4.0f * he[0] * he[2] : 4.0f * he[0] * he[1]);
                const uint32_t m = runs(area / sp2, 512.0f);
                const int u_ax = axis == 0 ? 1 : 0, w_ax = axis == 2 ? 1 : 2;
                for (uint32_t k = 0; k < m; ++k) {
                    const float t = ((float)k + 0.5f) / (float)m;
                    float l[3] = {0.0f, 0.0f, fixed};
                    l[u_ax] = (2.0f * t - 1.0f) * he[u_ax];
                    l[w_ax] = (2.0f * std::fmod(t * 7.0f, 1.0f) - 1.0f) * he[w_ax];
                    put(b, l[0], l[1], l[2]);
                }
            }
        } else if (b.shape == 2u) {  // capsule rings and end points, fluid3.rs:792-826
            const float r = b.a, hl = b.b;
            const uint32_t m = runs((4.0f * kPiF * r * hl) / sp2, 256.0f);
            for (float sy : {-hl * 0.5f, hl * 0.5f})
                for (uint32_t k = 0; k < m; ++k) {
                    const float a = 2.0f * kPiF * (float)k / (float)m;
                    put(b, std::cos(a) * r, sy, std::sin(a) * r);
                }
            for (float sg : {-1.0f, 1.0f}) {
                const float e = hl + r;
                put(b, 0.0f * sg * e, 1.0f * sg * e, 0.0f * sg * e);  // Vec3::Y * s * (hl + r)
            }
        }
    }
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
        if (b.flags & BODY_DYNAMIC) continue;  // they move; the domain follows the particles
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
        for (auto ev : vl_ev)
            if (ev) (void)hipEventDestroy(ev);
        if (vl_host) (void)hipHostFree(vl_host);
        if (rs_host) (void)hipHostFree(rs_host);
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
    if (const char *e = std::getenv("PHIENG_RESAMPLE_STORE_INITIAL")) rs_items_initial = std::strtoull(e, nullptr, 10);
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
    nl_cnt.ensure(want);
    nl_idx.ensure((size_t)(SL_CAP + 4) * want);
    if (list_cells) nl_cell.ensure(want);
    alloc_particles = want;
    if (rec_flags) ensure_record();
}

void Engine::set_ghosts(const std::vector<float4> &gen, float hk) {
    init_device();
    const uint32_t g = (uint32_t)gen.size();
    struct Cell {
        int32_t x, y, z;
    };
    // rebuild_boundary_grid (fluid3.rs:843-857): cells at the kernel h; a
    // ghost with a non-finite x is in no cell (the 2D rule also checks y,
    // fluid.rs:749; ghosts of finite bodies are finite either way)
    std::vector<Cell> cell(g);
    std::vector<uint32_t> order;
    order.reserve(g);
    for (uint32_t k = 0; k < g; ++k) {
        const float4 p = gen[k];
        if (!std::isfinite(p.x) || !std::isfinite(p.y)) continue;
        cell[k] = {sat_i32_h(std::floor(p.x / hk)), sat_i32_h(std::floor(p.y / hk)), sat_i32_h(std::floor(p.z / hk))};
        order.push_back(k);
    }
    // cell by cell, ghost index ascending inside a cell
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
        const Cell &p = cell[a], &q = cell[b];
        if (p.z != q.z) return p.z < q.z;
        if (p.y != q.y) return p.y < q.y;
        return p.x < q.x;
    });
    const uint32_t m = (uint32_t)order.size();
    std::vector<float4> pos(std::max(m, 1u));
    for (uint32_t k = 0; k < m; ++k) pos[k] = gen[order[k]];
    uint32_t ncell = 0;
    for (uint32_t k = 0; k < m; ++k)
        if (k == 0 || std::memcmp(&cell[order[k]], &cell[order[k - 1]], sizeof(Cell)) != 0) ++ncell;
    uint32_t cap = 16;
    while (cap < 2 * ncell) cap <<= 1;
    std::vector<int4> key(cap, make_int4(0, 0, 0, 0));
    std::vector<uint2> range(cap, make_uint2(0, 0));
    for (uint32_t k = 0; k < m;) {
        uint32_t e = k + 1;
        while (e < m && std::memcmp(&cell[order[e]], &cell[order[k]], sizeof(Cell)) == 0) ++e;
        const Cell c = cell[order[k]];
        uint32_t slot = ghost_hash(c.x, c.y, c.z) & (cap - 1);
        while (range[slot].y != 0) slot = (slot + 1) & (cap - 1);
        key[slot] = make_int4(c.x, c.y, c.z, 0);
        range[slot] = make_uint2(k, e);
        k = e;
    }
    g_pos.ensure(pos.size());
    g_state.ensure(pos.size());
    g_key.ensure(cap);
    g_range.ensure(cap);
    PH_CK(hipMemcpyAsync(g_pos.p, pos.data(), sizeof(float4) * pos.size(), hipMemcpyHostToDevice, stream));
    PH_CK(hipMemcpyAsync(g_key.p, key.data(), sizeof(int4) * cap, hipMemcpyHostToDevice, stream));
    PH_CK(hipMemcpyAsync(g_range.p, range.data(), sizeof(uint2) * cap, hipMemcpyHostToDevice, stream));
    PH_CK(hipStreamSynchronize(stream));  // the host vectors go out of scope
    g_count = m;
    g_mask = cap - 1;
    g_order = std::move(order);
}

GhostDev Engine::ghost_dev() {
    GhostDev d;
    d.pos = g_pos.p;
    d.state = g_state.p;
    d.key = g_key.p;
    d.range = g_range.p;
    return d;
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
    if (vl_enabled && n_total >= vl_min_particles && !(scenes.size() == 1 && scenes[0].ghost_count)) {
        if (vl_list_left == 0 && vl_k >= kVlLag && (vl_k - kVlLag) % 4 == 0) {
            const uint64_t u = vl_k - kVlLag;  // a multiple of 4: it recorded an event
            PH_CK(hipEventSynchronize(vl_ev[(u / 4) % 8]));
            if (vl_host[2 * (u % 32)] > vl_repair_limit()) {
                vl_list_left = vl_backoff;
                vl_backoff = std::min(2u * vl_backoff, 1024u);
            } else if (vl_k >= 512) {
                vl_backoff = 16;
            }
        }
        if (vl_list_left == 0) {
            substep_verlet();
            return;
        }
        --vl_list_left;
        vl_k = 0;
        ++vl_list_substeps;
    }
    sort_phase(nullptr);
    plain_phase(nullptr);
    finish_substep();
}

void Engine::substep_verlet() {
    const uint32_t nt = (uint32_t)tiles.size();
    const uint32_t n = n_total;
    const size_t m = alloc_particles;
    const bool rc = rec_flags & PHIENG_RECORD_CELLS, rd = rec_flags & PHIENG_RECORD_DENSITY,
               ra = rec_flags & PHIENG_RECORD_ACCEL;
    // buffers first, so no allocation failure can cut the substep short
    vl_cnt.ensure(m);
    vl_idx.ensure((size_t)VL_CAP * m);
    vl_word[0].ensure(m);
    vl_word[1].ensure(m);
    vl_row16.ensure((size_t)(VL_CAP / 8) * m);
    vl_state.ensure(1);
    const size_t ntl = std::max<size_t>(nt, 1);
    vl_halo.ensure(ntl * VL_HALO_CAP);
    vl_halo_n.ensure(ntl);
    const uint32_t np = scan_partials_needed(nbuckets);
    if (vl_partial.n < np) {
        vl_partial.ensure(np);
        vl_partial_dirty = true;
    }
    if (!vl_state_ready) {
        run(KIND_FILL, [&] {
            return launch_fill_u32(stream, reinterpret_cast<uint32_t *>(vl_state.p),
                                   (uint32_t)(sizeof(VerletDev) / sizeof(uint32_t)), 0u);
        });
        vl_state_ready = true;
    }
    if (!vl_host) {
        PH_CK(hipHostMalloc((void **)&vl_host, 64 * sizeof(uint32_t), hipHostMallocDefault));
        for (auto &ev : vl_ev) PH_CK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    }
    const uint32_t slot = (uint32_t)(vl_k % 32);
    float hmin = INFINITY;
    for (const SceneDev &s : scenes) hmin = std::min(hmin, s.h_grid);
    const float skin = std::min(std::max(vl_skin, 0.0f), 0.5f);
    const float thr = 0.45f * skin * hmin;
    // Rows built for another particle set or layout cannot be walked; the
    // cell record is the sorted state's, so it sorts every substep.
    const bool force = !vl_valid || vl_epoch != layout_epoch || vl_n != n || rc;
    const uint32_t par = (uint32_t)(vl_k & 1u);
    const int ns = 1 - cs, nt_ = 1 - ct;
    const uint32_t *gate = reinterpret_cast<const uint32_t *>(vl_state.p);  // VerletDev::rebuild
    vl_valid = false;  // until this substep is queued whole
    if (counts_dirty)
        run(KIND_FILL, [&] { return launch_fill_u32(stream, counts.p, (uint32_t)alloc_buckets, 0u); });
    if (vl_partial_dirty)
        run(KIND_FILL, [&] { return launch_fill_u32(stream, vl_partial.p, (uint32_t)vl_partial.n, 0u); });
    counts_dirty = true;
    vl_partial_dirty = true;
    ++sorts_done;
    // the cell-key pass adds up the scan's block sums (no reduce launch);
    // the scatter zeroes them again
    run(KIND_CELL_KEY, [&] {
        return launch_cell_key(stream, nt, d_tiles.p, d_scenes.p, posm[cs].p, tag[ct].p, keys.p, slots.p, counts.p,
                               rc ? rec_cells.p : nullptr, vl_state.p,
                               VlStep{force ? 1u : 0u, par, thr, vl_host + 2 * slot}, vl_word[0].p, vl_partial.p);
    });
    run(KIND_SCAN_APPLY, [&] { return launch_scan_apply_sum(stream, counts.p, nbuckets, vl_partial.p, start.p, gate); });
    counts_dirty = false;
    run(KIND_SCATTER, [&] {
        return launch_scatter(stream, n, keys.p, slots.p, start.p, tmp.p, gate, vl_partial.p, np);
    });
    vl_partial_dirty = false;
    run(KIND_RANK_GATHER, [&] {
        return launch_rank_gather(stream, nt, d_tiles.p, d_scenes.p, tmp.p, keys.p, start.p, posm[cs].p, vel[cs].p,
                                  tag[ct].p, posm[ns].p, vel[ns].p, tag[nt_].p, rc ? rec_sorted.p : nullptr, nullptr,
                                  vl_state.p, vl_word[0].p, vl_word[1].p);
    });
    VerletArgs va;
    va.vword = vl_word[1].p;
    va.vcnt = vl_cnt.p;
    va.vidx = vl_idx.p;
    va.vstride = list_stride();
    va.f2 = (1.0f + skin) * (1.0f + skin);
    va.skin = skin;
    va.vs = vl_state.p;
    va.par = par;
    va.tag_mirror = tag[ct].p;
    va.halo = vl_halo.p;
    va.halo_n = vl_halo_n.p;
    va.rows16 = vl_row16.p;
    va.lidx = nl_idx.p;
    va.lstride = list_stride();
    // The walk and the force pass read generation cs: between rebuilds it
    // holds the state in the kept order (the gather leaves it), and a
    // rebuild's density pass copies the sorted state (generation ns) there.
    // The force pass writes generation ns, which becomes the state.
    va.keep_posm = posm[cs].p;
    va.sorted_vel = vel[ns].p;
    va.keep_vel = vel[cs].p;
    run(KIND_DENSITY, [&] {
        return launch_density_verlet(stream, nt, d_tiles.p, d_scenes.p, start.p, posm[ns].p, posm[cs].p,
                                     tag[nt_].p, rq.p, rd ? rec_rho.p : nullptr, rd ? rec_p.p : nullptr,
                                     rd ? rec_count.p : nullptr, nl_cnt.p, va);
    });
    gen_open = true;
    run(KIND_FORCE, [&] {
        return launch_force_verlet(stream, nt, d_tiles.p, d_scenes.p, d_bodies.p, start.p, posm[cs].p, vel[cs].p,
                                   rq.p, tag[nt_].p, nl_cnt.p, nl_idx.p, list_stride(), posm[ns].p, vel[ns].p,
                                   ra ? rec_acc.p : nullptr, d_bbox.p, GhostDev{}, vl_state.p, vl_word[1].p,
                                   vl_halo.p, vl_halo_n.p, par);
    });
    cs = ns;  // the force pass was queued: its output is the state
    finish_substep();
    if (vl_k % 4 == 0) PH_CK(hipEventRecord(vl_ev[(vl_k / 4) % 8], stream));
    ++vl_k;
    vl_valid = true;
    vl_epoch = layout_epoch;
    vl_n = n;
}

// K1 .. K3: keys, bucket scan, scatter, rank + gather into generation 1-cs.
void Engine::sort_phase(uint32_t *perm_out) {
    vl_valid = false;  // reorders particles without keeping the rebuild slots
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
    run(KIND_SCAN_APPLY, [&] { return launch_scan_apply_sum(stream, counts.p, nbuckets, partial.p, start.p); });
    counts_dirty = false;
    run(KIND_SCATTER, [&] { return launch_scatter(stream, n, keys.p, slots.p, start.p, tmp.p); });
    run(KIND_RANK_GATHER, [&] {
        return launch_rank_gather(stream, nt, d_tiles.p, d_scenes.p, tmp.p, keys.p, start.p, posm[cs].p, vel[cs].p,
                                  tag[ct].p, posm[ns].p, vel[ns].p, tag[nt_].p, rc ? rec_sorted.p : nullptr, perm_out);
    });
}

// K4 + K5 of the plain step on generation 1-cs, writing generation cs.
void Engine::plain_phase(float4 *acc_sorted, const AuxFuse *af, bool density_done) {
    const uint32_t nt = (uint32_t)tiles.size();
    const bool rd = rec_flags & PHIENG_RECORD_DENSITY, ra = rec_flags & PHIENG_RECORD_ACCEL;
    const int ns = 1 - cs, nt_ = 1 - ct;
    if (list_cells) nl_cell.ensure(alloc_particles);
    ListOut lo;
    lo.cell = list_cells ? nl_cell.p : nullptr;
    lo.cnt = nl_cnt.p;
    lo.idx = nl_idx.p;
    lo.skin_factor2 = (1.0f + list_skin) * (1.0f + list_skin);
    if (list_cells && list_coef > 0.0f) {
        tile_vmax.ensure(std::max<size_t>(nt, 1));
        list_f2.ensure(1);
        run(KIND_SUR_AUX, [&] {
            return launch_list_radius(stream, nt, d_tiles.p, d_scenes.p, vel[ns].p, tile_vmax.p, list_coef, list_skin,
                                      list_f2.p, list_vmax_out);
        });
        lo.skin_factor2_dev = list_f2.p;
        lo.slow = list_slow;
    }
    if (!density_done)
        run(KIND_DENSITY, [&] {
            return launch_density_list(stream, nt, d_tiles.p, d_scenes.p, start.p, posm[ns].p, tag[nt_].p, rq.p,
                                       list_stride(), rd ? rec_rho.p : nullptr,
                                       rd ? rec_p.p : nullptr, rd ? rec_count.p : nullptr, lo);
        });
    // boundary ghosts (fluid3.rs:1018-1027): their density and pressure from
    // the fluid of this substep, before the force pass reads them
    const bool ghosts = scenes.size() == 1 && scenes[0].ghost_count;
    if (ghosts)
        run(KIND_GHOST, [&] {
            return launch_ghost_density(stream, d_scenes.p, 0u, start.p, posm[ns].p, ghost_dev(), g_count);
        });
    run(KIND_FORCE, [&] {
        return launch_force_list(stream, nt, d_tiles.p, d_scenes.p, d_bodies.p, start.p, posm[ns].p, vel[ns].p, rq.p,
                                 tag[nt_].p, nl_cnt.p, nl_idx.p, list_stride(), posm[cs].p, vel[cs].p,
                                 ra ? rec_acc.p : nullptr, d_bbox.p, acc_sorted, ghosts ? ghost_dev() : GhostDev{},
                                 af ? *af : AuxFuse{});
    });
    gen_open = true;
}

void Engine::density_phase() {
    const uint32_t nt = (uint32_t)tiles.size();
    const bool rd = rec_flags & PHIENG_RECORD_DENSITY;
    const int ns = 1 - cs, nt_ = 1 - ct;
    ListOut lo;
    lo.cnt = nl_cnt.p;
    lo.idx = nl_idx.p;
    lo.skin_factor2 = 1.0f;
    run(KIND_DENSITY, [&] {
        return launch_density_list(stream, nt, d_tiles.p, d_scenes.p, start.p, posm[ns].p, tag[nt_].p, rq.p,
                                   list_stride(), rd ? rec_rho.p : nullptr,
                                   rd ? rec_p.p : nullptr, rd ? rec_count.p : nullptr, lo);
    });
}

uint32_t Engine::resample_phase(float merge_hi, float split_lo, float min_mass, float h, uint32_t cap,
                                bool &changed) {
    const uint32_t n = n_total;
    const int ns = 1 - cs, nt_ = 1 - ct;
    // room for every split the cap allows: an accepted split has
    // kept + eligible before it + 1 < cap, so at most cap / 2 are accepted
    if ((size_t)n + std::min<size_t>(n, cap / 2 + 1) > alloc_particles)
        throw std::runtime_error("resample: the particle arrays cannot hold the split particles");
    const size_t m = alloc_particles;
    rs_ctr.ensure(4);
    if (!rs_host) PH_CK(hipHostMalloc((void **)&rs_host, 4 * sizeof(unsigned long long), hipHostMallocDefault));
    rs_flag.ensure(m);
    rs_minT.ensure(m);
    rs_pick.ensure(m);
    rs_state.ensure(m);
    rs_cK.ensure(m);
    rs_cE.ensure(m);
    rs_rK.ensure(m + 1);
    rs_rE.ensure(m + 1);
    rs_slot.ensure(m);
    rs_partial.ensure(scan_partials_needed((uint32_t)m) + 1);
    rs_takers.ensure(m);
    rs_mv.ensure(m);
    if (!rs_items.p) rs_items.ensure(rs_items_initial ? rs_items_initial : m);
    ResampleArgs a;
    a.tiles = d_tiles.p;
    a.scenes = d_scenes.p;
    a.start = start.p;
    a.posm = posm[ns].p;
    a.vel = vel[ns].p;
    a.tag = tag[nt_].p;
    a.rq = rq.p;
    a.n = n;
    a.merge_hi = merge_hi;
    a.split_lo = split_lo;
    a.min_mass = min_mass;
    const float q = 0.25f * h;
    a.md2 = q * q;
    a.off = q;
    a.cap = cap;
    a.ctr = rs_ctr.p;
    a.flag = rs_flag.p;
    a.takers = rs_takers.p;
    a.minT = rs_minT.p;
    a.pick = rs_pick.p;
    a.state = rs_state.p;
    a.mv = rs_mv.p;
    a.cK = rs_cK.p;
    a.cE = rs_cE.p;
    a.rK = rs_rK.p;
    a.rE = rs_rE.p;
    a.slot_of = rs_slot.p;
    a.partial = rs_partial.p;
    a.posm_out = posm[cs].p;
    a.vel_out = vel[cs].p;
    a.tag_out = tag[ct].p;
    for (;;) {
        a.items = rs_items.p;
        a.pcap = (uint32_t)std::min<size_t>(rs_items.n, 0xFFFFFFFFu);
        run(KIND_RESAMPLE, [&] { return launch_resample(stream, a); });
        PH_CK(hipMemcpyAsync(rs_host, rs_ctr.p, 4 * sizeof(unsigned long long), hipMemcpyDeviceToHost, stream));
        PH_CK(hipStreamSynchronize(stream));
        if (rs_host[1] <= a.pcap) break;
        if (rs_host[1] > 0xFFFFFFFFull) throw std::runtime_error("resample: more than 2^32 merge candidates");
        rs_items.ensure((size_t)rs_host[1]);  // the candidate store was too small: grow, run again
    }
    const uint64_t kept = rs_host[3], splits = rs_host[2];
    changed = kept != n || splits != 0;
    return (uint32_t)(kept + splits);
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
    vl_valid = false;
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
