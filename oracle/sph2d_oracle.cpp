// sph2d_oracle.cpp — CPU restatement of the PhiEng 2D SPH fluid step.
//
// TEST INFRASTRUCTURE ONLY.  Like sph_oracle.cpp it is a checker and a CPU
// baseline: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
// legs load it, and the product library never links it.
//
// It follows crates/dynamics/src/fluid.rs (FluidState2) in f32 with the
// reference's operation order (built with -ffp-contract=off): the grid of
// rebuild_grid (:330-363), neighbors (:368-388), poly6 / spiky_grad /
// pressure / artificial_pressure / cohesion (:265-327), cfl_substeps
// (:562-589), compute_density_pressure (:592-614) and substep (:898-1136)
// with the in-place Gauss-Seidel integrate of the reference (a Jacobi mode
// reads pre-step state only) and the static circle / box boundary of
// collide_particle_body (:463-525, identity rotation).  Buoyancy (:1076-1135)
// only acts on dynamic bodies and is not restated.
//
// What pins it (tests/test_oracle2d.py): the reference's kernel sanity and
// Tait-EOS known answers (fluid.rs:1209-1223, 1311-1331) including the
// bit-level linear-EOS result, the 2D kernel normalisation, O(N^2)
// brute-force neighbour counts and densities, the lattice closed form, the
// two-particle force closed form and momentum antisymmetry, and the
// reference's behaviour tests restated (dam break contained :1225-1256,
// determinism :1292-1309, hydrostatic band :1676-1719).  The optional pair
// terms (artificial viscosity, cohesion, XSPH) are pinned by the reference's
// sign tests (fluid.rs:1333-1360) and two-particle closed forms of a whole
// substep in tests/test_oracle_pair_terms.py.
// Parity unpinned: glam 0.29.3 Vec2 association order ((x*x) + (y*y)) and
// the last ulp of libm powf (taken in f64 and rounded, as in the 3D oracle).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <unordered_map>
#include <utility>
#include <vector>

namespace {

constexpr float kEps = 1e-6f;     // crates/math/src/math.rs:3
constexpr float kEpsSq = 1e-12f;  // crates/math/src/math.rs:4
constexpr float kPiF = 3.14159265358979323846f;

inline float rmax(float a, float b) {  // Rust f32::max
    if (a != a) return b;
    if (b != b) return a;
    return a > b ? a : b;
}
inline float rclamp(float v, float lo, float hi) {  // Rust f32::clamp
    if (v < lo) v = lo;
    if (v > hi) v = hi;
    return v;
}
inline int32_t sat_i32(float v) {  // Rust `f32 as i32`
    if (v != v) return 0;
    if (v >= 2147483648.0f) return INT32_MAX;
    if (v < -2147483648.0f) return INT32_MIN;
    return (int32_t)v;
}
inline bool finite2(float x, float y) { return std::isfinite(x) && std::isfinite(y); }
inline float len_sq(float x, float y) { return x * x + y * y; }  // glam Vec2::length_squared

}  // namespace

extern "C" {

// Mirror of FluidConfig2 (fluid.rs:55-140) fields the step reads.
struct Oracle2Cfg {
    float rest_density, smoothing_length, stiffness, gamma, artificial_viscosity, surface_tension, cfl;
    uint32_t max_substeps;
    float viscosity, wall_friction;
    float phase_rest_density[4];
    float max_velocity, particle_radius;
    uint32_t boundary_particles;  // FluidConfig2::boundary_particles
    float particle_mass;          // FluidConfig2::particle_mass (ghost mass, min split mass)
    uint32_t adaptive_resolution; // FluidConfig2::adaptive_resolution
    float split_density_ratio, merge_density_ratio;
};

// Static body, identity rotation: shape 0 circle (a = radius), 1 box
// (a, b = half extents).
struct Oracle2Body {
    uint32_t shape;
    float px, py, a, b;
};

}  // extern "C"

namespace {

struct P2 {
    float x, y, vx, vy, rho, p, m;
    uint8_t phase;
};

struct Ghost2 {
    float x, y, rho, p, m;
};

struct Fluid2 {
    Oracle2Cfg cfg;
    std::vector<P2> ps;
    std::unordered_map<uint64_t, std::vector<uint32_t>> grid;
    std::vector<float> fx, fy;
    std::vector<Ghost2> ghosts;
    std::unordered_map<uint64_t, std::vector<uint32_t>> ghost_grid;
    uint32_t cap = 4096;  // MAX_PARTICLES_2D (fluid.rs:26), the split cap of resample()

    static uint64_t key(int32_t cx, int32_t cy) { return ((uint64_t)(uint32_t)cx << 32) | (uint32_t)cy; }

    // fluid.rs:265-273
    static float poly6(float r2, float h) {
        if (r2 >= h * h || r2 < 0.0f) return 0.0f;
        const float h2 = h * h;
        const float t = h2 - r2;
        return 4.0f / (kPiF * h2 * h2 * h2 * h2) * t * t * t;
    }
    // fluid.rs:278-285
    static void spiky_grad(float rx, float ry, float r, float h, float &gx, float &gy) {
        if (r >= h || r < kEps) {
            gx = gy = 0.0f;
            return;
        }
        const float s = -30.0f / (kPiF * h * h * h * h * h) * (h - r) * (h - r);
        gx = rx / r * s;
        gy = ry / r * s;
    }
    // fluid.rs:290-297; powf taken in f64 and rounded
    static float pressure(float density, float rest, float k, float gamma) {
        if (gamma == 1.0f) return rmax(k * (density - rest), 0.0f);
        const float rho0 = rmax(rest, kEps);
        const float pw = (float)std::pow((double)(density / rho0), (double)gamma);
        return rmax(k * rho0 * (pw - 1.0f), 0.0f);
    }
    // fluid.rs:302-317
    static float artificial_pressure(float dx, float dy, float vx, float vy, float r_sq, float h, float rho_bar, float c,
                                     float alpha) {
        const float vd = vx * dx + vy * dy;
        if (vd >= 0.0f) return 0.0f;
        const float mu = h * vd / (r_sq + 0.01f * h * h);
        return -alpha * c * mu / rmax(rho_bar, kEps);
    }
    // fluid.rs:320-327
    static float cohesion(float dist, float h) { return dist >= h ? 0.0f : 1.0f - dist / h; }

    // fluid.rs:330-363 (cells hold indices in index order)
    void rebuild_grid() {
        grid.clear();
        const float h = cfg.smoothing_length;
        for (uint32_t i = 0; i < ps.size(); ++i) {
            const P2 &p = ps[i];
            if (!finite2(p.x, p.y)) continue;
            grid[key(sat_i32(std::floor(p.x / h)), sat_i32(std::floor(p.y / h)))].push_back(i);
        }
    }
    // fluid.rs:368-388: cells oy, ox in -1..=1, members in index order,
    // d2 < h*h with the raw smoothing length; includes i.
    void neighbors(uint32_t i, std::vector<uint32_t> &out) const { neighbors_in(ps, i, out); }
    // the same against another copy of the particles (the Jacobi mode's pre-step state)
    void neighbors_in(const std::vector<P2> &a, uint32_t i, std::vector<uint32_t> &out) const {
        out.clear();
        const float h = cfg.smoothing_length;
        const float px = a[i].x, py = a[i].y;
        const int32_t cx = sat_i32(std::floor(px / h)), cy = sat_i32(std::floor(py / h));
        for (int oy = -1; oy <= 1; ++oy)
            for (int ox = -1; ox <= 1; ++ox) {
                const int64_t x = (int64_t)cx + ox, y = (int64_t)cy + oy;
                if (x < INT32_MIN || x > INT32_MAX || y < INT32_MIN || y > INT32_MAX) continue;
                auto it = grid.find(key((int32_t)x, (int32_t)y));
                if (it == grid.end()) continue;
                for (uint32_t j : it->second) {
                    const float d2 = len_sq(a[j].x - px, a[j].y - py);
                    if (d2 < h * h) out.push_back(j);
                }
            }
    }
    // fluid.rs:666-743 for the bodies the C ABI creates (static circle and
    // box at angle 0: transform_point(p) = p + position up to zero signs).
    void refresh_boundaries(const Oracle2Body *bodies, uint32_t nb, float spacing_in) {
        ghosts.clear();
        const float rest = cfg.rest_density, gm = rmax(cfg.particle_mass, kEps);
        const float spacing = rmax(spacing_in, kEps);
        auto runs = [](float v) {  // (v).ceil() as usize, then .max(1)
            const float c = std::ceil(v);
            return c >= 1.0f ? (uint32_t)c : 1u;
        };
        for (uint32_t bi = 0; bi < nb; ++bi) {
            const Oracle2Body &b = bodies[bi];
            if (b.shape == 0u) {  // ring
                const uint32_t n = runs((2.0f * kPiF * b.a) / spacing);
                for (uint32_t k = 0; k < n; ++k) {
                    const float a = 2.0f * kPiF * (float)k / (float)n;
                    ghosts.push_back({std::cos(a) * b.a + b.px, std::sin(a) * b.a + b.py, rest, 0.0f, gm});
                }
            } else {  // edges: (origin, tangent) as listed in the reference
                const float org[4][2] = {{-b.a, -b.b}, {-b.a, b.b}, {-b.a, -b.b}, {b.a, -b.b}};
                const float tan[4][2] = {{0.0f, 1.0f}, {0.0f, 1.0f}, {1.0f, 0.0f}, {1.0f, 0.0f}};
                for (int e = 0; e < 4; ++e) {
                    const float extent = tan[e][0] == 1.0f ? b.a : b.b;
                    const uint32_t n = runs((2.0f * extent) / spacing);
                    for (uint32_t k = 0; k < n; ++k) {
                        const float t = ((float)k + 0.5f) / (float)n;
                        const float d = 2.0f * extent * t;
                        ghosts.push_back({org[e][0] + tan[e][0] * d + b.px, org[e][1] + tan[e][1] * d + b.py, rest,
                                          0.0f, gm});
                    }
                }
            }
        }
    }
    // fluid.rs:746-758, 762-780, 784-802, 804-818
    void update_boundaries(const Oracle2Body *bodies, uint32_t nb, float h) {
        refresh_boundaries(bodies, nb, h * 0.5f);
        ghost_grid.clear();
        for (uint32_t g = 0; g < ghosts.size(); ++g) {
            if (!finite2(ghosts[g].x, ghosts[g].y)) continue;
            ghost_grid[key(sat_i32(std::floor(ghosts[g].x / h)), sat_i32(std::floor(ghosts[g].y / h)))].push_back(g);
        }
        std::vector<uint32_t> nbr;
        for (Ghost2 &gh : ghosts) {
            near_cells(grid, [&](uint32_t j) { return std::make_pair(ps[j].x, ps[j].y); }, gh.x, gh.y, nbr);
            float rho = 0.0f;
            for (uint32_t j : nbr) rho += ps[j].m * poly6(len_sq(ps[j].x - gh.x, ps[j].y - gh.y), h);
            gh.rho = rclamp(rho, 0.0f, 3.0f * rmax(cfg.rest_density, kEps));
            gh.p = pressure(rho, cfg.rest_density, cfg.stiffness, cfg.gamma);
        }
    }
    // The 9 cells around a point at the raw smoothing length, oy -> ox,
    // members in index order, d2 < h * h (boundary_near / fluid_near).
    template <class At>
    void near_cells(const std::unordered_map<uint64_t, std::vector<uint32_t>> &g, At at, float px, float py,
                    std::vector<uint32_t> &out) const {
        out.clear();
        const float h = cfg.smoothing_length;
        const int32_t cx = sat_i32(std::floor(px / h)), cy = sat_i32(std::floor(py / h));
        for (int oy = -1; oy <= 1; ++oy)
            for (int ox = -1; ox <= 1; ++ox) {
                const int64_t x = (int64_t)cx + ox, y = (int64_t)cy + oy;
                if (x < INT32_MIN || x > INT32_MAX || y < INT32_MIN || y > INT32_MAX) continue;
                auto it = g.find(key((int32_t)x, (int32_t)y));
                if (it == g.end()) continue;
                for (uint32_t j : it->second) {
                    const std::pair<float, float> q = at(j);
                    if (len_sq(q.first - px, q.second - py) < h * h) out.push_back(j);
                }
            }
    }

    // fluid.rs:562-589
    uint32_t cfl_substeps(float dt) const {
        const uint32_t mx = cfg.max_substeps > 1 ? cfg.max_substeps : 1;
        if (cfg.cfl <= 0.0f || mx <= 1) return 1;
        const float h = rmax(cfg.smoothing_length, kEps);
        const float rho0 = rmax(cfg.rest_density, kEps);
        const float c0 = cfg.gamma == 1.0f ? std::sqrt(rmax(cfg.stiffness, 0.0f))
                                           : std::sqrt(cfg.gamma * rmax(cfg.stiffness, 0.0f) * rho0);
        float vmax = 0.0f;
        for (const P2 &p : ps) vmax = rmax(vmax, std::sqrt(len_sq(p.vx, p.vy)));
        const float speed = c0 + vmax;
        if (!std::isfinite(speed) || speed <= kEps) return 1;
        const float dt_max = cfg.cfl * h / speed;
        if (!std::isfinite(dt_max) || dt_max <= 0.0f) return 1;
        const float q = std::ceil(dt / dt_max);
        int64_t s;
        if (q != q) s = 0;
        else if (q >= 9.2233720368547758e18f) s = INT64_MAX;
        else if (q <= -9.2233720368547758e18f) s = INT64_MIN;
        else s = (int64_t)q;
        if (s < 1) s = 1;
        if (s > (int64_t)mx) s = mx;
        return (uint32_t)s;
    }
    // fluid.rs:592-614
    void density_pressure(float h, std::vector<uint32_t> &nb, uint32_t *count) {
        for (uint32_t i = 0; i < ps.size(); ++i) {
            neighbors(i, nb);
            const float px = ps[i].x, py = ps[i].y;
            float rho = 0.0f;
            for (uint32_t j : nb) rho += ps[j].m * poly6(len_sq(ps[j].x - px, ps[j].y - py), h);
            ps[i].rho = rclamp(rho, 0.0f, 3.0f * rmax(cfg.rest_density, kEps));
            if (count) count[i] = (uint32_t)nb.size();
        }
        for (P2 &p : ps) {
            const float rho0 = cfg.phase_rest_density[p.phase < 3 ? p.phase : 3];
            p.p = pressure(p.rho, rho0, cfg.stiffness, cfg.gamma);
        }
    }
    // fluid.rs:463-525 for identity-rotation circle and box; false when separated.
    static bool collide(float px, float py, float pr, const Oracle2Body &b, float &nx, float &ny, float &depth) {
        if (b.shape == 0u) {
            const float dx = px - b.px, dy = py - b.py;
            const float dist = std::sqrt(len_sq(dx, dy));
            if (dist >= b.a + pr) return false;
            if (dist > kEps) {
                nx = dx / dist;
                ny = dy / dist;
            } else {
                nx = 1.0f;
                ny = 0.0f;
            }
            depth = b.a + pr - dist;
            return true;
        }
        const float rx = px - b.px, ry = py - b.py;
        // local = R^T rel with R = identity (x_axis (1, 0), y_axis (0, 1))
        const float lx = 1.0f * rx + 0.0f * ry, ly = 0.0f * rx + 1.0f * ry;
        const float cx = rclamp(lx, -b.a, b.a), cy = rclamp(ly, -b.b, b.b);
        const float ddx = lx - cx, ddy = ly - cy;
        float lnx, lny;
        if (len_sq(ddx, ddy) > kEpsSq) {
            const float dist = std::sqrt(len_sq(ddx, ddy));
            if (dist >= pr) return false;
            lnx = ddx / dist;
            lny = ddy / dist;
            depth = pr - dist;
        } else {
            // least-penetration axis; signum().copysign(1.0) is +1 (fluid.rs:507-512)
            const float dx = b.a - std::fabs(lx), dy = b.b - std::fabs(ly);
            if (dx < dy) {
                lnx = 1.0f;
                lny = 0.0f;
                depth = pr + dx;
            } else {
                lnx = 0.0f;
                lny = 1.0f;
                depth = pr + dy;
            }
        }
        // normal = R local_n (glam Mat2 * Vec2: x_axis * n.x + y_axis * n.y)
        nx = 1.0f * lnx + 0.0f * lny;
        ny = 0.0f * lnx + 1.0f * lny;
        return true;
    }

    // fluid.rs:826-891, adaptive resolution (the 3D restatement documents
    // the rules; the split offset is (h/4, 0)).
    void resample(float h) {
        const float rest = rmax(cfg.rest_density, kEps);
        const float split_lo = cfg.split_density_ratio * rest;
        const float merge_hi = cfg.merge_density_ratio * rest;
        const float min_mass = rmax(cfg.particle_mass, kEps) * 0.125f;
        const size_t n = ps.size();
        std::vector<uint8_t> absorbed(n, 0);
        const float merge_dist_sq = (0.25f * h) * (0.25f * h);
        for (size_t i = 0; i < n; ++i) {
            if (absorbed[i] || ps[i].rho <= merge_hi) continue;
            for (size_t j = i; j-- > 0;) {
                if (absorbed[j] || ps[j].rho <= merge_hi) continue;
                if (len_sq(ps[i].x - ps[j].x, ps[i].y - ps[j].y) < merge_dist_sq) {
                    const float mi = ps[i].m, mj = ps[j].m, mt = mi + mj;
                    if (mt > kEps) {
                        ps[i].vx = (ps[i].vx * mi + ps[j].vx * mj) / mt;
                        ps[i].vy = (ps[i].vy * mi + ps[j].vy * mj) / mt;
                        ps[i].m = mt;
                    }
                    absorbed[j] = 1;
                    break;
                }
            }
        }
        std::vector<P2> out;
        out.reserve(n);
        for (size_t k = 0; k < n; ++k) {
            if (absorbed[k]) continue;
            const P2 p = ps[k];
            if (p.rho < split_lo && p.m > min_mass && out.size() + 1 < cap) {
                P2 a = p, b = p;
                a.m = b.m = p.m * 0.5f;
                const float off = 0.25f * h;
                a.x = p.x + off;
                a.y = p.y + 0.0f;
                b.x = p.x - off;
                b.y = p.y - 0.0f;
                out.push_back(a);
                out.push_back(b);
            } else {
                out.push_back(p);
            }
        }
        ps.swap(out);
    }

    // fluid.rs:898-1136 (no turbulence or dynamic bodies).
    // scale_out: the cancellation-aware magnitude sum_j |m_j term grad W| / rho_i
    // (plus the cohesion terms) that bounds the acceleration's rounding.
    void substep(float gx, float gy, const Oracle2Body *bodies, uint32_t nbodies, float dt, bool jacobi,
                 uint32_t *count, float *acc_out, double *scale_out = nullptr) {
        const float h = rmax(cfg.smoothing_length, kEps);
        rebuild_grid();
        std::vector<uint32_t> nb;
        density_pressure(h, nb, cfg.adaptive_resolution ? nullptr : count);
        // fluid.rs:918-923: resample, rebuild, recompute
        if (cfg.adaptive_resolution) {
            resample(h);
            rebuild_grid();
            density_pressure(h, nb, count);
        }
        // fluid.rs:929-938: the ghosts are a pure function of the static
        // bodies and the config, rebuilt here every substep
        if (cfg.boundary_particles) update_boundaries(bodies, nbodies, h);
        const float c_s = std::sqrt(rmax(cfg.stiffness, 0.0f));
        const size_t n = ps.size();
        fx.assign(n, 0.0f);
        fy.assign(n, 0.0f);
        for (uint32_t i = 0; i < n; ++i) {
            neighbors(i, nb);
            const P2 pi = ps[i];
            const float rho_i = rmax(pi.rho, kEps);
            const float p_i = pi.p;
            float f_x = 0.0f, f_y = 0.0f;
            double mag = 0.0;
            for (uint32_t j : nb) {
                if (j == i) continue;
                const P2 &pj = ps[j];
                const float rho_j = rmax(pj.rho, kEps);
                const float dx = pi.x - pj.x, dy = pi.y - pj.y;
                const float dist = std::sqrt(len_sq(dx, dy));
                float gxx, gyy;
                spiky_grad(dx, dy, dist, h, gxx, gyy);
                float term = p_i / (rho_i * rho_i) + pj.p / (rho_j * rho_j);
                if (cfg.artificial_viscosity > 0.0f && dist > kEps) {
                    const float rho_bar = 0.5f * (rho_i + rho_j);
                    term += artificial_pressure(dx, dy, pi.vx - pj.vx, pi.vy - pj.vy, dist * dist, h, rho_bar, c_s,
                                                cfg.artificial_viscosity);
                }
                const float s = pj.m * term;
                f_x = f_x - gxx * s;
                f_y = f_y - gyy * s;
                mag += std::sqrt((double)gxx * gxx + (double)gyy * gyy) * std::fabs((double)s);
                if (cfg.surface_tension > 0.0f && dist > kEps) {
                    const float coh = cohesion(dist, h);
                    if (coh > 0.0f) {
                        const float c = cfg.surface_tension * pi.m * pj.m * coh;
                        f_x = f_x - dx / dist * c;
                        f_y = f_y - dy / dist * c;
                        mag += std::fabs((double)c);
                    }
                }
            }
            // fluid.rs:985-995: ghost pressure after the fluid neighbours
            if (cfg.boundary_particles) {
                std::vector<uint32_t> gb;
                near_cells(ghost_grid, [&](uint32_t j) { return std::make_pair(ghosts[j].x, ghosts[j].y); }, pi.x,
                           pi.y, gb);
                for (uint32_t k : gb) {
                    const Ghost2 &gh = ghosts[k];
                    const float dx = pi.x - gh.x, dy = pi.y - gh.y;
                    const float dist = std::sqrt(len_sq(dx, dy));
                    float gxx, gyy;
                    spiky_grad(dx, dy, dist, h, gxx, gyy);
                    const float rho_b = rmax(gh.rho, kEps);
                    const float s = gh.m * (p_i / (rho_i * rho_i) + gh.p / (rho_b * rho_b));
                    f_x = f_x - gxx * s;
                    f_y = f_y - gyy * s;
                    mag += std::sqrt((double)gxx * gxx + (double)gyy * gyy) * std::fabs((double)s);
                }
            }
            fx[i] = f_x;
            fy[i] = f_y;
            if (scale_out) scale_out[i] = mag / (double)rho_i;
        }
        // Integrate: in place (reference) or from the pre-step state (Jacobi).
        const std::vector<P2> pre = jacobi ? ps : std::vector<P2>();
        const std::vector<P2> &src = jacobi ? pre : ps;
        for (uint32_t i = 0; i < n; ++i) {
            // neighbour list against the positions the reference reads here
            // (in place: already-moved j < i; Jacobi: the pre-step state)
            neighbors_in(src, i, nb);
            const P2 &me = src[i];
            const float rho_i = rmax(me.rho, kEps);
            const float ax = gx + fx[i] / rho_i, ay = gy + fy[i] / rho_i;
            if (acc_out) {
                acc_out[2 * i] = ax;
                acc_out[2 * i + 1] = ay;
            }
            float vx = me.vx + ax * dt, vy = me.vy + ay * dt;
            if (cfg.viscosity > 0.0f) {
                float sx = 0.0f, sy = 0.0f;
                for (uint32_t j : nb) {
                    if (j == i) continue;
                    const P2 &pj = src[j];
                    const float rho_bar = 0.5f * (rho_i + rmax(pj.rho, kEps));
                    const float q = pj.m / rho_bar;
                    const float w = poly6(len_sq(pj.x - me.x, pj.y - me.y), h);
                    sx = sx + (pj.vx - me.vx) * q * w;
                    sy = sy + (pj.vy - me.vy) * q * w;
                }
                vx = vx + sx * cfg.viscosity;
                vy = vy + sy * cfg.viscosity;
            }
            const float speed = std::sqrt(len_sq(vx, vy));
            if (speed > cfg.max_velocity) {
                const float r = cfg.max_velocity / speed;
                vx = vx * r;
                vy = vy * r;
            }
            if (!finite2(vx, vy)) vx = vy = 0.0f;
            float x = me.x + vx * dt, y = me.y + vy * dt;
            if (!finite2(x, y)) {
                x = me.x;
                y = me.y;
            }
            ps[i].vx = vx;
            ps[i].vy = vy;
            ps[i].x = x;
            ps[i].y = y;
        }
        // Boundaries (fluid.rs:1028-1071): push-out accumulates over bodies.
        const float pr = rmax(cfg.particle_radius, 0.0f);
        for (uint32_t i = 0; i < n; ++i)
            for (uint32_t b = 0; b < nbodies; ++b) {
                float nx, ny, depth;
                if (!collide(ps[i].x, ps[i].y, pr, bodies[b], nx, ny, depth)) continue;
                if (depth <= 0.0f) continue;
                ps[i].x = ps[i].x + nx * depth;
                ps[i].y = ps[i].y + ny * depth;
                const float vn = ps[i].vx * nx + ps[i].vy * ny;
                if (vn < 0.0f) {
                    ps[i].vx = ps[i].vx - nx * vn;
                    ps[i].vy = ps[i].vy - ny * vn;
                    if (cfg.wall_friction > 0.0f) {
                        const float mu = rclamp(cfg.wall_friction, 0.0f, 1.0f);
                        // static wall: velocity 0, so wall_vt = 0 - n * 0
                        const float wvx = 0.0f - nx * 0.0f, wvy = 0.0f - ny * 0.0f;
                        const float dvx = (wvx - ps[i].vx) * mu, dvy = (wvy - ps[i].vy) * mu;
                        ps[i].vx = ps[i].vx + dvx;
                        ps[i].vy = ps[i].vy + dvy;
                    }
                }
            }
    }
};

Fluid2 make(uint32_t n, const float *pos, const float *vel, const float *mass, const uint8_t *phase,
            const Oracle2Cfg *cfg) {
    Fluid2 f;
    f.cfg = *cfg;
    f.ps.resize(n);
    for (uint32_t i = 0; i < n; ++i) {
        P2 &p = f.ps[i];
        p.x = pos[2 * i];
        p.y = pos[2 * i + 1];
        p.vx = vel ? vel[2 * i] : 0.0f;
        p.vy = vel ? vel[2 * i + 1] : 0.0f;
        p.phase = phase ? phase[i] : 0;
        p.m = mass[i];
        p.rho = cfg->phase_rest_density[p.phase < 3 ? p.phase : 3];
        p.p = 0.0f;
    }
    return f;
}

void store(const Fluid2 &f, float *pos, float *vel, float *density, float *pressure) {
    for (size_t i = 0; i < f.ps.size(); ++i) {
        pos[2 * i] = f.ps[i].x;
        pos[2 * i + 1] = f.ps[i].y;
        if (vel) {
            vel[2 * i] = f.ps[i].vx;
            vel[2 * i + 1] = f.ps[i].vy;
        }
        if (density) density[i] = f.ps[i].rho;
        if (pressure) pressure[i] = f.ps[i].p;
    }
}

}  // namespace

extern "C" {

// One substep; pos/vel (2n, updated in place), density/pressure/count/acc
// (2n) and the acceleration's cancellation-aware scale of that substep out
// (each may be null).
int oracle2_substep_f32(uint32_t n, float *pos, float *vel, const float *mass, const uint8_t *phase,
                        const Oracle2Cfg *cfg, float gx, float gy, const Oracle2Body *bodies, uint32_t nbodies,
                        float dt, int jacobi, float *density, float *pressure, uint32_t *count, float *acc,
                        double *scale) {
    if (!n || !pos || !mass || !cfg) return 0;
    Fluid2 f = make(n, pos, vel, mass, phase, cfg);
    f.substep(gx, gy, bodies, nbodies, dt, jacobi != 0, count, acc, scale);
    store(f, pos, vel, density, pressure);
    return 1;
}

// `steps` World2 steps of dt (fluid.rs:537-556: CFL substeps of dt / n).
int oracle2_run_f32(uint32_t n, float *pos, float *vel, const float *mass, const uint8_t *phase,
                    const Oracle2Cfg *cfg, float gx, float gy, const Oracle2Body *bodies, uint32_t nbodies, float dt,
                    uint32_t steps, int jacobi, float *density, float *pressure) {
    if (!n || !pos || !mass || !cfg) return 0;
    Fluid2 f = make(n, pos, vel, mass, phase, cfg);
    for (uint32_t s = 0; s < steps; ++s) {
        if (dt <= 0.0f) break;
        const uint32_t sub = f.cfl_substeps(dt);
        const float sdt = dt / (float)sub;
        for (uint32_t k = 0; k < sub; ++k) f.substep(gx, gy, bodies, nbodies, sdt, jacobi != 0, nullptr, nullptr);
    }
    store(f, pos, vel, density, pressure);
    return 1;
}

// The boundary ghosts of `bodies` for a fluid state (fluid.rs:929-938):
// positions (2 per ghost), density and pressure in generation order, up to
// max.  Returns the ghost count.
uint32_t oracle2_boundary_f32(uint32_t n, const float *pos, const float *mass, const Oracle2Cfg *cfg,
                              const Oracle2Body *bodies, uint32_t nbodies, float *xy, float *density,
                              float *pressure, uint32_t max) {
    Fluid2 f = make(n, pos, nullptr, mass, nullptr, cfg);
    f.rebuild_grid();
    f.update_boundaries(bodies, nbodies, rmax(cfg->smoothing_length, kEps));
    const uint32_t m = std::min<uint32_t>(max, (uint32_t)f.ghosts.size());
    for (uint32_t g = 0; g < m; ++g) {
        if (xy) {
            xy[2 * g] = f.ghosts[g].x;
            xy[2 * g + 1] = f.ghosts[g].y;
        }
        if (density) density[g] = f.ghosts[g].rho;
        if (pressure) pressure[g] = f.ghosts[g].p;
    }
    return (uint32_t)f.ghosts.size();
}

// The variable-count forms (adaptive resolution): *n in and out, every
// array holds `room` particles (pos, vel 2 each), mass and phase updated;
// a result larger than room is not written (0 returned, *n the count).
static void store_all(const Fluid2 &f, float *pos, float *vel, float *mass, uint8_t *phase, float *density) {
    store(f, pos, vel, density, nullptr);
    for (size_t i = 0; i < f.ps.size(); ++i) {
        if (mass) mass[i] = f.ps[i].m;
        if (phase) phase[i] = f.ps[i].phase;
    }
}

// resample() alone (fluid.rs:826-891) on a state whose densities are given.
int oracle2_resample_f32(uint32_t *n, uint32_t cap, uint32_t room, float *pos, float *vel, float *mass, uint8_t *phase,
                         float *density, const Oracle2Cfg *cfg, float h) {
    Fluid2 f = make(*n, pos, vel, mass, phase, cfg);
    f.cap = cap;
    for (uint32_t i = 0; i < *n; ++i) f.ps[i].rho = density[i];
    f.resample(h);
    *n = (uint32_t)f.ps.size();
    if (f.ps.size() > room) return 0;
    store_all(f, pos, vel, mass, phase, density);
    return 1;
}

// `steps` World2 steps (as oracle2_run_f32) with a variable count.
int oracle2_run_cap_f32(uint32_t *n, uint32_t cap, uint32_t room, float *pos, float *vel, float *mass, uint8_t *phase,
                        const Oracle2Cfg *cfg, float gx, float gy, const Oracle2Body *bodies, uint32_t nbodies,
                        float dt, uint32_t steps, int jacobi) {
    Fluid2 f = make(*n, pos, vel, mass, phase, cfg);
    f.cap = cap;
    for (uint32_t s = 0; s < steps; ++s) {
        if (dt <= 0.0f || f.ps.empty()) break;
        const uint32_t sub = f.cfl_substeps(dt);
        const float sdt = dt / (float)sub;
        for (uint32_t k = 0; k < sub; ++k) f.substep(gx, gy, bodies, nbodies, sdt, jacobi != 0, nullptr, nullptr);
    }
    *n = (uint32_t)f.ps.size();
    if (f.ps.size() > room) return 0;
    store_all(f, pos, vel, mass, phase, nullptr);
    return 1;
}

float oracle2_poly6(float r2, float h) { return Fluid2::poly6(r2, h); }
void oracle2_spiky_grad(float rx, float ry, float r, float h, float *out2) {
    Fluid2::spiky_grad(rx, ry, r, h, out2[0], out2[1]);
}
float oracle2_pressure(float density, float rest, float k, float gamma) {
    return Fluid2::pressure(density, rest, k, gamma);
}
// fluid.rs:302-327 (pinned by fluid.rs:1333-1360 and two-particle closed forms).
float oracle2_artificial_pressure(float dx, float dy, float vx, float vy, float r_sq, float h, float rho_bar, float c,
                                  float alpha) {
    return Fluid2::artificial_pressure(dx, dy, vx, vy, r_sq, h, rho_bar, c, alpha);
}
float oracle2_cohesion(float dist, float h) { return Fluid2::cohesion(dist, h); }

}  // extern "C"
