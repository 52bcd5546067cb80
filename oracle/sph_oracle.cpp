// sph_oracle.cpp — CPU restatement of the PhiEng 3D SPH fluid step.
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker for the GPU path:
// only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
// it.  The product library (phi_eng_amd/libphieng_hip.so) never links it.
//
// It follows crates/dynamics/src/fluid3.rs line by line, in f32 with the
// reference's operation order (compiled with -ffp-contract=off), plus an
// FP64 twin of the same algorithm for error reporting.  Citations are
// file:line in the reference tree.
//
// What pins each function (tests/test_oracle.py):
//   poly6 / spiky_grad / pressure  — closed forms, kernel zeros and signs
//       (fluid3.rs:1296-1311, 1418-1455), the bit-level linear-EOS known
//       answer (fluid.rs:1326-1329), exact-lattice kernel sums.
//   cells / sorted order           — saturating-cast definition, brute force,
//       std::stable_sort equivalence.
//   density / neighbour counts     — O(N^2) brute force on small inputs,
//       lattice closed form (interior rho = m * sum W).
//   force / acceleration           — two-particle closed form, momentum
//       antisymmetry sum_i m_i F_i = 0, FP64 twin agreement.
//   artificial_pressure / cohesion / XSPH — the reference's sign tests
//       (fluid3.rs:1426-1453), two-particle closed forms of each term over a
//       whole substep (unequal masses; Jacobi and in-place integrate),
//       sum m_i F_i = 0 (AV) and sum F_i = 0 (cohesion) on a cloud
//       (tests/test_oracle_pair_terms.py).
//   integrate + boundary clamp     — containment, count preserved,
//       run-twice determinism, hydrostatic band (fluid3.rs:1313-1347,
//       1398-1416, 1560-1605), single-particle closed forms.
//   boundary ghosts                — counts and positions against the
//       closed forms of the generation rules (Fibonacci sphere, box face
//       sweep, capsule rings), ghost density against brute force, the
//       reference's contained / persisted tests (fluid3.rs:1490-1555),
//       ghost pressure pushing a particle away from a wall.
//   Parity unpinned: glam 0.29.3 Vec3 association order and libm powf
//   (the Tait power is taken in f64 and rounded to f32 here and on the GPU).
#include <cmath>
#include <cstdint>
#include <cstring>
#include <omp.h>

#include <algorithm>
#include <unordered_map>
#include <vector>

#include "../include/phieng_fluid3.h"

namespace {

// crates/math/src/math.rs:3-4
constexpr float kEps = 1e-6f;
constexpr float kEpsSq = 1e-12f;
constexpr float kPiF = 3.14159265358979323846f;  // std::f32::consts::PI

struct Body {  // static body, identity rotation (the C ABI cannot rotate)
    uint32_t shape;
    float px, py, pz;
    float a, b, c;
};

// Rust `f32::max`: NaN operands lose; ties keep the second operand.
template <class T>
inline T rmax(T a, T b) {
    if (a != a) return b;
    if (b != b) return a;
    return a > b ? a : b;
}
// Rust `f32::clamp` (fluid3.rs:649): NaN passes through.
template <class T>
inline T rclamp(T v, T lo, T hi) {
    if (v < lo) v = lo;
    if (v > hi) v = hi;
    return v;
}
// Rust `f32 as i32`: saturating, NaN -> 0.
inline int32_t sat_i32(double v) {
    if (v != v) return 0;
    if (v >= 2147483648.0) return INT32_MAX;
    if (v <= -2147483649.0) return INT32_MIN;
    return (int32_t)v;
}

template <class T>
struct V3 {
    T x, y, z;
};
template <class T>
inline V3<T> sub(V3<T> a, V3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class T>
inline V3<T> add(V3<T> a, V3<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class T>
inline V3<T> muls(V3<T> a, T s) { return {a.x * s, a.y * s, a.z * s}; }
template <class T>
inline V3<T> divs(V3<T> a, T s) { return {a.x / s, a.y / s, a.z / s}; }
// glam Vec3::dot: (x*x' + y*y') + z*z' (parity unpinned, see header).
template <class T>
inline T dot(V3<T> a, V3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class T>
inline T len(V3<T> a) { return std::sqrt(dot(a, a)); }
template <class T>
inline bool finite3(V3<T> a) { return std::isfinite(a.x) && std::isfinite(a.y) && std::isfinite(a.z); }

struct Key {
    int32_t x, y, z;
    bool operator==(const Key &o) const { return x == o.x && y == o.y && z == o.z; }
};
struct KeyHash {
    size_t operator()(const Key &k) const {
        uint64_t h = (uint64_t)(uint32_t)k.x * 0x9E3779B97F4A7C15ull;
        h ^= (uint64_t)(uint32_t)k.y * 0xC2B2AE3D27D4EB4Full + (h << 6) + (h >> 2);
        h ^= (uint64_t)(uint32_t)k.z * 0x165667B19E3779F9ull + (h << 6) + (h >> 2);
        return (size_t)h;
    }
};

template <class T>
struct Consts;
template <>
struct Consts<float> {
    static float pi() { return kPiF; }
};
template <>
struct Consts<double> {
    static double pi() { return 3.14159265358979323846; }
};

// A boundary ghost (fluid3.rs:716-841): a static particle on a body's
// surface with its own density and pressure from the fluid around it.
template <class T>
struct Ghost {
    V3<T> p;
    T density, pressure, mass;
};

template <class T>
struct Fluid {
    // FluidConfig3 (fluid3.rs:51-135) as the kernels read it.
    const FluidConfig3Ffi &cfg;
    uint32_t n;
    std::vector<V3<T>> pos, vel;
    std::vector<T> density, pressure, mass;
    std::vector<uint8_t> phase;
    std::unordered_map<Key, std::vector<uint32_t>, KeyHash> grid;
    std::vector<V3<T>> forces;
    std::vector<Ghost<T>> ghosts;
    std::unordered_map<Key, std::vector<uint32_t>, KeyHash> ghost_grid;
    // MAX_PARTICLES_3D (fluid3.rs:28), the split cap of resample(); a world
    // whose capacity was lifted passes its capacity instead.
    uint32_t cap = 4096;
    // With dynamic bodies: replaces the static clamp at the end of a
    // substep (the coupled clamp, the second grid rebuild and buoyancy).
    void (*tail)(void *ctx, float dt) = nullptr;
    void *tail_ctx = nullptr;
    // Jacobi substeps only: run the per-particle loops of the density, the
    // forces, the integrate and the static clamp on OpenMP threads (each
    // particle's result is independent of the others', so they equal the
    // one-thread results bit for bit).
    bool par = false;

    explicit Fluid(const FluidConfig3Ffi &c) : cfg(c), n(0) {}

    T h_kernel() const { return rmax<T>((T)cfg.smoothing_length, (T)kEps); }

    // fluid3.rs:220-228
    static T poly6(T r2, T h) {
        if (r2 >= h * h || r2 < (T)0) return (T)0;
        T h2 = h * h;
        T t = h2 - r2;
        return (T)315.0 / ((T)64.0 * Consts<T>::pi() * h2 * h2 * h2 * h2 * h) * t * t * t;
    }
    // fluid3.rs:232-239
    static V3<T> spiky_grad(V3<T> r_vec, T r, T h) {
        if (r >= h || r < (T)kEps) return {0, 0, 0};
        T s = (T)-45.0 / (Consts<T>::pi() * h * h * h * h * h * h) * (h - r) * (h - r);
        return muls(divs(r_vec, r), s);
    }
    // fluid3.rs:244-251 (Tait power in f64, rounded for f32)
    static T eos(T density, T rest, T k, T gamma) {
        if (gamma == (T)1.0) return rmax<T>(k * (density - rest), (T)0);
        T rho0 = rmax<T>(rest, (T)kEps);
        T pw = (T)std::pow((double)(density / rho0), (double)gamma);
        return rmax<T>(k * rho0 * (pw - (T)1.0), (T)0);
    }
    // fluid3.rs:254-269
    static T artificial_pressure(V3<T> delta, V3<T> v_rel, T r_sq, T h, T rho_bar, T c, T alpha) {
        T vd = dot(v_rel, delta);
        if (vd >= (T)0) return (T)0;
        T mu = h * vd / (r_sq + (T)0.01 * h * h);
        return -alpha * c * mu / rmax<T>(rho_bar, (T)kEps);
    }
    // fluid3.rs:272-278
    static T cohesion(T dist, T h) { return dist >= h ? (T)0 : (T)1.0 - dist / h; }

    Key cell_of(V3<T> p) const {
        T h = (T)cfg.smoothing_length;
        return {sat_i32((double)std::floor(p.x / h)), sat_i32((double)std::floor(p.y / h)),
                sat_i32((double)std::floor(p.z / h))};
    }
    // fluid3.rs:282-315 (non-finite positions skipped, index order kept)
    void rebuild_grid() {
        for (auto &kv : grid) kv.second.clear();
        for (uint32_t i = 0; i < n; ++i) {
            if (!finite3(pos[i])) continue;
            grid[cell_of(pos[i])].push_back(i);
        }
    }
    // fluid3.rs:319-340: cells oz -> oy -> ox, ascending index inside a cell,
    // raw smoothing_length for the cut-off, self included.
    void neighbors(uint32_t i, std::vector<uint32_t> &out) const {
        out.clear();
        T h = (T)cfg.smoothing_length;
        V3<T> p = pos[i];
        Key c = cell_of(p);
        for (int oz = -1; oz <= 1; ++oz)
            for (int oy = -1; oy <= 1; ++oy)
                for (int ox = -1; ox <= 1; ++ox) {
                    // i32 wrap cannot be reached from the C ABI (fluid_add
                    // rejects non-finite input); use 64-bit sums to stay defined.
                    int64_t kx = (int64_t)c.x + ox, ky = (int64_t)c.y + oy, kz = (int64_t)c.z + oz;
                    if (kx > INT32_MAX || kx < INT32_MIN || ky > INT32_MAX || ky < INT32_MIN ||
                        kz > INT32_MAX || kz < INT32_MIN)
                        continue;
                    auto it = grid.find(Key{(int32_t)kx, (int32_t)ky, (int32_t)kz});
                    if (it == grid.end()) continue;
                    for (uint32_t j : it->second) {
                        T d2 = dot(sub(pos[j], p), sub(pos[j], p));
                        if (d2 < h * h) out.push_back(j);
                    }
                }
    }
    // fluid3.rs:716-841 for the shapes the C ABI creates (static, identity
    // rotation: transform_point(p) = p + position up to zero signs).  cos,
    // sin and sqrt are the platform libm's in f32, as Rust's f32 methods
    // (parity unpinned across libms).
    void refresh_boundaries(const Body *bodies, uint32_t nb, float spacing_in) {
        ghosts.clear();
        const float rest = cfg.rest_density;
        const float gm = rmax<float>(cfg.particle_mass, kEps);
        const float spacing = rmax<float>(spacing_in, kEps);
        const float s2 = spacing * spacing;  // spacing.powi(2)
        auto push = [&](const Body &b, float lx, float ly, float lz) {
            ghosts.push_back({{(T)(lx + b.px), (T)(ly + b.py), (T)(lz + b.pz)}, (T)rest, (T)0, (T)gm});
        };
        auto count = [](float v, float cap) {  // v.ceil().clamp(1.0, cap) as usize
            float c = std::ceil(v);
            if (c < 1.0f) c = 1.0f;
            if (c > cap) c = cap;
            return (uint32_t)c;
        };
        for (uint32_t bi = 0; bi < nb; ++bi) {
            const Body &b = bodies[bi];
            if (b.shape == PHIENG_SHAPE3_SPHERE) {  // Fibonacci sphere, :725-748
                const float radius = b.a;
                const uint32_t m = count((4.0f * kPiF * radius * radius) / s2, 512.0f);
                const float golden = kPiF * (3.0f - std::sqrt(5.0f));
                for (uint32_t k = 0; k < m; ++k) {
                    const float z = 1.0f - 2.0f * ((float)k + 0.5f) / (float)m;
                    const float r = std::sqrt(1.0f - z * z);
                    const float theta = golden * (float)k;
                    push(b, std::cos(theta) * r * radius, z * radius, std::sin(theta) * r * radius);
                }
            } else if (b.shape == PHIENG_SHAPE3_BOX) {  // :749-791 (the reference's face layout kept as is)
                const float he[3] = {b.a, b.b, b.c};
                const int axes[6] = {0, 0, 1, 1, 2, 2};
                const float fixed[6] = {-he[1], -he[1], -he[0], -he[0], -he[2], -he[2]};
                for (int f = 0; f < 6; ++f) {
                    const int axis = axes[f];
                    const float area = axis == 0 ? 4.0f * he[1] * he[2] : (axis == 1 ? 4.0f * he[0] * he[2]
                                                                                     : 4.0f * he[0] * he[1]);
                    const uint32_t m = count(area / s2, 512.0f);
                    for (uint32_t k = 0; k < m; ++k) {
                        const float t = ((float)k + 0.5f) / (float)m;
                        const float u = 2.0f * t - 1.0f, w = 2.0f * std::fmod(t * 7.0f, 1.0f) - 1.0f;
                        float l[3] = {0.0f, 0.0f, fixed[f]};
                        if (axis == 0) {
                            l[1] = u * he[1];
                            l[2] = w * he[2];
                        } else if (axis == 1) {
                            l[0] = u * he[0];
                            l[2] = w * he[2];
                        } else {
                            l[0] = u * he[0];
                            l[1] = w * he[1];
                        }
                        push(b, l[0], l[1], l[2]);
                    }
                }
            } else if (b.shape == PHIENG_SHAPE3_CAPSULE) {  // :792-826
                const float r = b.a, hl = b.b;
                const float side = 4.0f * kPiF * r * hl;
                const uint32_t m = count(side / s2, 256.0f);
                const float sys[2] = {-hl * 0.5f, hl * 0.5f};
                for (float sy : sys)
                    for (uint32_t k = 0; k < m; ++k) {
                        const float a = 2.0f * kPiF * (float)k / (float)m;
                        push(b, std::cos(a) * r, sy, std::sin(a) * r);
                    }
                for (float sg : {-1.0f, 1.0f}) push(b, 0.0f * sg * (hl + r), sg * (hl + r), 0.0f * sg * (hl + r));
            }
        }
    }
    // fluid3.rs:843-857: ghost cells at the kernel h, non-finite x skipped.
    void rebuild_boundary_grid(T h) {
        ghost_grid.clear();
        for (uint32_t g = 0; g < ghosts.size(); ++g) {
            const V3<T> p = ghosts[g].p;
            if (!std::isfinite(p.x)) continue;
            ghost_grid[Key{sat_i32((double)std::floor(p.x / h)), sat_i32((double)std::floor(p.y / h)),
                           sat_i32((double)std::floor(p.z / h))}]
                .push_back(g);
        }
    }
    // fluid3.rs:860-880 (boundary_near) and :883-903 (fluid_near): the 27
    // cells around `point` at the raw smoothing length, oz -> oy -> ox,
    // members in index order, d2 < h * h.
    template <class Pos>
    void near_in(const std::unordered_map<Key, std::vector<uint32_t>, KeyHash> &g, const Pos &at, V3<T> point,
                 std::vector<uint32_t> &out) const {
        out.clear();
        const T h = (T)cfg.smoothing_length;
        const Key c = cell_of(point);
        for (int oz = -1; oz <= 1; ++oz)
            for (int oy = -1; oy <= 1; ++oy)
                for (int ox = -1; ox <= 1; ++ox) {
                    const int64_t kx = (int64_t)c.x + ox, ky = (int64_t)c.y + oy, kz = (int64_t)c.z + oz;
                    if (kx > INT32_MAX || kx < INT32_MIN || ky > INT32_MAX || ky < INT32_MIN || kz > INT32_MAX ||
                        kz < INT32_MIN)
                        continue;
                    auto it = g.find(Key{(int32_t)kx, (int32_t)ky, (int32_t)kz});
                    if (it == g.end()) continue;
                    for (uint32_t j : it->second) {
                        const V3<T> d = sub(at(j), point);
                        if (dot(d, d) < h * h) out.push_back(j);
                    }
                }
    }
    // fluid3.rs:906-920: the ghost's density from the fluid (clamped) and
    // its pressure from the unclamped sum with the base rest density.
    void compute_boundary_density_pressure(T h) {
        std::vector<uint32_t> nb;
        for (Ghost<T> &gh : ghosts) {
            near_in(grid, [&](uint32_t j) { return pos[j]; }, gh.p, nb);
            T rho = 0;
            for (uint32_t j : nb) rho += mass[j] * poly6(dot(sub(pos[j], gh.p), sub(pos[j], gh.p)), h);
            gh.density = rclamp<T>(rho, (T)0, (T)3.0 * rmax<T>((T)cfg.rest_density, (T)kEps));
            gh.pressure = eos(rho, (T)cfg.rest_density, (T)cfg.stiffness, (T)cfg.gamma);
        }
    }
    // fluid3.rs:1018-1027: the ghosts are a pure function of the static
    // bodies and the config (the fingerprint only caches them), so they are
    // rebuilt here on every substep.
    void update_boundaries(const Body *bodies, uint32_t nb, T h) {
        refresh_boundaries(bodies, nb, (float)h * 0.5f);
        rebuild_boundary_grid(h);
        compute_boundary_density_pressure(h);
    }

    // fluid3.rs:639-655
    void compute_density_pressure(T h, uint32_t *out_count) {
#pragma omp parallel if (par)
        {
        std::vector<uint32_t> nb;
#pragma omp for schedule(static, 256)
        for (uint32_t i = 0; i < n; ++i) {
            neighbors(i, nb);
            V3<T> pi = pos[i];
            T rho = 0;
            for (uint32_t j : nb) {
                T r2 = dot(sub(pos[j], pi), sub(pos[j], pi));
                rho += mass[j] * poly6(r2, h);
            }
            if (out_count) out_count[i] = (uint32_t)nb.size();
            density[i] = rclamp<T>(rho, (T)0, (T)3.0 * rmax<T>((T)cfg.rest_density, (T)kEps));
        }
        }
#pragma omp parallel for if (par) schedule(static, 1024)
        for (uint32_t i = 0; i < n; ++i) {
            T rho0 = (T)cfg.phase_rest_density[std::min<uint32_t>(phase[i], 3u)];
            pressure[i] = eos(density[i], rho0, (T)cfg.stiffness, (T)cfg.gamma);
        }
    }
    // fluid3.rs:1029-1086 (ghost term omitted: boundary_particles rejected).
    // Also returns the cancellation-aware scale of the acceleration.
    void compute_forces(T h, double *scale_out) {
        forces.assign(n, V3<T>{0, 0, 0});
        T c_s = std::sqrt(rmax<T>((T)cfg.stiffness, (T)0));
#pragma omp parallel if (par)
        {
        std::vector<uint32_t> nb;
#pragma omp for schedule(static, 256)
        for (uint32_t i = 0; i < n; ++i) {
            neighbors(i, nb);
            V3<T> pi = pos[i], vi = vel[i];
            T rho_i = rmax<T>(density[i], (T)kEps);
            T p_i = pressure[i];
            V3<T> f{0, 0, 0};
            double mag = 0;
            for (uint32_t j : nb) {
                if (j == i) continue;
                T rho_j = rmax<T>(density[j], (T)kEps);
                V3<T> delta = sub(pi, pos[j]);
                T dist = len(delta);
                V3<T> grad = spiky_grad(delta, dist, h);
                T term = p_i / (rho_i * rho_i) + pressure[j] / (rho_j * rho_j);
                if ((T)cfg.artificial_viscosity > (T)0 && dist > (T)kEps) {
                    T rho_bar = (T)0.5 * (rho_i + rho_j);
                    term += artificial_pressure(delta, sub(vi, vel[j]), dist * dist, h, rho_bar, c_s,
                                                (T)cfg.artificial_viscosity);
                }
                T s = mass[j] * term;
                f = sub(f, muls(grad, s));
                mag += std::sqrt((double)dot(grad, grad)) * std::fabs((double)s);
                if ((T)cfg.surface_tension > (T)0 && dist > (T)kEps) {
                    T coh = cohesion(dist, h);
                    if (coh > (T)0) {
                        T cs = (T)cfg.surface_tension * mass[i] * mass[j] * coh;
                        f = sub(f, muls(divs(delta, dist), cs));
                        mag += std::fabs((double)cs);
                    }
                }
            }
            // fluid3.rs:1074-1084: ghost pressure after the fluid neighbours
            if (cfg.boundary_particles) {
                std::vector<uint32_t> gb;
                near_in(ghost_grid, [&](uint32_t j) { return ghosts[j].p; }, pi, gb);
                for (uint32_t k : gb) {
                    const Ghost<T> &gh = ghosts[k];
                    V3<T> delta = sub(pi, gh.p);
                    T dist = len(delta);
                    V3<T> grad = spiky_grad(delta, dist, h);
                    T rho_b = rmax<T>(gh.density, (T)kEps);
                    T s = gh.mass * (p_i / (rho_i * rho_i) + gh.pressure / (rho_b * rho_b));
                    f = sub(f, muls(grad, s));
                    mag += std::sqrt((double)dot(grad, grad)) * std::fabs((double)s);
                }
            }
            forces[i] = f;
            if (scale_out) scale_out[i] = mag / (double)rho_i;
        }
        }
    }

    // fluid3.rs:471-585 for the shapes the C ABI creates (identity rotation,
    // so glam's quaternion products reduce to the identity up to zero signs).
    bool collide(V3<T> p, T pr, const Body &b, V3<T> &normal, T &depth) const {
        V3<T> c{(T)b.px, (T)b.py, (T)b.pz};
        if (b.shape == PHIENG_SHAPE3_SPHERE) {
            T radius = (T)b.a;
            V3<T> delta = sub(p, c);
            T dist = len(delta);
            if (dist >= radius + pr) return false;
            normal = dist > (T)kEps ? divs(delta, dist) : V3<T>{1, 0, 0};
            depth = radius + pr - dist;
            return true;
        }
        if (b.shape == PHIENG_SHAPE3_BOX) {
            V3<T> he{(T)b.a, (T)b.b, (T)b.c};
            V3<T> local = sub(p, c);
            V3<T> cl{rclamp(local.x, -he.x, he.x), rclamp(local.y, -he.y, he.y),
                     rclamp(local.z, -he.z, he.z)};
            V3<T> diff = sub(local, cl);
            if (dot(diff, diff) > (T)kEpsSq) {
                T dist = len(diff);
                if (dist >= pr) return false;
                normal = divs(diff, dist);
                depth = pr - dist;
                return true;
            }
            // Centre inside: eject along the least-penetration axis; the
            // reference's `signum().copysign(1.0)` is always +1 (fluid3.rs:514-520).
            T dx = he.x - std::fabs(local.x);
            T dy = he.y - std::fabs(local.y);
            T dz = he.z - std::fabs(local.z);
            if (dx <= dy && dx <= dz) {
                normal = {1, 0, 0};
                depth = pr + dx;
            } else if (dy <= dz) {
                normal = {0, 1, 0};
                depth = pr + dy;
            } else {
                normal = {0, 0, 1};
                depth = pr + dz;
            }
            return true;
        }
        if (b.shape == PHIENG_SHAPE3_CAPSULE) {
            T r = (T)b.a, hl = (T)b.b;
            V3<T> local = sub(p, c);
            T sy = rclamp(local.y, -hl, hl);
            V3<T> diff{local.x - (T)0, local.y - sy, local.z - (T)0};
            if (dot(diff, diff) > (T)kEpsSq) {
                T dist = len(diff);
                if (dist >= r + pr) return false;
                normal = divs(diff, dist);
                depth = r + pr - dist;
                return true;
            }
            V3<T> radial{local.x, 0, local.z};
            if (dot(radial, radial) > (T)kEpsSq) {
                // glam normalize_or_zero: v * (1 / length)
                T rcp = (T)1.0 / len(radial);
                normal = (std::isfinite(rcp) && rcp > (T)0) ? muls(radial, rcp) : V3<T>{0, 0, 0};
            } else {
                normal = {1, 0, 0};
            }
            depth = pr + r - len(diff);
            return true;
        }
        return false;
    }

    // fluid3.rs:1141-1179 for one particle (static bodies: impulses are
    // no-ops, body.rs:1349-1352; static body velocity is zero).
    void clamp_particle(uint32_t i, const Body *bodies, uint32_t nb) {
        T pr = rmax<T>((T)cfg.particle_radius, (T)0);
        for (uint32_t k = 0; k < nb; ++k) {
            V3<T> ppos = pos[i];
            V3<T> normal;
            T depth;
            if (!collide(ppos, pr, bodies[k], normal, depth)) continue;
            if (depth <= (T)0) continue;
            pos[i] = add(ppos, muls(normal, depth));
            T vn = dot(vel[i], normal);
            if (vn < (T)0) {
                vel[i] = sub(vel[i], muls(normal, vn));
                if ((T)cfg.wall_friction > (T)0) {
                    T mu = rclamp<T>((T)cfg.wall_friction, (T)0, (T)1);
                    V3<T> v_t = vel[i];
                    V3<T> wall_v{0, 0, 0};
                    V3<T> wall_vt = sub(wall_v, muls(normal, dot(wall_v, normal)));
                    V3<T> delta_v = muls(sub(wall_vt, v_t), mu);
                    vel[i] = add(vel[i], delta_v);
                }
            }
        }
    }

    // One integrate update of particle i from the given neighbour list,
    // reading neighbours from (rpos, rvel, density) (fluid3.rs:1092-1139).
    void integrate_one(uint32_t i, const std::vector<uint32_t> &nb, const std::vector<V3<T>> &rpos,
                       const std::vector<V3<T>> &rvel, V3<T> g, T dt, T h, V3<T> &out_v, V3<T> &out_x,
                       V3<T> *acc_out) const {
        T rho_i = rmax<T>(density[i], (T)kEps);
        V3<T> acc = add(g, divs(forces[i], rho_i));
        if (acc_out) *acc_out = acc;
        V3<T> vi = rvel[i];
        V3<T> v = add(vi, muls(acc, dt));
        if ((T)cfg.viscosity > (T)0) {
            V3<T> pi = rpos[i];
            V3<T> smooth{0, 0, 0};
            for (uint32_t j : nb) {
                if (j == i) continue;
                T rho_bar = (T)0.5 * (rho_i + rmax<T>(density[j], (T)kEps));
                T w = poly6(dot(sub(rpos[j], pi), sub(rpos[j], pi)), h);
                smooth = add(smooth, muls(muls(sub(rvel[j], vi), mass[j] / rho_bar), w));
            }
            v = add(v, muls(smooth, (T)cfg.viscosity));
        }
        T speed = len(v);
        if (speed > (T)cfg.max_velocity) v = muls(v, (T)cfg.max_velocity / speed);
        if (!finite3(v)) v = {0, 0, 0};
        V3<T> x = add(rpos[i], muls(v, dt));
        if (!finite3(x)) x = rpos[i];
        out_v = v;
        out_x = x;
    }

    // fluid3.rs:923-987, adaptive resolution.  Merge: in index order, a
    // particle denser than merge_hi absorbs the highest-index earlier dense,
    // not yet absorbed particle within h/4 (mass summed, velocity
    // mass-weighted), and the absorbed ones are removed in order.  Split: in
    // the new order, a particle thinner than split_lo and heavier than
    // min_mass becomes two halves at -+ h/4 along x while the output stays
    // under the cap.  `density` is this substep's (clamped) density.
    void resample(T h) {
        const T rest = rmax<T>((T)cfg.rest_density, (T)kEps);
        const T split_lo = (T)cfg.split_density_ratio * rest;
        const T merge_hi = (T)cfg.merge_density_ratio * rest;
        const T min_mass = rmax<T>((T)cfg.particle_mass, (T)kEps) * (T)0.125;
        std::vector<uint8_t> absorbed(n, 0);
        const T merge_dist_sq = ((T)0.25 * h) * ((T)0.25 * h);
        bool any = false;
        for (uint32_t i = 0; i < n; ++i) {
            if (absorbed[i] || density[i] <= merge_hi) continue;
            for (uint32_t j = i; j-- > 0;) {
                if (absorbed[j] || density[j] <= merge_hi) continue;
                const V3<T> d = sub(pos[i], pos[j]);
                if (dot(d, d) < merge_dist_sq) {
                    const T mi = mass[i], mj = mass[j], mt = mi + mj;
                    if (mt > (T)kEps) {
                        vel[i] = divs(add(muls(vel[i], mi), muls(vel[j], mj)), mt);
                        mass[i] = mt;
                    }
                    absorbed[j] = 1;
                    any = true;
                    break;
                }
            }
        }
        std::vector<V3<T>> np_, nv;
        std::vector<T> nrho, npr, nm;
        std::vector<uint8_t> nph;
        auto push = [&](uint32_t k, V3<T> x, T m) {
            np_.push_back(x);
            nv.push_back(vel[k]);
            nrho.push_back(density[k]);
            npr.push_back(pressure[k]);
            nm.push_back(m);
            nph.push_back(phase[k]);
        };
        const V3<T> off{(T)0.25 * h, (T)0, (T)0};
        for (uint32_t k = 0; k < n; ++k) {
            if (any && absorbed[k]) continue;
            if (density[k] < split_lo && mass[k] > min_mass && np_.size() + 1 < cap) {
                const T half = mass[k] * (T)0.5;
                push(k, add(pos[k], off), half);
                push(k, sub(pos[k], off), half);
            } else {
                push(k, pos[k], mass[k]);
            }
        }
        pos.swap(np_);
        vel.swap(nv);
        density.swap(nrho);
        pressure.swap(npr);
        mass.swap(nm);
        phase.swap(nph);
        n = (uint32_t)pos.size();
    }

    // fluid3.rs:990-1184.  jacobi = false replays the reference's in-place
    // Gauss-Seidel integrate (neighbours j < i already advanced, stale grid);
    // jacobi = true reads only pre-step state (the GPU contract).
    void substep(V3<T> g, T dt, bool jacobi, const Body *bodies, uint32_t nbodies, T *acc_out,
                 uint32_t *count_out, double *scale_out) {
        T h = h_kernel();
        rebuild_grid();
        compute_density_pressure(h, cfg.adaptive_resolution ? nullptr : count_out);
        // fluid3.rs:1007-1012: resample, rebuild, recompute
        if (cfg.adaptive_resolution) {
            resample(h);
            rebuild_grid();
            compute_density_pressure(h, count_out);
        }
        if (cfg.boundary_particles) update_boundaries(bodies, nbodies, h);
        compute_forces(h, scale_out);
        std::vector<uint32_t> nb;
        if (jacobi) {
            std::vector<V3<T>> npos(n), nvel(n);
#pragma omp parallel if (par)
            {
            std::vector<uint32_t> lnb;
#pragma omp for schedule(static, 256)
            for (uint32_t i = 0; i < n; ++i) {
                neighbors(i, lnb);
                V3<T> a;
                integrate_one(i, lnb, pos, vel, g, dt, h, nvel[i], npos[i], &a);
                if (acc_out) {
                    acc_out[3 * i] = a.x;
                    acc_out[3 * i + 1] = a.y;
                    acc_out[3 * i + 2] = a.z;
                }
            }
            }
            pos.swap(npos);
            vel.swap(nvel);
        } else {
            for (uint32_t i = 0; i < n; ++i) {
                neighbors(i, nb);
                V3<T> a, v, x;
                integrate_one(i, nb, pos, vel, g, dt, h, v, x, &a);
                vel[i] = v;
                pos[i] = x;
                if (acc_out) {
                    acc_out[3 * i] = a.x;
                    acc_out[3 * i + 1] = a.y;
                    acc_out[3 * i + 2] = a.z;
                }
            }
        }
        if (tail) {
            tail(tail_ctx, (float)dt);
            return;
        }
#pragma omp parallel for if (par && jacobi) schedule(static, 1024)
        for (uint32_t i = 0; i < n; ++i) clamp_particle(i, bodies, nbodies);
        // fluid3.rs:1184-1213: the second grid rebuild only feeds buoyancy of
        // dynamic bodies, which this path does not have.
    }

    // fluid3.rs:610-636
    uint32_t cfl_substeps(T dt) const {
        uint32_t mx = std::max<uint32_t>(cfg.max_substeps, 1u);
        if ((T)cfg.cfl <= (T)0 || mx <= 1) return 1;
        T h = rmax<T>((T)cfg.smoothing_length, (T)kEps);
        T rho0 = rmax<T>((T)cfg.rest_density, (T)kEps);
        T c0 = (T)cfg.gamma == (T)1.0 ? std::sqrt(rmax<T>((T)cfg.stiffness, (T)0))
                                      : std::sqrt((T)cfg.gamma * rmax<T>((T)cfg.stiffness, (T)0) * rho0);
        T vmax = 0;
        for (uint32_t i = 0; i < n; ++i) vmax = rmax<T>(vmax, len(vel[i]));
        T speed = c0 + vmax;
        if (!std::isfinite(speed) || speed <= (T)kEps) return 1;
        T dt_max = (T)cfg.cfl * h / speed;
        if (!std::isfinite(dt_max) || dt_max <= (T)0) return 1;
        double q = std::ceil((double)(dt / dt_max));
        int64_t s;
        if (q != q) s = 0;
        else if (q >= 9.2233720368547758e18) s = INT64_MAX;
        else if (q <= -9.2233720368547758e18) s = INT64_MIN;
        else s = (int64_t)q;
        return (uint32_t)std::min<int64_t>(std::max<int64_t>(s, 1), (int64_t)mx);
    }

    void load(uint32_t count, const float *p, const float *v, const float *m, const uint8_t *ph,
              const float *rho) {
        n = count;
        pos.resize(n);
        vel.resize(n);
        density.resize(n);
        pressure.assign(n, 0);
        mass.resize(n);
        phase.resize(n);
        for (uint32_t i = 0; i < n; ++i) {
            pos[i] = {(T)p[3 * i], (T)p[3 * i + 1], (T)p[3 * i + 2]};
            vel[i] = v ? V3<T>{(T)v[3 * i], (T)v[3 * i + 1], (T)v[3 * i + 2]} : V3<T>{0, 0, 0};
            mass[i] = (T)m[i];
            phase[i] = ph ? ph[i] : 0;
            density[i] = rho ? (T)rho[i] : (T)cfg.phase_rest_density[std::min<uint32_t>(phase[i], 3u)];
        }
    }
};

// ---------------------------------------------------------------- dynamic bodies
// The rigid side of the two-way coupling (fluid3.rs:1141-1213), restated
// from the reference with the f32 operations of its math library, glam
// 0.29.3 (not vendored; Vec3 and Mat3 in their scalar forms, Quat in its
// SSE2 form, association orders kept):
//   Quat * Vec3   v (w w - b.b) + b (2 (b.v)) + (b x v)(2 w),  b = (x, y, z);
//   Quat * Quat   per lane (w-term + x-term) + (y-term + z-term);
//   normalize     q / sqrt((xx + zz) + (yy + ww));
//   Mat3          columns; M v = (c0 vx + c1 vy) + c2 vz; inverse from the
//                 cross products of the columns and det = c2 . (c0 x c1).
// A body whose quaternion is exactly (0, 0, 0, 1) skips the rotation (the
// product reproduces its operand up to the sign of zero components).

struct Quat {
    float x, y, z, w;
};
struct Mat3 {
    V3<float> c0, c1, c2;
};

inline V3<float> crossf(V3<float> a, V3<float> b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline V3<float> quat_apply(Quat q, V3<float> v) {
    V3<float> b{q.x, q.y, q.z};
    float b2 = dot(b, b);
    return add(add(muls(v, q.w * q.w - b2), muls(b, dot(b, v) * 2.0f)), muls(crossf(b, v), q.w * 2.0f));
}
inline Quat quat_mul(Quat a, Quat b) {
    return {(a.w * b.x + a.x * b.w) + (a.y * b.z - a.z * b.y), (a.w * b.y - a.x * b.z) + (a.y * b.w + a.z * b.x),
            (a.w * b.z + a.x * b.y) + (a.z * b.w - a.y * b.x), (a.w * b.w - a.x * b.x) - (a.y * b.y + a.z * b.z)};
}
inline Quat quat_normalize(Quat q) {
    float l = std::sqrt((q.x * q.x + q.z * q.z) + (q.y * q.y + q.w * q.w));
    return {q.x / l, q.y / l, q.z / l, q.w / l};
}
inline V3<float> mat_apply(const Mat3 &m, V3<float> v) {
    return add(add(muls(m.c0, v.x), muls(m.c1, v.y)), muls(m.c2, v.z));
}
inline Mat3 mat_mul(const Mat3 &a, const Mat3 &b) { return {mat_apply(a, b.c0), mat_apply(a, b.c1), mat_apply(a, b.c2)}; }
inline Mat3 mat_t(const Mat3 &m) {
    return {{m.c0.x, m.c1.x, m.c2.x}, {m.c0.y, m.c1.y, m.c2.y}, {m.c0.z, m.c1.z, m.c2.z}};
}
inline Mat3 mat_inverse(const Mat3 &m) {
    V3<float> t0 = crossf(m.c1, m.c2), t1 = crossf(m.c2, m.c0), t2 = crossf(m.c0, m.c1);
    float inv = 1.0f / dot(m.c2, t2);
    return mat_t({muls(t0, inv), muls(t1, inv), muls(t2, inv)});
}
inline Mat3 mat_from_quat(Quat q) {
    float x2 = q.x + q.x, y2 = q.y + q.y, z2 = q.z + q.z;
    float xx = q.x * x2, xy = q.x * y2, xz = q.x * z2, yy = q.y * y2, yz = q.y * z2, zz = q.z * z2;
    float wx = q.w * x2, wy = q.w * y2, wz = q.w * z2;
    return {{1.0f - (yy + zz), xy + wz, xz - wy}, {xy - wz, 1.0f - (xx + zz), yz + wx}, {xz + wy, yz - wx, 1.0f - (xx + yy)}};
}

// crates/math/src/trig.rs:129-164: reduction by pi/2 (two-part constant,
// f64), sin and cos Taylor polynomials to x^11 / x^12 on [-pi/4, pi/4],
// quadrant from the integer n mod 4.
inline void trig_cos_sin(float angle, float &c, float &s) {
    if (!std::isfinite(angle)) {
        c = s = NAN;
        return;
    }
    const double a = angle;
    const double n = std::round(a * (2.0 / 3.14159265358979323846));
    const double r = (a - n * 1.570796012878418) - n * 3.13916478589249e-07;
    const double z = r * r;
    const double sp = -1.0 / 6.0 + z * (1.0 / 120.0 + z * (-1.0 / 5040.0 + z * (1.0 / 362880.0 + z * (-1.0 / 39916800.0))));
    const double sv = r * (1.0 + z * sp);
    const double cq =
        -0.5 + z * (1.0 / 24.0 + z * (-1.0 / 720.0 + z * (1.0 / 40320.0 + z * (-1.0 / 3628800.0 + z * (1.0 / 479001600.0)))));
    const double cv = 1.0 + z * cq;
    double qd = std::fmod(n, 4.0);
    if (qd < 0.0) qd += 4.0;
    const int quadrant = (int)qd;
    switch (quadrant) {
        case 1: c = (float)-sv; s = (float)cv; break;
        case 2: c = (float)-cv; s = (float)-sv; break;
        case 3: c = (float)sv; s = (float)-cv; break;
        default: c = (float)cv; s = (float)sv; break;
    }
}

struct RigidO {
    uint32_t shape = 1;
    float a = 0, b = 0, c = 0;
    bool is_static = true, sleeping = false;
    V3<float> pos{0, 0, 0};
    Quat rot{0, 0, 0, 1};
    V3<float> vel{0, 0, 0}, ang{0, 0, 0};
    float mass = 1, inv_mass = 1;
    Mat3 inv_inertia{{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    float linear_damping = 0.01f, angular_damping = 0.01f, gravity_scale = 1.0f;

    bool identity() const { return rot.x == 0.0f && rot.y == 0.0f && rot.z == 0.0f && rot.w == 1.0f; }
    V3<float> rotate(V3<float> v) const { return identity() ? v : quat_apply(rot, v); }
    V3<float> unrotate(V3<float> v) const { return identity() ? v : quat_apply({-rot.x, -rot.y, -rot.z, rot.w}, v); }
    V3<float> to_world(V3<float> l) const { return add(rotate(l), pos); }

    // body.rs:1266-1300, mass_properties_3d (body.rs:811-841),
    // InertiaTensor3::capsule (math.rs:148-160)
    void shape_mass(float density) {
        const float PI = kPiF, k43 = 4.0f / 3.0f;
        float m, ix, iy, iz;
        if (shape == 0) {
            m = k43 * PI * a * a * a * density;
            ix = iy = iz = 0.4f * m * a * a;
        } else if (shape == 1) {
            float w = a * 2.0f, h = b * 2.0f, d = c * 2.0f;
            m = w * h * d * density;
            ix = m * (h * h + d * d) / 12.0f;
            iy = m * (w * w + d * d) / 12.0f;
            iz = m * (w * w + h * h) / 12.0f;
        } else {
            float r = a, hl = b, len2 = hl * 2.0f;
            m = (PI * r * r * len2 + k43 * PI * r * r * r) * density;
            float vc = PI * r * r * len2, vs = k43 * PI * r * r * r;
            float den = m / rmax<float>(vc + vs, 1e-12f);
            float mc = den * vc, ms = den * vs;
            ix = iz = mc * (3.0f * r * r + len2 * len2) / 12.0f + ms * (0.4f * r * r + hl * hl);
            iy = 0.5f * mc * r * r + 0.4f * ms * r * r;
        }
        mass = rmax<float>(m, kEps);
        if (is_static) {
            inv_mass = 0.0f;
            inv_inertia = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
        } else {
            inv_mass = 1.0f / mass;
            inv_inertia = mat_inverse({{ix, 0, 0}, {0, iy, 0}, {0, 0, iz}});
        }
    }
    Mat3 world_inv_inertia() const {
        Mat3 r = mat_from_quat(rot);
        return mat_mul(mat_mul(r, inv_inertia), mat_t(r));
    }
    void apply_impulse(V3<float> j, V3<float> point) {  // body.rs:1349-1361
        if (is_static) return;
        sleeping = false;
        vel = add(vel, muls(j, inv_mass));
        ang = add(ang, mat_apply(world_inv_inertia(), crossf(sub(point, pos), j)));
    }
    void apply_force(V3<float> f, V3<float> point, float dt) {  // body.rs:1334-1346
        if (is_static) return;
        sleeping = false;
        vel = add(vel, muls(muls(f, inv_mass), dt));
        ang = add(ang, muls(mat_apply(world_inv_inertia(), crossf(sub(point, pos), f)), dt));
    }
    void integrate(float dt) {  // body.rs:1390-1406, 1372-1387
        if (is_static || sleeping) return;
        vel = muls(vel, 1.0f - std::fmin(linear_damping * dt, 1.0f));
        ang = muls(ang, 1.0f - std::fmin(angular_damping * dt, 1.0f));
        pos = add(pos, muls(vel, dt));
        float angle = len(ang) * dt;
        if (angle > 1e-6f) {
            V3<float> axis = divs(ang, len(ang));
            float ch, sh;
            trig_cos_sin(angle * 0.5f, ch, sh);
            rot = quat_normalize(quat_mul({axis.x * sh, axis.y * sh, axis.z * sh, ch}, rot));
        }
    }
    Body as_body() const { return {shape, pos.x, pos.y, pos.z, a, shape == 0 ? 0.0f : b, shape == 1 ? c : 0.0f}; }
};

// fluid3.rs:471-562 with the body's rotation; also the contact point.
inline bool collide_rigid(V3<float> p, float pr, const RigidO &bd, V3<float> &normal, float &depth,
                          V3<float> &point) {
    if (bd.shape == PHIENG_SHAPE3_SPHERE) {
        V3<float> delta = sub(p, bd.pos);
        float dist = len(delta);
        if (dist >= bd.a + pr) return false;
        normal = dist > kEps ? divs(delta, dist) : V3<float>{1, 0, 0};
        depth = bd.a + pr - dist;
        point = add(bd.pos, muls(normal, bd.a));
        return true;
    }
    V3<float> local = bd.unrotate(sub(p, bd.pos));
    if (bd.shape == PHIENG_SHAPE3_BOX) {
        V3<float> cl{rclamp(local.x, -bd.a, bd.a), rclamp(local.y, -bd.b, bd.b), rclamp(local.z, -bd.c, bd.c)};
        V3<float> diff = sub(local, cl);
        if (dot(diff, diff) > kEpsSq) {
            float dist = len(diff);
            if (dist >= pr) return false;
            normal = bd.rotate(divs(diff, dist));
            depth = pr - dist;
            point = bd.to_world(cl);
            return true;
        }
        float dx = bd.a - std::fabs(local.x), dy = bd.b - std::fabs(local.y), dz = bd.c - std::fabs(local.z);
        V3<float> axis;
        float d;
        if (dx <= dy && dx <= dz) {
            axis = {1, 0, 0};
            d = dx;
        } else if (dy <= dz) {
            axis = {0, 1, 0};
            d = dy;
        } else {
            axis = {0, 0, 1};
            d = dz;
        }
        normal = bd.rotate(axis);
        depth = pr + d;
        point = bd.to_world(cl);
        return true;
    }
    if (bd.shape == PHIENG_SHAPE3_CAPSULE) {
        float r = bd.a, hl = bd.b;
        float sy = rclamp(local.y, -hl, hl);
        V3<float> seg{0, sy, 0};
        V3<float> diff = sub(local, seg);
        if (dot(diff, diff) > kEpsSq) {
            float dist = len(diff);
            if (dist >= r + pr) return false;
            normal = bd.rotate(divs(diff, dist));
            depth = r + pr - dist;
            float rcp = 1.0f / len(diff);
            V3<float> u = (std::isfinite(rcp) && rcp > 0.0f) ? muls(diff, rcp) : V3<float>{0, 0, 0};
            point = bd.to_world(add(seg, muls(u, r)));
            return true;
        }
        V3<float> radial{local.x, 0, local.z};
        V3<float> dir{1, 0, 0};
        if (dot(radial, radial) > kEpsSq) {
            float rcp = 1.0f / len(radial);
            dir = (std::isfinite(rcp) && rcp > 0.0f) ? muls(radial, rcp) : V3<float>{0, 0, 0};
        }
        normal = bd.rotate(dir);
        depth = pr + r - len(diff);
        point = bd.to_world(add(seg, muls(dir, r)));
        return true;
    }
    return false;
}

struct SampleO {
    V3<float> point, normal;
    float area;
};

// fluid3.rs:375-465 (both faces of a box axis sample the plane at -half
// extent, as the reference's table has it)
inline void surface_samples_o(const RigidO &bd, std::vector<SampleO> &out) {
    out.clear();
    const float PI = kPiF;
    auto put = [&](V3<float> l, V3<float> nrm, float area) { out.push_back({bd.to_world(l), bd.rotate(nrm), area}); };
    if (bd.shape == 0) {
        V3<float> d[14] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
        int k = 6;
        for (float sx : {1.0f, -1.0f})
            for (float sy : {1.0f, -1.0f})
                for (float sz : {1.0f, -1.0f}) {
                    V3<float> v{sx, sy, sz};
                    float rcp = 1.0f / len(v);
                    d[k++] = muls(v, rcp);
                }
        float area = 4.0f * PI * bd.a * bd.a / 14.0f;
        for (int q = 0; q < 14; ++q) put(muls(d[q], bd.a), d[q], area);
    } else if (bd.shape == 1) {
        float area = 8.0f * (bd.a * bd.b + bd.b * bd.c + bd.c * bd.a) / 24.0f;
        const float t[4][2] = {{-0.75f, -0.75f}, {0.75f, -0.75f}, {-0.75f, 0.75f}, {0.75f, 0.75f}};
        for (int f = 0; f < 6; ++f) {
            int axis = f / 2;
            float sgn = (f % 2) ? -1.0f : 1.0f;
            V3<float> nrm = axis == 0 ? V3<float>{0, sgn, 0} : axis == 1 ? V3<float>{sgn, 0, 0} : V3<float>{0, 0, sgn};
            for (int k = 0; k < 4; ++k) {
                V3<float> l = axis == 0   ? V3<float>{bd.a * t[k][0], -bd.b, bd.c * t[k][1]}
                              : axis == 1 ? V3<float>{-bd.a, bd.b * t[k][0], bd.c * t[k][1]}
                                          : V3<float>{bd.a * t[k][0], bd.b * t[k][1], -bd.c};
                put(l, nrm, area);
            }
        }
    } else if (bd.shape == 2) {
        float r = bd.a, hl = bd.b;
        float side = 4.0f * PI * r * hl / 16.0f, cap = 2.0f * PI * r * r;
        for (float sy : {-hl * 0.5f, hl * 0.5f})
            for (int k = 0; k < 8; ++k) {
                float an = 2.0f * PI * (float)k / 8.0f;
                V3<float> dir{std::cos(an), 0.0f, std::sin(an)};
                put(add(muls(dir, r), muls(V3<float>{0, 1, 0}, sy)), dir, side);
            }
        for (float sg : {-1.0f, 1.0f}) put(muls(muls(V3<float>{0, 1, 0}, sg), hl + r), muls(V3<float>{0, 1, 0}, sg), cap);
    }
}

// The world and its coupled fluid.
struct CoupledWorld {
    Fluid<float> f;
    std::vector<RigidO> bodies;
    std::vector<Body> statics;  // for the ghosts (static bodies only, fluid3.rs:724)
    V3<float> g;
    std::vector<SampleO> samp;
    explicit CoupledWorld(const FluidConfig3Ffi &c) : f(c) {
        f.tail = &CoupledWorld::tail_fn;
        f.tail_ctx = this;
    }
    static void tail_fn(void *ctx, float dt) { static_cast<CoupledWorld *>(ctx)->coupled_tail(dt); }

    // fluid3.rs:345-369 on the current grid
    float sample_pressure(V3<float> point) const {
        const float h = f.h_kernel();
        const int32_t cx = sat_i32((double)std::floor(point.x / h)), cy = sat_i32((double)std::floor(point.y / h)),
                      cz = sat_i32((double)std::floor(point.z / h));
        float p = 0.0f;
        for (int oz = -1; oz <= 1; ++oz)
            for (int oy = -1; oy <= 1; ++oy)
                for (int ox = -1; ox <= 1; ++ox) {
                    auto it = f.grid.find(Key{cx + ox, cy + oy, cz + oz});
                    if (it == f.grid.end()) continue;
                    for (uint32_t j : it->second) {
                        V3<float> d = sub(f.pos[j], point);
                        float r2 = dot(d, d);
                        if (r2 < h * h) p += f.mass[j] * (f.pressure[j] / rmax<float>(f.density[j], kEps)) *
                                             Fluid<float>::poly6(r2, h);
                    }
                }
        return rmax<float>(p, 0.0f);
    }

    // fluid3.rs:1141-1213 after the integrate of a substep
    void coupled_tail(float dt) {
        const float pr = rmax<float>(f.cfg.particle_radius, 0.0f);
        for (uint32_t i = 0; i < f.n; ++i) {
            for (RigidO &bd : bodies) {
                V3<float> nrm, pt;
                float depth;
                if (!collide_rigid(f.pos[i], pr, bd, nrm, depth, pt)) continue;
                if (depth <= 0.0f) continue;
                f.pos[i] = add(f.pos[i], muls(nrm, depth));
                float vn = dot(f.vel[i], nrm);
                if (vn < 0.0f) {
                    f.vel[i] = sub(f.vel[i], muls(nrm, vn));
                    if (!bd.is_static) bd.sleeping = false;
                    bd.apply_impulse(muls(nrm, vn * f.mass[i]), pt);
                    if (f.cfg.wall_friction > 0.0f) {
                        float mu = rclamp<float>(f.cfg.wall_friction, 0.0f, 1.0f);
                        V3<float> vt = f.vel[i];
                        V3<float> wv = bd.vel;
                        V3<float> wvt = sub(wv, muls(nrm, dot(wv, nrm)));
                        V3<float> dv = muls(sub(wvt, vt), mu);
                        f.vel[i] = add(f.vel[i], dv);
                        bd.apply_impulse(muls(V3<float>{-dv.x, -dv.y, -dv.z}, f.mass[i]), pt);
                    }
                }
            }
        }
        f.rebuild_grid();
        const float drag = rmax<float>(f.cfg.drag, 0.0f);
        for (RigidO &bd : bodies) {
            if (bd.is_static) continue;
            surface_samples_o(bd, samp);
            if (samp.empty()) continue;
            bool wet = false;
            for (const SampleO &sm : samp) {
                float p = sample_pressure(sm.point);
                if (p <= 0.0f) continue;
                wet = true;
                V3<float> r = sub(sm.point, bd.pos);
                V3<float> vrel = sub(add(bd.vel, crossf(bd.ang, r)), V3<float>{0, 0, 0});
                float speed = len(vrel);
                V3<float> force = muls(V3<float>{-sm.normal.x, -sm.normal.y, -sm.normal.z}, p * sm.area);
                if (speed > kEps) force = sub(force, muls(vrel, drag * sm.area * speed));
                bd.apply_force(force, sm.point, dt);
            }
            if (wet) bd.sleeping = false;
        }
    }

    // World3::step_inner (world.rs:2766-2812) without the rigid solver
    void step(float dt, bool jacobi) {
        for (RigidO &bd : bodies)
            if (!bd.is_static && !bd.sleeping) bd.vel = add(bd.vel, muls(g, bd.gravity_scale * dt));
        if (f.cfg.enabled && f.n > 0 && dt > 0.0f) {
            statics.clear();
            for (const RigidO &bd : bodies)
                if (bd.is_static) statics.push_back(bd.as_body());
            uint32_t sub = f.cfl_substeps(dt);
            float sdt = dt / (float)sub;
            for (uint32_t k = 0; k < sub; ++k)
                f.substep(g, sdt, jacobi, statics.data(), (uint32_t)statics.size(), nullptr, nullptr, nullptr);
        }
        for (RigidO &bd : bodies) bd.integrate(dt);
    }
};

}  // namespace

extern "C" {

typedef struct OracleBody3 {
    uint32_t shape;
    float px, py, pz;
    float a, b, c;
} OracleBody3;

// A body of the coupled world step: shape (0 sphere, 1 box, 2 capsule),
// density (the mass properties follow as RigidBody3::set_shape), static
// flag and the motion state (rotation x, y, z, w), read and written back.
typedef struct OracleRigid3 {
    uint32_t shape;
    float a, b, c;
    float density;
    uint32_t is_static;
    float pos[3], rot[4], vel[3], ang[3];
    float linear_damping, angular_damping, gravity_scale;
} OracleRigid3;

static RigidO rigid_in(const OracleRigid3 &o) {
    RigidO r;
    r.shape = o.shape;
    r.a = o.a;
    r.b = o.shape == 0 ? 0.0f : o.b;
    r.c = o.shape == 1 ? o.c : 0.0f;
    r.is_static = o.is_static != 0;
    r.pos = {o.pos[0], o.pos[1], o.pos[2]};
    r.rot = {o.rot[0], o.rot[1], o.rot[2], o.rot[3]};
    r.vel = {o.vel[0], o.vel[1], o.vel[2]};
    r.ang = {o.ang[0], o.ang[1], o.ang[2]};
    r.linear_damping = o.linear_damping;
    r.angular_damping = o.angular_damping;
    r.gravity_scale = o.gravity_scale;
    r.shape_mass(o.density);
    return r;
}

// World3::step (world.rs:2670-2697, 2766-2812) for `steps` steps of dt
// without the rigid contact solver: the gravity kick of the dynamic bodies,
// the fluid (fluid3.rs:591-607) coupled to every body (fluid3.rs:1141-1213),
// the bodies' integration.  Particles and bodies are updated in place.
// Returns the substeps of the last step.
int oracle3_world_run_f32(uint32_t n, float *pos, float *vel, const float *mass, const uint8_t *phase,
                          const FluidConfig3Ffi *cfg, const float *g3, OracleRigid3 *bodies, uint32_t nbodies,
                          float dt, uint32_t steps, int jacobi) {
    CoupledWorld w(*cfg);
    w.f.load(n, pos, vel, mass, phase, nullptr);
    w.g = {g3[0], g3[1], g3[2]};
    for (uint32_t k = 0; k < nbodies; ++k) w.bodies.push_back(rigid_in(bodies[k]));
    uint32_t sub = 0;
    for (uint32_t s = 0; s < steps; ++s) {
        sub = (cfg->enabled && n > 0 && dt > 0.0f) ? w.f.cfl_substeps(dt) : 0;
        w.step(dt, jacobi != 0);
    }
    for (uint32_t i = 0; i < n; ++i) {
        pos[3 * i] = w.f.pos[i].x;
        pos[3 * i + 1] = w.f.pos[i].y;
        pos[3 * i + 2] = w.f.pos[i].z;
        vel[3 * i] = w.f.vel[i].x;
        vel[3 * i + 1] = w.f.vel[i].y;
        vel[3 * i + 2] = w.f.vel[i].z;
    }
    for (uint32_t k = 0; k < nbodies; ++k) {
        const RigidO &r = w.bodies[k];
        OracleRigid3 &o = bodies[k];
        o.pos[0] = r.pos.x;
        o.pos[1] = r.pos.y;
        o.pos[2] = r.pos.z;
        o.rot[0] = r.rot.x;
        o.rot[1] = r.rot.y;
        o.rot[2] = r.rot.z;
        o.rot[3] = r.rot.w;
        o.vel[0] = r.vel.x;
        o.vel[1] = r.vel.y;
        o.vel[2] = r.vel.z;
        o.ang[0] = r.ang.x;
        o.ang[1] = r.ang.y;
        o.ang[2] = r.ang.z;
    }
    return (int)sub;
}

void oracle3_cos_sin(float angle, float *c, float *s) { trig_cos_sin(angle, *c, *s); }
void oracle3_quat_mul(const float *a, const float *b, float *out) {
    Quat q = quat_mul({a[0], a[1], a[2], a[3]}, {b[0], b[1], b[2], b[3]});
    out[0] = q.x;
    out[1] = q.y;
    out[2] = q.z;
    out[3] = q.w;
}
void oracle3_quat_normalize(const float *a, float *out) {
    Quat q = quat_normalize({a[0], a[1], a[2], a[3]});
    out[0] = q.x;
    out[1] = q.y;
    out[2] = q.z;
    out[3] = q.w;
}
void oracle3_quat_rotate(const float *q, const float *v, float *out) {
    V3<float> r = quat_apply({q[0], q[1], q[2], q[3]}, {v[0], v[1], v[2]});
    out[0] = r.x;
    out[1] = r.y;
    out[2] = r.z;
}
// Mass, inverse mass and the body-frame inverse inertia (9, columns) of a
// dynamic body of the shape (RigidBody3::set_shape).
void oracle3_body_mass(uint32_t shape, float a, float b, float c, float density, float *mass, float *inv_mass,
                       float *inv_inertia9) {
    OracleRigid3 o{};
    o.shape = shape;
    o.a = a;
    o.b = b;
    o.c = c;
    o.density = density;
    o.rot[3] = 1.0f;
    RigidO r = rigid_in(o);
    *mass = r.mass;
    *inv_mass = r.inv_mass;
    const V3<float> cols[3] = {r.inv_inertia.c0, r.inv_inertia.c1, r.inv_inertia.c2};
    for (int k = 0; k < 3; ++k) {
        inv_inertia9[3 * k] = cols[k].x;
        inv_inertia9[3 * k + 1] = cols[k].y;
        inv_inertia9[3 * k + 2] = cols[k].z;
    }
}
// The body's buoyancy samples (fluid3.rs:375-465): points, outward
// normals (3 floats each) and areas, up to max; returns the count.
uint32_t oracle3_surface_samples(const OracleRigid3 *body, float *points, float *normals, float *areas, uint32_t max) {
    RigidO r = rigid_in(*body);
    std::vector<SampleO> ss;
    surface_samples_o(r, ss);
    for (uint32_t k = 0; k < ss.size() && k < max; ++k) {
        points[3 * k] = ss[k].point.x;
        points[3 * k + 1] = ss[k].point.y;
        points[3 * k + 2] = ss[k].point.z;
        normals[3 * k] = ss[k].normal.x;
        normals[3 * k + 1] = ss[k].normal.y;
        normals[3 * k + 2] = ss[k].normal.z;
        areas[k] = ss[k].area;
    }
    return (uint32_t)ss.size();
}
// sample_pressure (fluid3.rs:345-369) at a point for a given particle
// state (density and pressure as the density pass left them).
float oracle3_sample_pressure(uint32_t n, const float *pos, const float *mass, const float *density,
                              const float *pressure, const FluidConfig3Ffi *cfg, const float *point3) {
    CoupledWorld w(*cfg);
    w.f.load(n, pos, nullptr, mass, nullptr, density);
    for (uint32_t i = 0; i < n; ++i) w.f.pressure[i] = pressure[i];
    w.f.rebuild_grid();
    return w.sample_pressure({point3[0], point3[1], point3[2]});
}

// One f32 substep in place (fluid3.rs:990-1184).  jacobi = 0 reproduces the
// reference's Gauss-Seidel integrate.  Optional outputs (null to skip):
// acc3 = g + F/rho per particle, count = neighbour count incl. self,
// scale = cancellation-aware magnitude sum(|m_j term grad W|)/rho_i.
int oracle3_substep_f32(uint32_t n, float *pos, float *vel, float *density, float *pressure,
                        const float *mass, const uint8_t *phase, const FluidConfig3Ffi *cfg,
                        const float *g3, const OracleBody3 *bodies, uint32_t nbodies, float dt,
                        int jacobi, float *acc3, uint32_t *count, double *scale) {
    Fluid<float> f(*cfg);
    f.load(n, pos, vel, mass, phase, density);
    f.substep({g3[0], g3[1], g3[2]}, dt, jacobi != 0, reinterpret_cast<const Body *>(bodies), nbodies,
              acc3, count, scale);
    for (uint32_t i = 0; i < n; ++i) {
        pos[3 * i] = f.pos[i].x;
        pos[3 * i + 1] = f.pos[i].y;
        pos[3 * i + 2] = f.pos[i].z;
        vel[3 * i] = f.vel[i].x;
        vel[3 * i + 1] = f.vel[i].y;
        vel[3 * i + 2] = f.vel[i].z;
        if (density) density[i] = f.density[i];
        if (pressure) pressure[i] = f.pressure[i];
    }
    return 1;
}

// The boundary ghosts of the bodies for a fluid state (fluid3.rs:1018-1027
// on that state): positions (3 per ghost), density and pressure, in the
// reference's generation order, up to max.  Returns the ghost count.
uint32_t oracle3_boundary_f32(uint32_t n, const float *pos, const float *mass, const FluidConfig3Ffi *cfg,
                              const OracleBody3 *bodies, uint32_t nbodies, float *xyz, float *density,
                              float *pressure, uint32_t max) {
    Fluid<float> f(*cfg);
    f.load(n, pos, nullptr, mass, nullptr, nullptr);
    f.rebuild_grid();
    f.update_boundaries(reinterpret_cast<const Body *>(bodies), nbodies, f.h_kernel());
    const uint32_t m = std::min<uint32_t>(max, (uint32_t)f.ghosts.size());
    for (uint32_t g = 0; g < m; ++g) {
        if (xyz) {
            xyz[3 * g] = f.ghosts[g].p.x;
            xyz[3 * g + 1] = f.ghosts[g].p.y;
            xyz[3 * g + 2] = f.ghosts[g].p.z;
        }
        if (density) density[g] = f.ghosts[g].density;
        if (pressure) pressure[g] = f.ghosts[g].pressure;
    }
    return (uint32_t)f.ghosts.size();
}

// FluidState3::step_fluid (fluid3.rs:591-607) for `steps` world steps of dt
// (dt already clamped by the caller as phieng_world3_step does,
// ffi.rs:3221-3226).  State is carried across steps in a persistent Fluid
// so the hash grid is reused as in the reference.  Returns the number of
// substeps taken by the last step.
int oracle3_run_f32(uint32_t n, float *pos, float *vel, const float *mass, const uint8_t *phase,
                    const FluidConfig3Ffi *cfg, const float *g3, const OracleBody3 *bodies,
                    uint32_t nbodies, float dt, uint32_t steps, int jacobi) {
    Fluid<float> f(*cfg);
    f.load(n, pos, vel, mass, phase, nullptr);
    uint32_t sub = 0;
    for (uint32_t s = 0; s < steps; ++s) {
        if (!cfg->enabled || n == 0 || !(dt > 0.0f)) break;
        sub = f.cfl_substeps(dt);
        float sdt = dt / (float)sub;
        for (uint32_t k = 0; k < sub; ++k)
            f.substep({g3[0], g3[1], g3[2]}, sdt, jacobi != 0, reinterpret_cast<const Body *>(bodies),
                      nbodies, nullptr, nullptr, nullptr);
    }
    for (uint32_t i = 0; i < n; ++i) {
        pos[3 * i] = f.pos[i].x;
        pos[3 * i + 1] = f.pos[i].y;
        pos[3 * i + 2] = f.pos[i].z;
        vel[3 * i] = f.vel[i].x;
        vel[3 * i + 1] = f.vel[i].y;
        vel[3 * i + 2] = f.vel[i].z;
    }
    return (int)sub;
}

// oracle3_run_f32 with the Jacobi integrate on `threads` OpenMP threads
// (0: the OpenMP default).  The grid rebuild stays on one thread.
int oracle3_run_f32_par(uint32_t n, float *pos, float *vel, const float *mass, const uint8_t *phase,
                        const FluidConfig3Ffi *cfg, const float *g3, const OracleBody3 *bodies, uint32_t nbodies,
                        float dt, uint32_t steps, int threads) {
    if (threads > 0) omp_set_num_threads(threads);
    Fluid<float> f(*cfg);
    f.par = true;
    f.load(n, pos, vel, mass, phase, nullptr);
    uint32_t sub = 0;
    for (uint32_t s = 0; s < steps; ++s) {
        if (!cfg->enabled || n == 0 || !(dt > 0.0f)) break;
        sub = f.cfl_substeps(dt);
        float sdt = dt / (float)sub;
        for (uint32_t k = 0; k < sub; ++k)
            f.substep({g3[0], g3[1], g3[2]}, sdt, true, reinterpret_cast<const Body *>(bodies), nbodies, nullptr,
                      nullptr, nullptr);
    }
    for (uint32_t i = 0; i < n; ++i) {
        pos[3 * i] = f.pos[i].x;
        pos[3 * i + 1] = f.pos[i].y;
        pos[3 * i + 2] = f.pos[i].z;
        vel[3 * i] = f.vel[i].x;
        vel[3 * i + 1] = f.vel[i].y;
        vel[3 * i + 2] = f.vel[i].z;
    }
    return (int)sub;
}
int oracle3_max_threads(void) { return omp_get_max_threads(); }

// The variable-count forms (adaptive resolution changes the particle
// count).  *n is the count in and out; every array holds `room` particles
// (pos, vel 3 each); mass and phase are updated with the state; cap is the
// split cap of fluid3.rs:972 (MAX_PARTICLES_3D = 4096 in the reference, the
// world's capacity on the GPU).  The cap check lets the count pass the
// cap, so a result larger than room is not written: the call returns 0
// with *n set to the count it needs.
static void store_state(const Fluid<float> &f, float *pos, float *vel, float *mass, uint8_t *phase,
                        float *density, float *pressure) {
    for (uint32_t i = 0; i < f.n; ++i) {
        pos[3 * i] = f.pos[i].x;
        pos[3 * i + 1] = f.pos[i].y;
        pos[3 * i + 2] = f.pos[i].z;
        if (vel) {
            vel[3 * i] = f.vel[i].x;
            vel[3 * i + 1] = f.vel[i].y;
            vel[3 * i + 2] = f.vel[i].z;
        }
        if (mass) mass[i] = f.mass[i];
        if (phase) phase[i] = f.phase[i];
        if (density) density[i] = f.density[i];
        if (pressure) pressure[i] = f.pressure[i];
    }
}

// resample() alone (fluid3.rs:923-987) on a state whose densities are given.
int oracle3_resample_f32(uint32_t *n, uint32_t cap, uint32_t room, float *pos, float *vel, float *mass, uint8_t *phase,
                         float *density, const FluidConfig3Ffi *cfg, float h) {
    Fluid<float> f(*cfg);
    f.cap = cap;
    f.load(*n, pos, vel, mass, phase, density);
    f.resample(h);
    *n = f.n;
    if (f.n > room) return 0;
    store_state(f, pos, vel, mass, phase, density, nullptr);
    return 1;
}

// One substep (as oracle3_substep_f32) with a variable count.
int oracle3_substep_cap_f32(uint32_t *n, uint32_t cap, uint32_t room, float *pos, float *vel, float *mass, uint8_t *phase,
                            float *density, float *pressure, const FluidConfig3Ffi *cfg, const float *g3,
                            const OracleBody3 *bodies, uint32_t nbodies, float dt, int jacobi, float *acc3,
                            uint32_t *count, double *scale) {
    Fluid<float> f(*cfg);
    f.cap = cap;
    f.load(*n, pos, vel, mass, phase, nullptr);
    f.substep({g3[0], g3[1], g3[2]}, dt, jacobi != 0, reinterpret_cast<const Body *>(bodies), nbodies, acc3,
              count, scale);
    *n = f.n;
    if (f.n > room) return 0;
    store_state(f, pos, vel, mass, phase, density, pressure);
    return 1;
}

// `steps` world steps (as oracle3_run_f32) with a variable count.
int oracle3_run_cap_f32(uint32_t *n, uint32_t cap, uint32_t room, float *pos, float *vel, float *mass, uint8_t *phase,
                        const FluidConfig3Ffi *cfg, const float *g3, const OracleBody3 *bodies, uint32_t nbodies,
                        float dt, uint32_t steps, int jacobi) {
    Fluid<float> f(*cfg);
    f.cap = cap;
    f.load(*n, pos, vel, mass, phase, nullptr);
    for (uint32_t s = 0; s < steps; ++s) {
        if (!cfg->enabled || f.n == 0 || !(dt > 0.0f)) break;
        const uint32_t sub = f.cfl_substeps(dt);
        const float sdt = dt / (float)sub;
        for (uint32_t k = 0; k < sub; ++k)
            f.substep({g3[0], g3[1], g3[2]}, sdt, jacobi != 0, reinterpret_cast<const Body *>(bodies), nbodies,
                      nullptr, nullptr, nullptr);
    }
    *n = f.n;
    if (f.n > room) return 0;
    store_state(f, pos, vel, mass, phase, nullptr, nullptr);
    return 1;
}

// FP64 twin of one Jacobi substep from an f32 state.  Outputs in f64.
int oracle3_substep_f64(uint32_t n, const float *pos, const float *vel, const float *mass,
                        const uint8_t *phase, const FluidConfig3Ffi *cfg, const float *g3,
                        const OracleBody3 *bodies, uint32_t nbodies, float dt, double *out_rho,
                        double *out_p, double *out_acc3, double *out_pos3, double *out_vel3, double *out_scale) {
    Fluid<double> f(*cfg);
    f.par = true;  // Jacobi: the per-particle loops are independent
    f.load(n, pos, vel, mass, phase, nullptr);
    std::vector<double> acc(3 * (size_t)n);
    f.substep({g3[0], g3[1], g3[2]}, (double)dt, true, reinterpret_cast<const Body *>(bodies), nbodies,
              acc.data(), nullptr, out_scale);
    for (uint32_t i = 0; i < n; ++i) {
        if (out_rho) out_rho[i] = f.density[i];
        if (out_p) out_p[i] = f.pressure[i];
        if (out_acc3) {
            out_acc3[3 * i] = acc[3 * i];
            out_acc3[3 * i + 1] = acc[3 * i + 1];
            out_acc3[3 * i + 2] = acc[3 * i + 2];
        }
        if (out_pos3) {
            out_pos3[3 * i] = f.pos[i].x;
            out_pos3[3 * i + 1] = f.pos[i].y;
            out_pos3[3 * i + 2] = f.pos[i].z;
        }
        if (out_vel3) {
            out_vel3[3 * i] = f.vel[i].x;
            out_vel3[3 * i + 1] = f.vel[i].y;
            out_vel3[3 * i + 2] = f.vel[i].z;
        }
    }
    return 1;
}

// Cell key triples (fluid3.rs:305-309) with Rust's saturating cast.
void oracle3_cells(uint32_t n, const float *pos, float h, int32_t *cells3) {
    for (uint32_t i = 0; i < n; ++i) {
        cells3[3 * i] = sat_i32((double)std::floor(pos[3 * i] / h));
        cells3[3 * i + 1] = sat_i32((double)std::floor(pos[3 * i + 1] / h));
        cells3[3 * i + 2] = sat_i32((double)std::floor(pos[3 * i + 2] / h));
    }
}

// The grid's bucket contents laid out in z-major cell order: cells sorted
// by (cz, cy, cx), members in ascending index (fluid3.rs:301-314);
// non-finite positions, which the reference leaves out of the grid, last
// in index order.  Equal to a stable sort of 0..n by that key.
void oracle3_sorted_ids(uint32_t n, const float *pos, float h, uint32_t *ids) {
    std::vector<int32_t> c(3 * (size_t)n);
    oracle3_cells(n, pos, h, c.data());
    std::vector<uint8_t> fin(n);
    for (uint32_t i = 0; i < n; ++i)
        fin[i] = std::isfinite(pos[3 * i]) && std::isfinite(pos[3 * i + 1]) && std::isfinite(pos[3 * i + 2]);
    std::vector<uint32_t> idx(n);
    for (uint32_t i = 0; i < n; ++i) idx[i] = i;
    std::stable_sort(idx.begin(), idx.end(), [&](uint32_t a, uint32_t b) {
        if (fin[a] != fin[b]) return fin[a] > fin[b];
        if (!fin[a]) return false;
        if (c[3 * a + 2] != c[3 * b + 2]) return c[3 * a + 2] < c[3 * b + 2];
        if (c[3 * a + 1] != c[3 * b + 1]) return c[3 * a + 1] < c[3 * b + 1];
        return c[3 * a] < c[3 * b];
    });
    std::memcpy(ids, idx.data(), sizeof(uint32_t) * n);
}

// Scalar kernels for closed-form tests (fluid3.rs:220-251).
float oracle3_poly6(float r2, float h) { return Fluid<float>::poly6(r2, h); }
void oracle3_spiky_grad(const float *r_vec3, float r, float h, float *out3) {
    V3<float> g = Fluid<float>::spiky_grad({r_vec3[0], r_vec3[1], r_vec3[2]}, r, h);
    out3[0] = g.x;
    out3[1] = g.y;
    out3[2] = g.z;
}
float oracle3_pressure(float density, float rest, float k, float gamma) {
    return Fluid<float>::eos(density, rest, k, gamma);
}
double oracle3_poly6_f64(double r2, double h) { return Fluid<double>::poly6(r2, h); }
// fluid3.rs:254-269 and :272-278 (pinned by the reference's own sign test,
// fluid3.rs:1418-1455, and by two-particle closed forms of the whole step).
float oracle3_artificial_pressure(const float *delta3, const float *vrel3, float r_sq, float h, float rho_bar,
                                  float c, float alpha) {
    return Fluid<float>::artificial_pressure({delta3[0], delta3[1], delta3[2]}, {vrel3[0], vrel3[1], vrel3[2]}, r_sq,
                                             h, rho_bar, c, alpha);
}
float oracle3_cohesion(float dist, float h) { return Fluid<float>::cohesion(dist, h); }
uint32_t oracle3_cfl_substeps(uint32_t n, const float *vel, const FluidConfig3Ffi *cfg, float dt) {
    Fluid<float> f(*cfg);
    std::vector<float> p(3 * (size_t)n, 0.0f), m(n, 1.0f);
    f.load(n, p.data(), vel, m.data(), nullptr, nullptr);
    return f.cfl_substeps(dt);
}

}  // extern "C"
