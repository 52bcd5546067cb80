// sph_device.h — device-side types and arithmetic of the SPH fluid step.
//
// Every floating-point expression here keeps the operand order of the
// reference (crates/dynamics/src/fluid3.rs) and the library is built with
// -ffp-contract=off and IEEE division/sqrt, so a particle's sums come out
// bit-identical to the f32 CPU path when neighbours are visited in the
// reference order (cells oz -> oy -> ox, ascending original index).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define PH_EPS 1e-6f    // crates/math/src/math.rs:3
#define PH_EPS_SQ 1e-12f  // crates/math/src/math.rs:4
#define PH_ID_MASK 0x00FFFFFFu  // low 24 bits of a tag: original index
#define PH_TILE 256             // particles per tile (= threads per block)
#define PH_MAX_BODIES 64

#define PH_HD __host__ __device__ __forceinline__

// Collision body.  Static bodies have the identity rotation (the C ABI
// cannot rotate one); a dynamic body carries its rotation (x, y, z, w),
// flagged BODY_ROTATED once it is not exactly the identity.
#define BODY_DYNAMIC 1u
#define BODY_ROTATED 2u
struct BodyDev {
    uint32_t shape;  // 0 sphere, 1 box, 2 capsule (ffi.rs:3193-3195); 3 box of a 2D world
    float px, py, pz;
    float a, b, c;
    float qx, qy, qz, qw;
    uint32_t flags;
};

// Per-scene parameters: grid domain, bucket range, particle range and the
// f32 constants the kernels read.  One scene = one World3 fluid.
struct SceneDev {
    int32_t x0, y0, z0;  // domain origin (cell coordinates)
    int32_t nx, ny, nz;  // domain extent (cells)
    uint32_t bucket_base;  // first global bucket; ncells cells, then outlier, then non-finite
    uint32_t ncells;
    uint32_t part_begin, part_end;  // particle slots of this scene
    float h_grid;   // raw smoothing_length: cells and the d2 < h*h cut (fluid3.rs:321)
    float h_grid2;  // h_grid * h_grid
    float h;        // smoothing_length.max(EPS): kernels (fluid3.rs:999)
    float h2;       // h * h
    float poly6_c;  // 315 / (64 pi h2 h2 h2 h2 h)      (fluid3.rs:227)
    float spiky_c;  // -45 / (pi h h h h h h)           (fluid3.rs:237)
    float rho_max;  // 3 * rest_density.max(EPS)        (fluid3.rs:649)
    float rest_density;  // the config's base rest density (ghost pressure, fluid3.rs:917-918)
    uint32_t ghost_count;  // boundary ghosts of this scene (0: none)
    uint32_t ghost_mask;   // ghost cell table size - 1 (a power of two minus one)
    float stiffness, gamma;
    float phase_rho0[4];
    float av_alpha, c_s, kappa;
    float viscosity, max_velocity, particle_radius, wall_mu, wall_friction;
    float gx, gy, gz;
    float dt;
    uint32_t body_begin, body_count;
    uint32_t tile_begin, tile_count;  // this scene's run of the tile table
};

struct TileDev {
    uint32_t scene;
    uint32_t first;  // first particle slot; the tile covers up to PH_TILE slots
};

// Rust f32::max: a NaN operand loses.
PH_HD float rmaxf(float a, float b) {
    if (a != a) return b;
    if (b != b) return a;
    return a > b ? a : b;
}
// rmaxf(a, b) for b > 0 (not NaN): IEEE maxNum, which takes b for a NaN a
// and max(a, b) otherwise, is the same function.
PH_HD float rmax_pos(float a, float b) { return fmaxf(a, b); }
// Rust f32::clamp: NaN passes through.
PH_HD float rclampf(float v, float lo, float hi) {
    if (v < lo) v = lo;
    if (v > hi) v = hi;
    return v;
}
// Rust `f32 as i32`: saturating, NaN -> 0.
__device__ __forceinline__ int32_t sat_i32(float v) {
    if (v != v) return 0;
    if (v >= 2147483648.0f) return INT32_MAX;
    if (v < -2147483648.0f) return INT32_MIN;
    return (int32_t)v;
}
PH_HD bool finite3(float x, float y, float z) {
    return __builtin_isfinite(x) && __builtin_isfinite(y) && __builtin_isfinite(z);
}
// glam Vec3::length_squared: (x*x + y*y) + z*z.
PH_HD float dot3(float x, float y, float z) { return x * x + y * y + z * z; }

// sqrtf(r2) rounded to nearest, for a pair's squared distance (finite,
// >= 0).  For r2 >= 2^-96 the hardware root is within one ulp, and its two
// one-ulp neighbours are tested with exact FMA residuals: the correction
// the compiler emits for an IEEE sqrtf, without its rescaling of inputs
// below 2^-96 and its select for 0 and infinity.  Below 2^-96 (0 included)
// the result is some value below 2^-40; every caller tests dist < EPS
// (1e-6) before using the root, which such a pair fails either way.
__device__ __forceinline__ float sqrt_pair(float r2) {
    const float s = __builtin_amdgcn_sqrtf(r2);
    const float sd = __uint_as_float(__float_as_uint(s) - 1u);
    const float su = __uint_as_float(__float_as_uint(s) + 1u);
    float r = __builtin_fmaf(-sd, s, r2) <= 0.0f ? sd : s;
    r = __builtin_fmaf(-su, s, r2) > 0.0f ? su : r;
    return r;
}

// AMD's IEEE f32 division a / b is: scale both operands (v_div_scale),
// y = rcp(b) refined by one Newton step, q = a y, two residual corrections
// q += (a - b q) y, then a fix-up of special operands (v_div_fixup).  The
// scaling and the fix-up are identities when a != 0, |a| >= 2^-100,
// |a| <= 2 b, b is normal and 1 / b and a / b are normal; then the middle
// part alone gives the IEEE quotient bit for bit, and the refined
// reciprocal can be shared by every numerator over the same b.
__device__ __forceinline__ float div_recip(float b) {
    const float y0 = __builtin_amdgcn_rcpf(b);
    const float e = __builtin_fmaf(-b, y0, 1.0f);
    return __builtin_fmaf(e, y0, y0);
}
__device__ __forceinline__ float div_by_recip(float a, float b, float y) {
    float q = a * y;
    float r = __builtin_fmaf(-b, q, a);
    q = __builtin_fmaf(r, y, q);
    r = __builtin_fmaf(-b, q, a);
    return __builtin_fmaf(r, y, q);
}

// fluid3.rs:220-228 (coefficient precomputed on the host in f32).
__device__ __forceinline__ float poly6(float r2, const SceneDev &s) {
    if (r2 >= s.h2 || r2 < 0.0f) return 0.0f;
    float t = s.h2 - r2;
    return s.poly6_c * t * t * t;
}

// poly6 for 0 <= r2 < h_grid^2 <= h^2 (a neighbour of the cut-off test),
// where the kernel's range test cannot fail.
__device__ __forceinline__ float poly6_within(float r2, const SceneDev &s) {
    float t = s.h2 - r2;
    return s.poly6_c * t * t * t;
}

// (float)x^gamma from the f64 power.  For gamma = 7 (the Tait exponent of
// every configuration here) x^7 is four f64 products (relative error below
// 2^-50), which round to the same f32 as the f64 pow unless x^7 lies within
// 2^-48 (relative) of an f32 rounding midpoint; there the f64 pow decides.
__device__ __forceinline__ float tait_pow(float x, float gamma) {
    const double xd = (double)x;
    if (gamma == 7.0f) {
        const double x2 = xd * xd, x3 = x2 * xd, x6 = x3 * x3, x7 = x6 * xd;
        const float f = (float)x7;
        const double fd = (double)f;
        const float g = nextafterf(f, x7 > fd ? INFINITY : -INFINITY);
        const double mid = 0.5 * (fd + (double)g);  // exact: two f32 values
        if (isfinite(x7) && fabs(x7 - mid) > fabs(x7) * 0x1p-48) return f;
    }
    return (float)pow(xd, (double)gamma);
}

// fluid3.rs:244-251; the Tait power is taken in f64 and rounded.
__device__ __forceinline__ float eos(float density, float rest, float k, float gamma) {
    if (gamma == 1.0f) return rmaxf(k * (density - rest), 0.0f);
    float rho0 = rmaxf(rest, PH_EPS);
    float pw = tait_pow(density / rho0, gamma);
    return rmaxf(k * rho0 * (pw - 1.0f), 0.0f);
}

__device__ __forceinline__ void cell_of(float x, float y, float z, float h, int32_t &cx, int32_t &cy,
                                        int32_t &cz) {
    cx = sat_i32(floorf(x / h));
    cy = sat_i32(floorf(y / h));
    cz = sat_i32(floorf(z / h));
}

__device__ __forceinline__ bool in_domain(const SceneDev &s, int64_t x, int64_t y, int64_t z) {
    return x >= s.x0 && x < (int64_t)s.x0 + s.nx && y >= s.y0 && y < (int64_t)s.y0 + s.ny &&
           z >= s.z0 && z < (int64_t)s.z0 + s.nz;
}

// Visit every grid candidate of a particle in the reference's order
// (fluid3.rs:326-338): cells oz, oy, ox in -1..=1, members by ascending
// original index.  With a z-major bucket key the three cells ox = -1..1 of
// one (oy, oz) row are adjacent buckets, so a row is one contiguous slot
// range of the sorted arrays.  Particles whose 27 cells leave the domain
// while the scene has outliers take the per-cell path, which also scans
// the outlier bucket (kept sorted by (cell z, y, x), then index).
template <class F>
__device__ __forceinline__ void for_each_candidate(const SceneDev &s, const uint32_t *__restrict__ start,
                                                   const float4 *__restrict__ posm, int32_t cx,
                                                   int32_t cy, int32_t cz, F &&f) {
    const uint32_t base = s.bucket_base;
    const uint32_t ob = start[base + s.ncells], oe = start[base + s.ncells + 1];
    bool inside = in_domain(s, (int64_t)cx - 1, (int64_t)cy - 1, (int64_t)cz - 1) &&
                  in_domain(s, (int64_t)cx + 1, (int64_t)cy + 1, (int64_t)cz + 1);
    if (inside || ob == oe) {
        // Fast path: at most 9 contiguous row ranges.  Cells outside the
        // domain are empty when the scene has no outliers.
        const int64_t xlo = max((int64_t)cx - 1, (int64_t)s.x0);
        const int64_t xhi = min((int64_t)cx + 1, (int64_t)s.x0 + s.nx - 1);
        if (xlo > xhi) return;
        for (int oz = -1; oz <= 1; ++oz) {
            const int64_t z = (int64_t)cz + oz;
            if (z < s.z0 || z >= (int64_t)s.z0 + s.nz) continue;
            for (int oy = -1; oy <= 1; ++oy) {
                const int64_t y = (int64_t)cy + oy;
                if (y < s.y0 || y >= (int64_t)s.y0 + s.ny) continue;
                const uint32_t row = (uint32_t)(((z - s.z0) * s.ny + (y - s.y0)) * s.nx);
                const uint32_t b = start[base + row + (uint32_t)(xlo - s.x0)];
                const uint32_t e = start[base + row + (uint32_t)(xhi - s.x0) + 1];
                for (uint32_t j = b; j < e; ++j) f(j);
            }
        }
        return;
    }
    // Slow path: cell by cell, outlier cells from the outlier bucket.
    for (int oz = -1; oz <= 1; ++oz)
        for (int oy = -1; oy <= 1; ++oy)
            for (int ox = -1; ox <= 1; ++ox) {
                const int64_t x = (int64_t)cx + ox, y = (int64_t)cy + oy, z = (int64_t)cz + oz;
                if (in_domain(s, x, y, z)) {
                    const uint32_t k = (uint32_t)(((z - s.z0) * s.ny + (y - s.y0)) * s.nx + (x - s.x0));
                    const uint32_t b = start[base + k], e = start[base + k + 1];
                    for (uint32_t j = b; j < e; ++j) f(j);
                } else {
                    for (uint32_t j = ob; j < oe; ++j) {
                        const float4 q = posm[j];
                        int32_t qx, qy, qz;
                        cell_of(q.x, q.y, q.z, s.h_grid, qx, qy, qz);
                        if (qx == x && qy == y && qz == z) f(j);
                    }
                }
            }
}

// glam 0.29 Quat * Vec3 (the SSE2 form the reference's x86-64 build runs,
// crates/dynamics via glam's Quat): v (w w - b.b) + b (2 (b.v)) + (b x v)(2 w),
// b = (x, y, z), summed left to right.
PH_HD void quat_rotate(float qx, float qy, float qz, float qw, float vx, float vy, float vz, float &ox, float &oy,
                       float &oz) {
    const float b2 = qx * qx + qy * qy + qz * qz;
    const float s0 = qw * qw - b2;
    const float d2 = (qx * vx + qy * vy + qz * vz) * 2.0f;
    const float w2 = qw * 2.0f;
    const float cx = qy * vz - qz * vy, cy = qz * vx - qx * vz, cz = qx * vy - qy * vx;
    ox = vx * s0 + qx * d2 + cx * w2;
    oy = vy * s0 + qy * d2 + cy * w2;
    oz = vz * s0 + qz * d2 + cz * w2;
}

// The body frame: local = R^-1 (p - c), and back (world = R l + c for a
// point, R l for a direction).  Without BODY_ROTATED the rotation is the
// identity, whose quaternion product reproduces its operand (up to the
// sign of zero components), so it is skipped.
struct BodyFrame {
    const BodyDev &b;
    PH_HD bool rot() const { return (b.flags & BODY_ROTATED) != 0u; }
    PH_HD void to_local(float px, float py, float pz, float &lx, float &ly, float &lz) const {
        lx = px - b.px;
        ly = py - b.py;
        lz = pz - b.pz;
        if (rot()) quat_rotate(-b.qx, -b.qy, -b.qz, b.qw, lx, ly, lz, lx, ly, lz);
    }
    PH_HD void dir(float &x, float &y, float &z) const {
        if (rot()) quat_rotate(b.qx, b.qy, b.qz, b.qw, x, y, z, x, y, z);
    }
    PH_HD void point(float lx, float ly, float lz, float &x, float &y, float &z) const {
        dir(lx, ly, lz);
        x = lx + b.px;
        y = ly + b.py;
        z = lz + b.pz;
    }
};

// fluid3.rs:471-585 (collide_particle_body): outward normal, depth and,
// when cp is given, the contact point on the body surface (cp[0..2]).
// Returns false when separated.
PH_HD bool collide(float px, float py, float pz, float pr, const BodyDev &b, float &nx, float &ny, float &nz,
                   float &depth, float *cp = nullptr) {
    const BodyFrame f{b};
    if (b.shape == 0u) {  // sphere
        const float dx = px - b.px, dy = py - b.py, dz = pz - b.pz;
        const float dist = sqrtf(dot3(dx, dy, dz));
        if (dist >= b.a + pr) return false;
        if (dist > PH_EPS) {
            nx = dx / dist;
            ny = dy / dist;
            nz = dz / dist;
        } else {
            nx = 1.0f;
            ny = 0.0f;
            nz = 0.0f;
        }
        depth = b.a + pr - dist;
        if (cp) {
            cp[0] = b.px + nx * b.a;
            cp[1] = b.py + ny * b.a;
            cp[2] = b.pz + nz * b.a;
        }
        return true;
    }
    if (b.shape == 3u) {  // 2D box in the z = 0 plane (fluid.rs:481-517), identity rotation
        const float lx = px - b.px, ly = py - b.py;
        const float cxl = rclampf(lx, -b.a, b.a), cyl = rclampf(ly, -b.b, b.b);
        const float dx = lx - cxl, dy = ly - cyl;
        nz = 0.0f;
        if (dx * dx + dy * dy > PH_EPS_SQ) {
            const float dist = sqrtf(dx * dx + dy * dy);
            if (dist >= pr) return false;
            nx = dx / dist;
            ny = dy / dist;
            depth = pr - dist;
            return true;
        }
        // Centre inside: least-penetration axis, ties to y, always +axis (fluid.rs:505-512).
        const float ex = b.a - fabsf(lx), ey = b.b - fabsf(ly);
        if (ex < ey) {
            nx = 1.0f;
            ny = 0.0f;
            depth = pr + ex;
        } else {
            nx = 0.0f;
            ny = 1.0f;
            depth = pr + ey;
        }
        return true;
    }
    if (b.shape == 1u) {  // box
        float lx, ly, lz;
        f.to_local(px, py, pz, lx, ly, lz);
        const float cxl = rclampf(lx, -b.a, b.a), cyl = rclampf(ly, -b.b, b.b), czl = rclampf(lz, -b.c, b.c);
        const float dx = lx - cxl, dy = ly - cyl, dz = lz - czl;
        if (dot3(dx, dy, dz) > PH_EPS_SQ) {
            const float dist = sqrtf(dot3(dx, dy, dz));
            if (dist >= pr) return false;
            nx = dx / dist;
            ny = dy / dist;
            nz = dz / dist;
            f.dir(nx, ny, nz);
            depth = pr - dist;
            if (cp) f.point(cxl, cyl, czl, cp[0], cp[1], cp[2]);
            return true;
        }
        // Centre inside: least-penetration axis, always +axis (fluid3.rs:514-520).
        const float ex = b.a - fabsf(lx), ey = b.b - fabsf(ly), ez = b.c - fabsf(lz);
        nx = ny = nz = 0.0f;
        if (ex <= ey && ex <= ez) {
            nx = 1.0f;
            depth = pr + ex;
        } else if (ey <= ez) {
            ny = 1.0f;
            depth = pr + ey;
        } else {
            nz = 1.0f;
            depth = pr + ez;
        }
        f.dir(nx, ny, nz);
        if (cp) f.point(cxl, cyl, czl, cp[0], cp[1], cp[2]);
        return true;
    }
    if (b.shape == 2u) {  // capsule along local y
        float lx, ly, lz;
        f.to_local(px, py, pz, lx, ly, lz);
        const float sy = rclampf(ly, -b.b, b.b);
        const float dx = lx - 0.0f, dy = ly - sy, dz = lz - 0.0f;
        if (dot3(dx, dy, dz) > PH_EPS_SQ) {
            const float dist = sqrtf(dot3(dx, dy, dz));
            if (dist >= b.a + pr) return false;
            nx = dx / dist;
            ny = dy / dist;
            nz = dz / dist;
            f.dir(nx, ny, nz);
            depth = b.a + pr - dist;
            if (cp) {
                // seg + diff.normalize_or_zero() * r (glam: diff * (1 / |diff|))
                const float rcp = 1.0f / dist;
                const bool ok = __builtin_isfinite(rcp) && rcp > 0.0f;
                const float ux = ok ? dx * rcp : 0.0f, uy = ok ? dy * rcp : 0.0f, uz = ok ? dz * rcp : 0.0f;
                f.point(0.0f + ux * b.a, sy + uy * b.a, 0.0f + uz * b.a, cp[0], cp[1], cp[2]);
            }
            return true;
        }
        float rx = 1.0f, ry = 0.0f, rz = 0.0f;
        if (dot3(lx, 0.0f, lz) > PH_EPS_SQ) {
            const float rcp = 1.0f / sqrtf(dot3(lx, 0.0f, lz));  // glam normalize_or_zero
            if (__builtin_isfinite(rcp) && rcp > 0.0f) {
                rx = lx * rcp;
                ry = 0.0f * rcp;
                rz = lz * rcp;
            } else {
                rx = ry = rz = 0.0f;
            }
        }
        nx = rx;
        ny = ry;
        nz = rz;
        f.dir(nx, ny, nz);
        depth = pr + b.a - sqrtf(dot3(dx, dy, dz));
        if (cp) f.point(0.0f + rx * b.a, sy + ry * b.a, 0.0f + rz * b.a, cp[0], cp[1], cp[2]);
        return true;
    }
    return false;
}

// fluid3.rs:1141-1179 for one particle against the scene's bodies, static
// ones (their impulses are no-ops, body.rs:1349-1352, and their velocity
// is zero).  A dynamic body in contact (depth > 0) takes this particle's
// impulses in particle order and its friction reads the body's running
// velocity: with `dynamic` given the loop stops there and sets *dynamic
// (the caller keeps the state from before the loop and hands the particle
// to the ordered host pass); without it, a dynamic body acts as a static one.
__device__ __forceinline__ void clamp_to_bodies(const SceneDev &s, const BodyDev *__restrict__ bodies,
                                                float &x, float &y, float &z, float &vx, float &vy,
                                                float &vz, bool *dynamic = nullptr) {
    const float pr = rmaxf(s.particle_radius, 0.0f);
    for (uint32_t k = 0; k < s.body_count; ++k) {
        const BodyDev b = bodies[s.body_begin + k];
        float nx, ny, nz, depth;
        if (!collide(x, y, z, pr, b, nx, ny, nz, depth)) continue;
        if (depth <= 0.0f) continue;
        if (dynamic && (b.flags & BODY_DYNAMIC)) {
            *dynamic = true;
            return;
        }
        x = x + nx * depth;
        y = y + ny * depth;
        z = z + nz * depth;
        const float vn = vx * nx + vy * ny + vz * nz;
        if (vn < 0.0f) {
            vx = vx - nx * vn;
            vy = vy - ny * vn;
            vz = vz - nz * vn;
            if (s.wall_friction > 0.0f) {
                const float mu = s.wall_mu;
                // Static wall: velocity zero, so wall_vt = 0 - n * (0 . n).
                const float wd = 0.0f * nx + 0.0f * ny + 0.0f * nz;
                const float wtx = 0.0f - nx * wd, wty = 0.0f - ny * wd, wtz = 0.0f - nz * wd;
                const float dvx = (wtx - vx) * mu, dvy = (wty - vy) * mu, dvz = (wtz - vz) * mu;
                vx = vx + dvx;
                vy = vy + dvy;
                vz = vz + dvz;
            }
        }
    }
}

// ---------------------------------------------------------------- rows
// True when the particle's neighbourhood leaves the domain while the scene
// has outliers (then the per-cell path of for_each_candidate is needed).
__device__ __forceinline__ bool needs_cell_path(const SceneDev &s, const uint32_t *__restrict__ start, int32_t cx,
                                                int32_t cy, int32_t cz) {
    const bool inside = in_domain(s, (int64_t)cx - 1, (int64_t)cy - 1, (int64_t)cz - 1) &&
                        in_domain(s, (int64_t)cx + 1, (int64_t)cy + 1, (int64_t)cz + 1);
    return !inside && start[s.bucket_base + s.ncells] != start[s.bucket_base + s.ncells + 1];
}

// Slot range [b, e) of row r = 3 (oz + 1) + (oy + 1) (empty outside the domain).
__device__ __forceinline__ void row_range(const SceneDev &s, const uint32_t *__restrict__ start, int32_t cx,
                                          int32_t cy, int32_t cz, int r, uint32_t &b, uint32_t &e) {
    const int64_t xlo = max((int64_t)cx - 1, (int64_t)s.x0);
    const int64_t xhi = min((int64_t)cx + 1, (int64_t)s.x0 + s.nx - 1);
    const int64_t z = (int64_t)cz + (r / 3) - 1, y = (int64_t)cy + (r % 3) - 1;
    b = e = 0;
    if (xlo > xhi || z < s.z0 || z >= (int64_t)s.z0 + s.nz || y < s.y0 || y >= (int64_t)s.y0 + s.ny) return;
    const uint32_t row = s.bucket_base + (uint32_t)(((z - s.z0) * s.ny + (y - s.y0)) * s.nx);
    b = start[row + (uint32_t)(xlo - s.x0)];
    e = start[row + (uint32_t)(xhi - s.x0) + 1];
}

#define PH_ROW_BITS 32u  // candidates per row tested into one mask word

// Neighbour list (the plain force pass's, or the surrogate's skin list):
// entries per particle at most, and the count word's flags.
#define SL_CAP 96u
#define SLF_COUNT 0x00FFFFFFu
#define SLF_SLOW 0x01000000u  // sweep the cells (outliers near the domain, or > SL_CAP candidates)
#define SLF_NONE 0x02000000u  // non-finite position: no neighbours

// ---------------------------------------------------------------- Verlet list
// A plain world keeps, per particle, every particle within (1 + skin) h of
// it at the last rebuild (a Verlet list), so a substep tests ~45 listed
// candidates instead of the ~216 of its 27 cells.  Between rebuilds the
// particles keep their slots (the sort is skipped), so a row holds slots
// directly.  The reference's neighbour condition and order are recovered
// from one word per particle, its neighbour word: the original index (24
// bits) and the cell coordinates mod 4.  Two particles within h lie at most
// two cells apart per axis, so the difference mod 4 tells the offset
// exactly: 2 means not among the 27 cells, otherwise -1, 0, 1; and the
// reference visits neighbours by offset (oz, oy, ox), then by index
// (fluid3.rs:301-314, 326-338).
#define VL_CAP 96u                 // entries per row (4 to a 16-byte group)
#define VLF_SLOW 0x01000000u       // row unusable (longer than VL_CAP, or a per-cell particle)
#define VL_FAR 0x1p24f             // |x / h| from which cells are not compared mod 4
#define VL_NOT27 0xFFFFFFFFu
// A tile's halo: the union of its rows, staged in LDS by the walk.  At most
// VL_HALO_CAP particles over a slot span of at most VL_HALO_SPAN; a tile
// beyond either walks from global memory (VL_HALO_NONE).
#define VL_HALO_CAP 1792u
#define VL_HALO_SPAN 262144u  // 2^18 slots: cell layers of up to ~2^17 particles
#define VL_HALO_NONE 0xFFFFFFFFu

// State of the Verlet rows.  A field with two entries is indexed by the
// substep's parity p: substep t writes entry t & 1 and substep t + 1 reads
// it, so each substep's decision reads what no kernel of that substep
// writes.
struct VerletDev {
    uint32_t rebuild;         // this substep sorts and rebuilds every row
    uint32_t pad;
    float dacc[2];            // bound on any particle's displacement since the last rebuild, before substep t moved
    uint32_t slow[2];         // a particle needs the per-cell path: substep t + 1 rebuilds too
    uint32_t dmax2[2];        // largest squared displacement of substep t (float bits, >= 0)
    uint32_t qsum[2];         // rows put back in order in substep t
    unsigned long long rebuilds, repaired, substeps;  // statistics
};

// What the host hands a Verlet substep's first kernel: whether it must
// rebuild (the first of a run of Verlet substeps, or after the particles
// changed), its parity, the displacement bound, and where the statistics
// go (pinned host memory: [0] rows put back in order in the last substep,
// [1] this substep's rebuild flag).
struct VlStep {
    uint32_t forced;
    uint32_t par;
    float thr;
    uint32_t *host_out;
};

// This substep's decision, the same in every block that takes it: rebuild
// when forced, when a particle needed the per-cell path, or when the bound
// on any particle's displacement since the last rebuild (the sum over the
// substeps of the largest displacement) passes thr (a NaN bound rebuilds).
__device__ __forceinline__ bool vl_decide(const VerletDev *vs, const VlStep &st, float &bound) {
    const uint32_t q = st.par ^ 1u;
    bound = vs->dacc[q] + sqrtf(__uint_as_float(vs->dmax2[q]));
    return st.forced != 0u || vs->slow[q] != 0u || !(bound <= st.thr);
}

// One thread of the substep's first kernel: the decision for the later
// kernels, the entries this substep accumulates into cleared, the
// statistics.
__device__ __forceinline__ void vl_commit(VerletDev *vs, const VlStep &st, bool rb, float bound) {
    const uint32_t p = st.par, q = p ^ 1u;
    const uint32_t qs = st.forced ? 0u : vs->qsum[q];
    vs->rebuild = rb ? 1u : 0u;
    vs->dacc[p] = rb ? 0.0f : bound;
    vs->slow[p] = 0u;
    vs->dmax2[p] = 0u;
    vs->qsum[p] = 0u;
    vs->rebuilds += rb ? 1ull : 0ull;
    vs->repaired += qs;
    vs->substeps += 1ull;
    st.host_out[0] = qs;
    st.host_out[1] = rb ? 1u : 0u;
}

// Neighbour word of a particle: its original index (24 bits) and its cell
// coordinates mod 4, x at bit 24, y at bit 27, z at bit 30, a zero bit
// above the x and y fields.
__device__ __forceinline__ uint32_t vl_word(uint32_t id, int32_t cx, int32_t cy, int32_t cz) {
    return (id & PH_ID_MASK) | (((uint32_t)cx & 3u) << 24) | (((uint32_t)cy & 3u) << 27) |
           (((uint32_t)cz & 3u) << 30);
}
#define VL_KEY_MASK 0xDBFFFFFFu  // the index and the three fields (bits 26 and 29 clear)
#define VL_FIELD_LO 0x49000000u  // the low bit of each field
// Particle i's addend: (1 - c_i) mod 4 in each field.  Added to j's word it
// leaves (c_j - c_i + 1) mod 4 in each field (a field's carry goes to the
// zero bit above it, or out of the word): 0, 1, 2 for the cell offsets -1,
// 0, +1, and 3 for an offset of 2 (two particles within h lie at most two
// cells apart per axis).
__device__ __forceinline__ uint32_t vl_addend(uint32_t wi) {
    return (((1u - (wi >> 24)) & 3u) << 24) | (((1u - (wi >> 27)) & 3u) << 27) | (((1u - (wi >> 30)) & 3u) << 30);
}
// Visiting key of j among i's neighbours: the offsets (z, y, x) ranked 0..2,
// then the index, so keys increase in the reference order (fluid3.rs:326-338);
// VL_NOT27 when j's cell is not one of i's 27.  Valid for pairs within h.
__device__ __forceinline__ uint32_t vl_key(uint32_t ai, uint32_t wj) {
    const uint32_t t = (wj + ai) & VL_KEY_MASK;
    return (t & (t >> 1) & VL_FIELD_LO) ? VL_NOT27 : t;
}
__device__ __forceinline__ bool vl_far(float x, float y, float z, float h) {
    const float f = VL_FAR * h;
    return !(fabsf(x) < f && fabsf(y) < f && fabsf(z) < f);
}

// Blocks are dispatched round-robin over the 8 XCDs (block b -> XCD b % 8),
// each with its own L2.  Give every XCD one contiguous run of tiles, so
// the neighbour rows a tile reads were mostly fetched by tiles on the same
// XCD.  A bijection on [0, nt).
__device__ __forceinline__ uint32_t xcd_tile(uint32_t b, uint32_t nt) {
    const uint32_t x = b & 7u, k = b >> 3, per = nt >> 3, rem = nt & 7u;
    return x * per + min(x, rem) + k;
}

// ---------------------------------------------------------------- ghosts
// Boundary ghosts of one scene (fluid3.rs:716-920): positions sorted by
// ghost cell with the ghost index ascending inside a cell (xyz, mass), the
// per-substep state (clamped density, pressure, q = p / max(rho, EPS)^2),
// and an open-addressing table from a cell to its run of ghosts (an entry
// with range.y == 0 is empty; the table is at most half full).
struct GhostDev {
    const float4 *pos = nullptr;
    float4 *state = nullptr;
    const int4 *key = nullptr;
    const uint2 *range = nullptr;
};

__host__ __device__ inline uint32_t ghost_hash(int32_t x, int32_t y, int32_t z) {
    uint32_t h = (uint32_t)x * 0x9E3779B1u;
    h ^= (uint32_t)y * 0x85EBCA77u;
    h ^= (uint32_t)z * 0xC2B2AE3Du;
    return h ^ (h >> 15);
}

// Run [b, e) of the ghosts in cell (x, y, z); false when the cell has none.
// The probe visits at most mask + 1 slots and the table always has an
// empty one, so it ends.
__device__ __forceinline__ bool ghost_cell(const GhostDev &g, uint32_t mask, int32_t x, int32_t y, int32_t z,
                                           uint32_t &b, uint32_t &e) {
    uint32_t slot = ghost_hash(x, y, z) & mask;
    for (uint32_t probe = 0; probe <= mask; ++probe) {
        const uint2 r = g.range[slot];
        if (r.y == 0u) return false;
        const int4 k = g.key[slot];
        if (k.x == x && k.y == y && k.z == z) {
            b = r.x;
            e = r.y;
            return true;
        }
        slot = (slot + 1u) & mask;
    }
    return false;
}

// Ghost pressure on a fluid particle after its fluid neighbours
// (fluid3.rs:1074-1084): the ghosts of the 27 cells around it at the raw
// smoothing length (boundary_near, :860-880: oz -> oy -> ox, ascending ghost
// index, d2 < h * h), each f -= grad W * m_b (q_i + q_b).
__device__ __forceinline__ void ghost_force(const SceneDev &s, const GhostDev &g, const float4 &pi, float q_i,
                                            float &fx, float &fy, float &fz) {
    int32_t cx, cy, cz;
    cell_of(pi.x, pi.y, pi.z, s.h_grid, cx, cy, cz);
    const float h = s.h;
    for (int oz = -1; oz <= 1; ++oz)
        for (int oy = -1; oy <= 1; ++oy)
            for (int ox = -1; ox <= 1; ++ox) {
                const int64_t x = (int64_t)cx + ox, y = (int64_t)cy + oy, z = (int64_t)cz + oz;
                if (x < INT32_MIN || x > INT32_MAX || y < INT32_MIN || y > INT32_MAX || z < INT32_MIN ||
                    z > INT32_MAX)
                    continue;
                uint32_t b, e;
                if (!ghost_cell(g, s.ghost_mask, (int32_t)x, (int32_t)y, (int32_t)z, b, e)) continue;
                for (uint32_t k = b; k < e; ++k) {
                    const float4 gp = g.pos[k];
                    const float d2 = dot3(gp.x - pi.x, gp.y - pi.y, gp.z - pi.z);
                    if (!(d2 < s.h_grid2)) continue;
                    // delta = x_i - x_b; |delta|^2 is d2 exactly
                    const float ex = pi.x - gp.x, ey = pi.y - gp.y, ez = pi.z - gp.z;
                    const float dist = sqrtf(d2);
                    float gx = 0.0f, gy = 0.0f, gz = 0.0f;
                    if (!(dist >= h || dist < PH_EPS)) {
                        const float sc = s.spiky_c * (h - dist) * (h - dist);
                        gx = ex / dist * sc;
                        gy = ey / dist * sc;
                        gz = ez / dist * sc;
                    }
                    const float sm = gp.w * (q_i + g.state[k].z);
                    fx = fx - gx * sm;
                    fy = fy - gy * sm;
                    fz = fz - gz * sm;
                }
            }
}

// Surrogate neighbour record word: P = p / rho^2 with the phase (0..3) in
// its two lowest mantissa bits (a relative change below 2^-21 of P, which
// the surrogate sweep reads as is).
__device__ __forceinline__ float pack_pq(float q, uint32_t phase) {
    return __uint_as_float((__float_as_uint(q) & ~3u) | min(phase, 3u));
}

// The other 8 bytes of a neighbour record: the velocity as three halves
// (rounded toward zero, so magnitudes saturate at 65504 instead of
// overflowing) and dP as a bfloat16 (rounded to nearest even; it keeps the
// f32 exponent range, which a pressure quotient of a light fluid needs).
typedef _Float16 half2_t __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint2 pack_vd(float vx, float vy, float vz, float dq) {
    const uint32_t u = __float_as_uint(dq);
    const uint32_t bq = (u + 0x7fffu + ((u >> 16) & 1u)) & 0xffff0000u;
    const uint32_t hz = __builtin_bit_cast(uint32_t, __builtin_amdgcn_cvt_pkrtz(vz, 0.0f)) & 0xffffu;
    return make_uint2(__builtin_bit_cast(uint32_t, __builtin_amdgcn_cvt_pkrtz(vx, vy)), hz | bq);
}
__device__ __forceinline__ void unpack_vd(uint2 w, float &vx, float &vy, float &vz, float &dq) {
    const half2_t a = __builtin_bit_cast(half2_t, w.x), b = __builtin_bit_cast(half2_t, w.y);
    vx = (float)a.x;
    vy = (float)a.y;
    vz = (float)b.x;
    dq = __uint_as_float(w.y & 0xffff0000u);
}

// The judge's sample of slots: runs of 64 consecutive slots (one wave's
// worth, so the sample keeps the neighbour locality of the slot order),
// every stride-th run from run off.  Sample k is slot judge_slot(k); slot
// o is in the sample when judge_sample(o) returns true, with its k.
__host__ __device__ __forceinline__ uint32_t judge_slot(uint32_t k, uint32_t stride, uint32_t off) {
    return ((k >> 6) * stride + off) * 64u + (k & 63u);
}
__host__ __device__ __forceinline__ bool judge_sample(uint32_t o, uint32_t stride, uint32_t off, uint32_t &k) {
    const uint32_t run = o >> 6;
    if (run % stride != off) return false;
    k = (run / stride) * 64u + (o & 63u);
    return true;
}

// Neighbour records of one generation of the surrogate's state: per slot
// j, pos[j] = (x, y, z, P | phase) and vd[j] = pack_vd(v, dP).
struct SurRec {
    float4 *pos = nullptr;
    uint2 *vd = nullptr;
};

// What a judging plain substep of a surrogate world also writes from the
// force pass (one pass instead of a history gather and an update kernel):
// the new history (rho, P, P - P_prev) of each sorted slot, with P_prev read
// through the sort's permutation from the previous order, the neighbour
// records of the advanced state, and the judge's per-tile sums.
struct AuxFuse {
    const uint32_t *perm = nullptr;   // sorted slot -> slot before the sort (null: no sort, the same slot)
    const float4 *aux_in = nullptr;   // history in the order before the sort
    float4 *aux_out = nullptr;        // null: nothing fused
    SurRec rec_out;                   // pos may be null: no records
    // per tile: (sum |rho/rho0 - 1|, count) of the bulk, and over the
    // particles the surrogate was evaluated on before this substep
    // (previous slot o in the judge sample of (sur_stride, sur_off), its
    // acceleration at sur_acc[k], k its sample index; sur_acc null: none)
    // the sums of |a_plain - a_sur| and |a_plain| (accelerations without
    // gravity)
    float4 *tile_err = nullptr;
    const float4 *sur_acc = nullptr;
    uint32_t sur_stride = 16u, sur_off = 0u;
    // particles in contact with a dynamic body: defer[0] counts them,
    // defer[1 + k] (k < defer_cap) holds their slots; their state is left
    // as before the boundary clamp (null: no dynamic bodies)
    uint32_t *defer = nullptr;
    uint32_t defer_cap = 0;
};

