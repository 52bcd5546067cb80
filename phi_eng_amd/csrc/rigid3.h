// rigid3.h — the rigid-body side of the two-way fluid coupling (host).
//
// A dynamic body's state and the f32 operations the reference applies to
// it in a world step (crates/dynamics/src/world.rs:2766-2812 as far as the
// fluid reaches it): the gravity kick (gravity.rs:348-364), the particles'
// impulses and the buoyancy forces of the fluid step (fluid3.rs:1141-1213,
// body.rs:1334-1361) and the integration (body.rs:1372-1406).  Vector,
// quaternion and matrix arithmetic restates glam 0.29.3 (the reference's
// math library, not vendored): Vec3 and Mat3 are its scalar forms, Quat its
// SSE2 form (x86-64), each operation's association order kept.
#pragma once
#include <cmath>
#include <cstdint>
#include <vector>

#include "sph_device.h"

namespace ph {

struct F3 {
    float x = 0.0f, y = 0.0f, z = 0.0f;
};
inline F3 f3(float x, float y, float z) {
    F3 r;
    r.x = x;
    r.y = y;
    r.z = z;
    return r;
}
inline F3 operator+(F3 a, F3 b) { return f3(a.x + b.x, a.y + b.y, a.z + b.z); }
inline F3 operator-(F3 a, F3 b) { return f3(a.x - b.x, a.y - b.y, a.z - b.z); }
inline F3 operator*(F3 a, float s) { return f3(a.x * s, a.y * s, a.z * s); }
inline F3 operator/(F3 a, float s) { return f3(a.x / s, a.y / s, a.z / s); }
inline F3 operator-(F3 a) { return f3(-a.x, -a.y, -a.z); }
inline float dot(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline F3 cross(F3 a, F3 b) { return f3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
inline float length(F3 a) { return std::sqrt(dot(a, a)); }
// glam normalize_or_zero: v * (1 / |v|) when that factor is finite and > 0
inline F3 normalize_or_zero(F3 a) {
    const float rcp = 1.0f / length(a);
    return (std::isfinite(rcp) && rcp > 0.0f) ? a * rcp : F3{};
}

struct Q4 {
    float x = 0.0f, y = 0.0f, z = 0.0f, w = 1.0f;
};
inline F3 rotate(Q4 q, F3 v) {
    F3 r;
    quat_rotate(q.x, q.y, q.z, q.w, v.x, v.y, v.z, r.x, r.y, r.z);
    return r;
}
// glam Quat * Quat, SSE2 form: the four partial products of each lane are
// summed as (w-term + x-term) + (y-term + z-term).
inline Q4 qmul(Q4 a, Q4 b) {
    Q4 r;
    r.x = (a.w * b.x + a.x * b.w) + (a.y * b.z - a.z * b.y);
    r.y = (a.w * b.y - a.x * b.z) + (a.y * b.w + a.z * b.x);
    r.z = (a.w * b.z + a.x * b.y) + (a.z * b.w - a.y * b.x);
    r.w = (a.w * b.w - a.x * b.x) - (a.y * b.y + a.z * b.z);
    return r;
}
// glam Quat::normalize (Vec4, SSE2): divide by sqrt((xx + zz) + (yy + ww)).
inline Q4 qnormalize(Q4 q) {
    const float len = std::sqrt((q.x * q.x + q.z * q.z) + (q.y * q.y + q.w * q.w));
    Q4 r;
    r.x = q.x / len;
    r.y = q.y / len;
    r.z = q.z / len;
    r.w = q.w / len;
    return r;
}

struct M3 {  // columns
    F3 c0 = f3(1.0f, 0.0f, 0.0f), c1 = f3(0.0f, 1.0f, 0.0f), c2 = f3(0.0f, 0.0f, 1.0f);
};
inline F3 operator*(const M3 &m, F3 v) { return (m.c0 * v.x + m.c1 * v.y) + m.c2 * v.z; }
inline M3 operator*(const M3 &a, const M3 &b) {
    M3 r;
    r.c0 = a * b.c0;
    r.c1 = a * b.c1;
    r.c2 = a * b.c2;
    return r;
}
inline M3 transpose(const M3 &m) {
    M3 r;
    r.c0 = f3(m.c0.x, m.c1.x, m.c2.x);
    r.c1 = f3(m.c0.y, m.c1.y, m.c2.y);
    r.c2 = f3(m.c0.z, m.c1.z, m.c2.z);
    return r;
}
inline M3 diag(float a, float b, float c) {
    M3 r;
    r.c0 = f3(a, 0.0f, 0.0f);
    r.c1 = f3(0.0f, b, 0.0f);
    r.c2 = f3(0.0f, 0.0f, c);
    return r;
}
// glam Mat3::inverse: cofactor columns from cross products, det = c2 . (c0 x c1).
inline M3 inverse(const M3 &m) {
    const F3 t0 = cross(m.c1, m.c2), t1 = cross(m.c2, m.c0), t2 = cross(m.c0, m.c1);
    const float inv = 1.0f / dot(m.c2, t2);
    M3 r;
    r.c0 = t0 * inv;
    r.c1 = t1 * inv;
    r.c2 = t2 * inv;
    return transpose(r);
}
// glam Mat3::from_quat
inline M3 from_quat(Q4 q) {
    const float x2 = q.x + q.x, y2 = q.y + q.y, z2 = q.z + q.z;
    const float xx = q.x * x2, xy = q.x * y2, xz = q.x * z2;
    const float yy = q.y * y2, yz = q.y * z2, zz = q.z * z2;
    const float wx = q.w * x2, wy = q.w * y2, wz = q.w * z2;
    M3 r;
    r.c0 = f3(1.0f - (yy + zz), xy + wz, xz - wy);
    r.c1 = f3(xy - wz, 1.0f - (xx + zz), yz + wx);
    r.c2 = f3(xz + wy, yz - wx, 1.0f - (xx + yy));
    return r;
}

// crates/math/src/trig.rs:129-164 compute_cos_sin: Cody-Waite reduction by
// pi/2 in f64, Taylor polynomials of sin and cos on [-pi/4, pi/4], quadrant
// swap, each rounded to f32.
inline void cos_sin(float angle, float &c, float &s) {
    if (!std::isfinite(angle)) {
        c = s = NAN;
        return;
    }
    const double kInvPio2 = 2.0 / 3.14159265358979323846;  // FRAC_2_PI
    const double kPio2Hi = 1.570796012878418, kPio2Lo = 3.13916478589249e-07;
    const double a = (double)angle;
    const double n = std::round(a * kInvPio2);
    const double r = (a - n * kPio2Hi) - n * kPio2Lo;
    const double z = r * r;
    double p = -1.0 / 39916800.0;
    p = 1.0 / 362880.0 + z * p;
    p = -1.0 / 5040.0 + z * p;
    p = 1.0 / 120.0 + z * p;
    p = -1.0 / 6.0 + z * p;
    const double sd = r * (1.0 + z * p);
    double q = 1.0 / 479001600.0;
    q = -1.0 / 3628800.0 + z * q;
    q = 1.0 / 40320.0 + z * q;
    q = -1.0 / 720.0 + z * q;
    q = 1.0 / 24.0 + z * q;
    q = -0.5 + z * q;
    const double cd = 1.0 + z * q;
    double rem = std::fmod(n, 4.0);
    if (rem < 0.0) rem += 4.0;
    if (rem == 1.0) {
        c = (float)-sd;
        s = (float)cd;
    } else if (rem == 2.0) {
        c = (float)-cd;
        s = (float)-sd;
    } else if (rem == 3.0) {
        c = (float)sd;
        s = (float)-cd;
    } else {
        c = (float)cd;
        s = (float)sd;
    }
}

struct Rigid3 {
    uint32_t shape = PHIENG_SHAPE3_BOX_TAG;
    float a = 0.0f, b = 0.0f, c = 0.0f;
    bool is_static = true;
    bool sleeping = false;
    F3 pos;
    Q4 rot;
    F3 vel, ang;
    float mass = 1.0f, inv_mass = 1.0f;  // RigidBody3::new (body.rs:1227-1252)
    M3 inv_inertia;                      // body frame
    float linear_damping = 0.01f, angular_damping = 0.01f, gravity_scale = 1.0f;

    static constexpr uint32_t PHIENG_SHAPE3_BOX_TAG = 1u;

    // body.rs:1266-1300 set_shape + mass_properties_3d (body.rs:811-841)
    // and InertiaTensor3::capsule (crates/math/src/math.rs:148-160).
    void set_shape(uint32_t tag, float sa, float sb, float sc, float density) {
        shape = tag;
        a = sa;
        b = sb;
        c = sc;
        const float pi = 3.14159265358979323846f;
        const float four_thirds = 4.0f / 3.0f;
        float m = 0.0f, ix = 1.0f, iy = 1.0f, iz = 1.0f;
        if (tag == 0u) {
            const float volume = four_thirds * pi * sa * sa * sa;
            m = volume * density;
            ix = iy = iz = 0.4f * m * sa * sa;
        } else if (tag == 1u) {
            const float w = sa * 2.0f, h = sb * 2.0f, d = sc * 2.0f;
            const float volume = w * h * d;
            m = volume * density;
            ix = m * (h * h + d * d) / 12.0f;
            iy = m * (w * w + d * d) / 12.0f;
            iz = m * (w * w + h * h) / 12.0f;
        } else {
            const float r = sa, hl = sb;
            const float length = hl * 2.0f;
            const float volume = pi * r * r * length + four_thirds * pi * r * r * r;
            m = volume * density;
            const float v_cyl = pi * r * r * length;
            const float v_sph = four_thirds * pi * r * r * r;
            const float den = m / std::fmax(v_cyl + v_sph, 1e-12f);
            const float m_cyl = den * v_cyl, m_sph = den * v_sph;
            const float transverse =
                m_cyl * (3.0f * r * r + length * length) / 12.0f + m_sph * (0.4f * r * r + hl * hl);
            const float axial = 0.5f * m_cyl * r * r + 0.4f * m_sph * r * r;
            ix = iz = transverse;
            iy = axial;
        }
        mass = m > PH_EPS ? m : PH_EPS;  // mass.max(MIN_DENSITY), MIN_DENSITY = EPS (material.rs:11)
        if (is_static) {
            inv_mass = 0.0f;
            inv_inertia = diag(0.0f, 0.0f, 0.0f);
        } else {
            inv_mass = 1.0f / mass;
            inv_inertia = inverse(diag(ix, iy, iz));
        }
    }

    // body.rs:1328-1332: R I^-1 R^T
    M3 world_inv_inertia() const {
        const M3 r = from_quat(rot);
        return (r * inv_inertia) * transpose(r);
    }
    // body.rs:1349-1361
    void apply_impulse(F3 j, F3 point) {
        if (is_static) return;
        sleeping = false;
        vel = vel + j * inv_mass;
        const F3 torque = cross(point - pos, j);
        ang = ang + world_inv_inertia() * torque;
    }
    // body.rs:1334-1346
    void apply_force(F3 f, F3 point, float dt) {
        if (is_static) return;
        sleeping = false;
        vel = vel + f * inv_mass * dt;
        const F3 torque = cross(point - pos, f);
        ang = ang + world_inv_inertia() * torque * dt;
    }
    // gravity.rs:355-364 (uniform field, Euler kick)
    void gravity_kick(F3 g, float h) {
        if (is_static || sleeping) return;
        vel = vel + g * (gravity_scale * h);
    }
    // body.rs:1390-1406 and 1372-1387
    void integrate(float dt) {
        if (is_static || sleeping) return;
        const float damping = linear_damping * dt;
        vel = vel * (1.0f - std::fmin(damping, 1.0f));
        const float ang_damping = angular_damping * dt;
        ang = ang * (1.0f - std::fmin(ang_damping, 1.0f));
        pos = pos + vel * dt;
        const float angle = length(ang) * dt;
        if (angle > 1e-6f) {
            const F3 axis = ang / length(ang);
            const float half = angle * 0.5f;
            float ch, sh;
            cos_sin(half, ch, sh);
            Q4 d;
            d.x = axis.x * sh;
            d.y = axis.y * sh;
            d.z = axis.z * sh;
            d.w = ch;
            rot = qnormalize(qmul(d, rot));
        }
    }

    // The device record: position, rotation (flagged when not exactly the
    // identity), dynamic flag.
    BodyDev dev() const {
        BodyDev d{};
        d.shape = shape;
        d.px = pos.x;
        d.py = pos.y;
        d.pz = pos.z;
        d.a = a;
        d.b = shape == 0u ? 0.0f : b;
        d.c = shape == 1u ? c : 0.0f;
        d.qx = rot.x;
        d.qy = rot.y;
        d.qz = rot.z;
        d.qw = rot.w;
        const bool ident = rot.x == 0.0f && rot.y == 0.0f && rot.z == 0.0f && rot.w == 1.0f;
        d.flags = (is_static ? 0u : BODY_DYNAMIC) | (ident ? 0u : BODY_ROTATED);
        return d;
    }
    // A radius about pos that holds the whole shape.
    float bound() const {
        if (shape == 0u) return a;
        if (shape == 1u) return std::sqrt(a * a + b * b + c * c);
        return a + b;
    }
};

struct Sample3 {
    F3 point, normal;
    float area;
};

// fluid3.rs:375-465 surface_samples: 14 sphere directions, 4 samples per
// box face, 2 rings of 8 and the poles of a capsule.  The box's samples
// of both faces of an axis lie in the face plane at -half extent (the
// reference's table gives both the same fixed coordinate).  The capsule
// ring directions use the platform cosf / sinf, as the reference's
// f32::cos / f32::sin do.
inline void surface_samples(const Rigid3 &bd, std::vector<Sample3> &out) {
    out.clear();
    const float pi = 3.14159265358979323846f;
    // the identity rotation is skipped, as the collision test does
    const bool ident = bd.rot.x == 0.0f && bd.rot.y == 0.0f && bd.rot.z == 0.0f && bd.rot.w == 1.0f;
    auto push = [&](F3 local, F3 n, float area) {
        Sample3 s;
        s.point = (ident ? local : rotate(bd.rot, local)) + bd.pos;
        s.normal = ident ? n : rotate(bd.rot, n);
        s.area = area;
        out.push_back(s);
    };
    if (bd.shape == 0u) {
        const F3 dirs[14] = {f3(1, 0, 0), f3(-1, 0, 0), f3(0, 1, 0), f3(0, -1, 0), f3(0, 0, 1), f3(0, 0, -1),
                             normalize_or_zero(f3(1, 1, 1)), normalize_or_zero(f3(1, 1, -1)),
                             normalize_or_zero(f3(1, -1, 1)), normalize_or_zero(f3(1, -1, -1)),
                             normalize_or_zero(f3(-1, 1, 1)), normalize_or_zero(f3(-1, 1, -1)),
                             normalize_or_zero(f3(-1, -1, 1)), normalize_or_zero(f3(-1, -1, -1))};
        const float area = 4.0f * pi * bd.a * bd.a / 14.0f;
        for (const F3 &d : dirs) push(d * bd.a, d, area);
    } else if (bd.shape == 1u) {
        const float hx = bd.a, hy = bd.b, hz = bd.c;
        const float area = 8.0f * (hx * hy + hy * hz + hz * hx) / 24.0f;
        struct Face {
            F3 n;
            float fixed;
            int axis;
        };
        const Face faces[6] = {{f3(0, 1, 0), -hy, 0},  {f3(0, -1, 0), -hy, 0}, {f3(1, 0, 0), -hx, 1},
                               {f3(-1, 0, 0), -hx, 1}, {f3(0, 0, 1), -hz, 2},  {f3(0, 0, -1), -hz, 2}};
        const float ts[4][2] = {{-0.75f, -0.75f}, {0.75f, -0.75f}, {-0.75f, 0.75f}, {0.75f, 0.75f}};
        for (const Face &fc : faces)
            for (int k = 0; k < 4; ++k) {
                const float t0 = ts[k][0], t1 = ts[k][1];
                F3 local;
                if (fc.axis == 0) local = f3(hx * t0, fc.fixed, hz * t1);
                else if (fc.axis == 1) local = f3(fc.fixed, hy * t0, hz * t1);
                else local = f3(hx * t0, hy * t1, fc.fixed);
                push(local, fc.n, area);
            }
    } else if (bd.shape == 2u) {
        const float r = bd.a, hl = bd.b;
        const float side_area = 4.0f * pi * r * hl / 16.0f;
        const float cap_area = 2.0f * pi * r * r;
        const float sys[2] = {-hl * 0.5f, hl * 0.5f};
        for (float sy : sys)
            for (int k = 0; k < 8; ++k) {
                const float an = 2.0f * pi * (float)k / 8.0f;
                const F3 dir = f3(std::cos(an), 0.0f, std::sin(an));
                push(dir * r + f3(0.0f, 1.0f, 0.0f) * sy, dir, side_area);
            }
        const float ss[2] = {-1.0f, 1.0f};
        for (float sgn : ss) push(f3(0.0f, 1.0f, 0.0f) * (sgn * (hl + r)), f3(0.0f, 1.0f, 0.0f) * sgn, cap_area);
    }
}

}  // namespace ph
