// snapshot.hip — save / load of a GPU world in the reference's snapshot
// formats (crates/dynamics/src/snapshot.rs; C ABI crates/engine/src/ffi.rs:
// 4036-4117).  Host code only.
//
// Both formats carry the reference's World3Dto (snapshot.rs:3189-3250):
//   JSON    serde's field names, optional fields elided at their defaults
//           exactly where the reference elides them;
//   binary  the PHE1 envelope (magic, schema u32, kind u8 = 3, payload
//           length u32, payload, CRC-32 of the payload; :5570-5616) around
//           the payload of encode_world3 (:7920-8126), little-endian.
// A world is written at the lowest schema its features need, as world3_dto
// does (:4230-4318): 3, or 4 with a gravity configuration, 5 with any of
// the fluid options or ambient forces, 6 with CCD, 7 with the solver
// switches.  Every feature of a higher schema (thermal, collision tiers,
// nuclear, sensors, quantum, motors, electromagnetism) lies outside the
// GPU fluid path, so a file of schema 8 or more is refused, as are
// rotated static bodies, dynamic bodies of a 2D world, hulls, joints and
// contacts.  Fields the GPU
// step does not read (solver, validation, CCD, ...) are kept verbatim.
// Dynamic bodies of a 3D world keep their rotation and velocities.
// Load resets density to rest and pressure to 0 (apply_fluid3,
// :4545-4615), which the GPU step recomputes anyway, and gives a particle
// of mass 0 the config's particle_mass.  Unlike the reference, a file may
// hold more than MAX_PARTICLES_3D particles (the capacity is lifted).
#include <hip/hip_runtime.h>

#include <cerrno>
#include <charconv>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/phieng_fluid2.h"
#include "../../include/phieng_fluid3.h"
#include "world3_internal.h"

void snapshot_set_error(const char *where, const std::string &what);

namespace {

constexpr uint32_t kSchemaMin = 3, kSchemaGravity = 4, kSchemaAmbient = 5, kSchemaCcd = 6, kSchemaSolverFlags = 7,
                   kSchemaMax = 17, kSchemaFirstUnsupported = 8;
constexpr uint8_t kKindWorld2 = 2, kKindWorld3 = 3;

struct SnapError : std::runtime_error {
    int code;  // -1 IO, -2 encode / parse / corrupt
    SnapError(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};
[[noreturn]] void fail(const std::string &m) { throw SnapError(-2, m); }

// CRC-32 (IEEE 802.3, reflected polynomial 0xEDB88320), the checksum the
// reference takes with crc32fast.
uint32_t crc32(const uint8_t *p, size_t n) {
    static uint32_t table[256];
    static bool init = false;
    if (!init) {
        for (uint32_t i = 0; i < 256; ++i) {
            uint32_t c = i;
            for (int k = 0; k < 8; ++k) c = (c & 1u) ? 0xEDB88320u ^ (c >> 1) : c >> 1;
            table[i] = c;
        }
        init = true;
    }
    uint32_t c = 0xFFFFFFFFu;
    for (size_t i = 0; i < n; ++i) c = table[(c ^ p[i]) & 0xFFu] ^ (c >> 8);
    return c ^ 0xFFFFFFFFu;
}

bool finite3(const float *v) { return std::isfinite(v[0]) && std::isfinite(v[1]) && std::isfinite(v[2]); }

// ---------------------------------------------------------------- schema

bool has_ambient(const SnapWorld &w) {
    const SnapMeta &m = w.meta;
    const FluidConfig3Ffi &c = w.cfg;
    bool scales = false;
    for (const SnapBody &b : w.bodies) scales |= b.ambient_scale != 1.0f;
    return m.spring_on || m.drag_on || m.turb_on || scales || c.gamma != 1.0f || c.artificial_viscosity != 0.0f ||
           c.surface_tension != 0.0f || c.cfl != 0.0f || c.max_substeps != 1 || c.adaptive_resolution ||
           c.wall_friction != 0.0f || c.phase_count > 1 || c.boundary_particles;
}
bool gravity_default(const SnapMeta &m) {
    return m.g_mode == 0 && m.g_integrator == 0 && m.big_g == 1.0f && m.softening == 0.0f && m.attractor[0] == 0.0f &&
           m.attractor[1] == 0.0f && m.attractor[2] == 0.0f && m.attractor_mu == 0.0f && m.theta == 0.5f;
}
bool spring_default(const SnapMeta &m) {
    return !m.spring_on && m.spring_anchor[0] == 0.0f && m.spring_anchor[1] == 0.0f && m.spring_anchor[2] == 0.0f &&
           m.spring_k == 0.0f && m.spring_c == 0.0f && m.spring_max == 0.0f;
}
bool drag_default(const SnapMeta &m) {
    return !m.drag_on && m.drag_linear == 0.0f && m.drag_quadratic == 0.0f && m.drag_air == 1.0f &&
           m.drag_angular == 0.0f && m.drag_wind[0] == 0.0f && m.drag_wind[1] == 0.0f && m.drag_wind[2] == 0.0f;
}
bool turb_default(const SnapMeta &m) {
    return !m.turb_on && m.turb_amplitude == 0.0f && m.turb_frequency == 1.0f && m.turb_time_scale == 1.0f &&
           m.turb_seed == 0 && m.turb_octaves == 3 && !m.turb_div_free;
}
bool ambient_default(const SnapMeta &m) { return spring_default(m) && drag_default(m) && turb_default(m); }
bool ccd_default(const SnapMeta &m) {
    return !m.ccd_on && m.ccd_tier == 0 && m.ccd_max_movement == 0.1f && m.ccd_max_substeps == 8;
}

// world3_dto's choice (snapshot.rs:4291-4318) for the features a GPU world
// can hold.
uint32_t schema_of(const SnapWorld &w) {
    bool grav = !gravity_default(w.meta);
    for (const SnapBody &b : w.bodies) grav |= b.gravity_scale != 1.0f;
    if (!w.meta.friction || !w.meta.restitution) return kSchemaSolverFlags;
    if (!ccd_default(w.meta)) return kSchemaCcd;
    if (has_ambient(w)) return kSchemaAmbient;
    if (grav) return kSchemaGravity;
    return kSchemaMin;
}

// ---------------------------------------------------------------- JSON out

struct JsonOut {
    std::string s;
    int depth = 0;
    bool first = true;
    void nl() {
        s += '\n';
        s.append(2 * (size_t)depth, ' ');
    }
    void key(const char *k) {
        if (!first) s += ',';
        first = false;
        nl();
        s += '"';
        s += k;
        s += "\": ";
    }
    void open(char c) {
        s += c;
        ++depth;
        first = true;
    }
    void close(char c) {
        --depth;
        if (!first) nl();
        s += c;
        first = false;
    }
    void f(float v) {
        char b[64];
        auto r = std::to_chars(b, b + sizeof(b), v);
        s.append(b, r.ptr);
    }
    void u(uint64_t v) { s += std::to_string(v); }
    void b(bool v) { s += v ? "true" : "false"; }
    void kf(const char *k, float v) {
        key(k);
        f(v);
    }
    void ku(const char *k, uint64_t v) {
        key(k);
        u(v);
    }
    void kb(const char *k, bool v) {
        key(k);
        b(v);
    }
    int dims = 3;  // components of every vector written
    void kv3(const char *k, const float *v) {
        key(k);
        open('{');
        kf("x", v[0]);
        kf("y", v[1]);
        if (dims == 3) kf("z", v[2]);
        close('}');
    }
    void item() {
        if (!first) s += ',';
        first = false;
        nl();
    }
};

std::string to_json(const SnapWorld &w) {
    const SnapMeta &m = w.meta;
    const FluidConfig3Ffi &c = w.cfg;
    JsonOut o;
    o.dims = w.dims;
    o.open('{');
    o.ku("schema", schema_of(w));
    o.kv3("gravity", w.gravity);
    if (!gravity_default(m)) {
        o.key("gravity_config");
        o.open('{');
        o.ku("mode", m.g_mode);
        o.ku("integrator", m.g_integrator);
        o.kf("big_g", m.big_g);
        o.kf("softening", m.softening);
        o.kv3("attractor", m.attractor);
        o.kf("attractor_mu", m.attractor_mu);
        o.kf("theta", m.theta);
        o.close('}');
    }
    if (!ambient_default(m)) {
        o.key("ambient");
        o.open('{');
        o.key("spring");
        o.open('{');
        o.kb("enabled", m.spring_on);
        o.kv3("anchor", m.spring_anchor);
        o.kf("stiffness", m.spring_k);
        o.kf("damping", m.spring_c);
        o.kf("max_displacement", m.spring_max);
        o.close('}');
        o.key("drag");
        o.open('{');
        o.kb("enabled", m.drag_on);
        o.kf("linear", m.drag_linear);
        o.kf("quadratic", m.drag_quadratic);
        o.kf("air_density", m.drag_air);
        o.kf("angular", m.drag_angular);
        o.kv3("wind", m.drag_wind);
        o.close('}');
        o.key("turbulence");
        o.open('{');
        o.kb("enabled", m.turb_on);
        o.kf("amplitude", m.turb_amplitude);
        o.kf("frequency", m.turb_frequency);
        o.kf("time_scale", m.turb_time_scale);
        o.ku("seed", m.turb_seed);
        o.ku("octaves", m.turb_octaves);
        o.kb("divergence_free", m.turb_div_free);
        o.close('}');
        o.close('}');
    }
    const float time = has_ambient(w) ? w.time : 0.0f;
    if (time != 0.0f) o.kf("time", time);
    if (!ccd_default(m)) {
        o.key("ccd");
        o.open('{');
        o.kb("enabled", m.ccd_on);
        o.ku("tier", m.ccd_tier);
        o.kf("max_movement", m.ccd_max_movement);
        o.ku("max_substeps", m.ccd_max_substeps);
        o.close('}');
    }
    o.kf("dt", m.dt);
    o.kf("linear_sleep_threshold", m.linear_sleep);
    o.kf("angular_sleep_threshold", m.angular_sleep);
    o.key("solver");
    o.open('{');
    o.ku("iterations", m.iterations);
    o.kf("slop", m.slop);
    o.kf("baumgarte", m.baumgarte);
    o.kf("early_exit_tolerance", m.early_exit);
    if (!m.friction) o.kb("friction", false);
    if (!m.restitution) o.kb("restitution", false);
    o.close('}');
    o.key("validation");
    o.open('{');
    o.kb("check_finite", m.check_finite);
    o.kf("max_penetration", m.max_penetration);
    o.kf("max_mass_ratio", m.max_mass_ratio);
    o.kf("max_coordinate", m.max_coordinate);
    o.kf("max_velocity", m.max_velocity);
    o.kf("max_angular_velocity", m.max_angular_velocity);
    o.kb("log_warnings", m.log_warnings);
    o.kb("auto_correct", m.auto_correct);
    o.close('}');
    o.ku("next_body_id", w.bodies.size());
    o.ku("next_joint_id", 0);
    if (w.dims == 2) o.kf("cell_size", w.cell_size);
    o.key("bodies");
    o.open('[');
    const float zero3[3] = {0.0f, 0.0f, 0.0f};
    for (size_t k = 0; k < w.bodies.size(); ++k) {
        const SnapBody &b = w.bodies[k];
        o.item();
        o.open('{');
        o.ku("id", k);
        o.kb("is_static", b.is_static);
        o.kb("is_sleeping", b.sleeping);
        o.kf("linear_damping", b.linear_damping);
        o.kf("angular_damping", b.angular_damping);
        o.key("material");
        o.open('{');
        o.kf("friction", b.friction);
        o.kf("restitution", b.restitution);
        o.kf("density", b.density);
        if (b.specific_heat != 900.0f) o.kf("specific_heat", b.specific_heat);
        if (b.thermal_conductivity != 1.0f) o.kf("thermal_conductivity", b.thermal_conductivity);
        if (b.emissivity != 0.9f) o.kf("emissivity", b.emissivity);
        o.close('}');
        o.key("shape");
        o.open('{');
        o.key("type");
        if (w.dims == 2 && b.shape == 0) {  // Shape2Dto::Circle
            o.s += "\"circle\"";
            o.kf("radius", b.a);
        } else if (w.dims == 2) {  // Shape2Dto::Box
            o.s += "\"box\"";
            const float he[3] = {b.a, b.b, 0.0f};
            o.kv3("half_extents", he);
        } else if (b.shape == PHIENG_SHAPE3_SPHERE) {
            o.s += "\"sphere\"";
            o.kf("radius", b.a);
        } else if (b.shape == PHIENG_SHAPE3_BOX) {
            o.s += "\"box\"";
            const float he[3] = {b.a, b.b, b.c};
            o.kv3("half_extents", he);
        } else {
            o.s += "\"capsule\"";
            o.kf("radius", b.a);
            o.kf("half_length", b.b);
        }
        o.close('}');
        o.kv3("position", b.pos);
        o.key("rotation");
        // 2D: the columns of Mat2::IDENTITY (body2_dto); 3D: the quaternion x, y, z, w
        if (w.dims == 2 || b.is_static) {
            o.s += w.dims == 2 ? "[1.0, 0.0, 0.0, 1.0]" : "[0.0, 0.0, 0.0, 1.0]";
        } else {
            o.s += '[';
            for (int q = 0; q < 4; ++q) {
                if (q) o.s += ", ";
                o.f(b.rot[q]);
            }
            o.s += ']';
        }
        o.kv3("velocity", w.dims == 2 ? zero3 : b.vel);
        if (w.dims == 2)
            o.kf("angular_velocity", 0.0f);
        else
            o.kv3("angular_velocity", b.angvel);
        if (b.gravity_scale != 1.0f) o.kf("gravity_scale", b.gravity_scale);
        if (b.ambient_scale != 1.0f) o.kf("ambient_scale", b.ambient_scale);
        o.close('}');
    }
    o.close(']');
    o.key("joints");
    o.s += "[]";
    o.key("contacts");
    o.s += "[]";
    o.key("fluid");
    o.open('{');
    o.key("config");
    o.open('{');
    o.kf("particle_mass", c.particle_mass);
    o.kf("rest_density", c.rest_density);
    o.kf("smoothing_length", c.smoothing_length);
    o.kf("stiffness", c.stiffness);
    if (c.gamma != 1.0f) o.kf("gamma", c.gamma);
    if (c.artificial_viscosity != 0.0f) o.kf("artificial_viscosity", c.artificial_viscosity);
    if (c.surface_tension != 0.0f) o.kf("surface_tension", c.surface_tension);
    if (c.cfl != 0.0f) o.kf("cfl", c.cfl);
    if (c.max_substeps != 1) o.ku("max_substeps", c.max_substeps);
    if (c.adaptive_resolution) o.kb("adaptive_resolution", true);
    if (c.split_density_ratio != 0.5f) o.kf("split_density_ratio", c.split_density_ratio);
    if (c.merge_density_ratio != 1.6f) o.kf("merge_density_ratio", c.merge_density_ratio);
    if (c.boundary_particles) o.kb("boundary_particles", true);
    o.kf("viscosity", c.viscosity);
    o.kf("drag", c.drag);
    if (c.wall_friction != 0.0f) o.kf("wall_friction", c.wall_friction);
    o.kf("max_velocity", c.max_velocity);
    o.kf("particle_radius", c.particle_radius);
    o.kb("enabled", c.enabled != 0);
    if (c.phase_count > 1) {
        o.key("phases");
        o.open('{');
        o.ku("count", c.phase_count);
        o.key("rest_density");
        o.open('[');
        for (int k = 0; k < 4; ++k) {
            o.item();
            o.f(c.phase_rest_density[k]);
        }
        o.close(']');
        o.key("mass");
        o.open('[');
        for (int k = 0; k < 4; ++k) {
            o.item();
            o.f(c.phase_mass[k]);
        }
        o.close(']');
        o.close('}');
    }
    o.close('}');
    o.key("particles");
    o.open('[');
    for (size_t i = 0; i < w.mass.size(); ++i) {
        o.item();
        o.open('{');
        o.kv3("position", &w.pos[3 * i]);
        o.kv3("velocity", &w.vel[3 * i]);
        o.kf("mass", w.mass[i]);
        o.ku("phase", w.phase[i]);
        o.close('}');
    }
    o.close(']');
    o.close('}');
    o.close('}');
    return o.s;
}

// ---------------------------------------------------------------- JSON in

struct JVal {
    enum Type { NUL, BOOL, NUM, STR, ARR, OBJ } t = NUL;
    bool b = false;
    std::string text;  // number literal or string
    std::vector<JVal> a;
    std::vector<std::pair<std::string, JVal>> o;
    const JVal *get(const char *k) const {
        for (const auto &kv : o)
            if (kv.first == k) return &kv.second;
        return nullptr;
    }
};

struct JParser {
    const char *p, *e;
    int depth = 0;
    [[noreturn]] void bad(const char *m) { fail(std::string("JSON: ") + m); }
    void ws() {
        while (p < e && (*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r')) ++p;
    }
    void lit(const char *w) {
        const size_t n = std::strlen(w);
        if ((size_t)(e - p) < n || std::memcmp(p, w, n)) bad("bad literal");
        p += n;
    }
    std::string str() {
        if (p >= e || *p != '"') bad("expected a string");
        ++p;
        std::string out;
        while (p < e && *p != '"') {
            if (*p == '\\') {
                if (++p >= e) bad("bad escape");
                switch (*p) {
                    case '"': out += '"'; break;
                    case '\\': out += '\\'; break;
                    case '/': out += '/'; break;
                    case 'b': out += '\b'; break;
                    case 'f': out += '\f'; break;
                    case 'n': out += '\n'; break;
                    case 'r': out += '\r'; break;
                    case 't': out += '\t'; break;
                    case 'u':
                        if (e - p < 5) bad("bad escape");
                        out += '?';  // no key or value this loader reads holds non-ASCII text
                        p += 4;
                        break;
                    default: bad("bad escape");
                }
                ++p;
            } else {
                out += *p++;
            }
        }
        if (p >= e) bad("unterminated string");
        ++p;
        return out;
    }
    JVal val() {
        if (++depth > 64) bad("nesting too deep");
        ws();
        if (p >= e) bad("unexpected end");
        JVal v;
        if (*p == '{') {
            v.t = JVal::OBJ;
            ++p;
            ws();
            if (p < e && *p == '}') {
                ++p;
            } else {
                for (;;) {
                    ws();
                    std::string k = str();
                    ws();
                    if (p >= e || *p != ':') bad("expected ':'");
                    ++p;
                    v.o.emplace_back(std::move(k), val());
                    ws();
                    if (p < e && *p == ',') {
                        ++p;
                        continue;
                    }
                    if (p < e && *p == '}') {
                        ++p;
                        break;
                    }
                    bad("expected ',' or '}'");
                }
            }
        } else if (*p == '[') {
            v.t = JVal::ARR;
            ++p;
            ws();
            if (p < e && *p == ']') {
                ++p;
            } else {
                for (;;) {
                    v.a.push_back(val());
                    ws();
                    if (p < e && *p == ',') {
                        ++p;
                        continue;
                    }
                    if (p < e && *p == ']') {
                        ++p;
                        break;
                    }
                    bad("expected ',' or ']'");
                }
            }
        } else if (*p == '"') {
            v.t = JVal::STR;
            v.text = str();
        } else if (*p == 't') {
            lit("true");
            v.t = JVal::BOOL;
            v.b = true;
        } else if (*p == 'f') {
            lit("false");
            v.t = JVal::BOOL;
        } else if (*p == 'n') {
            lit("null");
        } else {
            const char *s = p;
            if (p < e && *p == '-') ++p;
            while (p < e && ((*p >= '0' && *p <= '9') || *p == '.' || *p == 'e' || *p == 'E' || *p == '+' || *p == '-'))
                ++p;
            if (p == s) bad("unexpected character");
            v.t = JVal::NUM;
            v.text.assign(s, p);
        }
        --depth;
        return v;
    }
};

const JVal &req(const JVal &o, const char *k) {
    if (o.t != JVal::OBJ) fail(std::string("JSON: expected an object holding '") + k + "'");
    const JVal *v = o.get(k);
    if (!v) fail(std::string("JSON: missing field '") + k + "'");
    return *v;
}
float jf(const JVal &v, const char *what) {
    if (v.t != JVal::NUM) fail(std::string("JSON: '") + what + "' is not a number");
    errno = 0;
    char *end = nullptr;
    const double d = std::strtod(v.text.c_str(), &end);
    if (end != v.text.c_str() + v.text.size()) fail(std::string("JSON: bad number in '") + what + "'");
    return (float)d;
}
uint64_t ju(const JVal &v, const char *what, uint64_t max) {
    if (v.t != JVal::NUM || v.text.empty() || v.text.find_first_not_of("0123456789") != std::string::npos)
        fail(std::string("JSON: '") + what + "' is not an unsigned integer");
    if (v.text.size() > 19) fail(std::string("JSON: '") + what + "' out of range");
    const uint64_t u = std::stoull(v.text);
    if (u > max) fail(std::string("JSON: '") + what + "' out of range");
    return u;
}
bool jb(const JVal &v, const char *what) {
    if (v.t != JVal::BOOL) fail(std::string("JSON: '") + what + "' is not a bool");
    return v.b;
}
float rf(const JVal &o, const char *k) { return jf(req(o, k), k); }
float of(const JVal &o, const char *k, float d) {
    const JVal *v = o.get(k);
    return v ? jf(*v, k) : d;
}
int g_dims = 3;  // dimension of the file being read (set by the loader)
void rv3(const JVal &o, const char *k, float *out) {
    const JVal &v = req(o, k);
    out[0] = rf(v, "x");
    out[1] = rf(v, "y");
    out[2] = g_dims == 3 ? rf(v, "z") : 0.0f;
}

void check_finite(float v, const char *what) {
    if (!std::isfinite(v)) fail(std::string("non-finite ") + what);
}

// Shared checks of a decoded world (world3_from_dto, load_body3,
// apply_fluid3: snapshot.rs:5436-5500, 4759-4830, 4545-4615) and the
// limits of the GPU path.
struct BodyIn {
    uint32_t id;
    bool is_static, sleeping, has_shape;
    uint32_t shape_tag;  // 0 sphere, 1 box, 2 capsule, 3 hull, 4 mesh
    float dims[3];
    float pos[3], rot[4], vel[3], angvel[3];
};

// BodyIn.shape_tag is the file's tag: 3D sphere 0, box 1, capsule 2, hull
// 3, mesh 4; 2D circle 0, box 1, polygon 2 (encode_shape2/3).
void add_body(SnapWorld &w, const BodyIn &in, SnapBody b) {
    check_finite(in.pos[0], "body.position");
    check_finite(in.pos[1], "body.position");
    check_finite(in.pos[2], "body.position");
    for (float r : in.rot) check_finite(r, "body.rotation");
    if (!finite3(in.vel) || !finite3(in.angvel)) fail("non-finite body velocity");
    check_finite(b.linear_damping, "body.linear_damping");
    check_finite(b.angular_damping, "body.angular_damping");
    check_finite(b.gravity_scale, "body.gravity_scale");
    check_finite(b.friction, "material.friction");
    check_finite(b.restitution, "material.restitution");
    check_finite(b.density, "material.density");
    if (in.id != w.bodies.size()) fail("body id out of order");
    if (!in.is_static && w.dims == 2) fail("a dynamic body in a 2D world: the GPU path couples 3D worlds only");
    if (!in.has_shape) fail("a body without a shape");
    if (in.shape_tag > (w.dims == 2 ? 1u : 2u)) fail("a hull, mesh or polygon body: not a GPU boundary shape");
    const bool identity = w.dims == 2 ? in.rot[0] == 1.0f && in.rot[1] == 0.0f && in.rot[2] == 0.0f && in.rot[3] == 1.0f
                                      : in.rot[0] == 0.0f && in.rot[1] == 0.0f && in.rot[2] == 0.0f && in.rot[3] > 0.0f;
    if (in.is_static && !identity) fail("a rotated static body: the GPU boundary has identity rotations only");
    if (in.is_static && (in.vel[0] != 0.0f || in.vel[1] != 0.0f || in.vel[2] != 0.0f || in.angvel[0] != 0.0f ||
                         in.angvel[1] != 0.0f || in.angvel[2] != 0.0f))
        fail("a moving static body");
    b.is_static = in.is_static;
    if (!in.is_static) {
        for (int k = 0; k < 4; ++k) b.rot[k] = in.rot[k];
        for (int k = 0; k < 3; ++k) {
            b.vel[k] = in.vel[k];
            b.angvel[k] = in.angvel[k];
        }
    }
    for (int k = 0; k < 3; ++k) check_finite(in.dims[k], "shape");
    const bool ok = in.shape_tag == 0   ? in.dims[0] > 0.0f
                    : in.shape_tag == 1 ? in.dims[0] > 0.0f && in.dims[1] > 0.0f && (w.dims == 2 || in.dims[2] > 0.0f)
                                        : in.dims[0] > 0.0f && in.dims[1] > 0.0f;
    if (!ok || !(b.density > 0.0f)) fail("invalid shape or density");
    // set_material sanitises friction and restitution (body.rs:1306-1317)
    if (!(b.friction >= 0.0f)) b.friction = 0.4f;
    b.restitution = std::min(std::max(b.restitution, 0.0f), 1.0f);
    if (w.dims == 2)  // a circle is a sphere in the z = 0 plane; the 2D box is its own shape
        b.shape = in.shape_tag == 0 ? 0u : 3u;
    else
        b.shape = in.shape_tag == 0   ? PHIENG_SHAPE3_SPHERE
                  : in.shape_tag == 1 ? PHIENG_SHAPE3_BOX
                                      : PHIENG_SHAPE3_CAPSULE;
    for (int k = 0; k < 3; ++k) b.pos[k] = in.pos[k];
    b.a = in.dims[0];
    b.b = in.shape_tag == 0 ? 0.0f : in.dims[1];
    b.c = w.dims == 3 && in.shape_tag == 1 ? in.dims[2] : 0.0f;
    b.sleeping = in.sleeping;
    w.bodies.push_back(b);
}

// FluidConfig3Dto -> FluidConfig3 (apply_fluid3, snapshot.rs:4547-4586).
struct FluidCfgIn {
    float particle_mass, rest_density, smoothing_length, stiffness, gamma = 1.0f, av = 0.0f, st = 0.0f, cfl = 0.0f;
    uint32_t max_substeps = 1;
    bool adaptive = false, boundary = false, enabled = true;
    float split = 0.5f, merge = 1.6f, viscosity, drag, wall_friction = 0.0f, max_velocity, particle_radius;
    bool phases = false;
    uint32_t phase_count = 1;
    float phase_rd[4] = {1.0f, 1.0f, 1.0f, 1.0f}, phase_m[4];
};

FluidConfig3Ffi fluid_config(const FluidCfgIn &c) {
    const float fin[] = {c.particle_mass, c.rest_density, c.smoothing_length, c.stiffness, c.gamma, c.av, c.st,
                         c.cfl, c.viscosity, c.drag, c.max_velocity, c.particle_radius};
    for (float v : fin) check_finite(v, "fluid config value");
    FluidConfig3Ffi f{};
    f.particle_mass = c.particle_mass;
    f.rest_density = c.rest_density;
    f.smoothing_length = c.smoothing_length;
    f.stiffness = c.stiffness;
    f.gamma = c.gamma;
    f.artificial_viscosity = c.av;
    f.surface_tension = c.st;
    f.cfl = c.cfl;
    f.max_substeps = c.max_substeps;
    f.adaptive_resolution = c.adaptive ? 1 : 0;
    f.split_density_ratio = c.split;
    f.merge_density_ratio = c.merge;
    f.boundary_particles = c.boundary ? 1 : 0;
    f.viscosity = c.viscosity;
    f.drag = c.drag;
    f.wall_friction = c.wall_friction;
    f.max_velocity = c.max_velocity;
    f.particle_radius = c.particle_radius;
    f.enabled = c.enabled ? 1 : 0;
    f.phase_count = c.phases ? std::min<uint32_t>(std::max<uint32_t>(c.phase_count, 1u), 4u) : 1u;
    for (int k = 0; k < 4; ++k) {
        f.phase_rest_density[k] = c.phases ? c.phase_rd[k] : 1.0f;
        f.phase_mass[k] = c.phases ? c.phase_m[k] : c.particle_mass;
    }
    return f;
}

void add_particle(SnapWorld &w, const float *p, const float *v, float m, uint8_t ph) {
    if (!finite3(p) || !finite3(v)) fail("non-finite fluid particle");
    for (int k = 0; k < 3; ++k) {
        w.pos.push_back(p[k]);
        w.vel.push_back(v[k]);
    }
    w.mass.push_back(m > 0.0f ? m : std::max(w.cfg.particle_mass, 0.0f));
    w.phase.push_back(ph);
}

void check_schema(uint32_t s) {
    if (s < kSchemaMin || s > kSchemaMax) fail("unsupported snapshot schema version " + std::to_string(s));
    if (s >= kSchemaFirstUnsupported)
        fail("schema " + std::to_string(s) + " carries features outside the GPU fluid path");
}

void world_checks(const SnapWorld &w) {
    if (!finite3(w.gravity)) fail("non-finite gravity");
    const SnapMeta &m = w.meta;
    check_finite(m.dt, "dt");
    check_finite(m.linear_sleep, "linear_sleep_threshold");
    check_finite(m.angular_sleep, "angular_sleep_threshold");
    check_finite(w.time, "time");
    check_finite(m.ccd_max_movement, "ccd.max_movement");
    if (m.ccd_max_movement <= 0.0f) fail("ccd.max_movement must be positive");
    if (m.ccd_max_substeps < 1) fail("ccd.max_substeps must be >= 1");
}

SnapWorld from_json(const std::string &text, int dims) {
    JParser jp{text.data(), text.data() + text.size()};
    const JVal root = jp.val();
    jp.ws();
    if (jp.p != jp.e) fail("JSON: trailing characters");
    g_dims = dims;
    SnapWorld w;
    w.dims = dims;
    SnapMeta &m = w.meta;
    check_schema((uint32_t)ju(req(root, "schema"), "schema", 0xFFFFFFFFull));
    rv3(root, "gravity", w.gravity);
    if (const JVal *g = root.get("gravity_config")) {
        m.g_mode = (uint32_t)ju(req(*g, "mode"), "mode", 0xFFFFFFFFull);
        m.g_integrator = (uint32_t)ju(req(*g, "integrator"), "integrator", 0xFFFFFFFFull);
        m.big_g = rf(*g, "big_g");
        m.softening = rf(*g, "softening");
        rv3(*g, "attractor", m.attractor);
        m.attractor_mu = rf(*g, "attractor_mu");
        m.theta = rf(*g, "theta");
    }
    if (const JVal *a = root.get("ambient")) {
        const JVal &sp = req(*a, "spring"), &dr = req(*a, "drag"), &tu = req(*a, "turbulence");
        m.spring_on = jb(req(sp, "enabled"), "enabled");
        rv3(sp, "anchor", m.spring_anchor);
        m.spring_k = rf(sp, "stiffness");
        m.spring_c = rf(sp, "damping");
        m.spring_max = rf(sp, "max_displacement");
        m.drag_on = jb(req(dr, "enabled"), "enabled");
        m.drag_linear = rf(dr, "linear");
        m.drag_quadratic = rf(dr, "quadratic");
        m.drag_air = rf(dr, "air_density");
        m.drag_angular = rf(dr, "angular");
        rv3(dr, "wind", m.drag_wind);
        m.turb_on = jb(req(tu, "enabled"), "enabled");
        m.turb_amplitude = rf(tu, "amplitude");
        m.turb_frequency = rf(tu, "frequency");
        m.turb_time_scale = rf(tu, "time_scale");
        m.turb_seed = (uint32_t)ju(req(tu, "seed"), "seed", 0xFFFFFFFFull);
        m.turb_octaves = (uint8_t)ju(req(tu, "octaves"), "octaves", 255);
        m.turb_div_free = jb(req(tu, "divergence_free"), "divergence_free");
    }
    w.time = of(root, "time", 0.0f);
    if (const JVal *c = root.get("ccd")) {
        m.ccd_on = jb(req(*c, "enabled"), "enabled");
        m.ccd_tier = (uint32_t)ju(req(*c, "tier"), "tier", 0xFFFFFFFFull);
        m.ccd_max_movement = rf(*c, "max_movement");
        m.ccd_max_substeps = (uint32_t)ju(req(*c, "max_substeps"), "max_substeps", 0xFFFFFFFFull);
    }
    m.dt = rf(root, "dt");
    m.linear_sleep = rf(root, "linear_sleep_threshold");
    m.angular_sleep = rf(root, "angular_sleep_threshold");
    const JVal &so = req(root, "solver");
    m.iterations = (uint32_t)ju(req(so, "iterations"), "iterations", 0xFFFFFFFFull);
    m.slop = rf(so, "slop");
    m.baumgarte = rf(so, "baumgarte");
    m.early_exit = rf(so, "early_exit_tolerance");
    if (const JVal *f = so.get("friction")) m.friction = jb(*f, "friction");
    if (const JVal *r = so.get("restitution")) m.restitution = jb(*r, "restitution");
    const JVal &va = req(root, "validation");
    m.check_finite = jb(req(va, "check_finite"), "check_finite");
    m.max_penetration = rf(va, "max_penetration");
    m.max_mass_ratio = rf(va, "max_mass_ratio");
    m.max_coordinate = rf(va, "max_coordinate");
    m.max_velocity = rf(va, "max_velocity");
    m.max_angular_velocity = rf(va, "max_angular_velocity");
    m.log_warnings = jb(req(va, "log_warnings"), "log_warnings");
    m.auto_correct = jb(req(va, "auto_correct"), "auto_correct");
    const uint64_t next_body = ju(req(root, "next_body_id"), "next_body_id", 0xFFFFFFFFull);
    if (dims == 2) {
        if (!root.get("next_joint_id")) fail("JSON: missing field 'next_joint_id'");
        w.cell_size = rf(root, "cell_size");
        check_finite(w.cell_size, "cell_size");
    }
    const JVal &bodies = req(root, "bodies");
    if (bodies.t != JVal::ARR) fail("JSON: 'bodies' is not an array");
    for (const JVal &bj : bodies.a) {
        BodyIn in{};
        SnapBody b;
        in.id = (uint32_t)ju(req(bj, "id"), "id", 0xFFFFFFFFull);
        in.is_static = jb(req(bj, "is_static"), "is_static");
        in.sleeping = jb(req(bj, "is_sleeping"), "is_sleeping");
        b.linear_damping = rf(bj, "linear_damping");
        b.angular_damping = rf(bj, "angular_damping");
        const JVal &ma = req(bj, "material");
        b.friction = rf(ma, "friction");
        b.restitution = rf(ma, "restitution");
        b.density = rf(ma, "density");
        b.specific_heat = of(ma, "specific_heat", 900.0f);
        b.thermal_conductivity = of(ma, "thermal_conductivity", 1.0f);
        b.emissivity = of(ma, "emissivity", 0.9f);
        const JVal *sh = bj.get("shape");
        in.has_shape = sh && sh->t != JVal::NUL;
        if (in.has_shape) {
            const JVal &ty = req(*sh, "type");
            if (ty.t != JVal::STR) fail("JSON: shape type");
            if (dims == 2) {
                if (ty.text == "circle") {
                    in.shape_tag = 0;
                    in.dims[0] = rf(*sh, "radius");
                } else if (ty.text == "box") {
                    in.shape_tag = 1;
                    rv3(*sh, "half_extents", in.dims);
                } else if (ty.text == "polygon") {
                    in.shape_tag = 2;
                } else {
                    fail("JSON: unknown shape type");
                }
            } else if (ty.text == "sphere") {
                in.shape_tag = 0;
                in.dims[0] = rf(*sh, "radius");
            } else if (ty.text == "box") {
                in.shape_tag = 1;
                rv3(*sh, "half_extents", in.dims);
            } else if (ty.text == "capsule") {
                in.shape_tag = 2;
                in.dims[0] = rf(*sh, "radius");
                in.dims[1] = rf(*sh, "half_length");
            } else if (ty.text == "convexhull" || ty.text == "convexmesh") {
                in.shape_tag = 3;
            } else {
                fail("JSON: unknown shape type");
            }
        }
        rv3(bj, "position", in.pos);
        const JVal &ro = req(bj, "rotation");
        if (ro.t != JVal::ARR || ro.a.size() != 4) fail("JSON: rotation");
        for (int k = 0; k < 4; ++k) in.rot[k] = jf(ro.a[k], "rotation");
        rv3(bj, "velocity", in.vel);
        if (dims == 2)
            in.angvel[0] = rf(bj, "angular_velocity");
        else
            rv3(bj, "angular_velocity", in.angvel);
        b.gravity_scale = of(bj, "gravity_scale", 1.0f);
        b.ambient_scale = of(bj, "ambient_scale", 1.0f);
        add_body(w, in, b);
    }
    if (next_body != w.bodies.size()) fail("next_body_id disagrees with body count");
    const JVal *joints = dims == 2 ? &req(root, "joints") : root.get("joints");
    if (joints && (joints->t != JVal::ARR || !joints->a.empty())) fail("joints are outside the GPU fluid path");
    if (const JVal *nj = root.get("next_joint_id"))
        if (ju(*nj, "next_joint_id", 0xFFFFFFFFull) != 0) fail("next_joint_id disagrees with joint count");
    const JVal &co = req(root, "contacts");
    if (co.t != JVal::ARR || !co.a.empty()) fail("contacts are outside the GPU fluid path");
    FluidCfgIn fc;
    const JVal *fl = root.get("fluid");
    if (fl) {
        const JVal &c = req(*fl, "config");
        fc.particle_mass = rf(c, "particle_mass");
        fc.rest_density = rf(c, "rest_density");
        fc.smoothing_length = rf(c, "smoothing_length");
        fc.stiffness = rf(c, "stiffness");
        fc.gamma = of(c, "gamma", 1.0f);
        fc.av = of(c, "artificial_viscosity", 0.0f);
        fc.st = of(c, "surface_tension", 0.0f);
        fc.cfl = of(c, "cfl", 0.0f);
        if (const JVal *v = c.get("max_substeps")) fc.max_substeps = (uint32_t)ju(*v, "max_substeps", 0xFFFFFFFFull);
        if (const JVal *v = c.get("adaptive_resolution")) fc.adaptive = jb(*v, "adaptive_resolution");
        fc.split = of(c, "split_density_ratio", 0.5f);
        fc.merge = of(c, "merge_density_ratio", 1.6f);
        if (const JVal *v = c.get("boundary_particles")) fc.boundary = jb(*v, "boundary_particles");
        fc.viscosity = rf(c, "viscosity");
        fc.drag = rf(c, "drag");
        fc.wall_friction = of(c, "wall_friction", 0.0f);
        fc.max_velocity = rf(c, "max_velocity");
        fc.particle_radius = rf(c, "particle_radius");
        fc.enabled = jb(req(c, "enabled"), "enabled");
        const JVal *ph = c.get("phases");
        if (ph && ph->t != JVal::NUL) {
            fc.phases = true;
            fc.phase_count = (uint32_t)ju(req(*ph, "count"), "count", 255);
            const JVal &rd = req(*ph, "rest_density"), &ms = req(*ph, "mass");
            if (rd.t != JVal::ARR || rd.a.size() != 4 || ms.t != JVal::ARR || ms.a.size() != 4)
                fail("JSON: phases arrays");
            for (int k = 0; k < 4; ++k) {
                fc.phase_rd[k] = jf(rd.a[k], "phases.rest_density");
                fc.phase_m[k] = jf(ms.a[k], "phases.mass");
            }
        }
    } else {
        // FluidConfig3Dto's Default is FluidConfig3::default (fluid3.rs:105-134)
        fc.particle_mass = 0.0625f;
        fc.rest_density = 1.0f;
        fc.smoothing_length = 0.5f;
        fc.stiffness = 100.0f;
        fc.viscosity = 0.3f;
        fc.drag = 4.0f;
        fc.max_velocity = 30.0f;
        fc.particle_radius = 0.125f;
    }
    w.cfg = fluid_config(fc);
    world_checks(w);
    if (fl) {
        const JVal &ps = req(*fl, "particles");
        if (ps.t != JVal::ARR) fail("JSON: 'particles' is not an array");
        if (ps.a.size() > PH_ID_MASK + 1u) fail("more than 2^24 fluid particles");
        for (const JVal &pj : ps.a) {
            float p[3], v[3];
            rv3(pj, "position", p);
            rv3(pj, "velocity", v);
            const float mass = of(pj, "mass", 0.0f);
            const JVal *phv = pj.get("phase");
            add_particle(w, p, v, mass, phv ? (uint8_t)ju(*phv, "phase", 255) : 0);
        }
    }
    return w;
}

// ---------------------------------------------------------------- binary

struct BinOut {
    std::vector<uint8_t> b;
    void u8(uint8_t v) { b.push_back(v); }
    void u32(uint32_t v) {
        for (int k = 0; k < 4; ++k) b.push_back((uint8_t)(v >> (8 * k)));
    }
    void f32(float v) {
        uint32_t u;
        std::memcpy(&u, &v, 4);
        u32(u);
    }
    int dims = 3;
    void v3(const float *v) {
        f32(v[0]);
        f32(v[1]);
        if (dims == 3) f32(v[2]);
    }
};

struct BinIn {
    const uint8_t *p;
    size_t n, pos = 0;
    const uint8_t *take(size_t k) {
        if (n - pos < k) fail("binary payload truncated");
        const uint8_t *q = p + pos;
        pos += k;
        return q;
    }
    uint8_t u8() { return *take(1); }
    uint32_t u32() {
        const uint8_t *q = take(4);
        return (uint32_t)q[0] | ((uint32_t)q[1] << 8) | ((uint32_t)q[2] << 16) | ((uint32_t)q[3] << 24);
    }
    float f32() {
        const uint32_t u = u32();
        float v;
        std::memcpy(&v, &u, 4);
        return v;
    }
    int dims = 3;
    void v3(float *v) {
        v[0] = f32();
        v[1] = f32();
        v[2] = dims == 3 ? f32() : 0.0f;
    }
};

// encode_world3 (snapshot.rs:7920-8126) for schemas 3 .. 7; encode_world2
// (:6033-6238) is the same payload with 2-vectors, the broad-phase cell
// size after next_joint_id, 2D shapes and a scalar angular velocity.
std::vector<uint8_t> to_binary(const SnapWorld &w) {
    const uint32_t schema = schema_of(w);
    const SnapMeta &m = w.meta;
    const FluidConfig3Ffi &c = w.cfg;
    BinOut o;
    o.dims = w.dims;
    o.v3(w.gravity);
    o.f32(m.dt);
    o.f32(m.linear_sleep);
    o.f32(m.angular_sleep);
    o.u32(m.iterations);  // encode_solver (:5618-5628)
    o.f32(m.slop);
    o.f32(m.baumgarte);
    o.f32(m.early_exit);
    if (schema >= kSchemaSolverFlags) {
        o.u8(m.friction);
        o.u8(m.restitution);
    }
    o.u8(m.check_finite);  // encode_validation (:5828-5837)
    o.f32(m.max_penetration);
    o.f32(m.max_mass_ratio);
    o.f32(m.max_coordinate);
    o.f32(m.max_velocity);
    o.f32(m.max_angular_velocity);
    o.u8(m.log_warnings);
    o.u8(m.auto_correct);
    if (schema >= kSchemaGravity) {  // encode_gravity_config_3 (:5660-5670)
        o.u32(m.g_mode);
        o.u32(m.g_integrator);
        o.f32(m.big_g);
        o.f32(m.softening);
        o.v3(m.attractor);
        o.f32(m.attractor_mu);
        o.f32(m.theta);
    }
    if (schema >= kSchemaAmbient) {  // encode_ambient_config_3 (:5746-5769) + clock
        o.u8(m.spring_on);
        o.v3(m.spring_anchor);
        o.f32(m.spring_k);
        o.f32(m.spring_c);
        o.f32(m.spring_max);
        o.u8(m.drag_on);
        o.f32(m.drag_linear);
        o.f32(m.drag_quadratic);
        o.f32(m.drag_air);
        o.f32(m.drag_angular);
        o.v3(m.drag_wind);
        o.u8(m.turb_on);
        o.f32(m.turb_amplitude);
        o.f32(m.turb_frequency);
        o.f32(m.turb_time_scale);
        o.u32(m.turb_seed);
        o.u8(m.turb_octaves);
        o.u8(m.turb_div_free);
        o.f32(has_ambient(w) ? w.time : 0.0f);
    }
    if (schema >= kSchemaCcd) {
        o.u8(m.ccd_on);
        o.u8((uint8_t)m.ccd_tier);
        o.f32(m.ccd_max_movement);
        o.u32(m.ccd_max_substeps);
    }
    o.u32((uint32_t)w.bodies.size());  // next_body_id
    o.u32(0);                          // next_joint_id
    if (w.dims == 2) o.f32(w.cell_size);
    o.u32((uint32_t)w.bodies.size());
    for (size_t k = 0; k < w.bodies.size(); ++k) {
        const SnapBody &b = w.bodies[k];
        o.u32((uint32_t)k);
        o.u8((b.is_static ? 1u : 0u) | (b.sleeping ? 2u : 0u));
        o.f32(b.linear_damping);
        o.f32(b.angular_damping);
        o.f32(b.friction);
        o.f32(b.restitution);
        o.f32(b.density);
        if (w.dims == 2) {  // encode_shape2 (:5852-5872): circle, box
            o.u8(b.shape == 0 ? 0 : 1);
            o.f32(b.a);
            if (b.shape != 0) o.f32(b.b);
        } else if (b.shape == PHIENG_SHAPE3_SPHERE) {  // encode_shape3 (:5911-5954)
            o.u8(0);
            o.f32(b.a);
        } else if (b.shape == PHIENG_SHAPE3_BOX) {
            o.u8(1);
            o.f32(b.a);
            o.f32(b.b);
            o.f32(b.c);
        } else {
            o.u8(2);
            o.f32(b.a);
            o.f32(b.b);
        }
        o.v3(b.pos);
        // rotation: 3D the quaternion x, y, z, w; 2D the columns of the matrix
        const float z3[3] = {0.0f, 0.0f, 0.0f};
        if (w.dims == 2) {
            o.f32(1.0f);
            o.f32(0.0f);
            o.f32(0.0f);
            o.f32(1.0f);
            o.v3(z3);     // velocity
            o.f32(0.0f);  // angular velocity
        } else {
            for (int q = 0; q < 4; ++q) o.f32(b.rot[q]);
            o.v3(b.vel);
            o.v3(b.angvel);
        }
        if (schema >= kSchemaGravity) o.f32(b.gravity_scale);
        if (schema >= kSchemaAmbient) o.f32(b.ambient_scale);
    }
    o.u32(0);  // joints
    o.u32(0);  // contacts
    // encode_fluid3 / encode_fluid_config3 (:6527-6677)
    o.f32(c.particle_mass);
    o.f32(c.rest_density);
    o.f32(c.smoothing_length);
    o.f32(c.stiffness);
    if (schema >= kSchemaAmbient) {
        o.f32(c.gamma);
        o.f32(c.artificial_viscosity);
        o.f32(c.surface_tension);
        o.f32(c.cfl);
        o.u32(c.max_substeps);
        o.u8(c.adaptive_resolution ? 1 : 0);
        o.f32(c.split_density_ratio);
        o.f32(c.merge_density_ratio);
        o.u8(c.boundary_particles ? 1 : 0);
        o.f32(c.wall_friction);
        if (c.phase_count > 1) {
            o.u8((uint8_t)c.phase_count);
            for (int k = 0; k < 4; ++k) o.f32(c.phase_rest_density[k]);
            for (int k = 0; k < 4; ++k) o.f32(c.phase_mass[k]);
        } else {
            o.u8(1);
            for (int k = 0; k < 4; ++k) o.f32(1.0f);
            for (int k = 0; k < 4; ++k) o.f32(c.particle_mass);
        }
    }
    o.f32(c.viscosity);
    o.f32(c.drag);
    o.f32(c.max_velocity);
    o.f32(c.particle_radius);
    o.u8(c.enabled ? 1 : 0);
    o.u32((uint32_t)w.mass.size());
    for (size_t i = 0; i < w.mass.size(); ++i) {
        o.v3(&w.pos[3 * i]);
        o.v3(&w.vel[3 * i]);
        if (schema >= kSchemaAmbient) {
            o.f32(w.mass[i]);
            o.u8(w.phase[i]);
        }
    }
    // the envelope (encode_envelope, :5570-5579)
    std::vector<uint8_t> out = {'P', 'H', 'E', '1'};
    BinOut h;
    h.u32(schema);
    h.u8(w.dims == 2 ? kKindWorld2 : kKindWorld3);
    h.u32((uint32_t)o.b.size());
    out.insert(out.end(), h.b.begin(), h.b.end());
    out.insert(out.end(), o.b.begin(), o.b.end());
    BinOut t;
    t.u32(crc32(o.b.data(), o.b.size()));
    out.insert(out.end(), t.b.begin(), t.b.end());
    return out;
}

// decode_envelope + world3_from_binary / world2_from_binary (snapshot.rs:
// 5585-5616, 8128-8433, 7720-7910).
SnapWorld from_binary(const std::vector<uint8_t> &bytes, int dims) {
    if (bytes.size() < 13) fail("binary snapshot too short");
    if (std::memcmp(bytes.data(), "PHE1", 4) != 0) fail("bad snapshot magic");
    BinIn hdr{bytes.data() + 4, 9};
    const uint32_t version = hdr.u32();
    const uint8_t kind = hdr.u8();
    const uint32_t len = hdr.u32();
    if (version < kSchemaMin || version > kSchemaMax) fail("unsupported snapshot schema version " + std::to_string(version));
    if (kind != (dims == 2 ? kKindWorld2 : kKindWorld3)) fail("snapshot kind mismatch");
    if (bytes.size() != 13 + (size_t)len + 4) fail("binary snapshot length mismatch");
    BinIn crc{bytes.data() + 13 + len, 4};
    if (crc32(bytes.data() + 13, len) != crc.u32()) fail("snapshot CRC mismatch");
    check_schema(version);
    BinIn r{bytes.data() + 13, len};
    r.dims = dims;
    SnapWorld w;
    w.dims = dims;
    SnapMeta &m = w.meta;
    r.v3(w.gravity);
    m.dt = r.f32();
    m.linear_sleep = r.f32();
    m.angular_sleep = r.f32();
    m.iterations = r.u32();
    m.slop = r.f32();
    m.baumgarte = r.f32();
    m.early_exit = r.f32();
    if (version >= kSchemaSolverFlags) {
        m.friction = r.u8() != 0;
        m.restitution = r.u8() != 0;
    }
    m.check_finite = r.u8() != 0;
    m.max_penetration = r.f32();
    m.max_mass_ratio = r.f32();
    m.max_coordinate = r.f32();
    m.max_velocity = r.f32();
    m.max_angular_velocity = r.f32();
    m.log_warnings = r.u8() != 0;
    m.auto_correct = r.u8() != 0;
    if (version >= kSchemaGravity) {
        m.g_mode = r.u32();
        m.g_integrator = r.u32();
        m.big_g = r.f32();
        m.softening = r.f32();
        r.v3(m.attractor);
        m.attractor_mu = r.f32();
        m.theta = r.f32();
    }
    if (version >= kSchemaAmbient) {
        m.spring_on = r.u8() != 0;
        r.v3(m.spring_anchor);
        m.spring_k = r.f32();
        m.spring_c = r.f32();
        m.spring_max = r.f32();
        m.drag_on = r.u8() != 0;
        m.drag_linear = r.f32();
        m.drag_quadratic = r.f32();
        m.drag_air = r.f32();
        m.drag_angular = r.f32();
        r.v3(m.drag_wind);
        m.turb_on = r.u8() != 0;
        m.turb_amplitude = r.f32();
        m.turb_frequency = r.f32();
        m.turb_time_scale = r.f32();
        m.turb_seed = r.u32();
        m.turb_octaves = r.u8();
        m.turb_div_free = r.u8() != 0;
        w.time = r.f32();
    }
    if (version >= kSchemaCcd) {
        m.ccd_on = r.u8() != 0;
        m.ccd_tier = r.u8();
        m.ccd_max_movement = r.f32();
        m.ccd_max_substeps = r.u32();
    }
    const uint32_t next_body = r.u32();
    const uint32_t next_joint = r.u32();
    if (dims == 2) {
        w.cell_size = r.f32();
        check_finite(w.cell_size, "cell_size");
    }
    const uint32_t nb = r.u32();
    if (nb > 1000000u) fail("body count exceeds 1,000,000");
    for (uint32_t k = 0; k < nb; ++k) {
        BodyIn in{};
        SnapBody b;
        in.id = r.u32();
        const uint8_t flags = r.u8();
        if (flags & ~3u) fail("bad body flags");
        in.is_static = (flags & 1u) != 0;
        in.sleeping = (flags & 2u) != 0;
        b.linear_damping = r.f32();
        b.angular_damping = r.f32();
        b.friction = r.f32();
        b.restitution = r.f32();
        b.density = r.f32();
        const uint8_t tag = r.u8();
        in.has_shape = tag != 255;
        in.shape_tag = tag;
        if (dims == 2) {
            if (tag == 0) {
                in.dims[0] = r.f32();
            } else if (tag == 1) {
                in.dims[0] = r.f32();
                in.dims[1] = r.f32();
            } else if (tag == 2) {
                fail("a polygon body: not a GPU boundary shape");
            } else if (tag != 255) {
                fail("unknown 2D shape tag");
            }
        } else if (tag == 0) {
            in.dims[0] = r.f32();
        } else if (tag == 1) {
            r.v3(in.dims);
        } else if (tag == 2) {
            in.dims[0] = r.f32();
            in.dims[1] = r.f32();
        } else if (tag == 3 || tag == 4) {
            fail("a convex hull or mesh body: not a GPU boundary shape");
        } else if (tag != 255) {
            fail("unknown 3D shape tag");
        }
        r.v3(in.pos);
        for (int q = 0; q < 4; ++q) in.rot[q] = r.f32();
        r.v3(in.vel);
        if (dims == 2)
            in.angvel[0] = r.f32();
        else
            r.v3(in.angvel);
        if (version >= kSchemaGravity) b.gravity_scale = r.f32();
        if (version >= kSchemaAmbient) b.ambient_scale = r.f32();
        add_body(w, in, b);
    }
    if (next_body != w.bodies.size()) fail("next_body_id disagrees with body count");
    if (r.u32() != 0 || next_joint != 0) fail("joints are outside the GPU fluid path");
    if (r.u32() != 0) fail("contacts are outside the GPU fluid path");
    FluidCfgIn fc;
    fc.particle_mass = r.f32();
    fc.rest_density = r.f32();
    fc.smoothing_length = r.f32();
    fc.stiffness = r.f32();
    if (version >= kSchemaAmbient) {
        fc.gamma = r.f32();
        fc.av = r.f32();
        fc.st = r.f32();
        fc.cfl = r.f32();
        fc.max_substeps = r.u32();
        fc.adaptive = r.u8() != 0;
        fc.split = r.f32();
        fc.merge = r.f32();
        fc.boundary = r.u8() != 0;
        fc.wall_friction = r.f32();
        fc.phase_count = r.u8();
        for (int k = 0; k < 4; ++k) fc.phase_rd[k] = r.f32();
        for (int k = 0; k < 4; ++k) fc.phase_m[k] = r.f32();
        fc.phases = fc.phase_count > 1;
    }
    fc.viscosity = r.f32();
    fc.drag = r.f32();
    fc.max_velocity = r.f32();
    fc.particle_radius = r.f32();
    const uint8_t en = r.u8();
    if (en > 1) fail("bad fluid enabled flag");
    fc.enabled = en != 0;
    w.cfg = fluid_config(fc);
    world_checks(w);
    const uint32_t np = r.u32();
    if (np > PH_ID_MASK + 1u) fail("more than 2^24 fluid particles");
    const size_t per = (dims == 2 ? 16 : 24) + (version >= kSchemaAmbient ? 5 : 0);
    if ((size_t)np * per > r.n - r.pos) fail("binary payload truncated");
    w.pos.reserve(3 * (size_t)np);
    w.vel.reserve(3 * (size_t)np);
    w.mass.reserve(np);
    w.phase.reserve(np);
    for (uint32_t i = 0; i < np; ++i) {
        float p[3], v[3];
        r.v3(p);
        r.v3(v);
        float mass = 0.0f;
        uint8_t ph = 0;
        if (version >= kSchemaAmbient) {
            mass = r.f32();
            ph = r.u8();
        }
        add_particle(w, p, v, mass, ph);
    }
    if (r.pos != r.n) fail("trailing bytes in snapshot payload");
    return w;
}

// ---------------------------------------------------------------- files

bool path_arg(const uint8_t *p, size_t len, std::string &out) {
    if (!p || len == 0 || len > 4096) return false;
    out.assign(reinterpret_cast<const char *>(p), len);
    // UTF-8 well-formedness (str::from_utf8 in ffi.rs:4058-4061)
    for (size_t i = 0; i < len;) {
        const uint8_t c = p[i];
        size_t k = c < 0x80 ? 1 : (c >> 5) == 6 ? 2 : (c >> 4) == 14 ? 3 : (c >> 3) == 30 ? 4 : 0;
        if (!k || i + k > len) return false;
        for (size_t j = 1; j < k; ++j)
            if ((p[i + j] & 0xC0) != 0x80) return false;
        i += k;
    }
    return out.find('\0') == std::string::npos;
}

void write_file(const std::string &path, const uint8_t *data, size_t n) {
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f) throw SnapError(-1, "cannot open " + path + " for writing");
    const size_t w = std::fwrite(data, 1, n, f);
    const int c = std::fclose(f);
    if (w != n || c != 0) throw SnapError(-1, "writing " + path + " failed");
}

std::vector<uint8_t> read_file(const std::string &path) {
    FILE *f = std::fopen(path.c_str(), "rb");
    if (!f) throw SnapError(-1, "cannot open " + path);
    std::vector<uint8_t> out;
    uint8_t buf[1 << 16];
    size_t k;
    while ((k = std::fread(buf, 1, sizeof(buf), f)) > 0) out.insert(out.end(), buf, buf + k);
    const bool err = std::ferror(f) != 0;
    std::fclose(f);
    if (err) throw SnapError(-1, "reading " + path + " failed");
    return out;
}

// phieng_world3_save / phieng_world_save (ffi.rs:4042-4082, 3107-3160).
int32_t save_world(const char *where, World3 *world, int dims, const uint8_t *path_ptr, size_t path_len,
                   uint32_t format, uint32_t mode) {
    std::string path;
    if (!world || !path_arg(path_ptr, path_len, path)) return 0;
    if (mode != PHIENG_SNAPSHOT_EXACT && mode != PHIENG_SNAPSHOT_FRESH) return 0;
    if (format != PHIENG_SNAPSHOT_JSON && format != PHIENG_SNAPSHOT_BINARY) return 0;
    try {
        SnapWorld s;
        world3_snapshot(world, s);
        if (s.dims != dims) fail("the world's dimension does not match the entry point");
        if (format == PHIENG_SNAPSHOT_JSON) {
            const std::string text = to_json(s);
            write_file(path, reinterpret_cast<const uint8_t *>(text.data()), text.size());
        } else {
            const std::vector<uint8_t> b = to_binary(s);
            write_file(path, b.data(), b.size());
        }
        return 1;
    } catch (const SnapError &e) {
        snapshot_set_error(where, e.what());
        return e.code;
    } catch (const std::exception &e) {
        snapshot_set_error(where, e.what());
        return -2;
    }
}

// phieng_world3_load / phieng_world_load (ffi.rs:4084-4111, 3162-3190):
// JSON when the first non-blank byte is '{', else binary.
World3 *load_world(const char *where, int dims, const uint8_t *path_ptr, size_t path_len, uint32_t mode) {
    std::string path;
    if (!path_arg(path_ptr, path_len, path)) return nullptr;
    if (mode != PHIENG_SNAPSHOT_EXACT && mode != PHIENG_SNAPSHOT_FRESH) return nullptr;
    try {
        const std::vector<uint8_t> bytes = read_file(path);
        size_t i = 0;
        while (i < bytes.size() && (bytes[i] == ' ' || bytes[i] == '\t' || bytes[i] == '\n' || bytes[i] == '\r' ||
                                    bytes[i] == '\f'))
            ++i;
        SnapWorld s = i < bytes.size() && bytes[i] == '{'
                          ? from_json(std::string(bytes.begin(), bytes.end()), dims)
                          : from_binary(bytes, dims);
        std::string err;
        World3 *w = world3_from_snapshot(s, err);
        if (!w) snapshot_set_error(where, err);
        return w;
    } catch (const std::exception &e) {
        snapshot_set_error(where, e.what());
        return nullptr;
    }
}

}  // namespace

extern "C" {

int32_t phieng_world3_save(const World3 *world, const uint8_t *path_ptr, size_t path_len, uint32_t format,
                           uint32_t mode) {
    return save_world("phieng_world3_save", const_cast<World3 *>(world), 3, path_ptr, path_len, format, mode);
}

World3 *phieng_world3_load(const uint8_t *path_ptr, size_t path_len, uint32_t mode) {
    return load_world("phieng_world3_load", 3, path_ptr, path_len, mode);
}

int32_t phieng_world_save(const World2 *world, const uint8_t *path_ptr, size_t path_len, uint32_t format,
                          uint32_t mode) {
    return save_world("phieng_world_save", world ? world2_base(const_cast<World2 *>(world)) : nullptr, 2, path_ptr,
                      path_len, format, mode);
}

World2 *phieng_world_load(const uint8_t *path_ptr, size_t path_len, uint32_t mode) {
    World3 *w = load_world("phieng_world_load", 2, path_ptr, path_len, mode);
    return w ? world2_of(w) : nullptr;
}

}  // extern "C"
