// surrogate_step.hip — the learned pressure term in the fluid step.
//
// A surrogate substep replaces the plain substep (cell sort, density pass,
// pressure force + XSPH pass; fluid3.rs:990-1124) by ONE neighbour sweep
// over the particle order and cell table of the last plain substep, using
// a skin list built there (candidates within (1 + skin) h), which
//   * sums the density rho_i over the current positions, and
//   * aggregates pairwise messages m_j grad W_ij times (1, P_j, dP_j), where
//     P_j = p_j / rho_j^2 and dP_j its last change are the neighbour's
//     values from the previous substep, plus XSPH with previous densities;
// then a 2-layer MLP (BF16 MFMA) maps invariants of the aggregates to
// coefficients c of a correction of the lagged-pressure acceleration
//   a0 = -(P_i sum m grad W + sum m P_j grad W) / rho_i,
//   a  = a0 + c0 t + c1 a0 + c2 10 G^,   t = -(sum m dP_j grad W) / rho_i.
// With c = 0 this is the network-free method.  Plain substeps every few
// steps judge the result (bulk density deviation), refresh the history and
// rebuild the skin list.
#include <hip/hip_runtime.h>

#include "mlp.h"
#include "sph_device.h"
#include "surrogate.h"
#include "surrogate_step.h"

// Indexed loads of a structured buffer: the record stride sits in the
// resource and the slot is the index, so no offset arithmetic is spent per
// load, and an index at or past the resource's count reads 0.  These are
// LLVM's buffer-load intrinsics, which clang exposes as builtins only in
// their raw (offset) form.
typedef float sur_f4v __attribute__((ext_vector_type(4)));
typedef uint32_t sur_u2v __attribute__((ext_vector_type(2)));
__device__ sur_f4v sur_load_idx_b128(__amdgpu_buffer_rsrc_t, int, int, int, int) __asm(
    "llvm.amdgcn.struct.ptr.buffer.load.v4f32");
__device__ sur_u2v sur_load_idx_b64(__amdgpu_buffer_rsrc_t, int, int, int, int) __asm(
    "llvm.amdgcn.struct.ptr.buffer.load.v2i32");

namespace {

struct SurAcc {
    float rho = 0.0f;
    uint32_t cnt = 0;
    float gx = 0.0f, gy = 0.0f, gz = 0.0f;  // sum m_j grad W_ij
    float hx = 0.0f, hy = 0.0f, hz = 0.0f;  // sum m_j P_j grad W_ij
    float dx = 0.0f, dy = 0.0f, dz = 0.0f;  // sum m_j dP_j grad W_ij
    float sx = 0.0f, sy = 0.0f, sz = 0.0f;  // sum m_j W_ij v_j (the XSPH sum is this - v_i rho)
};

// One candidate of particle i, branch-free: contributes when ok and
// r2 < h^2.  ra = (x, y, z, P|phase), v and dq = dP of the candidate; its
// mass is the phase's entry of mt.  Candidates come from
// grid cells, which hold finite positions only, so offsets and data are
// finite and a zero weight zeroes a term.  Fast math throughout (hardware
// rsqrt, fused multiply-adds): the surrogate is judged on its density
// error, not on bit parity.
template <bool UMASS = false>
__device__ __forceinline__ void sur_cand(const SceneDev &s, bool ok, const float4 &pi, float4 mt, const float4 &ra,
                                         float vx, float vy, float vz, float dq, SurAcc &a) {
    const float dx = ra.x - pi.x, dy = ra.y - pi.y, dz = ra.z - pi.z;
    const float r2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
    const bool in = ok && r2 < s.h_grid2;
    const float d = s.h2 - r2;
    float m = mt.x;  // one mass for every phase (UMASS), else the phase's
    if (!UMASS) {
        const uint32_t ph = __float_as_uint(ra.w) & 3u;
        const bool p1 = (ph & 1u) != 0u, p2 = (ph & 2u) != 0u;
        const float m01 = p1 ? mt.y : mt.x, m23 = p1 ? mt.w : mt.z;
        m = p2 ? m23 : m01;
    }
    const float mw = in ? (m * s.poly6_c) * (d * d * d) : 0.0f;
    a.cnt += in ? 1u : 0u;
    a.rho += mw;
    // the particle itself (and any coincident one) sits at r2 = 0: no gradient
    const bool nb = in && r2 > PH_EPS_SQ;
    const float inv_r = __builtin_amdgcn_rsqf(r2);
    const float hr = s.h - r2 * inv_r;
    const float mg = nb ? -(m * s.spiky_c) * (hr * hr * inv_r) : 0.0f;  // m_j grad W = mg (x_j - x_i)
    a.gx = fmaf(mg, dx, a.gx);
    a.gy = fmaf(mg, dy, a.gy);
    a.gz = fmaf(mg, dz, a.gz);
    const float mh = mg * ra.w, md = mg * dq;
    a.hx = fmaf(mh, dx, a.hx);
    a.hy = fmaf(mh, dy, a.hy);
    a.hz = fmaf(mh, dz, a.hz);
    a.dx = fmaf(md, dx, a.dx);
    a.dy = fmaf(md, dy, a.dy);
    a.dz = fmaf(md, dz, a.dz);
    // XSPH: sum m_j W_ij v_j (mw = 0 off the neighbourhood); the numerator
    // sum m_j W_ij (v_j - v_i) is this minus v_i rho, formed once per particle
    a.sx = fmaf(vx, mw, a.sx);
    a.sy = fmaf(vy, mw, a.sy);
    a.sz = fmaf(vz, mw, a.sz);
}

// Sweep over particle i's grid candidates (collection on the gathered
// state, records assembled from the state arrays).
__device__ __forceinline__ void sur_sweep(const SceneDev &s, const uint32_t *__restrict__ start,
                                          const float4 *__restrict__ posm, const float4 *__restrict__ vel,
                                          const float4 *__restrict__ aux, const uint32_t *__restrict__ tag,
                                          float4 mt, uint32_t i, const float4 &pi, SurAcc &a) {
    if (!finite3(pi.x, pi.y, pi.z)) return;
    int32_t cx, cy, cz;
    cell_of(pi.x, pi.y, pi.z, s.h_grid, cx, cy, cz);
    for_each_candidate(s, start, posm, cx, cy, cz, [&](uint32_t j) {
        const float4 pj = posm[j], vj = vel[j], aj = aux[j];
        const float4 ra = make_float4(pj.x, pj.y, pj.z, pack_pq(aj.y, tag[j] >> 24));
        sur_cand(s, true, pi, mt, ra, vj.x, vj.y, vj.z, aj.z, a);
    });
}

__device__ __forceinline__ float len3(float x, float y, float z) { return sqrtf(x * x + y * y + z * z); }

// Network inputs, basis vectors and the network-free acceleration.
struct SurFeat {
    float x[MLP_IN];
    float b[SUR_NCOEF][3];
    float a0[3];
    float rho, p, q;
};

__device__ __forceinline__ float fast_log1p(float x) { return __logf(1.0f + x); }

// Network inputs, basis vectors and the network-free acceleration of one
// particle from its aggregates (fast math; the Tait power via exp2/log2).
__device__ __forceinline__ void sur_features(const SceneDev &s, float rho0, float qs, const SurAcc &a,
                                             float rho_lag_i, float dq_lag_i, SurFeat &f) {
    const float rho = rclampf(a.rho, 0.0f, s.rho_max);
    const float ir0 = __builtin_amdgcn_rcpf(rho0);
    float p;
    if (s.gamma == 1.0f) {
        p = rmaxf(s.stiffness * (rho - rho0), 0.0f);
    } else {
        const float pw = __builtin_amdgcn_exp2f(s.gamma * __builtin_amdgcn_logf(rho * ir0));
        p = rmaxf(s.stiffness * rho0 * (pw - 1.0f), 0.0f);
    }
    const float rc = rmaxf(rho, PH_EPS);
    const float inv = __builtin_amdgcn_rcpf(rc);
    const float q = p * inv * inv;
    f.rho = rho;
    f.p = p;
    f.q = q;
    f.a0[0] = -fmaf(q, a.gx, a.hx) * inv;
    f.a0[1] = -fmaf(q, a.gy, a.hy) * inv;
    f.a0[2] = -fmaf(q, a.gz, a.hz) * inv;
    const float t0 = -a.dx * inv, t1 = -a.dy * inv, t2 = -a.dz * inv;
    const float gl = __builtin_amdgcn_sqrtf(fmaf(a.gz, a.gz, fmaf(a.gy, a.gy, a.gx * a.gx)));
    const float gi = gl > 0.0f ? __builtin_amdgcn_rcpf(gl) : 0.0f;
    const float a0l = __builtin_amdgcn_sqrtf(fmaf(f.a0[2], f.a0[2], fmaf(f.a0[1], f.a0[1], f.a0[0] * f.a0[0])));
    const float tl = __builtin_amdgcn_sqrtf(fmaf(t2, t2, fmaf(t1, t1, t0 * t0)));
    f.b[0][0] = t0;
    f.b[0][1] = t1;
    f.b[0][2] = t2;
    f.b[1][0] = f.a0[0];
    f.b[1][1] = f.a0[1];
    f.b[1][2] = f.a0[2];
    f.b[2][0] = 10.0f * a.gx * gi;
    f.b[2][1] = 10.0f * a.gy * gi;
    f.b[2][2] = 10.0f * a.gz * gi;
    const float tiny = 1e-20f, iqs = __builtin_amdgcn_rcpf(qs);
    f.x[0] = fmaf(rho, ir0, -1.0f) * 100.0f;
    f.x[1] = fmaf(rho_lag_i, ir0, -1.0f) * 100.0f;
    f.x[2] = fast_log1p(a0l * 0.1f);
    f.x[3] = fast_log1p(tl * 0.1f);
    f.x[4] = fmaf(f.a0[2], t2, fmaf(f.a0[1], t1, f.a0[0] * t0)) * __builtin_amdgcn_rcpf(fmaf(a0l, tl, tiny));
    f.x[5] = gl * s.h * ir0;
    f.x[6] = fmaf(f.a0[2], a.gz, fmaf(f.a0[1], a.gy, f.a0[0] * a.gx)) * gi * __builtin_amdgcn_rcpf(a0l + tiny);
    f.x[7] = (float)a.cnt * (1.0f / 30.0f) - 1.0f;
    f.x[8] = fast_log1p(q * iqs);
    f.x[9] = copysignf(fast_log1p(fabsf(dq_lag_i) * iqs), dq_lag_i);
    f.x[10] = 1.0f;
    for (int c = 11; c < MLP_IN; ++c) f.x[c] = 0.0f;
}

// Integrate one particle with acceleration g + ap (the plain substep's
// integrate, speed cap, finite guards and boundary clamp; fluid3.rs:1088-1179).
__device__ __forceinline__ void sur_integrate(const SceneDev &s, const BodyDev *__restrict__ bodies, const float4 &pi,
                                              const float4 &vi, const float ap[3], float xsx, float xsy, float xsz,
                                              float4 &po, float4 &vo) {
    const float ax = s.gx + ap[0], ay = s.gy + ap[1], az = s.gz + ap[2];
    float ux = vi.x + ax * s.dt, uy = vi.y + ay * s.dt, uz = vi.z + az * s.dt;
    if (s.viscosity > 0.0f) {
        ux += xsx * s.viscosity;
        uy += xsy * s.viscosity;
        uz += xsz * s.viscosity;
    }
    const float speed = sqrtf(dot3(ux, uy, uz));
    if (speed > s.max_velocity) {
        const float r = s.max_velocity / speed;
        ux *= r;
        uy *= r;
        uz *= r;
    }
    if (!finite3(ux, uy, uz)) ux = uy = uz = 0.0f;
    float x = pi.x + ux * s.dt, y = pi.y + uy * s.dt, z = pi.z + uz * s.dt;
    if (!finite3(x, y, z)) {
        x = pi.x;
        y = pi.y;
        z = pi.z;
    }
    clamp_to_bodies(s, bodies, x, y, z, ux, uy, uz);
    po = make_float4(x, y, z, pi.w);
    vo = make_float4(ux, uy, uz, 0.0f);
}

// The two record arrays of one generation, indexed by slot.
struct RecBuf {
    __amdgpu_buffer_rsrc_t p, v;
    __device__ RecBuf(const SurRec &r, uint32_t count)
        : p(__builtin_amdgcn_make_buffer_rsrc(r.pos, (short)16, (int)count, 0x00020000)),
          v(__builtin_amdgcn_make_buffer_rsrc(r.vd, (short)8, (int)count, 0x00020000)) {}
    __device__ __forceinline__ float4 pos(uint32_t j) const {
        return __builtin_bit_cast(float4, sur_load_idx_b128(p, (int)j, 0, 0, 0));
    }
    __device__ __forceinline__ uint2 vd(uint32_t j) const {
        return __builtin_bit_cast(uint2, sur_load_idx_b64(v, (int)j, 0, 0, 0));
    }
};

template <bool UMASS>
__device__ __forceinline__ void sur_rec_cand(const SceneDev &s, bool ok, const float4 &pi, float4 mt,
                                             const float4 &ra, uint2 rb, SurAcc &a) {
    float vx, vy, vz, dq;
    unpack_vd(rb, vx, vy, vz, dq);
    sur_cand<UMASS>(s, ok, pi, mt, ra, vx, vy, vz, dq, a);
}

// Sweep over the skin list of particle i: its recorded candidates, a group
// of four at a time with their eight record loads issued together while the
// next group's slots are already in flight; the per-cell path when the list
// could not hold them.  Group g (candidates 4g .. 4g + 3) of particle i is
// the 16 bytes at grp[g * mstride + i]; the list holds one spare group row,
// so the look-ahead load after the last group stays inside it, and every
// load is unconditional (slots past the count are replaced by i and masked).
template <bool UMASS>
__device__ __forceinline__ void sur_list_sweep(const SceneDev &s, const uint32_t *__restrict__ start,
                                               const float4 *__restrict__ posm_p, const RecBuf &rec, float4 mt,
                                               uint32_t cw, const int4 &cell, const uint4 *__restrict__ grp_p,
                                               uint32_t mstride, uint32_t i, const float4 &pi, SurAcc &a) {
    if (cw & SLF_NONE) return;
    if (cw & SLF_SLOW) {
        for_each_candidate(s, start, posm_p, cell.x, cell.y, cell.z, [&](uint32_t j) {
            sur_rec_cand<UMASS>(s, true, pi, mt, rec.pos(j), rec.vd(j), a);
        });
        return;
    }
    const uint32_t n = cw & SLF_COUNT;
    // one group row (16 mstride bytes) per buffer resource, its base moved a
    // row per group with scalar arithmetic: lane offsets stay 32-bit at any
    // particle count
    const uint32_t off = 16u * i;
    const uint4 *row = grp_p;
    auto load_row = [&](const uint4 *r, uint32_t o) {
        const __amdgpu_buffer_rsrc_t lr =
            __builtin_amdgcn_make_buffer_rsrc(const_cast<uint4 *>(r), (short)0, (int)(16u * mstride), 0x00020000);
        return __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(lr, o, 0, 0));
    };
    // rows past the last group read as zero (offset beyond the resource)
    auto row_off = [&](uint32_t k) { return k < n ? off : 0xfffffff0u; };
    struct Recs {
        float4 p0, p1, p2, p3;
        uint2 v0, v1, v2, v3;
    };
    auto fetch = [&](uint4 g, uint32_t k) {
        const uint32_t j0 = k < n ? g.x : i, j1 = k + 1 < n ? g.y : i, j2 = k + 2 < n ? g.z : i,
                       j3 = k + 3 < n ? g.w : i;
        Recs r;
        r.p0 = rec.pos(j0);
        r.v0 = rec.vd(j0);
        r.p1 = rec.pos(j1);
        r.v1 = rec.vd(j1);
        r.p2 = rec.pos(j2);
        r.v2 = rec.vd(j2);
        r.p3 = rec.pos(j3);
        r.v3 = rec.vd(j3);
        return r;
    };
    uint4 grp = load_row(row, off);
#pragma unroll 1
    for (uint32_t k = 0; k < n; k += 4) {
        const Recs cur = fetch(grp, k);
        row += mstride;
        const uint4 nxt = load_row(row, row_off(k + 4));
        sur_rec_cand<UMASS>(s, true, pi, mt, cur.p0, cur.v0, a);
        sur_rec_cand<UMASS>(s, k + 1 < n, pi, mt, cur.p1, cur.v1, a);
        sur_rec_cand<UMASS>(s, k + 2 < n, pi, mt, cur.p2, cur.v2, a);
        sur_rec_cand<UMASS>(s, k + 3 < n, pi, mt, cur.p3, cur.v3, a);
        grp = nxt;
    }
}

// LDS exchange among the lanes of one wave: its LDS operations complete in
// issue order, so ordering the compiler's memory operations is enough.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

}  // namespace

// After a plain substep, on the same gathered state: training samples
// (target = the plain pressure + viscous acceleration acc_target) for the
// particles with id % stride == phase, sample id / stride.
__global__ __launch_bounds__(PH_TILE) void k_sur_collect(SurCollectArgs a) {
    const uint32_t tile = xcd_tile(blockIdx.x, gridDim.x);
    const TileDev t = a.tiles[tile];
    const SceneDev s = a.scenes[t.scene];
    const uint32_t i = t.first + threadIdx.x;
    if (i >= s.part_end) return;
    const uint32_t tg = a.tag[i];
    const uint32_t id = tg & PH_ID_MASK;
    if (id % a.stride != a.phase) return;
    const uint32_t k = a.sample_base + id / a.stride;
    if (k >= a.max_samples) return;
    const float4 pi = a.posm[i], ai = a.aux[i];
    SurAcc acc;
    sur_sweep(s, a.start, a.posm, a.vel, a.aux, a.tag, a.mass, i, pi, acc);
    const float rho0 = a.scenes[t.scene].phase_rho0[min(tg >> 24, 3u)];
    SurFeat f;
    sur_features(s, rho0, a.qs, acc, ai.x, ai.z, f);
    float *o = a.samples + (size_t)k * SUR_SAMPLE;
    for (int c = 0; c < MLP_IN; ++c) o[SUR_X + c] = f.x[c];
    for (int c = 0; c < SUR_NCOEF; ++c)
        for (int d = 0; d < 3; ++d) o[SUR_B + 3 * c + d] = f.b[c][d];
    const float4 at = a.acc_target[i];
    o[SUR_A0] = f.a0[0];
    o[SUR_A0 + 1] = f.a0[1];
    o[SUR_A0 + 2] = f.a0[2];
    o[SUR_AT] = at.x;
    o[SUR_AT + 1] = at.y;
    o[SUR_AT + 2] = at.z;
    o[SUR_S] = len3(at.x, at.y, at.z) + 10.0f;
}

// One surrogate substep: the skin-list sweep over the neighbour records,
// the network on MFMA (computed transposed, H^T = W1^T X^T and
// Y^T = W2^T H^T, so the hidden tiles that two K = 16 products leave in a
// lane are the operand the K = 32 product needs there), and the integrate;
// writes the advanced state, the new history and the records of the next
// substep.  JUDGE: the same evaluation for the judge's sample (sample
// k = blockIdx.x * PH_TILE + threadIdx.x is slot judge_slot(k)), writing
// only the acceleration without gravity to judge_acc[k] (one scene).
template <bool UMASS, bool JUDGE>
__global__ __launch_bounds__(PH_TILE, 7) void k_sur_step(SurStepArgs a) {
    __shared__ uint16_t xs[PH_TILE][MLP_IN];
    __shared__ float ys[PH_TILE][4];
    __shared__ mlp_bf16x4 w1p[8][64];
    __shared__ mlp_bf16x8 w2q[4][64];
    __shared__ __align__(16) float b1s[MLP_HID];
    const uint32_t tile = JUDGE ? 0u : xcd_tile(blockIdx.x, gridDim.x);
    const uint32_t scene = JUDGE ? 0u : a.tiles[tile].scene;
    const SceneDev s = a.scenes[scene];
    const uint32_t sample = blockIdx.x * PH_TILE + threadIdx.x;
    const uint32_t i = JUDGE ? judge_slot(sample, a.judge_stride, a.judge_off) : a.tiles[tile].first + threadIdx.x;
    const bool active = i < s.part_end;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, r = lane & 15, q = lane >> 4;
    if (a.use_net) {
        // lane-ready weight tiles: w1p[nt][l] = W1[4q+i][16 nt + r] (K = 16);
        // w2q[p][l] = W2[hid(p, q, e)][r] (K = 32), hid(p, q, e) =
        // 16 (2p + e / 4) + 4q + e % 4: the hidden units that the first
        // layer's tiles 2p and 2p + 1 leave in lane l
        for (int e = threadIdx.x; e < 8 * 64; e += PH_TILE) {
            const int nt = e >> 6, l = e & 63, lr = l & 15, lq = l >> 4;
            mlp_bf16x4 v1;
            for (int k = 0; k < 4; ++k) v1[k] = (short)a.mlp.w1[(4 * lq + k) * MLP_HID + nt * 16 + lr];
            w1p[nt][l] = v1;
            if (nt < 4) {
                mlp_bf16x8 v2;
                for (int k = 0; k < 8; ++k)
                    v2[k] = (short)a.mlp.w2[(16 * (2 * nt + (k >> 2)) + 4 * lq + (k & 3)) * MLP_OUT + lr];
                w2q[nt][l] = v2;
            }
        }
        if (threadIdx.x < MLP_HID) b1s[threadIdx.x] = a.mlp.b1[threadIdx.x];
        __syncthreads();  // the only block-wide wait: each wave's MLP below uses its own LDS rows
    }
    SurFeat f;
    float4 pi = make_float4(0.0f, 0.0f, 0.0f, 0.0f), vi = pi, ai = pi;
    float xsx = 0.0f, xsy = 0.0f, xsz = 0.0f;
    uint32_t tg = 0;
    if (active) {
        pi = a.posm[i];
        vi = a.vel[i];
        ai = a.aux[i];
        tg = a.tag[i];
        const uint32_t cw = a.cnt[i];
        const int4 cl = (cw & SLF_SLOW) ? a.cell[i] : make_int4(0, 0, 0, 0);
        const RecBuf rec(a.rec, s.part_end);
        const float4 mt = make_float4(a.mass.x, a.mass.y, a.mass.z, a.mass.w);
        SurAcc acc;
        sur_list_sweep<UMASS>(s, a.start, a.posm, rec, mt, cw, cl, reinterpret_cast<const uint4 *>(a.idx), a.mstride, i,
                              pi, acc);
        const float rho0 = a.scenes[scene].phase_rho0[min(tg >> 24, 3u)];
        sur_features(s, rho0, a.qs, acc, ai.x, ai.z, f);
        const float irl = __builtin_amdgcn_rcpf(rmaxf(ai.x, PH_EPS));
        xsx = fmaf(-vi.x, acc.rho, acc.sx) * irl;
        xsy = fmaf(-vi.y, acc.rho, acc.sy) * irl;
        xsz = fmaf(-vi.z, acc.rho, acc.sz) * irl;
    }
    float c[SUR_NCOEF] = {0.0f, 0.0f, 0.0f};
    if (a.use_net) {
        for (int k = 0; k < MLP_IN; ++k) xs[threadIdx.x][k] = active ? mlp_bf16_hw(f.x[k]) : (uint16_t)0;
        wave_sync();
#pragma unroll 1
        for (int mt = 0; mt < 4; ++mt) {
            const int row = w * 64 + mt * 16 + r;
            mlp_bf16x4 xb;
#pragma unroll
            for (int k = 0; k < 4; ++k) xb[k] = (short)xs[row][4 * q + k];
            mlp_f32x4 y = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll 1
            for (int p = 0; p < 4; ++p) {
                // hidden tiles 2p and 2p + 1 (+ b1), then one K = 32 product over both
                const float4 bv0 = *reinterpret_cast<const float4 *>(&b1s[32 * p + 4 * q]);
                const float4 bv1 = *reinterpret_cast<const float4 *>(&b1s[32 * p + 16 + 4 * q]);
                const mlp_f32x4 z0 = mlp_mfma(w1p[2 * p][lane], xb, mlp_f32x4{bv0.x, bv0.y, bv0.z, bv0.w});
                const mlp_f32x4 z1 = mlp_mfma(w1p[2 * p + 1][lane], xb, mlp_f32x4{bv1.x, bv1.y, bv1.z, bv1.w});
                mlp_bf16x8 hb;
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    hb[k] = (short)mlp_bf16_hw(mlp_gelu(z0[k], nullptr));
                    hb[4 + k] = (short)mlp_bf16_hw(mlp_gelu(z1[k], nullptr));
                }
                y = mlp_mfma32(w2q[p][lane], hb, y);
            }
            if (q == 0)
                *reinterpret_cast<float4 *>(ys[row]) =
                    make_float4(y[0] + a.mlp.b2[0], y[1] + a.mlp.b2[1], y[2] + a.mlp.b2[2], 0.0f);
        }
        wave_sync();
        for (int k = 0; k < SUR_NCOEF; ++k) c[k] = ys[threadIdx.x][k];
    }
    if (!active) return;
    float ap[3];
    for (int d = 0; d < 3; ++d) ap[d] = f.a0[d] + c[0] * f.b[0][d] + c[1] * f.b[1][d] + c[2] * f.b[2][d];
    if (JUDGE) {
        a.judge_acc[sample] = make_float4(ap[0], ap[1], ap[2], 0.0f);
        return;
    }
    float4 po, vo;
    sur_integrate(s, a.bodies, pi, vi, ap, xsx, xsy, xsz, po, vo);
    a.posm_out[i] = po;
    a.vel_out[i] = vo;
    const float dq = f.q - ai.y;
    a.aux_out[i] = make_float4(f.rho, f.q, dq, 0.0f);
    a.rec_out.pos[i] = make_float4(po.x, po.y, po.z, pack_pq(f.q, tg >> 24));
    a.rec_out.vd[i] = pack_vd(vo.x, vo.y, vo.z, dq);
}

// aux_out[d] = aux_in[perm[d]]: history follows its particle through the sort.
__global__ __launch_bounds__(256) void k_gather_aux(uint32_t n, const uint32_t *__restrict__ perm,
                                                    const float4 *__restrict__ aux_in, float4 *__restrict__ aux_out) {
    const uint32_t d = blockIdx.x * 256 + threadIdx.x;
    if (d < n) aux_out[d] = aux_in[perm[d]];
}

// After a plain substep: new history (rho, P, P - P_prev) and, per tile,
// the bulk density deviation sum |rho/rho0 - 1| over particles with
// rho >= 0.9 rho0, and their count (the judge's density measure; this path
// leaves the acceleration sums at 0).
__global__ __launch_bounds__(256) void k_aux_update(const TileDev *__restrict__ tiles, const SceneDev *__restrict__ scenes,
                                                    const float2 *__restrict__ rq, const float4 *__restrict__ aux_prev,
                                                    const uint32_t *__restrict__ tag, float4 *__restrict__ aux_out,
                                                    float4 *__restrict__ tile_err, const float4 *__restrict__ posm_new,
                                                    const float4 *__restrict__ vel_new, SurRec rec_out) {
    const uint32_t tile = blockIdx.x;
    const TileDev t = tiles[tile];
    const SceneDev &s = scenes[t.scene];
    const uint32_t i = t.first + threadIdx.x;
    float e = 0.0f, c = 0.0f;
    if (i < s.part_end) {
        const float2 v = rq[i];
        const float4 ap = aux_prev[i];
        const uint32_t tg = tag[i];
        aux_out[i] = make_float4(v.x, v.y, v.y - ap.y, 0.0f);
        if (rec_out.pos) {
            // the surrogate's neighbour records of the new state
            const float4 p = posm_new[i], u = vel_new[i];
            rec_out.pos[i] = make_float4(p.x, p.y, p.z, pack_pq(v.y, tg >> 24));
            rec_out.vd[i] = pack_vd(u.x, u.y, u.z, v.y - ap.y);
        }
        const float rho0 = s.phase_rho0[min(tg >> 24, 3u)];
        const float r = v.x / rho0;
        if (r >= 0.9f) {
            e = fabsf(r - 1.0f);
            c = 1.0f;
        }
    }
    for (int d = 32; d >= 1; d >>= 1) {
        e += __shfl_xor(e, d, 64);
        c += __shfl_xor(c, d, 64);
    }
    __shared__ float red[4][2];
    if ((threadIdx.x & 63) == 0) {
        red[threadIdx.x >> 6][0] = e;
        red[threadIdx.x >> 6][1] = c;
    }
    __syncthreads();
    if (threadIdx.x == 0)
        tile_err[tile] = make_float4(red[0][0] + red[1][0] + red[2][0] + red[3][0],
                                     red[0][1] + red[1][1] + red[2][1] + red[3][1], 0.0f, 0.0f);
}

// The judge's verdict from the per-tile sums: out[0] = mean bulk density
// deviation, out[1] = its particle count, out[2] = sum |a_plain - a_sur| /
// sum |a_plain| over the evaluated particles (-1 without any), out[3] = the
// denominator, out[4] = the count in slow (which is then cleared), out[5] =
// vmax[0].  One block, sums in double in a fixed order; out may be pinned
// host memory.
__global__ __launch_bounds__(1024) void k_err_reduce(const float4 *__restrict__ tile_err, uint32_t ntiles,
                                                     float *__restrict__ out, uint32_t *__restrict__ slow,
                                                     const float *__restrict__ vmax) {
    double v[4] = {0.0, 0.0, 0.0, 0.0};
    for (uint32_t k = threadIdx.x; k < ntiles; k += 1024) {
        const float4 t = tile_err[k];
        v[0] += t.x;
        v[1] += t.y;
        v[2] += t.z;
        v[3] += t.w;
    }
    for (int d = 32; d >= 1; d >>= 1)
        for (int c = 0; c < 4; ++c) v[c] += __shfl_xor(v[c], d, 64);
    __shared__ double red[16][4];
    if ((threadIdx.x & 63) == 0)
        for (int c = 0; c < 4; ++c) red[threadIdx.x >> 6][c] = v[c];
    __syncthreads();
    if (threadIdx.x == 0) {
        double t[4] = {0.0, 0.0, 0.0, 0.0};
        for (int w = 0; w < 16; ++w)
            for (int c = 0; c < 4; ++c) t[c] += red[w][c];
        out[0] = (float)(t[1] > 0 ? t[0] / t[1] : 0.0);
        out[1] = (float)t[1];
        out[2] = (float)(t[3] > 0 ? t[2] / t[3] : -1.0);
        out[3] = (float)t[3];
        out[4] = (float)slow[0];
        out[5] = vmax[0];
        slow[0] = 0u;
    }
}

// History of freshly enabled particles: rest density, no pressure.
__global__ __launch_bounds__(256) void k_aux_reset(uint32_t n, const uint32_t *__restrict__ tag, float4 rho0,
                                                   float4 *__restrict__ aux) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const uint32_t ph = min(tag[i] >> 24, 3u);
    const float r = ph == 0 ? rho0.x : (ph == 1 ? rho0.y : (ph == 2 ? rho0.z : rho0.w));
    aux[i] = make_float4(r, 0.0f, 0.0f, 0.0f);
}

namespace ph {
hipError_t launch_sur_collect(hipStream_t st, uint32_t ntiles, const SurCollectArgs &a) {
    hipLaunchKernelGGL(k_sur_collect, dim3(ntiles), dim3(PH_TILE), 0, st, a);
    return hipGetLastError();
}
hipError_t launch_sur_step(hipStream_t st, uint32_t ntiles, const SurStepArgs &a) {
    // one mass for every phase (the common single-phase fluid): the sweep
    // skips the per-candidate phase lookup
    const bool umass = a.mass.x == a.mass.y && a.mass.x == a.mass.z && a.mass.x == a.mass.w;
    if (umass)
        hipLaunchKernelGGL((k_sur_step<true, false>), dim3(ntiles), dim3(PH_TILE), 0, st, a);
    else
        hipLaunchKernelGGL((k_sur_step<false, false>), dim3(ntiles), dim3(PH_TILE), 0, st, a);
    return hipGetLastError();
}
hipError_t launch_sur_judge(hipStream_t st, uint32_t nsamples, const SurStepArgs &a) {
    if (nsamples == 0) return hipSuccess;
    const bool umass = a.mass.x == a.mass.y && a.mass.x == a.mass.z && a.mass.x == a.mass.w;
    const dim3 grid((nsamples + PH_TILE - 1) / PH_TILE);
    if (umass)
        hipLaunchKernelGGL((k_sur_step<true, true>), grid, dim3(PH_TILE), 0, st, a);
    else
        hipLaunchKernelGGL((k_sur_step<false, true>), grid, dim3(PH_TILE), 0, st, a);
    return hipGetLastError();
}
hipError_t launch_gather_aux(hipStream_t st, uint32_t n, const uint32_t *perm, const float4 *in, float4 *out) {
    hipLaunchKernelGGL(k_gather_aux, dim3((n + 255) / 256), dim3(256), 0, st, n, perm, in, out);
    return hipGetLastError();
}
hipError_t launch_aux_update(hipStream_t st, uint32_t ntiles, const TileDev *tiles, const SceneDev *scenes,
                             const float2 *rq, const float4 *aux_prev, const uint32_t *tag, float4 *aux_out,
                             float4 *tile_err, const float4 *posm_new, const float4 *vel_new, SurRec rec_out) {
    hipLaunchKernelGGL(k_aux_update, dim3(ntiles), dim3(256), 0, st, tiles, scenes, rq, aux_prev, tag, aux_out,
                       tile_err, posm_new, vel_new, rec_out);
    return hipGetLastError();
}
hipError_t launch_err_reduce(hipStream_t st, const float4 *tile_err, uint32_t ntiles, float *out, uint32_t *slow,
                             const float *vmax) {
    hipLaunchKernelGGL(k_err_reduce, dim3(1), dim3(1024), 0, st, tile_err, ntiles, out, slow, vmax);
    return hipGetLastError();
}
hipError_t launch_aux_reset(hipStream_t st, uint32_t n, const uint32_t *tag, float4 rho0, float4 *aux) {
    hipLaunchKernelGGL(k_aux_reset, dim3((n + 255) / 256), dim3(256), 0, st, n, tag, rho0, aux);
    return hipGetLastError();
}
}  // namespace ph
