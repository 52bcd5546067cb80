// surrogate_step.hip — the learned pressure term in the fluid step.
//
// A surrogate substep replaces the plain step's second neighbour sweep
// (pressure force + XSPH, fluid3.rs:1029-1124) with one sweep that
//   * sums the exact density (same candidates, order and arithmetic as
//     the density pass, so rho is the plain value), and
//   * aggregates pairwise messages m_j grad W_ij times (1, P_j, dP_j), where
//     P_j = p_j / rho_j^2 and dP_j its last change are the neighbour's
//     values from the previous substep, plus XSPH with previous densities;
// then a 2-layer MLP (BF16 MFMA) maps invariants of the aggregates to
// coefficients c of a correction of the lagged-pressure acceleration
//   a0 = -(P_i sum m grad W + sum m P_j grad W) / rho_i,
//   a  = a0 + c0 t + c1 a0 + c2 10 G^,   t = -(sum m dP_j grad W) / rho_i.
// With c = 0 this is the network-free method.  Plain substeps every few
// steps judge the result (bulk density deviation) and refresh history.
#include <hip/hip_runtime.h>

#include "mlp.h"
#include "sph_device.h"
#include "surrogate.h"
#include "surrogate_step.h"

namespace {

struct SurAcc {
    float rho = 0.0f;
    uint32_t cnt = 0;
    float gx = 0.0f, gy = 0.0f, gz = 0.0f;  // sum m_j grad W_ij
    float hx = 0.0f, hy = 0.0f, hz = 0.0f;  // sum m_j P_j grad W_ij
    float dx = 0.0f, dy = 0.0f, dz = 0.0f;  // sum m_j dP_j grad W_ij
    float sx = 0.0f, sy = 0.0f, sz = 0.0f;  // XSPH with previous densities
};

// One neighbour j (r2 < h^2, j may be i) of particle i.  The density term
// is the density pass's expression; the message terms use fast math.
__device__ __forceinline__ void sur_pair(const SceneDev &s, uint32_t i, uint32_t j, const float4 &pi, const float4 &vi,
                                         float rho_lag_i, const float4 &pj, const float4 *__restrict__ vel,
                                         const float4 *__restrict__ aux, float r2, SurAcc &a) {
    const float w = poly6(r2, s);
    ++a.cnt;
    a.rho += pj.w * w;
    if (j == i || r2 <= PH_EPS_SQ) return;
    const float4 aj = aux[j];
    const float4 vj = vel[j];
    const float inv_r = __frsqrt_rn(r2);
    const float r = r2 * inv_r;
    if (r < s.h) {
        const float hr = s.h - r;
        const float mg = pj.w * s.spiky_c * hr * hr * inv_r;  // m_j grad W = mg (x_i - x_j)
        const float ex = pi.x - pj.x, ey = pi.y - pj.y, ez = pi.z - pj.z;
        a.gx += mg * ex;
        a.gy += mg * ey;
        a.gz += mg * ez;
        const float mh = mg * aj.y, md = mg * aj.z;
        a.hx += mh * ex;
        a.hy += mh * ey;
        a.hz += mh * ez;
        a.dx += md * ex;
        a.dy += md * ey;
        a.dz += md * ez;
    }
    if (s.viscosity > 0.0f) {
        const float rb = 0.5f * (rmaxf(rho_lag_i, PH_EPS) + rmaxf(aj.x, PH_EPS));
        const float q = pj.w * w / rb;
        a.sx += (vj.x - vi.x) * q;
        a.sy += (vj.y - vi.y) * q;
        a.sz += (vj.z - vi.z) * q;
    }
}

// Sweep over particle i's neighbours (the density pass's traversal).
__device__ __forceinline__ void sur_sweep(const SceneDev &s, const uint32_t *__restrict__ start,
                                          const float4 *__restrict__ posm, const float4 *__restrict__ vel,
                                          const float4 *__restrict__ aux, uint32_t i, const float4 &pi, const float4 &vi,
                                          float rho_lag_i, SurAcc &a) {
    if (!finite3(pi.x, pi.y, pi.z)) return;
    int32_t cx, cy, cz;
    cell_of(pi.x, pi.y, pi.z, s.h_grid, cx, cy, cz);
    const float hg2 = s.h_grid2;
    if (needs_cell_path(s, start, cx, cy, cz)) {
        for_each_candidate(s, start, posm, cx, cy, cz, [&](uint32_t j) {
            const float4 pj = posm[j];
            const float r2 = dot3(pj.x - pi.x, pj.y - pi.y, pj.z - pi.z);
            if (r2 < hg2) sur_pair(s, i, j, pi, vi, rho_lag_i, pj, vel, aux, r2, a);
        });
        return;
    }
#pragma unroll 1
    for (int r = 0; r < 9; ++r) {
        uint32_t br, er;
        row_range(s, start, cx, cy, cz, r, br, er);
        const uint32_t lim = min(er, br + PH_ROW_BITS);
        uint32_t m = 0;
        uint32_t j = br;
        for (; j + 4 <= lim; j += 4) {
            const float4 q0 = posm[j], q1 = posm[j + 1], q2 = posm[j + 2], q3 = posm[j + 3];
            const uint32_t o = j - br;
            m |= (dot3(q0.x - pi.x, q0.y - pi.y, q0.z - pi.z) < hg2 ? 1u : 0u) << o;
            m |= (dot3(q1.x - pi.x, q1.y - pi.y, q1.z - pi.z) < hg2 ? 2u : 0u) << o;
            m |= (dot3(q2.x - pi.x, q2.y - pi.y, q2.z - pi.z) < hg2 ? 4u : 0u) << o;
            m |= (dot3(q3.x - pi.x, q3.y - pi.y, q3.z - pi.z) < hg2 ? 8u : 0u) << o;
        }
        for (; j < lim; ++j) {
            const float4 pj = posm[j];
            m |= (dot3(pj.x - pi.x, pj.y - pi.y, pj.z - pi.z) < hg2 ? 1u : 0u) << (j - br);
        }
        while (m) {
            const uint32_t k = __ffs(m) - 1;
            m &= m - 1;
            const float4 pj = posm[br + k];
            const float r2 = dot3(pj.x - pi.x, pj.y - pi.y, pj.z - pi.z);
            sur_pair(s, i, br + k, pi, vi, rho_lag_i, pj, vel, aux, r2, a);
        }
        for (j = lim; j < er; ++j) {
            const float4 pj = posm[j];
            const float r2 = dot3(pj.x - pi.x, pj.y - pi.y, pj.z - pi.z);
            if (r2 < hg2) sur_pair(s, i, j, pi, vi, rho_lag_i, pj, vel, aux, r2, a);
        }
    }
}

__device__ __forceinline__ float len3(float x, float y, float z) { return sqrtf(x * x + y * y + z * z); }

// Network inputs, basis vectors and the network-free acceleration.
struct SurFeat {
    float x[MLP_IN];
    float b[SUR_NCOEF][3];
    float a0[3];
    float rho, p, q;
};

__device__ __forceinline__ void sur_features(const SceneDev &s, float rho0, float qs, const SurAcc &a,
                                             float rho_lag_i, float dq_lag_i, SurFeat &f) {
    const float rho = rclampf(a.rho, 0.0f, s.rho_max);
    const float p = eos(rho, rho0, s.stiffness, s.gamma);
    const float rc = rmaxf(rho, PH_EPS);
    const float q = p / (rc * rc);
    f.rho = rho;
    f.p = p;
    f.q = q;
    const float inv = 1.0f / rc;
    f.a0[0] = -(q * a.gx + a.hx) * inv;
    f.a0[1] = -(q * a.gy + a.hy) * inv;
    f.a0[2] = -(q * a.gz + a.hz) * inv;
    const float t0 = -a.dx * inv, t1 = -a.dy * inv, t2 = -a.dz * inv;
    const float gl = len3(a.gx, a.gy, a.gz);
    const float gi = gl > 0.0f ? 1.0f / gl : 0.0f;
    const float a0l = len3(f.a0[0], f.a0[1], f.a0[2]), tl = len3(t0, t1, t2);
    f.b[0][0] = t0;
    f.b[0][1] = t1;
    f.b[0][2] = t2;
    f.b[1][0] = f.a0[0];
    f.b[1][1] = f.a0[1];
    f.b[1][2] = f.a0[2];
    f.b[2][0] = 10.0f * a.gx * gi;
    f.b[2][1] = 10.0f * a.gy * gi;
    f.b[2][2] = 10.0f * a.gz * gi;
    const float tiny = 1e-20f;
    f.x[0] = (rho / rho0 - 1.0f) * 100.0f;
    f.x[1] = (rho_lag_i / rho0 - 1.0f) * 100.0f;
    f.x[2] = log1pf(a0l * 0.1f);
    f.x[3] = log1pf(tl * 0.1f);
    f.x[4] = (f.a0[0] * t0 + f.a0[1] * t1 + f.a0[2] * t2) / (a0l * tl + tiny);
    f.x[5] = gl * s.h / rho0;
    f.x[6] = (f.a0[0] * a.gx + f.a0[1] * a.gy + f.a0[2] * a.gz) * gi / (a0l + tiny);
    f.x[7] = (float)a.cnt * (1.0f / 30.0f) - 1.0f;
    f.x[8] = log1pf(q / qs);
    f.x[9] = copysignf(log1pf(fabsf(dq_lag_i) / qs), dq_lag_i);
    f.x[10] = 1.0f;
    for (int c = 11; c < MLP_IN; ++c) f.x[c] = 0.0f;
}

}  // namespace

// mode SUR_MODE_STEP: surrogate substep (writes the advanced state and the
// new history).  mode SUR_MODE_COLLECT: after a plain substep on the same
// gathered state, write training samples (target = the plain pressure
// acceleration acc_target) for particles with id % stride == phase.
__global__ __launch_bounds__(MLP_ROWS) void k_sur_sweep(SurSweepArgs a) {
    __shared__ uint16_t xs[MLP_ROWS][MLP_IN];
    __shared__ uint16_t hs[4][16][MLP_HPAD];
    __shared__ float ys[MLP_ROWS][MLP_OUT];
    const uint32_t tile = xcd_tile(blockIdx.x, gridDim.x);
    const TileDev t = a.tiles[tile];
    const SceneDev s = a.scenes[t.scene];
    const uint32_t i = t.first + threadIdx.x;
    const bool active = i < s.part_end;
    SurFeat f;
    float4 pi = make_float4(0.0f, 0.0f, 0.0f, 0.0f), vi = pi, ai = pi;
    float xs_x = 0.0f, xs_y = 0.0f, xs_z = 0.0f;  // XSPH sum
    uint32_t tg = 0;
    if (active) {
        pi = a.posm[i];
        vi = a.vel[i];
        ai = a.aux[i];
        tg = a.tag[i];
        SurAcc acc;
        sur_sweep(s, a.start, a.posm, a.vel, a.aux, i, pi, vi, ai.x, acc);
        const float rho0 = a.scenes[t.scene].phase_rho0[min(tg >> 24, 3u)];
        sur_features(s, rho0, a.qs, acc, ai.x, ai.z, f);
        for (int c = 0; c < MLP_IN; ++c) xs[threadIdx.x][c] = mlp_bf16(f.x[c]);
        xs_x = acc.sx;
        xs_y = acc.sy;
        xs_z = acc.sz;
    } else {
        for (int c = 0; c < MLP_IN; ++c) xs[threadIdx.x][c] = 0;
    }
    if (a.mode == SUR_MODE_COLLECT) {
        if (active) {
            const uint32_t id = tg & PH_ID_MASK;
            if (id % a.stride == a.phase) {
                const uint32_t k = a.sample_base + id / a.stride;
                if (k < a.max_samples) {
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
            }
        }
        return;
    }
    // SUR_MODE_STEP: network coefficients for the 256 particles of the tile
    if (a.use_net) mlp_forward_block(a.mlp, xs, hs, ys);
    if (!active) return;
    float c[SUR_NCOEF];
    for (int k = 0; k < SUR_NCOEF; ++k) c[k] = a.use_net ? ys[threadIdx.x][k] : 0.0f;
    float ap[3];
    for (int d = 0; d < 3; ++d) ap[d] = f.a0[d] + c[0] * f.b[0][d] + c[1] * f.b[1][d] + c[2] * f.b[2][d];
    const float ax = s.gx + ap[0], ay = s.gy + ap[1], az = s.gz + ap[2];
    float ux = vi.x + ax * s.dt, uy = vi.y + ay * s.dt, uz = vi.z + az * s.dt;
    if (s.viscosity > 0.0f) {
        ux += xs_x * s.viscosity;
        uy += xs_y * s.viscosity;
        uz += xs_z * s.viscosity;
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
    clamp_to_bodies(s, a.bodies, x, y, z, ux, uy, uz);
    a.posm_out[i] = make_float4(x, y, z, pi.w);
    a.vel_out[i] = make_float4(ux, uy, uz, 0.0f);
    a.aux_out[i] = make_float4(f.rho, f.q, f.q - ai.y, 0.0f);
}

// aux_out[d] = aux_in[perm[d]]: history follows its particle through the sort.
__global__ __launch_bounds__(256) void k_gather_aux(uint32_t n, const uint32_t *__restrict__ perm,
                                                    const float4 *__restrict__ aux_in, float4 *__restrict__ aux_out) {
    const uint32_t d = blockIdx.x * 256 + threadIdx.x;
    if (d < n) aux_out[d] = aux_in[perm[d]];
}

// After a plain substep: new history (rho, P, P - P_prev) and, per tile,
// the bulk density deviation sum |rho/rho0 - 1| over particles with
// rho >= 0.9 rho0, and their count (the judge's measure).
__global__ __launch_bounds__(256) void k_aux_update(const TileDev *__restrict__ tiles, const SceneDev *__restrict__ scenes,
                                                    const float2 *__restrict__ rq, const float4 *__restrict__ aux_prev,
                                                    const uint32_t *__restrict__ tag, float4 *__restrict__ aux_out,
                                                    float2 *__restrict__ tile_err) {
    const uint32_t tile = blockIdx.x;
    const TileDev t = tiles[tile];
    const SceneDev &s = scenes[t.scene];
    const uint32_t i = t.first + threadIdx.x;
    float e = 0.0f, c = 0.0f;
    if (i < s.part_end) {
        const float2 v = rq[i];
        const float4 ap = aux_prev[i];
        aux_out[i] = make_float4(v.x, v.y, v.y - ap.y, 0.0f);
        const float rho0 = s.phase_rho0[min(tag[i] >> 24, 3u)];
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
        tile_err[tile] = make_float2(red[0][0] + red[1][0] + red[2][0] + red[3][0],
                                     red[0][1] + red[1][1] + red[2][1] + red[3][1]);
}

__global__ __launch_bounds__(256) void k_err_reduce(const float2 *__restrict__ tile_err, uint32_t ntiles,
                                                    float *__restrict__ out) {
    double e = 0.0, c = 0.0;
    for (uint32_t k = threadIdx.x; k < ntiles; k += 256) {
        e += tile_err[k].x;
        c += tile_err[k].y;
    }
    __shared__ double re[256], rc[256];
    re[threadIdx.x] = e;
    rc[threadIdx.x] = c;
    __syncthreads();
    for (int d = 128; d >= 1; d >>= 1) {
        if ((int)threadIdx.x < d) {
            re[threadIdx.x] += re[threadIdx.x + d];
            rc[threadIdx.x] += rc[threadIdx.x + d];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = (float)(rc[0] > 0 ? re[0] / rc[0] : 0.0);
        out[1] = (float)rc[0];
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
hipError_t launch_sur_sweep(hipStream_t st, uint32_t ntiles, const SurSweepArgs &a) {
    hipLaunchKernelGGL(k_sur_sweep, dim3(ntiles), dim3(MLP_ROWS), 0, st, a);
    return hipGetLastError();
}
hipError_t launch_gather_aux(hipStream_t st, uint32_t n, const uint32_t *perm, const float4 *in, float4 *out) {
    hipLaunchKernelGGL(k_gather_aux, dim3((n + 255) / 256), dim3(256), 0, st, n, perm, in, out);
    return hipGetLastError();
}
hipError_t launch_aux_update(hipStream_t st, uint32_t ntiles, const TileDev *tiles, const SceneDev *scenes,
                             const float2 *rq, const float4 *aux_prev, const uint32_t *tag, float4 *aux_out,
                             float2 *tile_err) {
    hipLaunchKernelGGL(k_aux_update, dim3(ntiles), dim3(256), 0, st, tiles, scenes, rq, aux_prev, tag, aux_out,
                       tile_err);
    return hipGetLastError();
}
hipError_t launch_err_reduce(hipStream_t st, const float2 *tile_err, uint32_t ntiles, float *out) {
    hipLaunchKernelGGL(k_err_reduce, dim3(1), dim3(256), 0, st, tile_err, ntiles, out);
    return hipGetLastError();
}
hipError_t launch_aux_reset(hipStream_t st, uint32_t n, const uint32_t *tag, float4 rho0, float4 *aux) {
    hipLaunchKernelGGL(k_aux_reset, dim3((n + 255) / 256), dim3(256), 0, st, n, tag, rho0, aux);
    return hipGetLastError();
}
}  // namespace ph
