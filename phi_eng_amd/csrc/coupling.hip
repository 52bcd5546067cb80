// coupling.hip — device side of the two-way coupling with dynamic bodies
// (fluid3.rs:1141-1213).
//
// The force pass leaves a particle that touches a dynamic body unclamped
// and lists its slot; these kernels move those particles' state to and
// from a compact staging array for the host pass that replays the
// reference's clamp loop in particle order, and gather the terms of the
// pressure the reference samples on each dynamic body's surface
// (sample_pressure, fluid3.rs:345-369) after the clamp.
#include <hip/hip_runtime.h>

#include "coupling.h"

namespace {

// Staging record of deferred particle k: (x, y, z, mass), (vx, vy, vz, tag bits).
__global__ __launch_bounds__(256) void k_defer_gather(const uint32_t *__restrict__ defer, uint32_t count,
                                                      const float4 *__restrict__ posm,
                                                      const float4 *__restrict__ vel,
                                                      const uint32_t *__restrict__ tag, float4 *__restrict__ out) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= count) return;
    const uint32_t i = defer[1 + k];
    out[2 * (size_t)k] = posm[i];
    float4 v = vel[i];
    v.w = __uint_as_float(tag[i]);
    out[2 * (size_t)k + 1] = v;
}

__global__ __launch_bounds__(256) void k_defer_scatter(const uint32_t *__restrict__ defer, uint32_t count,
                                                       const float4 *__restrict__ in, float4 *__restrict__ posm,
                                                       float4 *__restrict__ vel) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= count) return;
    const uint32_t i = defer[1 + k];
    const float4 p = in[2 * (size_t)k], v = in[2 * (size_t)k + 1];
    posm[i] = p;
    vel[i] = make_float4(v.x, v.y, v.z, 0.0f);
}

// One thread per particle.  For every sample point within h of it whose
// cell is one of the particle's 27 (the cells sample_pressure visits):
// the term m_j (p_j / max(rho_j, EPS)) poly6(r2, h) and its visiting key
// (cell offset rank oz, oy, ox, then the particle's original index).
__global__ __launch_bounds__(256) void k_sample_terms(ph::SampleArgs a) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const SceneDev &s = a.scenes[0];
    const float4 p = a.posm[i];
    if (!finite3(p.x, p.y, p.z)) return;  // rebuild_grid skips non-finite positions
    int32_t cx, cy, cz;
    cell_of(p.x, p.y, p.z, s.h_grid, cx, cy, cz);
    const uint32_t tg = a.tag[i];
    const float rho = a.rq[i].x;
    const float pres = eos(rho, s.phase_rho0[min(tg >> 24, 3u)], s.stiffness, s.gamma);
    const float c0 = p.w * (pres / rmaxf(rho, PH_EPS));
    for (uint32_t b = 0; b < a.nbodies; ++b) {
        const float4 bb = a.bound[b];
        if (!(dot3(p.x - bb.x, p.y - bb.y, p.z - bb.z) < bb.w)) continue;
        const uint2 range = a.range[b];
        for (uint32_t k = range.x; k < range.x + range.y; ++k) {
            const float4 q = a.spos[k];
            const float r2 = dot3(p.x - q.x, p.y - q.y, p.z - q.z);
            if (!(r2 < s.h2)) continue;
            const int4 sc = a.scell[k];
            const int64_t ox = (int64_t)cx - sc.x, oy = (int64_t)cy - sc.y, oz = (int64_t)cz - sc.z;
            if (ox < -1 || ox > 1 || oy < -1 || oy > 1 || oz < -1 || oz > 1) continue;
            const uint32_t rank = (uint32_t)(((oz + 1) * 3 + (oy + 1)) * 3 + (ox + 1));
            const uint32_t slot = atomicAdd(a.count, 1u);
            if (slot < a.cap)
                a.rec[slot] = make_uint4(k, (rank << 24) | (tg & PH_ID_MASK), __float_as_uint(c0 * poly6(r2, s)), 0u);
        }
    }
}

}  // namespace

namespace ph {

hipError_t launch_defer_gather(hipStream_t st, const uint32_t *defer, uint32_t count, const float4 *posm,
                               const float4 *vel, const uint32_t *tag, float4 *out) {
    if (!count) return hipSuccess;
    hipLaunchKernelGGL(k_defer_gather, dim3((count + 255) / 256), dim3(256), 0, st, defer, count, posm, vel, tag, out);
    return hipGetLastError();
}

hipError_t launch_defer_scatter(hipStream_t st, const uint32_t *defer, uint32_t count, const float4 *in,
                                float4 *posm, float4 *vel) {
    if (!count) return hipSuccess;
    hipLaunchKernelGGL(k_defer_scatter, dim3((count + 255) / 256), dim3(256), 0, st, defer, count, in, posm, vel);
    return hipGetLastError();
}

hipError_t launch_sample_terms(hipStream_t st, const SampleArgs &a) {
    if (!a.n) return hipSuccess;
    hipLaunchKernelGGL(k_sample_terms, dim3((a.n + 255) / 256), dim3(256), 0, st, a);
    return hipGetLastError();
}

}  // namespace ph
