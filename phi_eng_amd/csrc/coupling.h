// coupling.h — launchers of the dynamic-body coupling kernels (coupling.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "sph_device.h"

namespace ph {

// The slots in defer[1 .. count] (the force pass's list of particles that
// touch a dynamic body): their state to `out` (2 float4 each) and back.
hipError_t launch_defer_gather(hipStream_t st, const uint32_t *defer, uint32_t count, const float4 *posm,
                               const float4 *vel, const uint32_t *tag, float4 *out);
hipError_t launch_defer_scatter(hipStream_t st, const uint32_t *defer, uint32_t count, const float4 *in,
                                float4 *posm, float4 *vel);

// Terms of the sampled pressure (one scene, the state after the clamp).
struct SampleArgs {
    const SceneDev *scenes;
    uint32_t n;
    const float4 *posm;
    const float2 *rq;      // (clamped density, -) of this substep
    const uint32_t *tag;
    const float4 *bound;   // per dynamic body: centre, squared radius that holds its samples' h-balls
    const uint2 *range;    // per dynamic body: first sample, sample count
    uint32_t nbodies;
    const float4 *spos;    // sample points
    const int4 *scell;     // their cells at h = max(smoothing_length, EPS)
    uint4 *rec;            // (sample, key, term bits, 0)
    uint32_t *count;       // records wanted (may exceed cap: then grow and run again)
    uint32_t cap;
};
hipError_t launch_sample_terms(hipStream_t st, const SampleArgs &a);

}  // namespace ph
