// surrogate_step.h — launch interface of the surrogate substep kernels.
#pragma once
#include <hip/hip_runtime.h>

#include "mlp.h"
#include "sph_device.h"

#define SUR_MODE_STEP 0
#define SUR_MODE_COLLECT 1

struct SurSweepArgs {
    int mode;
    int use_net;  // 0: the network-free method (c = 0)
    const TileDev *tiles;
    const SceneDev *scenes;
    const BodyDev *bodies;
    const uint32_t *start;
    const float4 *posm, *vel, *aux;  // gathered state + previous-substep history
    const uint32_t *tag;
    MlpDev mlp;
    float qs;  // pressure-quotient scale of the inputs
    float4 *posm_out, *vel_out, *aux_out;
    // collect mode
    const float4 *acc_target;
    float *samples;
    uint32_t sample_base, stride, phase, max_samples;
};

namespace ph {
hipError_t launch_sur_sweep(hipStream_t st, uint32_t ntiles, const SurSweepArgs &a);
hipError_t launch_gather_aux(hipStream_t st, uint32_t n, const uint32_t *perm, const float4 *in, float4 *out);
hipError_t launch_aux_update(hipStream_t st, uint32_t ntiles, const TileDev *tiles, const SceneDev *scenes,
                             const float2 *rq, const float4 *aux_prev, const uint32_t *tag, float4 *aux_out,
                             float2 *tile_err);
hipError_t launch_err_reduce(hipStream_t st, const float2 *tile_err, uint32_t ntiles, float *out);
hipError_t launch_aux_reset(hipStream_t st, uint32_t n, const uint32_t *tag, float4 rho0, float4 *aux);
}  // namespace ph
