// surrogate_step.h — launch interface of the surrogate substep kernels.
#pragma once
#include <hip/hip_runtime.h>

#include "mlp.h"
#include "sph_device.h"

// Training samples from a plain substep: the surrogate's inputs on the
// gathered state and the plain pressure acceleration as the target.
struct SurCollectArgs {
    const TileDev *tiles;
    const SceneDev *scenes;
    const uint32_t *start;
    const float4 *posm, *vel, *aux;  // gathered state + previous-substep history
    const uint32_t *tag;
    float4 mass;  // particle mass of phases 0..3
    float qs;  // pressure-quotient scale of the inputs
    const float4 *acc_target;
    float *samples;
    uint32_t sample_base, stride, phase, max_samples;
};

// One surrogate substep (sweep, then update) over the order and cell
// table of the last sort.
struct SurStepArgs {
    int use_net;  // 0: the network-free method (c = 0)
    const TileDev *tiles;
    const SceneDev *scenes;
    const BodyDev *bodies;
    const uint32_t *start;
    const float4 *posm, *vel, *aux;
    const uint32_t *tag;
    // skin list written by the density pass of the last plain substep
    const int4 *cell;      // skin list: cell at list time
    const uint32_t *cnt;   // skin list: candidate count | SLF_ flags
    const uint32_t *idx;   // skin list: candidates 4g..4g+3 of particle i as uint4 g * mstride + i
    uint32_t mstride;
    // neighbour records of the current state: (x, y, z, P with the phase
    // in its 2 low bits) and (v, dP) as halves and a bfloat16
    SurRec rec;
    float4 mass;  // particle mass of phases 0..3
    MlpDev mlp;
    float qs;
    float4 *posm_out, *vel_out, *aux_out;
    SurRec rec_out;
    // the judge's evaluation (launch_sur_judge): sample k is slot
    // judge_slot(k, judge_stride, judge_off); its acceleration without
    // gravity goes to judge_acc[k]
    float4 *judge_acc;
    uint32_t judge_stride, judge_off;
};

namespace ph {
hipError_t launch_sur_collect(hipStream_t st, uint32_t ntiles, const SurCollectArgs &a);
hipError_t launch_sur_step(hipStream_t st, uint32_t ntiles, const SurStepArgs &a);
hipError_t launch_sur_judge(hipStream_t st, uint32_t nsamples, const SurStepArgs &a);
hipError_t launch_gather_aux(hipStream_t st, uint32_t n, const uint32_t *perm, const float4 *in, float4 *out);
hipError_t launch_aux_update(hipStream_t st, uint32_t ntiles, const TileDev *tiles, const SceneDev *scenes,
                             const float2 *rq, const float4 *aux_prev, const uint32_t *tag, float4 *aux_out,
                             float4 *tile_err, const float4 *posm_new = nullptr, const float4 *vel_new = nullptr,
                             SurRec rec_out = SurRec());
hipError_t launch_err_reduce(hipStream_t st, const float4 *tile_err, uint32_t ntiles, float *out, uint32_t *slow,
                             const float *vmax);
hipError_t launch_aux_reset(hipStream_t st, uint32_t n, const uint32_t *tag, float4 rho0, float4 *aux);
}  // namespace ph
