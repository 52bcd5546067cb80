// resample.h — adaptive resolution (FluidState3::resample, fluid3.rs:923-987;
// the 2D twin fluid.rs:826-891) on the device: merge of close dense pairs,
// removal of the absorbed particles, split of thin particles, renumbering.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "sph_device.h"

namespace ph {

struct ResampleArgs {
    // the sorted state of this substep (one scene, slots 0 .. n-1) and its
    // densities from the density pass
    const TileDev *tiles;
    const SceneDev *scenes;
    const uint32_t *start;
    const float4 *posm;
    const float4 *vel;
    const uint32_t *tag;
    const float2 *rq;
    uint32_t n;
    float merge_hi;  // merge_density_ratio * rest_density.max(EPS)
    float split_lo;  // split_density_ratio * rest_density.max(EPS)
    float min_mass;  // particle_mass.max(EPS) * 0.125
    float md2;       // (0.25 h) * (0.25 h), h = smoothing_length.max(EPS)
    float off;       // 0.25 h, the split offset along x
    uint32_t cap;    // split cap: the output stays below it (MAX_PARTICLES_3D)
    // work arrays (n each unless stated)
    unsigned long long *ctr;  // [0] takers, [1] merge candidates, [2] splits accepted, [3] particles kept
    uint32_t *flag;           // per slot: bit 0 absorbed; bits 2.. taker index + 1 (0: not a taker)
    uint4 *takers;            // per taker: slot, index, first candidate, candidate count
    uint32_t *items;          // candidate slots, one run per taker (pcap entries)
    uint32_t pcap;
    uint32_t *minT;           // per slot: lowest undecided taker index listing it
    uint32_t *pick;           // per taker: the absorbed slot, or ~0
    uint32_t *state;          // per taker: 0 undecided, 1 decided, 2 merged values final
    float4 *mv;               // per taker: velocity, mass after its merge chain
    uint32_t *cK, *cE;        // per index: kept, split-eligible (scan inputs, zeroed by the scans)
    uint32_t *rK, *rE;        // their exclusive scans (n + 1)
    uint32_t *slot_of;        // per index: its slot
    uint32_t *partial;        // scan partials
    // the new particle set, in its new index order
    float4 *posm_out;
    float4 *vel_out;
    uint32_t *tag_out;
};

// Launches the whole pass on the stream (no host synchronisation).  After
// it, ctr[1] > pcap means the candidate store was too small (grow it to
// ctr[1] and run again); else the new count is ctr[3] + ctr[2] and, unless
// nothing changed (ctr[3] == n and ctr[2] == 0), the new set is in
// posm_out / vel_out / tag_out.
hipError_t launch_resample(hipStream_t st, const ResampleArgs &a);

// Per particle, in index order: its mass and phase (n slots of the state).
hipError_t launch_export_mass(hipStream_t st, uint32_t n, const float4 *posm, const uint32_t *tag, float *mass,
                              uint8_t *phase);

}  // namespace ph
