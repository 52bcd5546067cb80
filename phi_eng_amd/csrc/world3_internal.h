// world3_internal.h — what the batch needs from a World3 (world3.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "../../include/phieng_fluid3.h"
#include "engine.h"

uint32_t world3_count(const World3 *w);
const FluidConfig3Ffi &world3_config(const World3 *w);
const std::vector<BodyDev> &world3_bodies(const World3 *w);
ph::SceneHost world3_scene_host(const World3 *w, float dt);
// Current device state of the world's particles (sorted order + tags).
void world3_device_state(const World3 *w, const float4 **posm, const float4 **vel, const uint32_t **tag,
                         hipStream_t *stream);
// Same, writable; marks the world's grid domain and records stale.
void world3_device_state_mut(World3 *w, float4 **posm, float4 **vel, uint32_t **tag, hipStream_t *stream);
