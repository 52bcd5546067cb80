// surrogate.h — the learned pressure term: Surrogate3 (weights, Adam state,
// training samples) and the sample layout.
#pragma once
#include <hip/hip_runtime.h>

#include <random>
#include <vector>

#include "engine.h"
#include "mlp.h"

// One training sample (floats): MLP input x[16], the SUR_NCOEF basis
// vectors b_k (3 each), the network-free acceleration a0, the target
// acceleration a_t (plain SPH pressure part), and the sample's scale s.
#define SUR_NCOEF 3
#define SUR_X 0
#define SUR_B 16
#define SUR_A0 (SUR_B + 3 * SUR_NCOEF)
#define SUR_AT (SUR_A0 + 3)
#define SUR_S (SUR_AT + 3)
#define SUR_SAMPLE 32
// Most samples one Surrogate3 holds (2 GiB of fp32 samples).
#define SUR_MAX_SAMPLES (1u << 24)

struct Surrogate3 {
    int device = -1;
    hipStream_t stream = nullptr;
    ph::DBuf<float> params, grad, mom, vel, loss, data;
    ph::DBuf<uint16_t> w1, w2;
    ph::DBuf<uint32_t> idx;
    std::vector<float> host;
    uint32_t nsamples = 0, sample_cap = 0;
    uint64_t adam_t = 0;

    ~Surrogate3();
    void init(uint32_t seed);
    void upload_params();
    void download_params();
    void zero(float *p, uint32_t n);
    void reserve_samples(uint32_t cap);
    // Room for `need` samples in total: geometric growth, at most
    // SUR_MAX_SAMPLES (throws beyond it).
    void grow_samples(uint64_t need);
    void run_epoch(uint32_t batch, float lr, std::mt19937 &rng, bool train, double &lsum, double &l0sum);
    MlpDev dev() const;
};
