// surrogate.hip — MLP training / evaluation kernels and the Surrogate3
// object (learned pressure term, no reference counterpart).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <mutex>
#include <random>
#include <string>
#include <vector>

#include "../../include/phieng_fluid3.h"
#include "engine.h"
#include "mlp.h"
#include "surrogate.h"

// ---------------------------------------------------------------- scratch
// Staging of phieng_surrogate3_mlp_forward_noalloc: the input rows, then the
// output rows, of at most SUR_NOALLOC_MAX_ROWS rows (8 MiB).  Each GPU the
// library runs on holds its own copy until the process ends.  Host code
// never passes this name to a kernel or a copy: it uses the device address
// that scratch_on_device() looked up for the current GPU.
#define SUR_NOALLOC_MAX_ROWS 65536u
static __device__ __align__(256) float scratch[(size_t)SUR_NOALLOC_MAX_ROWS * (MLP_IN + MLP_OUT)];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
float *g_scratch_dev[kMaxDevices] = {};

// The device address of `scratch` on `device`, which must be the current
// device; looked up once per device and kept.
float *scratch_on_device(int device) {
    if (device < 0 || device >= kMaxDevices) throw std::runtime_error("device index out of range for the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[device]) {
        void *p = nullptr;
        PH_CK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) throw std::runtime_error("hipGetSymbolAddress returned a null scratch address");
        g_scratch_dev[device] = static_cast<float *>(p);
    }
    return g_scratch_dev[device];
}
}  // namespace

// ---------------------------------------------------------------- kernels
// Plain MLP evaluation of `rows` input rows (fp32 in, rounded to bf16).
__global__ __launch_bounds__(MLP_ROWS) void k_mlp_forward(const float *__restrict__ x, uint32_t rows, MlpDev m,
                                                          float *__restrict__ y) {
    __shared__ uint16_t xs[MLP_ROWS][MLP_IN];
    __shared__ uint16_t hs[4][16][MLP_HPAD];
    __shared__ float ys[MLP_ROWS][MLP_OUT];
    const uint32_t r0 = blockIdx.x * MLP_ROWS, row = r0 + threadIdx.x;
    for (int c = 0; c < MLP_IN; ++c) xs[threadIdx.x][c] = mlp_bf16(row < rows ? x[(size_t)row * MLP_IN + c] : 0.0f);
    mlp_forward_block(m, xs, hs, ys);
    if (row < rows)
        for (int c = 0; c < MLP_OUT; ++c) y[(size_t)row * MLP_OUT + c] = ys[threadIdx.x][c];
}

// Parameter gradient of sum_rows dy . y(x) (dy fp32, rounded to bf16).
__global__ __launch_bounds__(MLP_ROWS) void k_mlp_backward(const float *__restrict__ x, const float *__restrict__ dy,
                                                           uint32_t rows, MlpDev m, float *__restrict__ grad) {
    __shared__ uint16_t xs[MLP_ROWS][MLP_IN];
    __shared__ uint16_t hs[4][16][MLP_HPAD];
    __shared__ uint16_t dzs[4][16][MLP_HPAD];
    __shared__ uint16_t dys[MLP_ROWS][MLP_OUT];
    const uint32_t row = blockIdx.x * MLP_ROWS + threadIdx.x;
    for (int c = 0; c < MLP_IN; ++c) xs[threadIdx.x][c] = mlp_bf16(row < rows ? x[(size_t)row * MLP_IN + c] : 0.0f);
    for (int c = 0; c < MLP_OUT; ++c)
        dys[threadIdx.x][c] = mlp_bf16(row < rows ? dy[(size_t)row * MLP_OUT + c] : 0.0f);
    mlp_backward_block(m, xs, hs, dzs, dys, grad);
}

// One training batch: forward, the acceleration loss of each sample
//   a_pred = a0 + sum_k c_k b_k,  L = mean_s |a_pred - a_t|^2 / s^2
// (c = the first SUR_NCOEF outputs), its output gradient, backward.
// Sample layout: SUR_SAMPLE floats (see surrogate.h).
__global__ __launch_bounds__(MLP_ROWS) void k_mlp_train(const float *__restrict__ data, const uint32_t *__restrict__ idx,
                                                        uint32_t batch, MlpDev m, float *__restrict__ grad,
                                                        float *__restrict__ loss, int with_backward) {
    __shared__ uint16_t xs[MLP_ROWS][MLP_IN];
    __shared__ uint16_t hs[4][16][MLP_HPAD];
    __shared__ uint16_t dzs[4][16][MLP_HPAD];
    __shared__ uint16_t dys[MLP_ROWS][MLP_OUT];
    __shared__ float ys[MLP_ROWS][MLP_OUT];
    __shared__ float lsum[MLP_ROWS / 64][2];
    const uint32_t k = blockIdx.x * MLP_ROWS + threadIdx.x;
    const bool valid = k < batch;
    const float *smp = valid ? data + (size_t)idx[k] * SUR_SAMPLE : nullptr;
    for (int c = 0; c < MLP_IN; ++c) xs[threadIdx.x][c] = mlp_bf16(valid ? smp[SUR_X + c] : 0.0f);
    mlp_forward_block(m, xs, hs, ys);
    float l = 0.0f, l0 = 0.0f;
    float dyv[MLP_OUT];
    for (int c = 0; c < MLP_OUT; ++c) dyv[c] = 0.0f;
    if (valid) {
        const float inv_s2 = 1.0f / (smp[SUR_S] * smp[SUR_S]);
        float d[3], d0[3];
        for (int a = 0; a < 3; ++a) {
            float pred = smp[SUR_A0 + a];
            for (int c = 0; c < SUR_NCOEF; ++c) pred += ys[threadIdx.x][c] * smp[SUR_B + 3 * c + a];
            d[a] = pred - smp[SUR_AT + a];
            d0[a] = smp[SUR_A0 + a] - smp[SUR_AT + a];
        }
        l = (d[0] * d[0] + d[1] * d[1] + d[2] * d[2]) * inv_s2;
        l0 = (d0[0] * d0[0] + d0[1] * d0[1] + d0[2] * d0[2]) * inv_s2;
        const float sc = 2.0f * inv_s2 / (float)batch;
        for (int c = 0; c < SUR_NCOEF; ++c)
            dyv[c] = sc * (d[0] * smp[SUR_B + 3 * c] + d[1] * smp[SUR_B + 3 * c + 1] + d[2] * smp[SUR_B + 3 * c + 2]);
    }
    for (int c = 0; c < MLP_OUT; ++c) dys[threadIdx.x][c] = mlp_bf16(dyv[c]);
    for (int d = 32; d >= 1; d >>= 1) {
        l += __shfl_xor(l, d, 64);
        l0 += __shfl_xor(l0, d, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        lsum[threadIdx.x >> 6][0] = l;
        lsum[threadIdx.x >> 6][1] = l0;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        atomicAdd(&loss[0], lsum[0][0] + lsum[1][0] + lsum[2][0] + lsum[3][0]);
        atomicAdd(&loss[1], lsum[0][1] + lsum[1][1] + lsum[2][1] + lsum[3][1]);
    }
    if (with_backward) mlp_backward_block(m, xs, hs, dzs, dys, grad);
}

// Adam on the fp32 master parameters, then the bf16 weight copies.
__global__ void k_adam(float *__restrict__ p, float *__restrict__ g, float *__restrict__ mo, float *__restrict__ ve,
                       uint32_t n, float lr, float b1, float b2, float eps, float bc1, float bc2,
                       uint16_t *__restrict__ w1, uint16_t *__restrict__ w2) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * mo[i] + (1.0f - b1) * gi;
    const float vi = b2 * ve[i] + (1.0f - b2) * gi * gi;
    mo[i] = mi;
    ve[i] = vi;
    const float pi = p[i] - lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
    p[i] = pi;
    g[i] = 0.0f;
    if (i < MLP_OFF_B1) w1[i - MLP_OFF_W1] = mlp_bf16(pi);
    else if (i >= MLP_OFF_W2 && i < MLP_OFF_B2) w2[i - MLP_OFF_W2] = mlp_bf16(pi);
}

__global__ void k_fill_f32(float *__restrict__ p, uint32_t n, float v) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// ---------------------------------------------------------------- host
namespace {
thread_local std::string g_sur_error;
void set_err(const char *w, const char *m) { g_sur_error = std::string(w) + ": " + m; }
}  // namespace

Surrogate3::~Surrogate3() {
    if (stream) {
        (void)hipStreamSynchronize(stream);
        (void)hipStreamDestroy(stream);
    }
}

void Surrogate3::init(uint32_t seed) {
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) throw std::runtime_error("no HIP device available");
    PH_CK(hipGetDevice(&device));
    PH_CK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    (void)scratch_on_device(device);  // set-up of the allocation-free forward on this device
    params.ensure(MLP_NPARAM);
    grad.ensure(MLP_NPARAM);
    mom.ensure(MLP_NPARAM);
    vel.ensure(MLP_NPARAM);
    w1.ensure(MLP_IN * MLP_HID);
    w2.ensure(MLP_HID * MLP_OUT);
    loss.ensure(2);
    std::mt19937 rng(seed);
    std::normal_distribution<float> nd(0.0f, 1.0f);
    host.assign(MLP_NPARAM, 0.0f);
    for (int i = 0; i < MLP_IN * MLP_HID; ++i) host[MLP_OFF_W1 + i] = nd(rng) * (1.0f / 4.0f);
    for (int i = 0; i < MLP_HID * MLP_OUT; ++i) host[MLP_OFF_W2 + i] = nd(rng) * 1e-3f;
    upload_params();
}

void Surrogate3::upload_params() {
    PH_CK(hipMemcpyAsync(params.p, host.data(), sizeof(float) * MLP_NPARAM, hipMemcpyHostToDevice, stream));
    std::vector<uint16_t> h1(MLP_IN * MLP_HID), h2(MLP_HID * MLP_OUT);
    for (size_t i = 0; i < h1.size(); ++i) h1[i] = mlp_bf16(host[MLP_OFF_W1 + i]);
    for (size_t i = 0; i < h2.size(); ++i) h2[i] = mlp_bf16(host[MLP_OFF_W2 + i]);
    PH_CK(hipMemcpyAsync(w1.p, h1.data(), 2 * h1.size(), hipMemcpyHostToDevice, stream));
    PH_CK(hipMemcpyAsync(w2.p, h2.data(), 2 * h2.size(), hipMemcpyHostToDevice, stream));
    zero(grad.p, MLP_NPARAM);
    zero(mom.p, MLP_NPARAM);
    zero(vel.p, MLP_NPARAM);
    adam_t = 0;
    PH_CK(hipStreamSynchronize(stream));
}

void Surrogate3::zero(float *p, uint32_t n) {
    hipLaunchKernelGGL(k_fill_f32, dim3((n + 255) / 256), dim3(256), 0, stream, p, n, 0.0f);
    PH_CK(hipGetLastError());
}

void Surrogate3::download_params() {
    host.resize(MLP_NPARAM);
    PH_CK(hipMemcpyAsync(host.data(), params.p, sizeof(float) * MLP_NPARAM, hipMemcpyDeviceToHost, stream));
    PH_CK(hipStreamSynchronize(stream));
}

MlpDev Surrogate3::dev() const {
    MlpDev m;
    m.w1 = w1.p;
    m.b1 = params.p + MLP_OFF_B1;
    m.w2 = w2.p;
    m.b2 = params.p + MLP_OFF_B2;
    return m;
}

void Surrogate3::reserve_samples(uint32_t cap) {
    if (cap <= sample_cap && data.p) return;
    ph::DBuf<float> nd;
    nd.ensure((size_t)cap * SUR_SAMPLE);
    if (nsamples)
        PH_CK(hipMemcpyAsync(nd.p, data.p, sizeof(float) * SUR_SAMPLE * nsamples, hipMemcpyDeviceToDevice, stream));
    PH_CK(hipStreamSynchronize(stream));
    std::swap(data.p, nd.p);
    std::swap(data.n, nd.n);
    sample_cap = cap;
}

void Surrogate3::grow_samples(uint64_t need) {
    if (need > SUR_MAX_SAMPLES) throw std::runtime_error("surrogate sample store full (at most 2^24 samples)");
    if (need <= sample_cap && data.p) return;
    const uint64_t cap = std::min<uint64_t>(std::max<uint64_t>(need, 2ull * sample_cap), SUR_MAX_SAMPLES);
    reserve_samples((uint32_t)cap);
}

// One pass over the dataset in shuffled batches (train) or in order (eval).
void Surrogate3::run_epoch(uint32_t batch, float lr, std::mt19937 &rng, bool train, double &lsum, double &l0sum) {
    std::vector<uint32_t> perm(nsamples);
    std::iota(perm.begin(), perm.end(), 0u);
    if (train) std::shuffle(perm.begin(), perm.end(), rng);
    idx.ensure(nsamples);
    PH_CK(hipMemcpyAsync(idx.p, perm.data(), sizeof(uint32_t) * nsamples, hipMemcpyHostToDevice, stream));
    std::vector<float> lh;
    for (uint32_t b0 = 0; b0 < nsamples; b0 += batch) {
        const uint32_t bs = std::min(batch, nsamples - b0);
        zero(loss.p, 2);
        hipLaunchKernelGGL(k_mlp_train, dim3((bs + MLP_ROWS - 1) / MLP_ROWS), dim3(MLP_ROWS), 0, stream, data.p,
                           idx.p + b0, bs, dev(), grad.p, loss.p, train ? 1 : 0);
        PH_CK(hipGetLastError());
        if (train) {
            ++adam_t;
            const float b1 = 0.9f, b2 = 0.999f;
            const float bc1 = 1.0f - std::pow(b1, (float)adam_t), bc2 = 1.0f - std::pow(b2, (float)adam_t);
            hipLaunchKernelGGL(k_adam, dim3((MLP_NPARAM + 255) / 256), dim3(256), 0, stream, params.p, grad.p, mom.p,
                               vel.p, (uint32_t)MLP_NPARAM, lr, b1, b2, 1e-8f, bc1, bc2, w1.p, w2.p);
            PH_CK(hipGetLastError());
        }
        float hl[2];
        PH_CK(hipMemcpyAsync(hl, loss.p, sizeof(hl), hipMemcpyDeviceToHost, stream));
        PH_CK(hipStreamSynchronize(stream));
        lsum += hl[0];
        l0sum += hl[1];
    }
}

extern "C" {

const char *phieng_surrogate3_last_error(void) { return g_sur_error.c_str(); }

uint32_t phieng_surrogate3_num_params(void) { return MLP_NPARAM; }

Surrogate3 *phieng_surrogate3_new(uint32_t seed) {
    Surrogate3 *s = nullptr;
    try {
        s = new Surrogate3();
        s->init(seed);
        return s;
    } catch (const std::exception &e) {
        set_err("phieng_surrogate3_new", e.what());
        delete s;
        return nullptr;
    }
}

void phieng_surrogate3_free(Surrogate3 *s) {
    try {
        delete s;
    } catch (...) {
    }
}

int32_t phieng_surrogate3_get_params(Surrogate3 *s, float *out, uint32_t n) {
    if (!s || !out || n < MLP_NPARAM) return 0;
    try {
        s->download_params();
        std::memcpy(out, s->host.data(), sizeof(float) * MLP_NPARAM);
        return 1;
    } catch (const std::exception &e) {
        set_err("phieng_surrogate3_get_params", e.what());
        return 0;
    }
}

int32_t phieng_surrogate3_set_params(Surrogate3 *s, const float *in, uint32_t n) {
    if (!s || !in || n < MLP_NPARAM) return 0;
    for (uint32_t i = 0; i < MLP_NPARAM; ++i)
        if (!std::isfinite(in[i])) return 0;
    try {
        s->host.assign(in, in + MLP_NPARAM);
        s->upload_params();
        return 1;
    } catch (const std::exception &e) {
        set_err("phieng_surrogate3_set_params", e.what());
        return 0;
    }
}

int32_t phieng_surrogate3_mlp_forward(Surrogate3 *s, const float *x, uint32_t rows, float *y) {
    if (!s || !x || !y || rows == 0) return 0;
    try {
        ph::DBuf<float> dx, dy;
        dx.ensure((size_t)rows * MLP_IN);
        dy.ensure((size_t)rows * MLP_OUT);
        PH_CK(hipMemcpyAsync(dx.p, x, sizeof(float) * rows * MLP_IN, hipMemcpyHostToDevice, s->stream));
        hipLaunchKernelGGL(k_mlp_forward, dim3((rows + MLP_ROWS - 1) / MLP_ROWS), dim3(MLP_ROWS), 0, s->stream, dx.p,
                           rows, s->dev(), dy.p);
        PH_CK(hipGetLastError());
        PH_CK(hipMemcpyAsync(y, dy.p, sizeof(float) * rows * MLP_OUT, hipMemcpyDeviceToHost, s->stream));
        PH_CK(hipStreamSynchronize(s->stream));
        return 1;
    } catch (const std::exception &e) {
        set_err("phieng_surrogate3_mlp_forward", e.what());
        return 0;
    }
}

int32_t phieng_surrogate3_mlp_forward_noalloc(Surrogate3 *s, const float *x, uint32_t rows, float *y) {
    if (!s || !x || !y || rows == 0) {
        set_err("phieng_surrogate3_mlp_forward_noalloc", "null argument or zero rows");
        return 0;
    }
    if (rows > SUR_NOALLOC_MAX_ROWS) {
        set_err("phieng_surrogate3_mlp_forward_noalloc",
                ("rows = " + std::to_string(rows) + " exceeds the scratch maximum of " +
                 std::to_string(SUR_NOALLOC_MAX_ROWS) + " rows").c_str());
        return 0;
    }
    try {
        // No allocation: inputs and outputs are staged in this GPU's static
        // scratch (looked up at Surrogate3 set-up).  Calls that overlap on
        // one GPU share that array and race.
        float *sx = scratch_on_device(s->device);
        float *sy = sx + (size_t)rows * MLP_IN;
        PH_CK(hipMemcpyAsync(sx, x, sizeof(float) * rows * MLP_IN, hipMemcpyHostToDevice, s->stream));
        hipLaunchKernelGGL(k_mlp_forward, dim3((rows + MLP_ROWS - 1) / MLP_ROWS), dim3(MLP_ROWS), 0, s->stream, sx,
                           rows, s->dev(), sy);
        PH_CK(hipGetLastError());
        PH_CK(hipMemcpyAsync(y, sy, sizeof(float) * rows * MLP_OUT, hipMemcpyDeviceToHost, s->stream));
        PH_CK(hipStreamSynchronize(s->stream));
        return 1;
    } catch (const std::exception &e) {
        set_err("phieng_surrogate3_mlp_forward_noalloc", e.what());
        return 0;
    }
}

int32_t phieng_surrogate3_mlp_backward(Surrogate3 *s, const float *x, const float *dy, uint32_t rows, float *grad) {
    if (!s || !x || !dy || !grad || rows == 0) return 0;
    try {
        ph::DBuf<float> dx, ddy, dg;
        dx.ensure((size_t)rows * MLP_IN);
        ddy.ensure((size_t)rows * MLP_OUT);
        dg.ensure(MLP_NPARAM);
        PH_CK(hipMemcpyAsync(dx.p, x, sizeof(float) * rows * MLP_IN, hipMemcpyHostToDevice, s->stream));
        PH_CK(hipMemcpyAsync(ddy.p, dy, sizeof(float) * rows * MLP_OUT, hipMemcpyHostToDevice, s->stream));
        s->zero(dg.p, MLP_NPARAM);
        hipLaunchKernelGGL(k_mlp_backward, dim3((rows + MLP_ROWS - 1) / MLP_ROWS), dim3(MLP_ROWS), 0, s->stream, dx.p,
                           ddy.p, rows, s->dev(), dg.p);
        PH_CK(hipGetLastError());
        PH_CK(hipMemcpyAsync(grad, dg.p, sizeof(float) * MLP_NPARAM, hipMemcpyDeviceToHost, s->stream));
        PH_CK(hipStreamSynchronize(s->stream));
        return 1;
    } catch (const std::exception &e) {
        set_err("phieng_surrogate3_mlp_backward", e.what());
        return 0;
    }
}

int32_t phieng_surrogate3_add_samples(Surrogate3 *s, const float *samples, uint32_t count) {
    if (!s || !samples) return 0;
    try {
        s->grow_samples((uint64_t)s->nsamples + count);
        PH_CK(hipMemcpyAsync(s->data.p + (size_t)s->nsamples * SUR_SAMPLE, samples, sizeof(float) * SUR_SAMPLE * count,
                             hipMemcpyHostToDevice, s->stream));
        PH_CK(hipStreamSynchronize(s->stream));
        s->nsamples += count;
        return 1;
    } catch (const std::exception &e) {
        set_err("phieng_surrogate3_add_samples", e.what());
        return 0;
    }
}

uint32_t phieng_surrogate3_num_samples(const Surrogate3 *s) { return s ? s->nsamples : 0; }

int32_t phieng_surrogate3_clear_samples(Surrogate3 *s) {
    if (!s) return 0;
    s->nsamples = 0;
    return 1;
}

int32_t phieng_surrogate3_get_samples(Surrogate3 *s, float *out, uint32_t first, uint32_t count) {
    if (!s || !out || first + count > s->nsamples) return 0;
    try {
        PH_CK(hipMemcpyAsync(out, s->data.p + (size_t)first * SUR_SAMPLE, sizeof(float) * SUR_SAMPLE * count,
                             hipMemcpyDeviceToHost, s->stream));
        PH_CK(hipStreamSynchronize(s->stream));
        return 1;
    } catch (const std::exception &e) {
        set_err("phieng_surrogate3_get_samples", e.what());
        return 0;
    }
}

// Train with Adam for `epochs` passes in shuffled batches; returns the mean
// sample loss of the first and of the last epoch (and of the network-free
// prediction a0 in *out_loss_base).
int32_t phieng_surrogate3_train(Surrogate3 *s, uint32_t epochs, uint32_t batch, float lr, uint32_t seed,
                                float *out_loss_first, float *out_loss_last, float *out_loss_base) {
    if (!s || s->nsamples == 0 || batch == 0 || !(lr > 0.0f)) return 0;
    try {
        std::mt19937 rng(seed);
        for (uint32_t e = 0; e < epochs; ++e) {
            double l = 0, l0 = 0;
            s->run_epoch(batch, lr, rng, true, l, l0);
            if (e == 0 && out_loss_first) *out_loss_first = (float)(l / s->nsamples);
            if (e + 1 == epochs && out_loss_last) *out_loss_last = (float)(l / s->nsamples);
            if (out_loss_base) *out_loss_base = (float)(l0 / s->nsamples);
        }
        return 1;
    } catch (const std::exception &e) {
        set_err("phieng_surrogate3_train", e.what());
        return 0;
    }
}

// Mean sample loss of the current network and of a0 alone (no training).
int32_t phieng_surrogate3_eval(Surrogate3 *s, float *out_loss, float *out_loss_base) {
    if (!s || s->nsamples == 0) return 0;
    try {
        std::mt19937 rng(0);
        double l = 0, l0 = 0;
        s->run_epoch(4096, 0.0f, rng, false, l, l0);
        if (out_loss) *out_loss = (float)(l / s->nsamples);
        if (out_loss_base) *out_loss_base = (float)(l0 / s->nsamples);
        return 1;
    } catch (const std::exception &e) {
        set_err("phieng_surrogate3_eval", e.what());
        return 0;
    }
}

}  // extern "C"
