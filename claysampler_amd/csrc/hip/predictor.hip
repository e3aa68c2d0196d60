// Host side of the learned voice-tail predictor (include/clay_hip.h,
// clay_mlp_* and clay_hip_mlp_*): standardisation, initialisation, the
// training loop (shuffled minibatches, one forward/backward and one
// reduce/Adam launch per step) and inference, on the renderer's stream.
#include <hip/hip_runtime.h>

#include "../../../include/clay_hip.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <numeric>
#include <stdexcept>
#include <string>
#include <vector>

template<bool Backward>
__global__ void clayMlpPass(const float*, const float*, const float*, const int*, int, int, float, float*, float*);
__global__ void clayMlpReduce(const float*, int, float*, float*, float*, float*, float*, int, float, float, float,
                              float, int);

hipStream_t clayStreamOf(clay_hip* h);
int clayDeviceOf(clay_hip* h);
void claySetError(clay_hip* h, const char* message);

namespace {

constexpr int kIn = CLAY_MLP_INPUTS, kHid = CLAY_MLP_HIDDEN, kParams = CLAY_MLP_PARAMS;
constexpr int kTile = 32, kThreads = 256;

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void check(hipError_t e, const char* what)
{
    if (e == hipErrorOutOfMemory)
        throw Error(CLAY_E_MEMORY, std::string(what) + ": out of memory");
    if (e != hipSuccess)
        throw Error(CLAY_E_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

template<class F>
int guarded(clay_hip* h, F&& body)
{
    try {
        body();
        return CLAY_OK;
    } catch (const Error& e) {
        if (h)
            claySetError(h, e.what());
        return e.code;
    } catch (const std::exception& e) {
        if (h)
            claySetError(h, e.what());
        return CLAY_E_INVALID;
    }
}

// SplitMix64 (Steele, Lea, Flood 2014) for initialisation and shuffles.
struct Rng {
    uint64_t state;
    uint64_t next()
    {
        uint64_t z = (state += 0x9e3779b97f4a7c15ull);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        return z ^ (z >> 31);
    }
    double uniform() { return double(next() >> 11) * (1.0 / 9007199254740992.0); }
    double normal()  // Box-Muller
    {
        const double u = std::max(uniform(), 1e-300), v = uniform();
        return std::sqrt(-2 * std::log(u)) * std::cos(6.283185307179586 * v);
    }
};

struct Buffers {
    float *x = nullptr, *y = nullptr, *params = nullptr, *m = nullptr, *v = nullptr, *partial = nullptr,
          *grad = nullptr, *loss = nullptr, *out = nullptr;
    int* index = nullptr;
    ~Buffers()
    {
        for (void* p : {(void*)x, (void*)y, (void*)params, (void*)m, (void*)v, (void*)partial, (void*)grad,
                        (void*)loss, (void*)out, (void*)index})
            if (p)
                (void)hipFree(p);
    }
};

std::vector<float> standardX(const clay_tail_model& m, const float* x, size_t n)
{
    std::vector<float> out(n * kIn);
    for (size_t s = 0; s < n; ++s)
        for (int i = 0; i < kIn; ++i)
            out[s * kIn + size_t(i)] = (x[s * kIn + size_t(i)] - m.input_mean[i]) * m.input_scale[i];
    return out;
}

std::vector<float> standardY(const clay_tail_model& m, const float* y, size_t n)
{
    std::vector<float> out(n);
    for (size_t s = 0; s < n; ++s)
        out[s] = (y[s] - m.output_mean) * m.output_scale;
    return out;
}

template<class T>
T* upload(const std::vector<T>& v, hipStream_t s)
{
    T* d = nullptr;
    check(hipMalloc(&d, std::max<size_t>(1, v.size()) * sizeof(T)), "predictor buffers");
    if (!v.empty())
        check(hipMemcpyAsync(d, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s), "predictor upload");
    return d;
}

int blocksFor(size_t n) { return int((n + kTile - 1) / kTile); }

// ---------------------------------------------------------------- no-alloc gradient
// clay_hip_mlp_gradient_noalloc keeps every device buffer of a gradient
// evaluation in one file-scope array: the standardised samples, the
// parameters, the per-block partial gradients, the summed gradient and the
// loss. Each device holds its own copy for the life of the process; host code
// uses the per-device address from hipGetSymbolAddress, never the array's name.
constexpr size_t kNoAllocSamples = 65536;
constexpr size_t kNoAllocBlocks = (kNoAllocSamples + kTile - 1) / kTile;
constexpr size_t kNoAllocX = 0;
constexpr size_t kNoAllocY = kNoAllocX + kNoAllocSamples * kIn;
constexpr size_t kNoAllocParams = kNoAllocY + kNoAllocSamples;
constexpr size_t kNoAllocGrad = kNoAllocParams + kParams;
constexpr size_t kNoAllocLoss = kNoAllocGrad + kParams;
constexpr size_t kNoAllocPartial = (kNoAllocLoss + 1 + 63) / 64 * 64;
constexpr size_t kNoAllocFloats = kNoAllocPartial + kNoAllocBlocks * (kParams + 1);

} // namespace

// Input and target standardisation of a model, by value to the kernel.
struct ClayMlpStandardisation {
    float inputMean[CLAY_MLP_INPUTS], inputScale[CLAY_MLP_INPUTS];
    float outputMean, outputScale;
};

static __device__ __align__(256) float scratch[kNoAllocFloats];

// In place: x' = (x - mean) * scale per input, y' = (y - mean) * scale; the
// same float operations as the host standardisation of clay_hip_mlp_gradient.
__global__ void clayMlpStandardize(float* __restrict__ x, float* __restrict__ y, int n, ClayMlpStandardisation st)
{
    const int s = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (s >= n)
        return;
    for (int i = 0; i < kIn; ++i)
        x[size_t(s) * kIn + size_t(i)] = (x[size_t(s) * kIn + size_t(i)] - st.inputMean[i]) * st.inputScale[i];
    y[s] = (y[s] - st.outputMean) * st.outputScale;
}

namespace {

// Address of `scratch` on `device`, looked up once per device with that
// device current (the library's calls run on the renderer's device).
float* scratchOn(int device)
{
    static std::mutex lock;
    static float* address[64] = {};
    if (device < 0 || device >= 64)
        throw Error(CLAY_E_DEVICE, "device index beyond the no-alloc scratch table");
    std::lock_guard<std::mutex> guard(lock);
    if (!address[device]) {
        void* p = nullptr;
        check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
        address[device] = static_cast<float*>(p);
    }
    return address[device];
}

} // namespace

extern "C" {

int clay_mlp_init(clay_tail_model* model, const float* x, const float* y, size_t n, uint64_t seed)
{
    if (!model || !x || !y || !n)
        return CLAY_E_INVALID;
    std::memset(model, 0, sizeof *model);
    for (int i = 0; i < kIn; ++i) {
        double sum = 0, sq = 0;
        for (size_t s = 0; s < n; ++s) {
            const double v = x[s * kIn + size_t(i)];
            sum += v;
            sq += v * v;
        }
        const double mean = sum / double(n), var = std::max(0.0, sq / double(n) - mean * mean);
        model->input_mean[i] = float(mean);
        model->input_scale[i] = var > 1e-12 ? float(1 / std::sqrt(var)) : 1.f;
    }
    double sum = 0, sq = 0;
    for (size_t s = 0; s < n; ++s) {
        sum += y[s];
        sq += double(y[s]) * y[s];
    }
    const double mean = sum / double(n), var = std::max(0.0, sq / double(n) - mean * mean);
    model->output_mean = float(mean);
    model->output_scale = var > 1e-12 ? float(1 / std::sqrt(var)) : 1.f;
    // He-normal weights for the ReLU layers, 1/sqrt(fan_in) for the output; zero biases.
    Rng rng{seed};
    float* p = model->params;
    for (int i = 0; i < kIn * kHid; ++i)
        p[i] = float(rng.normal() * std::sqrt(2.0 / kIn));
    p += kIn * kHid + kHid;
    for (int i = 0; i < kHid * kHid; ++i)
        p[i] = float(rng.normal() * std::sqrt(2.0 / kHid));
    p += kHid * kHid + kHid;
    for (int i = 0; i < kHid; ++i)
        p[i] = float(rng.normal() * std::sqrt(1.0 / kHid));
    return CLAY_OK;
}

int clay_hip_mlp_gradient(clay_hip* h, const clay_tail_model* model, const float* x, const float* y, size_t n,
                          float* gradient, float* loss)
{
    if (!h || !model || !x || !y || !n || !gradient || n > size_t(INT32_MAX))
        return CLAY_E_INVALID;
    return guarded(h, [&] {
        check(hipSetDevice(clayDeviceOf(h)), "hipSetDevice");
        hipStream_t s = clayStreamOf(h);
        Buffers b;
        b.x = upload(standardX(*model, x, n), s);
        b.y = upload(standardY(*model, y, n), s);
        b.params = upload(std::vector<float>(model->params, model->params + kParams), s);
        const int blocks = blocksFor(n);
        check(hipMalloc(&b.partial, size_t(blocks) * (kParams + 1) * sizeof(float)), "predictor buffers");
        check(hipMalloc(&b.grad, kParams * sizeof(float)), "predictor buffers");
        check(hipMalloc(&b.loss, sizeof(float)), "predictor buffers");
        hipLaunchKernelGGL(clayMlpPass<true>, dim3(blocks), dim3(kThreads), 0, s, b.params, b.x, b.y, nullptr,
                           int(n), 0, float(1.0 / double(n)), b.partial, nullptr);
        check(hipGetLastError(), "clayMlpPass");
        hipLaunchKernelGGL(clayMlpReduce, dim3((kParams + 1 + kThreads - 1) / kThreads), dim3(kThreads), 0, s,
                           b.partial, blocks, b.params, nullptr, nullptr, b.grad, b.loss, 0, 0.f, 0.f, 0.f, 0.f, 0);
        check(hipGetLastError(), "clayMlpReduce");
        check(hipMemcpyAsync(gradient, b.grad, kParams * sizeof(float), hipMemcpyDeviceToHost, s), "gradient");
        float l = 0;
        check(hipMemcpyAsync(&l, b.loss, sizeof(float), hipMemcpyDeviceToHost, s), "loss");
        check(hipStreamSynchronize(s), "gradient");
        if (loss)
            *loss = l;
    });
}

int clay_hip_mlp_gradient_noalloc(clay_hip* h, const clay_tail_model* model, const float* x, const float* y,
                                  size_t n, float* gradient, float* loss)
{
    if (!h || !model || !x || !y || !n || !gradient)
        return CLAY_E_INVALID;
    if (n > kNoAllocSamples) {
        claySetError(h, "clay_hip_mlp_gradient_noalloc: more samples than its scratch holds (65536)");
        return CLAY_E_INVALID;
    }
    return guarded(h, [&] {
        const int device = clayDeviceOf(h);
        check(hipSetDevice(device), "hipSetDevice");
        hipStream_t s = clayStreamOf(h);
        float* base = scratchOn(device);
        float *dx = base + kNoAllocX, *dy = base + kNoAllocY, *params = base + kNoAllocParams;
        float *grad = base + kNoAllocGrad, *dloss = base + kNoAllocLoss, *partial = base + kNoAllocPartial;
        check(hipMemcpyAsync(dx, x, n * kIn * sizeof(float), hipMemcpyHostToDevice, s), "samples");
        check(hipMemcpyAsync(dy, y, n * sizeof(float), hipMemcpyHostToDevice, s), "targets");
        check(hipMemcpyAsync(params, model->params, kParams * sizeof(float), hipMemcpyHostToDevice, s), "parameters");
        ClayMlpStandardisation st;
        std::memcpy(st.inputMean, model->input_mean, sizeof st.inputMean);
        std::memcpy(st.inputScale, model->input_scale, sizeof st.inputScale);
        st.outputMean = model->output_mean;
        st.outputScale = model->output_scale;
        hipLaunchKernelGGL(clayMlpStandardize, dim3((n + kThreads - 1) / kThreads), dim3(kThreads), 0, s, dx, dy,
                           int(n), st);
        check(hipGetLastError(), "clayMlpStandardize");
        // Every partial entry of every block is written by the pass, and the
        // reduce writes the whole gradient and the loss: nothing to clear.
        const int blocks = blocksFor(n);
        hipLaunchKernelGGL(clayMlpPass<true>, dim3(blocks), dim3(kThreads), 0, s, params, dx, dy, nullptr, int(n), 0,
                           float(1.0 / double(n)), partial, nullptr);
        check(hipGetLastError(), "clayMlpPass");
        hipLaunchKernelGGL(clayMlpReduce, dim3((kParams + 1 + kThreads - 1) / kThreads), dim3(kThreads), 0, s,
                           partial, blocks, params, nullptr, nullptr, grad, dloss, 0, 0.f, 0.f, 0.f, 0.f, 0);
        check(hipGetLastError(), "clayMlpReduce");
        float l = 0;
        check(hipMemcpyAsync(gradient, grad, kParams * sizeof(float), hipMemcpyDeviceToHost, s), "gradient");
        check(hipMemcpyAsync(&l, dloss, sizeof(float), hipMemcpyDeviceToHost, s), "loss");
        check(hipStreamSynchronize(s), "gradient");
        if (loss)
            *loss = l;
    });
}

int clay_hip_mlp_train(clay_hip* h, clay_tail_model* model, const float* x, const float* y, size_t n,
                       const clay_mlp_train_config* cfg, float* epoch_loss)
{
    if (!h || !model || !x || !y || !n || !cfg || cfg->epochs < 1 || cfg->batch < kTile || cfg->batch % kTile ||
        n > size_t(INT32_MAX))
        return CLAY_E_INVALID;
    return guarded(h, [&] {
        check(hipSetDevice(clayDeviceOf(h)), "hipSetDevice");
        hipStream_t s = clayStreamOf(h);
        Buffers b;
        b.x = upload(standardX(*model, x, n), s);
        b.y = upload(standardY(*model, y, n), s);
        b.params = upload(std::vector<float>(model->params, model->params + kParams), s);
        b.m = upload(std::vector<float>(kParams, 0.f), s);
        b.v = upload(std::vector<float>(kParams, 0.f), s);
        const size_t batch = size_t(cfg->batch);
        const int maxBlocks = blocksFor(batch);
        const size_t steps = (n + batch - 1) / batch;
        check(hipMalloc(&b.partial, size_t(maxBlocks) * (kParams + 1) * sizeof(float)), "predictor buffers");
        check(hipMalloc(&b.loss, steps * sizeof(float)), "predictor buffers");
        check(hipMalloc(&b.index, n * sizeof(int)), "predictor buffers");
        std::vector<int> order(n);
        std::vector<float> losses(steps);
        Rng rng{cfg->seed};
        int step = 0;
        for (int epoch = 0; epoch < cfg->epochs; ++epoch) {
            std::iota(order.begin(), order.end(), 0);
            for (size_t i = n - 1; i > 0; --i)  // Fisher-Yates
                std::swap(order[i], order[size_t(rng.next() % (i + 1))]);
            check(hipMemcpyAsync(b.index, order.data(), n * sizeof(int), hipMemcpyHostToDevice, s), "shuffle");
            for (size_t k = 0; k < steps; ++k) {
                const size_t first = k * batch, count = std::min(batch, n - first);
                const int blocks = blocksFor(count);
                hipLaunchKernelGGL(clayMlpPass<true>, dim3(blocks), dim3(kThreads), 0, s, b.params, b.x, b.y, b.index,
                                   int(first + count), int(first), float(1.0 / double(count)), b.partial, nullptr);
                check(hipGetLastError(), "clayMlpPass");
                ++step;
                hipLaunchKernelGGL(clayMlpReduce, dim3((kParams + 1 + kThreads - 1) / kThreads), dim3(kThreads), 0, s,
                                   b.partial, blocks, b.params, b.m, b.v, nullptr, b.loss + k, 1, cfg->lr, cfg->beta1,
                                   cfg->beta2, cfg->eps, step);
                check(hipGetLastError(), "clayMlpReduce");
            }
            check(hipMemcpyAsync(losses.data(), b.loss, steps * sizeof(float), hipMemcpyDeviceToHost, s), "losses");
            check(hipStreamSynchronize(s), "training epoch");  // the host order buffer is reused next epoch
            if (epoch_loss) {
                double sum = 0;
                for (size_t k = 0; k < steps; ++k)
                    sum += double(losses[k]) * double(std::min(batch, n - k * batch));
                epoch_loss[epoch] = float(sum / double(n));
            }
        }
        check(hipMemcpy(model->params, b.params, kParams * sizeof(float), hipMemcpyDeviceToHost), "parameters");
    });
}

int clay_hip_mlp_predict(clay_hip* h, const clay_tail_model* model, const float* x, size_t n, float* y)
{
    if (!h || !model || !x || !y || !n || n > size_t(INT32_MAX))
        return CLAY_E_INVALID;
    return guarded(h, [&] {
        check(hipSetDevice(clayDeviceOf(h)), "hipSetDevice");
        hipStream_t s = clayStreamOf(h);
        Buffers b;
        b.x = upload(standardX(*model, x, n), s);
        b.params = upload(std::vector<float>(model->params, model->params + kParams), s);
        check(hipMalloc(&b.out, n * sizeof(float)), "predictor buffers");
        hipLaunchKernelGGL(clayMlpPass<false>, dim3(blocksFor(n)), dim3(kThreads), 0, s, b.params, b.x, nullptr,
                           nullptr, int(n), 0, 1.f, nullptr, b.out);
        check(hipGetLastError(), "clayMlpPass");
        check(hipMemcpyAsync(y, b.out, n * sizeof(float), hipMemcpyDeviceToHost, s), "predictions");
        check(hipStreamSynchronize(s), "predictions");
        for (size_t i = 0; i < n; ++i)
            y[i] = y[i] / model->output_scale + model->output_mean;
    });
}

} // extern "C"
