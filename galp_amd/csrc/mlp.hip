// mlp.hip -- the learned dispatcher's network: a 4 -> H -> H -> 1 MLP
// (ReLU, ReLU, linear) in fp32, with hand-written inference, gradient and
// Adam kernels.  No reference counterpart: the reference dispatches by a
// fixed policy (-l, lfht.cuh:127-165).
//
// Inputs (all log2-scaled): degree, mean neighbour degree, max neighbour
// degree, distinct labels seen in the previous iteration (-1 if none).
// Output: log2(1 + distinct labels this iteration).  The prediction only
// chooses a kernel and sizes tables; labels never depend on it.
//
// Parameter layout (row-major, float): W1[H][4], b1[H], W2[H][H], b2[H],
// W3[H], b3.  Loss: mean over the batch of 0.5 * (y_hat - y)^2.
// Gradients are reduced per block from LDS-staged activations (no float
// atomics), then summed over blocks in a fixed order: training is
// deterministic for a given seed.

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

#include "device_common.hpp"
#include "internal.hpp"
#include "mlp.hpp"

namespace galp {

template <int H>
__global__ void __launch_bounds__(256)
k_mlp_predict(const float *__restrict__ params, const float *__restrict__ x, int64_t n, float *__restrict__ y) {
    __shared__ float sp[mlp_num_params(H)];
    for (int i = threadIdx.x; i < mlp_num_params(H); i += 256) sp[i] = params[i];
    __syncthreads();
    for (int64_t s = (int64_t)blockIdx.x * 256 + threadIdx.x; s < n; s += (int64_t)gridDim.x * 256)
        y[s] = mlp_forward<H>(sp, x[4 * s], x[4 * s + 1], x[4 * s + 2], x[4 * s + 3]);
}

// Per-block gradient of the batch loss: block b handles samples
// [b*S, b*S + S) of the batch, writes its partial gradient (P floats) and
// partial loss.  `scale` = 1 / batch size.
template <int H, int S>
__global__ void __launch_bounds__(S)
k_mlp_grad(const float *__restrict__ params, const float *__restrict__ x, const float *__restrict__ y,
           int64_t n, float scale, float *__restrict__ pgrad, float *__restrict__ ploss) {
    constexpr int P = mlp_num_params(H);
    __shared__ float sp[P];
    __shared__ float sx[S][4];
    __shared__ float sh1[S][H + 1];   // +1: spread LDS banks
    __shared__ float sh2[S][H + 1];
    __shared__ float sd1[S][H + 1];
    __shared__ float sd2[S][H + 1];
    __shared__ float sg[S];
    __shared__ float sl[S];
    for (int i = threadIdx.x; i < P; i += S) sp[i] = params[i];
    __syncthreads();
    const float *W1 = sp, *b1 = sp + 4 * H, *W2 = b1 + H, *b2 = W2 + H * H, *W3 = b2 + H, *b3 = W3 + H;
    const int t = threadIdx.x;
    const int64_t s = (int64_t)blockIdx.x * S + t;
    const bool ok = s < n;
    float xs[4] = {0.f, 0.f, 0.f, 0.f};
    if (ok)
        for (int k = 0; k < 4; ++k) xs[k] = x[4 * s + k];
    float h1[H], z2[H];
#pragma unroll
    for (int j = 0; j < H; ++j) {
        float a = b1[j] + W1[4 * j] * xs[0] + W1[4 * j + 1] * xs[1] + W1[4 * j + 2] * xs[2] + W1[4 * j + 3] * xs[3];
        h1[j] = a > 0.f ? a : 0.f;
    }
    float out = b3[0];
    for (int j = 0; j < H; ++j) {
        float a = b2[j];
#pragma unroll
        for (int k = 0; k < H; ++k) a += W2[H * j + k] * h1[k];
        z2[j] = a;
        out += W3[j] * (a > 0.f ? a : 0.f);
    }
    const float r = ok ? out - y[s] : 0.f;
    const float g = r * scale;
    sl[t] = 0.5f * r * r * scale;
    sg[t] = g;
    for (int k = 0; k < 4; ++k) sx[t][k] = xs[k];
    for (int j = 0; j < H; ++j) {
        sh1[t][j] = h1[j];
        sh2[t][j] = z2[j] > 0.f ? z2[j] : 0.f;
        sd2[t][j] = z2[j] > 0.f ? g * W3[j] : 0.f;
    }
    for (int k = 0; k < H; ++k) {
        float a = 0.f;
        for (int j = 0; j < H; ++j) a += W2[H * j + k] * sd2[t][j];
        sd1[t][k] = h1[k] > 0.f ? a : 0.f;
    }
    __syncthreads();
    // every parameter's gradient is a sum over the block's samples
    float *out_g = pgrad + (int64_t)blockIdx.x * P;
    for (int q = t; q < P; q += S) {
        float acc = 0.f;
        if (q < 4 * H) {
            const int j = q / 4, k = q % 4;
            for (int i = 0; i < S; ++i) acc += sd1[i][j] * sx[i][k];
        } else if (q < 5 * H) {
            const int j = q - 4 * H;
            for (int i = 0; i < S; ++i) acc += sd1[i][j];
        } else if (q < 5 * H + H * H) {
            const int j = (q - 5 * H) / H, k = (q - 5 * H) % H;
            for (int i = 0; i < S; ++i) acc += sd2[i][j] * sh1[i][k];
        } else if (q < 6 * H + H * H) {
            const int j = q - 5 * H - H * H;
            for (int i = 0; i < S; ++i) acc += sd2[i][j];
        } else if (q < 7 * H + H * H) {
            const int k = q - 6 * H - H * H;
            for (int i = 0; i < S; ++i) acc += sg[i] * sh2[i][k];
        } else {
            for (int i = 0; i < S; ++i) acc += sg[i];
        }
        out_g[q] = acc;
    }
    if (t == 0) {
        float a = 0.f;
        for (int i = 0; i < S; ++i) a += sl[i];
        ploss[blockIdx.x] = a;
    }
}

// Sums the per-block partial gradients in block order (deterministic) into
// grad, and the partial losses into *loss.
__global__ void k_mlp_reduce(const float *__restrict__ pgrad, const float *__restrict__ ploss, int nblocks, int P,
                             float *__restrict__ grad, float *__restrict__ loss) {
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < P; q += gridDim.x * blockDim.x) {
        float a = 0.f;
        for (int b = 0; b < nblocks; ++b) a += pgrad[(int64_t)b * P + q];
        grad[q] = a;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        float a = 0.f;
        for (int b = 0; b < nblocks; ++b) a += ploss[b];
        *loss = a;
    }
}

// A whole training run in one launch of one workgroup: `epochs` passes over
// the (already shuffled) samples in minibatches of B, each minibatch's
// gradient of the mean loss, then one Adam step (Kingma & Ba 2015).  The
// batch is walked in chunks of kTrainChunk samples: one thread per sample
// runs the forward and backward pass and stages its activations and deltas
// in LDS, then every thread adds the chunk's share to the gradient of the
// parameters it owns (kept in registers across the chunks, summed in sample
// order, so a run is deterministic).  The owners then apply Adam to their
// parameters, with first and second moments held in registers for the
// whole run, and publish them to LDS for the next minibatch.  One launch per
// run replaces three per minibatch (gradient, block reduction, Adam).
constexpr int kTrainThreads = 512;
constexpr int kTrainChunk = 256;
template <int H>
__global__ void __launch_bounds__(kTrainThreads)
k_mlp_train(float *__restrict__ params, const float *__restrict__ x, const float *__restrict__ y, int64_t n,
            int epochs, int B, float lr, float b1, float b2, float eps) {
    constexpr int P = mlp_num_params(H);
    constexpr int S = kTrainChunk;
    constexpr int Q = (P + kTrainThreads - 1) / kTrainThreads;   // parameters per thread
    __shared__ float sp[P];
    __shared__ float sx[S][4];
    __shared__ float sh1[S][H + 1];   // +1: spread LDS banks
    __shared__ float sh2[S][H + 1];
    __shared__ float sd1[S][H + 1];
    __shared__ float sd2[S][H + 1];
    __shared__ float sg[S];
    const int t = threadIdx.x;
    float pm[Q], pv[Q], pw[Q];
#pragma unroll
    for (int r = 0; r < Q; ++r) {
        const int q = t + r * kTrainThreads;
        pm[r] = 0.f;
        pv[r] = 0.f;
        pw[r] = q < P ? params[q] : 0.f;
        if (q < P) sp[q] = pw[r];
    }
    __syncthreads();
    const float *W1 = sp, *bb1 = sp + 4 * H, *W2 = bb1 + H, *bb2 = W2 + H * H, *W3 = bb2 + H, *bb3 = W3 + H;
    const float scale = 1.0f / (float)B;
    float pow1 = 1.f, pow2 = 1.f;   // b1^step, b2^step
    for (int e = 0; e < epochs; ++e) {
        for (int64_t off = 0; off + B <= n; off += B) {
            float acc[Q];
#pragma unroll
            for (int r = 0; r < Q; ++r) acc[r] = 0.f;
            for (int c0 = 0; c0 < B; c0 += S) {
                const int cs = min(S, B - c0);
                if (t < cs) {   // forward and backward of sample off + c0 + t
                    const int64_t s = off + c0 + t;
                    float xs[4];
#pragma unroll
                    for (int k = 0; k < 4; ++k) xs[k] = x[4 * s + k];
                    float h1[H], z2[H];
#pragma unroll
                    for (int j = 0; j < H; ++j) {
                        const float a = bb1[j] + W1[4 * j] * xs[0] + W1[4 * j + 1] * xs[1] + W1[4 * j + 2] * xs[2] +
                                        W1[4 * j + 3] * xs[3];
                        h1[j] = a > 0.f ? a : 0.f;
                    }
                    float out = bb3[0];
#pragma unroll 2
                    for (int j = 0; j < H; ++j) {
                        float a = bb2[j];
#pragma unroll
                        for (int k = 0; k < H; ++k) a += W2[H * j + k] * h1[k];
                        z2[j] = a;
                        out += W3[j] * (a > 0.f ? a : 0.f);
                    }
                    const float g = (out - y[s]) * scale;
                    sg[t] = g;
#pragma unroll
                    for (int k = 0; k < 4; ++k) sx[t][k] = xs[k];
                    for (int j = 0; j < H; ++j) {
                        sh1[t][j] = h1[j];
                        sh2[t][j] = z2[j] > 0.f ? z2[j] : 0.f;
                        sd2[t][j] = z2[j] > 0.f ? g * W3[j] : 0.f;
                    }
#pragma unroll 2
                    for (int k = 0; k < H; ++k) {
                        float a = 0.f;
#pragma unroll 8
                        for (int j = 0; j < H; ++j) a += W2[H * j + k] * sd2[t][j];
                        sd1[t][k] = h1[k] > 0.f ? a : 0.f;
                    }
                }
                __syncthreads();
#pragma unroll
                for (int r = 0; r < Q; ++r) {
                    const int q = t + r * kTrainThreads;
                    float a = acc[r];
                    if (q < 4 * H) {
                        const int j = q / 4, k = q % 4;
                        _Pragma("unroll 4") for (int i = 0; i < cs; ++i) a += sd1[i][j] * sx[i][k];
                    } else if (q < 5 * H) {
                        const int j = q - 4 * H;
                        _Pragma("unroll 4") for (int i = 0; i < cs; ++i) a += sd1[i][j];
                    } else if (q < 5 * H + H * H) {
                        const int j = (q - 5 * H) / H, k = (q - 5 * H) % H;
                        _Pragma("unroll 4") for (int i = 0; i < cs; ++i) a += sd2[i][j] * sh1[i][k];
                    } else if (q < 6 * H + H * H) {
                        const int j = q - 5 * H - H * H;
                        _Pragma("unroll 4") for (int i = 0; i < cs; ++i) a += sd2[i][j];
                    } else if (q < 7 * H + H * H) {
                        const int k = q - 6 * H - H * H;
                        _Pragma("unroll 4") for (int i = 0; i < cs; ++i) a += sg[i] * sh2[i][k];
                    } else if (q < P) {
                        _Pragma("unroll 4") for (int i = 0; i < cs; ++i) a += sg[i];
                    }
                    acc[r] = a;
                }
                __syncthreads();   // the chunk's LDS rows are rewritten next
            }
            pow1 *= b1;
            pow2 *= b2;
            const float c1 = 1.f - pow1, c2 = 1.f - pow2;
#pragma unroll
            for (int r = 0; r < Q; ++r) {
                const int q = t + r * kTrainThreads;
                const float g = acc[r];
                pm[r] = b1 * pm[r] + (1.f - b1) * g;
                pv[r] = b2 * pv[r] + (1.f - b2) * g * g;
                pw[r] -= lr * (pm[r] / c1) / (sqrtf(pv[r] / c2) + eps);
                if (q < P) sp[q] = pw[r];
            }
            __syncthreads();
        }
    }
#pragma unroll
    for (int r = 0; r < Q; ++r) {
        const int q = t + r * kTrainThreads;
        if (q < P) params[q] = pw[r];
    }
}

template <int H>
static int mlp_grad_launch(hipStream_t st, const float *dp, const float *dx, const float *dy, int64_t n,
                           float *dpgrad, float *dploss, float *dgrad, float *dloss) {
    constexpr int S = 64;
    const int nb = (int)((n + S - 1) / S);
    k_mlp_grad<H, S><<<nb, S, 0, st>>>(dp, dx, dy, n, 1.0f / (float)n, dpgrad, dploss);
    k_mlp_reduce<<<(mlp_num_params(H) + 255) / 256, 256, 0, st>>>(dpgrad, dploss, nb, mlp_num_params(H), dgrad,
                                                                  dloss);
    return hipGetLastError() == hipSuccess ? GALP_OK : GALP_EHIP;
}

int mlp_predict_dev(hipStream_t st, int H, const float *dparams, const float *dx, int64_t n, float *dy) {
    const int grid = (int)std::min<int64_t>((n + 255) / 256, 4096);
    if (n <= 0) return GALP_OK;
    switch (H) {
    case 8: k_mlp_predict<8><<<grid, 256, 0, st>>>(dparams, dx, n, dy); break;
    case 16: k_mlp_predict<16><<<grid, 256, 0, st>>>(dparams, dx, n, dy); break;
    case 32: k_mlp_predict<32><<<grid, 256, 0, st>>>(dparams, dx, n, dy); break;
    default: return GALP_EINVAL;
    }
    return hipGetLastError() == hipSuccess ? GALP_OK : GALP_EHIP;
}

static int mlp_train_dev(hipStream_t st, int H, float *dp, const float *dx, const float *dy, int64_t n, int epochs,
                         int B, float lr) {
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    switch (H) {
    case 8: k_mlp_train<8><<<1, kTrainThreads, 0, st>>>(dp, dx, dy, n, epochs, B, lr, b1, b2, eps); break;
    case 16: k_mlp_train<16><<<1, kTrainThreads, 0, st>>>(dp, dx, dy, n, epochs, B, lr, b1, b2, eps); break;
    case 32: k_mlp_train<32><<<1, kTrainThreads, 0, st>>>(dp, dx, dy, n, epochs, B, lr, b1, b2, eps); break;
    default: return GALP_EINVAL;
    }
    return hipGetLastError() == hipSuccess ? GALP_OK : GALP_EHIP;
}

static int mlp_grad_dev(hipStream_t st, int H, const float *dp, const float *dx, const float *dy, int64_t n,
                        float *dpgrad, float *dploss, float *dgrad, float *dloss) {
    switch (H) {
    case 8: return mlp_grad_launch<8>(st, dp, dx, dy, n, dpgrad, dploss, dgrad, dloss);
    case 16: return mlp_grad_launch<16>(st, dp, dx, dy, n, dpgrad, dploss, dgrad, dloss);
    case 32: return mlp_grad_launch<32>(st, dp, dx, dy, n, dpgrad, dploss, dgrad, dloss);
    default: return GALP_EINVAL;
    }
}

// RAII device buffer for the host-driven MLP entry points.
struct DevBuf {
    void *p = nullptr;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
    hipError_t alloc(size_t bytes) { return hipMalloc(&p, bytes ? bytes : 4); }
    template <class T>
    T *as() const {
        return (T *)p;
    }
};

// ---------------------------------------------------------------------------
// Allocation-free loss and gradient (galp_mlp_loss_grad_noalloc): the
// per-block partial gradients and losses of k_mlp_grad live in this
// file-scope device array instead of memory allocated per call.  One copy per
// device, held until the process ends.  Host code never passes the array's
// name to a kernel (on the host it is a stand-in, not a device address):
// mlp_scratch_setup() looks up each device's address once, with that device
// current, and the entry point passes that address.
constexpr int64_t kNoAllocMaxSamples = 32768;                    // per call
constexpr int64_t kNoAllocMaxBlocks = kNoAllocMaxSamples / 64;   // k_mlp_grad blocks of 64 samples
constexpr int64_t kNoAllocScratchFloats = kNoAllocMaxBlocks * mlp_num_params(32) + kNoAllocMaxBlocks;
static __device__ __align__(256) float scratch[kNoAllocScratchFloats];   // 2.56 MB

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static float *g_scratch_ptr[kMaxDevices] = {};

// Resolves the device address of `scratch` on `device` (which must be
// current), once per device.  Called by galp_open.
int mlp_scratch_setup(galp_ctx *ctx, int device) {
    if (device < 0 || device >= kMaxDevices) return fail(ctx, GALP_EINVAL, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[device]) return GALP_OK;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return fail(ctx, GALP_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[device] = (float *)p;
    return GALP_OK;
}

static float *scratch_of(int device) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    return (device >= 0 && device < kMaxDevices) ? g_scratch_ptr[device] : nullptr;
}

}  // namespace galp

using namespace galp;

#define GALP_CK(call)                                                                           \
    do {                                                                                        \
        hipError_t _e = (call);                                                                 \
        if (_e != hipSuccess) {                                                                 \
            if (st) {   /* drain queued work before the buffers are freed */                       \
                (void)hipStreamSynchronize(st);                                                 \
                (void)hipStreamDestroy(st);                                                     \
            }                                                                                   \
            return fail(ctx, GALP_EHIP, std::string(#call) + ": " + hipGetErrorString(_e));     \
        }                                                                                       \
    } while (0)

extern "C" {

GALP_API int64_t galp_mlp_num_params(int32_t hidden) { return mlp_num_params(hidden); }

GALP_API int galp_mlp_init(int32_t hidden, uint64_t seed, float *params) {
    if (!params || (hidden != 8 && hidden != 16 && hidden != 32)) return GALP_EINVAL;
    // He initialisation from a seeded SplitMix64 stream (Box-Muller normals)
    uint64_t s = seed * 0x9e3779b97f4a7c15ULL + 0x2545F4914F6CDD1DULL;
    auto next = [&]() {
        s += 0x9e3779b97f4a7c15ULL;
        uint64_t z = s;
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
        return z ^ (z >> 31);
    };
    auto normal = [&]() {
        double u1 = ((next() >> 11) + 1.0) / 9007199254740993.0, u2 = (next() >> 11) / 9007199254740992.0;
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    };
    const int H = hidden;
    float *W1 = params, *b1 = W1 + 4 * H, *W2 = b1 + H, *b2 = W2 + H * H, *W3 = b2 + H, *b3 = W3 + H;
    for (int i = 0; i < 4 * H; ++i) W1[i] = (float)(normal() * std::sqrt(2.0 / 4.0));
    for (int i = 0; i < H * H; ++i) W2[i] = (float)(normal() * std::sqrt(2.0 / H));
    for (int i = 0; i < H; ++i) W3[i] = (float)(normal() * std::sqrt(1.0 / H));
    for (int i = 0; i < H; ++i) b1[i] = b2[i] = 0.01f;
    b3[0] = 0.f;
    return GALP_OK;
}

GALP_API int galp_mlp_predict(galp_ctx *ctx, int32_t hidden, const float *params, const float *x, int64_t n,
                              float *yhat) {
    if (!ctx || !params || (n > 0 && (!x || !yhat)) || n < 0) return GALP_EINVAL;
    if (hidden != 8 && hidden != 16 && hidden != 32) return fail(ctx, GALP_EINVAL, "hidden must be 8, 16 or 32");
    hipStream_t st = nullptr;
    GALP_CK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    DevBuf dp, dx, dy;
    const int64_t P = mlp_num_params(hidden);
    GALP_CK(dp.alloc(4 * P));
    GALP_CK(dx.alloc(16 * (size_t)n));
    GALP_CK(dy.alloc(4 * (size_t)n));
    GALP_CK(hipMemcpyAsync(dp.p, params, 4 * P, hipMemcpyHostToDevice, st));
    if (n) GALP_CK(hipMemcpyAsync(dx.p, x, 16 * (size_t)n, hipMemcpyHostToDevice, st));
    if (mlp_predict_dev(st, hidden, dp.as<float>(), dx.as<float>(), n, dy.as<float>()) != GALP_OK)
        GALP_CK(hipGetLastError());
    if (n) GALP_CK(hipMemcpyAsync(yhat, dy.p, 4 * (size_t)n, hipMemcpyDeviceToHost, st));
    GALP_CK(hipStreamSynchronize(st));
    (void)hipStreamDestroy(st);
    return GALP_OK;
}

GALP_API int galp_mlp_loss_grad(galp_ctx *ctx, int32_t hidden, const float *params, const float *x, const float *y,
                                int64_t n, float *grad, float *loss) {
    if (!ctx || !params || !x || !y || !grad || !loss || n <= 0) return GALP_EINVAL;
    if (hidden != 8 && hidden != 16 && hidden != 32) return fail(ctx, GALP_EINVAL, "hidden must be 8, 16 or 32");
    hipStream_t st = nullptr;
    GALP_CK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    const int64_t P = mlp_num_params(hidden);
    const int64_t nb = (n + 63) / 64;
    DevBuf dp, dx, dy, dpg, dpl, dg, dl;
    GALP_CK(dp.alloc(4 * P));
    GALP_CK(dx.alloc(16 * (size_t)n));
    GALP_CK(dy.alloc(4 * (size_t)n));
    GALP_CK(dpg.alloc(4 * (size_t)(nb * P)));
    GALP_CK(dpl.alloc(4 * (size_t)nb));
    GALP_CK(dg.alloc(4 * P));
    GALP_CK(dl.alloc(4));
    GALP_CK(hipMemcpyAsync(dp.p, params, 4 * P, hipMemcpyHostToDevice, st));
    GALP_CK(hipMemcpyAsync(dx.p, x, 16 * (size_t)n, hipMemcpyHostToDevice, st));
    GALP_CK(hipMemcpyAsync(dy.p, y, 4 * (size_t)n, hipMemcpyHostToDevice, st));
    if (mlp_grad_dev(st, hidden, dp.as<float>(), dx.as<float>(), dy.as<float>(), n, dpg.as<float>(),
                     dpl.as<float>(), dg.as<float>(), dl.as<float>()) != GALP_OK)
        GALP_CK(hipGetLastError());
    GALP_CK(hipMemcpyAsync(grad, dg.p, 4 * P, hipMemcpyDeviceToHost, st));
    GALP_CK(hipMemcpyAsync(loss, dl.p, 4, hipMemcpyDeviceToHost, st));
    GALP_CK(hipStreamSynchronize(st));
    (void)hipStreamDestroy(st);
    return GALP_OK;
}

GALP_API int galp_mlp_loss_grad_noalloc(galp_ctx *ctx, int32_t hidden, const float *d_params, const float *d_x,
                                        const float *d_y, int64_t n, float *d_grad, float *d_loss, void *stream) {
    if (!ctx || !d_params || !d_x || !d_y || !d_grad || !d_loss || n <= 0) return GALP_EINVAL;
    if (hidden != 8 && hidden != 16 && hidden != 32) return fail(ctx, GALP_EINVAL, "hidden must be 8, 16 or 32");
    if (n > kNoAllocMaxSamples)
        return fail(ctx, GALP_EINVAL, "galp_mlp_loss_grad_noalloc: n = " + std::to_string(n) +
                                          " samples exceeds the scratch limit of " +
                                          std::to_string(kNoAllocMaxSamples) + " per call");
    int cur = -1;
    if (hipGetDevice(&cur) != hipSuccess || cur != ctx->device)
        return fail(ctx, GALP_EINVAL, "galp_mlp_loss_grad_noalloc: the current device is not the context's");
    float *sp = scratch_of(ctx->device);
    if (!sp) return fail(ctx, GALP_EHIP, "galp_mlp_loss_grad_noalloc: scratch not set up on this device");
    float *pgrad = sp;                                            // ceil(n / 64) * P floats
    float *ploss = sp + kNoAllocMaxBlocks * mlp_num_params(32);   // ceil(n / 64) floats
    // every slot the kernels read is written by them in this call: no clearing needed
    if (mlp_grad_dev((hipStream_t)stream, hidden, d_params, d_x, d_y, n, pgrad, ploss, d_grad, d_loss) != GALP_OK)
        return fail(ctx, GALP_EHIP, std::string("galp_mlp_loss_grad_noalloc: launch failed: ") +
                                        hipGetErrorString(hipGetLastError()));
    return GALP_OK;
}

GALP_API int galp_mlp_train(galp_ctx *ctx, int32_t hidden, float *params, const float *x, const float *y, int64_t n,
                            int32_t epochs, int32_t batch, float lr, uint64_t seed, float *final_loss) {
    if (!ctx || !params || !x || !y || n <= 0 || epochs < 0 || batch <= 0) return GALP_EINVAL;
    if (hidden != 8 && hidden != 16 && hidden != 32) return fail(ctx, GALP_EINVAL, "hidden must be 8, 16 or 32");
    hipStream_t st = nullptr;
    GALP_CK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    const int64_t P = mlp_num_params(hidden);
    // seeded shuffle of the samples once, then contiguous minibatches per epoch
    std::vector<int64_t> order((size_t)n);
    std::iota(order.begin(), order.end(), 0);
    uint64_t s = seed ^ 0x5DEECE66DULL;
    for (int64_t i = n - 1; i > 0; --i) {
        s += 0x9e3779b97f4a7c15ULL;
        uint64_t z = s;
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
        z ^= z >> 31;
        std::swap(order[(size_t)i], order[(size_t)(z % (uint64_t)(i + 1))]);
    }
    std::vector<float> xs((size_t)n * 4), ys((size_t)n);
    for (int64_t i = 0; i < n; ++i) {
        std::memcpy(&xs[(size_t)i * 4], &x[order[(size_t)i] * 4], 16);
        ys[(size_t)i] = y[order[(size_t)i]];
    }
    const int64_t B = std::min<int64_t>(batch, n);
    DevBuf dp, dx, dy, dg, dl;
    GALP_CK(dp.alloc(4 * P));
    GALP_CK(dx.alloc(16 * (size_t)n));
    GALP_CK(dy.alloc(4 * (size_t)n));
    GALP_CK(dg.alloc(4 * P));
    GALP_CK(dl.alloc(4));
    GALP_CK(hipMemcpyAsync(dp.p, params, 4 * P, hipMemcpyHostToDevice, st));
    GALP_CK(hipMemcpyAsync(dx.p, xs.data(), 16 * (size_t)n, hipMemcpyHostToDevice, st));
    GALP_CK(hipMemcpyAsync(dy.p, ys.data(), 4 * (size_t)n, hipMemcpyHostToDevice, st));
    if (epochs > 0 && mlp_train_dev(st, hidden, dp.as<float>(), dx.as<float>(), dy.as<float>(), n, epochs, (int)B,
                                    lr) != GALP_OK)
        GALP_CK(hipGetLastError());
    // loss of the final parameters over the whole set
    {
        const int64_t nbf = (n + 63) / 64;
        DevBuf fpg, fpl;
        GALP_CK(fpg.alloc(4 * (size_t)(nbf * P)));
        GALP_CK(fpl.alloc(4 * (size_t)nbf));
        if (mlp_grad_dev(st, hidden, dp.as<float>(), dx.as<float>(), dy.as<float>(), n, fpg.as<float>(),
                         fpl.as<float>(), dg.as<float>(), dl.as<float>()) != GALP_OK)
            GALP_CK(hipGetLastError());
        float fl = 0.f;
        GALP_CK(hipMemcpyAsync(&fl, dl.p, 4, hipMemcpyDeviceToHost, st));
        GALP_CK(hipMemcpyAsync(params, dp.p, 4 * P, hipMemcpyDeviceToHost, st));
        GALP_CK(hipStreamSynchronize(st));
        if (final_loss) *final_loss = fl;
    }
    (void)hipStreamDestroy(st);
    return GALP_OK;
}

}  // extern "C"
