// mlp.hpp -- the learned dispatcher's 4 -> H -> H -> 1 MLP: parameter count
// and the forward pass shared by the inference, dispatch and training code.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace galp {

__host__ __device__ constexpr int mlp_num_params(int H) { return 4 * H + H + H * H + H + H + 1; }

// Forward pass of one sample; p points at the parameters (LDS or global).
template <int H>
__device__ __forceinline__ float mlp_forward(const float *__restrict__ p, float x0, float x1, float x2, float x3) {
    const float *W1 = p, *b1 = p + 4 * H, *W2 = b1 + H, *b2 = W2 + H * H, *W3 = b2 + H, *b3 = W3 + H;
    float h1[H];
#pragma unroll
    for (int j = 0; j < H; ++j) {
        float a = b1[j] + W1[4 * j] * x0 + W1[4 * j + 1] * x1 + W1[4 * j + 2] * x2 + W1[4 * j + 3] * x3;
        h1[j] = a > 0.f ? a : 0.f;
    }
    float out = b3[0];
#pragma unroll 4
    for (int j = 0; j < H; ++j) {
        float a = b2[j];
#pragma unroll
        for (int k = 0; k < H; ++k) a += W2[H * j + k] * h1[k];
        out += W3[j] * (a > 0.f ? a : 0.f);
    }
    return out;
}

int mlp_predict_dev(hipStream_t st, int H, const float *dparams, const float *dx, int64_t n, float *dy);

}  // namespace galp
