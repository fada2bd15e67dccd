// pmh_mlp_dev.h -- the learned sizer's forward pass as a device function, shared
// by the training kernels (pmh_mlp.hip) and the sample-sketch block of the
// partitioned engine (pmh_hashagg.hip), which runs one query's inference
// where its features are computed.
//
// Parameter vector layout (floats): W1[F][64], b1[64], W2[64][64], b2[64],
// W3[64][2], b3[2]; F = PMH_MLP_IN.  One wavefront evaluates one sample:
// lane j is hidden unit j, and activations move between lanes by shuffles.
#pragma once

#include <hip/hip_runtime.h>

#include "pmh.h"

namespace pmh {
namespace mlpdev {

constexpr int H = PMH_MLP_HIDDEN;
constexpr int F = PMH_MLP_IN;
constexpr int O_W1 = 0, O_B1 = O_W1 + F * H, O_W2 = O_B1 + H, O_B2 = O_W2 + H * H, O_W3 = O_B2 + H,
              O_B3 = O_W3 + H * 2, O_END = O_B3 + 2;
static_assert(O_END == PMH_MLP_PARAMS, "parameter layout");
static_assert(H == 64, "one wavefront lane per hidden unit");

__device__ __forceinline__ float relu(float x) { return x > 0.f ? x : 0.f; }

struct Acts {
    float h1, h2;  // this lane's hidden units
    float o0, o1;  // outputs (uniform across the wave)
};

// p: the parameters (LDS or global); x: F features (uniform across the wave).
__device__ __forceinline__ Acts forward_lane(const float *p, const float *__restrict__ x, int lane) {
    Acts a;
    float z = p[O_B1 + lane];
#pragma unroll
    for (int i = 0; i < F; ++i) z += p[O_W1 + i * H + lane] * x[i];
    a.h1 = relu(z);
    z = p[O_B2 + lane];
#pragma unroll 8
    for (int k = 0; k < H; ++k) z += p[O_W2 + k * H + lane] * __shfl(a.h1, k);
    a.h2 = relu(z);
    float p0 = p[O_W3 + lane * 2 + 0] * a.h2, p1 = p[O_W3 + lane * 2 + 1] * a.h2;
#pragma unroll
    for (int o = 32; o; o >>= 1) {
        p0 += __shfl_xor(p0, o);
        p1 += __shfl_xor(p1, o);
    }
    a.o0 = p0 + p[O_B3 + 0];
    a.o1 = p1 + p[O_B3 + 1];
    return a;
}

}  // namespace mlpdev
}  // namespace pmh
