// pmh_mlp.hip -- the learned sizer: a 3-layer fp32 MLP (features -> 64 ->
// 64 -> 2) with hand-written forward, backward and Adam kernels.
//
// No counterpart in the reference (it has no cardinality estimator; its
// nearest precedents are fixed thresholds, crates/polars-metal-core/src/
// cost.rs:38 and crates/polars-metal-kernels/src/groupby.rs:493).
//
// Layout: one wavefront per sample, lane j = hidden unit j, so a layer is 64
// lanes each doing one dot product; activations move between lanes with
// shuffles.  The network is tiny (about 5.3K parameters), so the weights sit
// in LDS and there is no MFMA: the work is latency-bound, not compute-bound.
//
// Parameter vector layout (floats): W1[F][64], b1[64], W2[64][64], b2[64],
// W3[64][2], b3[2]; F = kMlpIn.  Outputs: o0 = log2(groups + 1) / 32,
// o1 = logit of "the sort engine is faster than the hash engine".
// Loss per sample: (o0 - y0)^2 + lambda * mask * BCE(sigmoid(o1), y1), where
// mask = (y1 >= 0) lets logs without a timing label train o0 only; the
// batch loss is the mean over samples.
#include <cmath>
#include <cstring>
#include <vector>

#include "pmh_mlp.h"
#include "pmh_mlp_dev.h"

namespace pmh {

namespace {

constexpr int H = kMlpHidden;
constexpr int F = kMlpIn;
using mlpdev::O_W1;
using mlpdev::O_B1;
using mlpdev::O_W2;
using mlpdev::O_B2;
using mlpdev::O_W3;
using mlpdev::O_B3;
using mlpdev::relu;
using Acts = mlpdev::Acts;
static_assert(mlpdev::O_END == kMlpParams, "parameter layout");

// Loads the parameters into LDS (every thread of the block helps).
__device__ __forceinline__ void load_params(float *sp, const float *__restrict__ p) {
    for (int i = threadIdx.x; i < kMlpParams; i += blockDim.x) sp[i] = p[i];
    __syncthreads();
}

__device__ __forceinline__ Acts forward_one(const float *sp, const float *__restrict__ x, int lane) {
    return mlpdev::forward_lane(sp, x, lane);
}

__global__ __launch_bounds__(256) void mlp_forward(const float *__restrict__ params, const float *__restrict__ x,
                                                   uint32_t n, float *__restrict__ y) {
    __shared__ float sp[kMlpParams];
    load_params(sp, params);
    const int lane = threadIdx.x & 63;
    const uint32_t waves = blockDim.x / 64;
    for (uint32_t s = blockIdx.x * waves + threadIdx.x / 64; s < n; s += gridDim.x * waves) {
        const Acts a = forward_one(sp, x + (uint64_t)s * F, lane);
        if (lane == 0) {
            y[2 * s] = a.o0;
            y[2 * s + 1] = a.o1;
        }
    }
}

// Gradient of the mean loss over samples [0, n) into grad (accumulated with
// atomics: grad must be zero on entry) and the summed loss into loss[0].
// Each wave keeps its partial gradient in registers (lane j owns column j
// of W1 and W2, and rows of W3) and flushes once.
__global__ __launch_bounds__(256) void mlp_backward(const float *__restrict__ params, const float *__restrict__ x,
                                                    const float *__restrict__ y, uint32_t n, const uint32_t *idx,
                                                    float lambda, float *__restrict__ grad, float *__restrict__ loss) {
    __shared__ float sp[kMlpParams];
    load_params(sp, params);
    const int lane = threadIdx.x & 63;
    const uint32_t waves = blockDim.x / 64;
    const float inv_n = 1.f / (float)n;
    float gW1[F], gW2[H];
#pragma unroll
    for (int i = 0; i < F; ++i) gW1[i] = 0.f;
#pragma unroll
    for (int k = 0; k < H; ++k) gW2[k] = 0.f;
    float gb1 = 0.f, gb2 = 0.f, gW3a = 0.f, gW3b = 0.f, gb3a = 0.f, gb3b = 0.f, lsum = 0.f;
    for (uint32_t t = blockIdx.x * waves + threadIdx.x / 64; t < n; t += gridDim.x * waves) {
        const uint32_t s = idx ? idx[t] : t;
        const float *xs = x + (uint64_t)s * F;
        const Acts a = forward_one(sp, xs, lane);
        const float y0 = y[2 * s], y1 = y[2 * s + 1];
        const float m = y1 >= 0.f ? 1.f : 0.f;
        const float sig = 1.f / (1.f + __expf(-a.o1));
        const float d0 = a.o0 - y0;
        // softplus(o1) - y1 * o1, in a form that does not overflow
        const float sp1 = a.o1 > 0.f ? a.o1 + log1pf(__expf(-a.o1)) : log1pf(__expf(a.o1));
        lsum += d0 * d0 + lambda * m * (sp1 - y1 * a.o1);
        const float g0 = 2.f * d0 * inv_n;
        const float g1 = lambda * m * (sig - y1) * inv_n;
        // layer 3
        gW3a += a.h2 * g0;
        gW3b += a.h2 * g1;
        gb3a += g0;
        gb3b += g1;
        const float dz2 = (sp[O_W3 + lane * 2] * g0 + sp[O_W3 + lane * 2 + 1] * g1) * (a.h2 > 0.f ? 1.f : 0.f);
        gb2 += dz2;
        // W2[k][lane] gets h1[k] * dz2[lane]; dh1[lane] = sum_j W2[lane][j] dz2[j]
        float dh1 = 0.f;
#pragma unroll 8
        for (int k = 0; k < H; ++k) gW2[k] += __shfl(a.h1, k) * dz2;
        // row `lane` of W2, walked from a lane-dependent start so the 64
        // lanes read 64 different LDS banks
#pragma unroll 8
        for (int k = 0; k < H; ++k) {
            const int kk = (k + lane) & (H - 1);
            dh1 += sp[O_W2 + lane * H + kk] * __shfl(dz2, kk);
        }
        const float dz1 = dh1 * (a.h1 > 0.f ? 1.f : 0.f);
        gb1 += dz1;
#pragma unroll
        for (int i = 0; i < F; ++i) gW1[i] += xs[i] * dz1;
    }
    // flush the wave's partial gradient
#pragma unroll
    for (int i = 0; i < F; ++i) atomicAdd(&grad[O_W1 + i * H + lane], gW1[i]);
    atomicAdd(&grad[O_B1 + lane], gb1);
#pragma unroll 8
    for (int k = 0; k < H; ++k) atomicAdd(&grad[O_W2 + k * H + lane], gW2[k]);
    atomicAdd(&grad[O_B2 + lane], gb2);
    atomicAdd(&grad[O_W3 + lane * 2], gW3a);
    atomicAdd(&grad[O_W3 + lane * 2 + 1], gW3b);
    // biases of layer 3 and the loss: uniform across the wave, lane 0 adds
    if (lane == 0) {
        atomicAdd(&grad[O_B3], gb3a);
        atomicAdd(&grad[O_B3 + 1], gb3b);
        atomicAdd(loss, lsum);
    }
}

// Adam (Kingma & Ba, 2015) with bias correction; zeroes the gradient.
__global__ void mlp_adam(float *__restrict__ p, float *__restrict__ g, float *__restrict__ m, float *__restrict__ v,
                         int step, float lr, float b1, float b2, float eps) {
    const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < kMlpParams; i += gridDim.x * blockDim.x) {
        const float gi = g[i];
        const float mi = b1 * m[i] + (1.f - b1) * gi;
        const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
        g[i] = 0.f;
    }
}

__global__ void mlp_zero(float *__restrict__ a, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) a[i] = 0.f;
}

void launch_check() { PMH_HIP(hipGetLastError()); }

// One device buffer of floats, freed with the model.
struct DBuf {
    float *p = nullptr;
    size_t n = 0;
    void ensure(size_t count) {
        if (count <= n) return;
        if (p) PMH_HIP(hipFree(p));
        PMH_HIP(hipMalloc(&p, count * sizeof(float)));
        n = count;
    }
    ~DBuf() {
        if (p) (void)hipFree(p);
    }
};

}  // namespace

struct Mlp {
    DBuf params, grad, m, v, x, y, tmp, loss;
    DBuf idx;  // uint32 sample order of an epoch (stored in float-sized slots)
    uint32_t n_data = 0;
    int step = 0;
};

Mlp *mlp_new(const float *init) {
    rt();
    auto *M = new Mlp();
    M->params.ensure(kMlpParams);
    M->grad.ensure(kMlpParams);
    M->m.ensure(kMlpParams);
    M->v.ensure(kMlpParams);
    M->loss.ensure(4);
    Runtime &R = rt();
    PMH_HIP(hipMemcpyAsync(M->params.p, init, kMlpParams * 4, hipMemcpyHostToDevice, R.stream));
    hipLaunchKernelGGL(mlp_zero, dim3(8), dim3(256), 0, R.stream, M->grad.p, kMlpParams);
    hipLaunchKernelGGL(mlp_zero, dim3(8), dim3(256), 0, R.stream, M->m.p, kMlpParams);
    hipLaunchKernelGGL(mlp_zero, dim3(8), dim3(256), 0, R.stream, M->v.p, kMlpParams);
    launch_check();
    PMH_HIP(hipStreamSynchronize(R.stream));
    return M;
}

void mlp_free(Mlp *M) { delete M; }

void mlp_get_params(Mlp *M, float *out) {
    Runtime &R = rt();
    PMH_HIP(hipMemcpyAsync(out, M->params.p, kMlpParams * 4, hipMemcpyDeviceToHost, R.stream));
    PMH_HIP(hipStreamSynchronize(R.stream));
}

void mlp_forward_host(Mlp *M, const float *x, uint32_t n, float *y) {
    Runtime &R = rt();
    if (!n) return;
    M->tmp.ensure((size_t)n * (F + 2));
    PMH_HIP(hipMemcpyAsync(M->tmp.p, x, (size_t)n * F * 4, hipMemcpyHostToDevice, R.stream));
    float *dy = M->tmp.p + (size_t)n * F;
    hipLaunchKernelGGL(mlp_forward, dim3((n + 3) / 4), dim3(256), 0, R.stream, M->params.p, M->tmp.p, n, dy);
    launch_check();
    PMH_HIP(hipMemcpyAsync(y, dy, (size_t)n * 2 * 4, hipMemcpyDeviceToHost, R.stream));
    PMH_HIP(hipStreamSynchronize(R.stream));
}

float mlp_grad_host(Mlp *M, const float *x, const float *y, uint32_t n, float lambda, float *grad) {
    Runtime &R = rt();
    if (!n) fail(PMH_E_VALUE, "polars_metal: mlp gradient over zero samples");
    M->tmp.ensure((size_t)n * (F + 2));
    PMH_HIP(hipMemcpyAsync(M->tmp.p, x, (size_t)n * F * 4, hipMemcpyHostToDevice, R.stream));
    PMH_HIP(hipMemcpyAsync(M->tmp.p + (size_t)n * F, y, (size_t)n * 2 * 4, hipMemcpyHostToDevice, R.stream));
    hipLaunchKernelGGL(mlp_zero, dim3(8), dim3(256), 0, R.stream, M->grad.p, kMlpParams);
    hipLaunchKernelGGL(mlp_zero, dim3(1), dim3(64), 0, R.stream, M->loss.p, 1);
    const uint32_t blocks = std::min<uint32_t>((n + 3) / 4, 1024);
    hipLaunchKernelGGL(mlp_backward, dim3(blocks), dim3(256), 0, R.stream, M->params.p, M->tmp.p,
                       M->tmp.p + (size_t)n * F, n, (const uint32_t *)nullptr, lambda, M->grad.p, M->loss.p);
    launch_check();
    float loss = 0;
    PMH_HIP(hipMemcpyAsync(grad, M->grad.p, kMlpParams * 4, hipMemcpyDeviceToHost, R.stream));
    PMH_HIP(hipMemcpyAsync(&loss, M->loss.p, 4, hipMemcpyDeviceToHost, R.stream));
    hipLaunchKernelGGL(mlp_zero, dim3(8), dim3(256), 0, R.stream, M->grad.p, kMlpParams);
    PMH_HIP(hipStreamSynchronize(R.stream));
    return loss / (float)n;
}

void mlp_set_data(Mlp *M, const float *x, const float *y, uint32_t n) {
    Runtime &R = rt();
    M->x.ensure((size_t)std::max<uint32_t>(n, 1) * F);
    M->y.ensure((size_t)std::max<uint32_t>(n, 1) * 2);
    if (n) {
        PMH_HIP(hipMemcpyAsync(M->x.p, x, (size_t)n * F * 4, hipMemcpyHostToDevice, R.stream));
        PMH_HIP(hipMemcpyAsync(M->y.p, y, (size_t)n * 2 * 4, hipMemcpyHostToDevice, R.stream));
    }
    M->n_data = n;
    PMH_HIP(hipStreamSynchronize(R.stream));
}

// Minibatch Adam over the uploaded data for `epochs` epochs; each epoch
// visits the samples in an order drawn from `seed` (host xorshift, uploaded
// once per epoch).  Returns the mean training loss of the last epoch.
float mlp_train(Mlp *M, uint32_t epochs, uint32_t batch, float lr, float lambda, uint32_t seed) {
    Runtime &R = rt();
    const uint32_t n = M->n_data;
    if (!n || !batch) fail(PMH_E_VALUE, "polars_metal: mlp training needs data and a batch size");
    M->idx.ensure(n);
    std::vector<uint32_t> order(n);
    uint64_t st = seed * 0x9E3779B97F4A7C15ull + 1;
    auto rnd = [&]() {
        st ^= st << 13;
        st ^= st >> 7;
        st ^= st << 17;
        return st;
    };
    float last = 0;
    for (uint32_t e = 0; e < epochs; ++e) {
        for (uint32_t i = 0; i < n; ++i) order[i] = i;
        for (uint32_t i = n - 1; i > 0; --i) std::swap(order[i], order[rnd() % (i + 1)]);
        PMH_HIP(hipMemcpyAsync(M->idx.p, order.data(), (size_t)n * 4, hipMemcpyHostToDevice, R.stream));
        hipLaunchKernelGGL(mlp_zero, dim3(1), dim3(64), 0, R.stream, M->loss.p, 1);
        for (uint32_t b0 = 0; b0 < n; b0 += batch) {
            const uint32_t nb = std::min<uint32_t>(batch, n - b0);
            const uint32_t blocks = std::min<uint32_t>((nb + 3) / 4, 256);
            hipLaunchKernelGGL(mlp_backward, dim3(blocks), dim3(256), 0, R.stream, M->params.p, M->x.p, M->y.p, nb,
                               (const uint32_t *)M->idx.p + b0, lambda, M->grad.p, M->loss.p);
            hipLaunchKernelGGL(mlp_adam, dim3(21), dim3(256), 0, R.stream, M->params.p, M->grad.p, M->m.p, M->v.p,
                               ++M->step, lr, 0.9f, 0.999f, 1e-8f);
        }
        launch_check();
        // the loss buffer holds the summed per-sample loss of the epoch
        float l = 0;
        PMH_HIP(hipMemcpyAsync(&l, M->loss.p, 4, hipMemcpyDeviceToHost, R.stream));
        PMH_HIP(hipStreamSynchronize(R.stream));
        last = l / (float)n;
    }
    return last;
}

// ---------------------------------------------------------------------------
// the installed sizer used by method "auto"
// ---------------------------------------------------------------------------
namespace {
struct Sizer {
    bool on = false;
    float host[kMlpParams];
    float *dev = nullptr;  // params + one feature row + outputs
};
Sizer g_sizer;
}  // namespace

void sizer_install(const float *params) {
    Runtime &R = rt();
    if (!params) {
        g_sizer.on = false;
        return;
    }
    memcpy(g_sizer.host, params, sizeof(g_sizer.host));
    if (!g_sizer.dev) PMH_HIP(hipMalloc(&g_sizer.dev, (kMlpParams + F + 2) * 4));
    PMH_HIP(hipMemcpyAsync(g_sizer.dev, params, kMlpParams * 4, hipMemcpyHostToDevice, R.stream));
    PMH_HIP(hipStreamSynchronize(R.stream));
    g_sizer.on = true;
}

bool sizer_installed() { return g_sizer.on; }

const float *sizer_device_params() { return g_sizer.on ? g_sizer.dev : nullptr; }

void sizer_predict_many(const float *feat, uint32_t n, float *out) {
    if (!n) return;
    Runtime &R = rt();
    static DBuf buf;
    buf.ensure((size_t)n * (F + 2));
    PMH_HIP(hipMemcpyAsync(buf.p, feat, (size_t)n * F * 4, hipMemcpyHostToDevice, R.stream));
    float *dy = buf.p + (size_t)n * F;
    hipLaunchKernelGGL(mlp_forward, dim3((n + 3) / 4), dim3(256), 0, R.stream, g_sizer.dev, buf.p, n, dy);
    launch_check();
    PMH_HIP(hipMemcpyAsync(out, dy, (size_t)n * 2 * 4, hipMemcpyDeviceToHost, R.stream));
    PMH_HIP(hipStreamSynchronize(R.stream));
}

}  // namespace pmh
