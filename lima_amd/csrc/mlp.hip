// mlp.hip -- the learned step-scale model: BF16 MFMA inference and training kernels (FP32
// accumulation, FP32 master weights, Adam), behind the lima_mlp_* C ABI.
//
// Model: s = 4 sigmoid(w3 . silu(W2^T silu(W1^T x + b1) + b2) + b3), x = 32 per-atom features.
// Training minimises 1/2 (s - t)^2 over samples (x, t) recorded from plain steepest-descent runs.
//
// Training step (one wave per 16 samples, one block per wave):
//   forward       mlp_forward16 (MFMA, weights = A operand)
//   dz3           (s - t) s (1 - s / 4)
//   dz2           dz3 w3 silu'(z2)                                  (D layout: rows hidden, cols samples)
//   dh1 = W2 dz2  MFMA with A = W2 (input-major copy), B = dz2 as held
//   dz1           dh1 silu'(z1)
//   dW2^T = dz2 h1^T, dW1^T = dz1 x^T   MFMA over k = samples; dz and h1 are transposed through LDS
//   bias / w3 / b3 gradients: sums over the 16 sample lanes
// Each block writes its whole gradient; k_mlp_adam sums the blocks in a fixed order (deterministic)
// and applies Adam, then refreshes the BF16 operand copies.
#include <chrono>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "engine.h"
#include "lima_mlp.h"

namespace lima {

struct Mlp {
    float* p = nullptr;            // FP32 parameters [MLP_NPARAM]
    float* m = nullptr;            // Adam moments
    float* v = nullptr;
    unsigned short* w1t = nullptr; // BF16 operand copies
    unsigned short* w2t = nullptr;
    unsigned short* w2 = nullptr;
    int64_t adam_t = 0;
    MlpWeights view() const { return MlpWeights{w1t, w2t, w2, p}; }
};

namespace {

inline int nblk(int64_t n, int b) { return (int)((n + b - 1) / b); }

__global__ void k_mlp_pack(const float* __restrict__ p, unsigned short* __restrict__ w1t, unsigned short* __restrict__ w2t,
                           unsigned short* __restrict__ w2) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < MLP_IN * MLP_H) {   // W1 [in][h] -> w1t [h][in]
        const int in = t / MLP_H, h = t % MLP_H;
        w1t[h * MLP_IN + in] = f2bf(p[MLP_OFF_W1 + t]);
    }
    if (t < MLP_H * MLP_H) {    // W2 [in][out] -> w2t [out][in], w2 [in][out]
        const int in = t / MLP_H, out = t % MLP_H;
        w2t[out * MLP_H + in] = f2bf(p[MLP_OFF_W2 + t]);
        w2[t] = f2bf(p[MLP_OFF_W2 + t]);
    }
}

// B operand of the inputs for 16 samples: lane's sample lane % 16, inputs 16 ks + 4 (lane / 16) + i
__device__ __forceinline__ void load_x16(const float* __restrict__ x, int64_t s0, int64_t n, bf16x4* xb) {
    const int lane = threadIdx.x & 63, r = lane & 15, g = lane >> 4;
    const int64_t s = s0 + r;
#pragma unroll
    for (int ks = 0; ks < 2; ks++) {
        if (s < n) {
            const float* row = x + s * MLP_IN + ks * 16 + 4 * g;
            xb[ks] = pack4(row[0], row[1], row[2], row[3]);
        } else {
            xb[ks] = pack4(0.f, 0.f, 0.f, 0.f);
        }
    }
}

__global__ __launch_bounds__(64) void k_mlp_forward(MlpWeights w, const float* __restrict__ x, int64_t n, float* __restrict__ out) {
    const int64_t s0 = (int64_t)blockIdx.x * 16;
    bf16x4 xb[2], h1[4];
    f32x4 z1[4], z2[4];
    load_x16(x, s0, n, xb);
    const float z3 = mlp_forward16(w, xb, z1, h1, z2);
    const int lane = threadIdx.x & 63;
    if (lane < 16 && s0 + lane < n) out[s0 + lane] = mlp_scale(z3);
}

// sum over the 16 sample lanes that share lane / 16
__device__ __forceinline__ float sum16(float v) {
    v += __shfl_xor(v, 1, 64);
    v += __shfl_xor(v, 2, 64);
    v += __shfl_xor(v, 4, 64);
    v += __shfl_xor(v, 8, 64);
    return v;
}

// One wave per block, 16 samples (idx[s0 .. s0+15] of the shuffled order); writes the block's whole
// gradient (un-normalised sum over its samples) to grad[blockIdx.x][MLP_NPARAM] and its loss sum.
__global__ __launch_bounds__(64) void k_mlp_grad(MlpWeights w, const float* __restrict__ x, const float* __restrict__ tgt,
                                                 const int* __restrict__ idx, int64_t nbatch, float* __restrict__ grad,
                                                 float* __restrict__ loss) {
    __shared__ unsigned short tz2[MLP_H * 16], th1[MLP_H * 16], tz1[MLP_H * 16];
    const int lane = threadIdx.x & 63, r = lane & 15, g = lane >> 4;
    const int64_t b0 = (int64_t)blockIdx.x * 16;
    const bool valid = b0 + r < nbatch;
    const int sample = valid ? idx[b0 + r] : 0;
    // gather this wave's samples: B operand of the inputs
    bf16x4 xb[2];
#pragma unroll
    for (int ks = 0; ks < 2; ks++) {
        const float* row = x + (int64_t)sample * MLP_IN + ks * 16 + 4 * g;
        xb[ks] = valid ? pack4(row[0], row[1], row[2], row[3]) : pack4(0.f, 0.f, 0.f, 0.f);
    }
    bf16x4 h1[4];
    f32x4 z1[4], z2[4];
    const float z3 = mlp_forward16(w, xb, z1, h1, z2);
    const float s = mlp_scale(z3);
    const float t = valid ? tgt[sample] : s;
    const float dz3 = valid ? (s - t) * s * (1.f - 0.25f * s) : 0.f;
    float* G = grad + (int64_t)blockIdx.x * MLP_NPARAM;
    // b3 and the loss (each sample appears in the 4 lanes of its column; count lane group 0)
    {
        const float l = (g == 0 && valid) ? 0.5f * (s - t) * (s - t) : 0.f;
        const float db3 = g == 0 ? dz3 : 0.f;
        float sl = sum16(l), sb = sum16(db3);
        if (lane == 0) { G[MLP_OFF_B3] = sb; loss[blockIdx.x] = sl; }
    }
    // dz2, w3 and b2 gradients; stage dz2 and h1 for the sample-major products
    f32x4 dz2[4];
    bf16x4 dz2b[4];
#pragma unroll
    for (int hb = 0; hb < 4; hb++) {
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int h = hb * 16 + 4 * g + i;
            const float a2 = silu(z2[hb][i]);
            const float d = dz3 * w.p[MLP_OFF_W3 + h] * silu_grad(z2[hb][i]);
            dz2[hb][i] = d;
            const float gw3 = sum16(dz3 * a2), gb2 = sum16(d);
            if (r == 0) { G[MLP_OFF_W3 + h] = gw3; G[MLP_OFF_B2 + h] = gb2; }
            tz2[h * 16 + r] = f2bf(d);
            th1[h * 16 + r] = (unsigned short)h1[hb][i];
        }
        dz2b[hb] = pack4(dz2[hb][0], dz2[hb][1], dz2[hb][2], dz2[hb][3]);
    }
    // dh1 = W2 dz2 (A = W2 input-major, B = dz2 as held), dz1, b1 gradient
#pragma unroll
    for (int ib = 0; ib < 4; ib++) {
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int ob = 0; ob < 4; ob++) {
            const bf16x4 a = *reinterpret_cast<const bf16x4*>(w.w2 + (ib * 16 + r) * MLP_H + ob * 16 + 4 * g);
            acc = mfma16(a, dz2b[ob], acc);
        }
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int h = ib * 16 + 4 * g + i;
            const float d1 = acc[i] * silu_grad(z1[ib][i]);
            const float gb1 = sum16(d1);
            if (r == 0) G[MLP_OFF_B1 + h] = gb1;
            tz1[h * 16 + r] = f2bf(d1);
        }
    }
    __syncthreads();
    // dW2^T [out][in] = sum_s dz2[out][s] h1[in][s]
#pragma unroll
    for (int ob = 0; ob < 4; ob++) {
        const bf16x4 a = *reinterpret_cast<const bf16x4*>(tz2 + (ob * 16 + r) * 16 + 4 * g);
#pragma unroll
        for (int ib = 0; ib < 4; ib++) {
            const bf16x4 b = *reinterpret_cast<const bf16x4*>(th1 + (ib * 16 + r) * 16 + 4 * g);
            const f32x4 d = mfma16(a, b, f32x4{0.f, 0.f, 0.f, 0.f});
#pragma unroll
            for (int i = 0; i < 4; i++) G[MLP_OFF_W2 + (ib * 16 + r) * MLP_H + ob * 16 + 4 * g + i] = d[i];
        }
    }
    // dW1^T [h][in] = sum_s dz1[h][s] x[in][s]; B: column input ks*16 + r, k = samples 4 g + i
    bf16x4 xt[2];
#pragma unroll
    for (int ks = 0; ks < 2; ks++) {
        float v4[4];
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int64_t bi = b0 + 4 * g + i;
            v4[i] = bi < nbatch ? x[(int64_t)idx[bi] * MLP_IN + ks * 16 + r] : 0.f;
        }
        xt[ks] = pack4(v4[0], v4[1], v4[2], v4[3]);
    }
#pragma unroll
    for (int hb = 0; hb < 4; hb++) {
        const bf16x4 a = *reinterpret_cast<const bf16x4*>(tz1 + (hb * 16 + r) * 16 + 4 * g);
#pragma unroll
        for (int ks = 0; ks < 2; ks++) {
            const f32x4 d = mfma16(a, xt[ks], f32x4{0.f, 0.f, 0.f, 0.f});
#pragma unroll
            for (int i = 0; i < 4; i++) G[MLP_OFF_W1 + (ks * 16 + r) * MLP_H + hb * 16 + 4 * g + i] = d[i];
        }
    }
}

// Sum the block gradients in order, mean over the batch, Adam (beta 0.9 / 0.999, eps 1e-8).
__global__ void k_mlp_adam(float* __restrict__ p, float* __restrict__ m, float* __restrict__ v, const float* __restrict__ grad,
                           int nblocks, float inv_batch, float lr, float bc1, float bc2) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= MLP_NPARAM) return;
    float gsum = 0.f;
    for (int b = 0; b < nblocks; b++) gsum += grad[(int64_t)b * MLP_NPARAM + k];
    const float gk = gsum * inv_batch;
    const float mk = 0.9f * m[k] + 0.1f * gk;
    const float vk = 0.999f * v[k] + 0.001f * gk * gk;
    m[k] = mk;
    v[k] = vk;
    p[k] -= lr * (mk / bc1) / (sqrtf(vk / bc2) + 1e-8f);
}

struct SplitMix {
    uint64_t s;
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

void mlp_refresh(Mlp& M) {
    hipLaunchKernelGGL(k_mlp_pack, dim3(nblk(MLP_H * MLP_H, 256)), dim3(256), 0, 0, M.p, M.w1t, M.w2t, M.w2);
    HIP_CHECK(hipGetLastError());
}

}  // namespace

}  // namespace lima

struct lima_mlp {
    lima::Mlp m;
};

// Staging array of lima_mlp_forward_noalloc: feature rows [LIMA_MLP_NOALLOC_ROWS][32], then the
// step scales [LIMA_MLP_NOALLOC_ROWS] (8.65 MB).  Every GPU the library runs on holds its own copy
// until the process ends.  The name is a host-side stand-in: host code uses the per-device address
// that scratch_address() obtains, never the name.
static __device__ __align__(256) float scratch[(size_t)LIMA_MLP_NOALLOC_ROWS * (lima::MLP_IN + 1)];

namespace {
template <class F>
lima_status mguard(F&& f) {
    try {
        f();
        return LIMA_OK;
    } catch (const lima::LimaFail& x) {
        lima::g_err = x.what();
        return x.code;
    } catch (const lima::HipError& x) {
        lima::g_err = x.what();
        return LIMA_ERR_HIP;
    } catch (const std::exception& x) {
        lima::g_err = x.what();
        return LIMA_ERR_BAD_INPUT;
    }
}
constexpr int SCRATCH_MAX_DEVICES = 64;
std::mutex g_scratch_mu;                            // guards g_scratch and every use of the array
float* g_scratch[SCRATCH_MAX_DEVICES] = {};         // device address of `scratch`, per device

// Address of the current device's copy of `scratch`, looked up once per device with that device
// current.  Call with g_scratch_mu held.
float* scratch_address() {
    int dev = -1;
    HIP_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= SCRATCH_MAX_DEVICES) throw lima::LimaFail(LIMA_ERR_NO_DEVICE, "device index out of range");
    if (!g_scratch[dev]) {
        void* p = nullptr;
        HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) throw lima::LimaFail(LIMA_ERR_HIP, "hipGetSymbolAddress returned a null address");
        g_scratch[dev] = static_cast<float*>(p);
    }
    return g_scratch[dev];
}
void mlp_free(lima::Mlp& M) {
    void* ptrs[] = {M.p, M.m, M.v, M.w1t, M.w2t, M.w2};
    for (void* p : ptrs) if (p) (void)hipFree(p);
}
}  // namespace

namespace lima {
MlpWeights mlp_view(const lima_mlp* h) { return h->m.view(); }
}

extern "C" {

int32_t lima_mlp_n_params(void) { return lima::MLP_NPARAM; }

lima_status lima_mlp_create(uint64_t seed, lima_mlp** out) {
    if (!out) { lima::g_err = "null output handle"; return LIMA_ERR_BAD_INPUT; }
    *out = nullptr;
    lima_mlp* h = new lima_mlp();
    const lima_status st = mguard([&] {
        using namespace lima;
        int ndev = 0;
        if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) throw LimaFail(LIMA_ERR_NO_DEVICE, "no HIP device visible");
        Mlp& M = h->m;
        HIP_CHECK(hipMalloc(&M.p, MLP_NPARAM * sizeof(float)));
        HIP_CHECK(hipMalloc(&M.m, MLP_NPARAM * sizeof(float)));
        HIP_CHECK(hipMalloc(&M.v, MLP_NPARAM * sizeof(float)));
        HIP_CHECK(hipMalloc(&M.w1t, MLP_IN * MLP_H * sizeof(unsigned short)));
        HIP_CHECK(hipMalloc(&M.w2t, MLP_H * MLP_H * sizeof(unsigned short)));
        HIP_CHECK(hipMalloc(&M.w2, MLP_H * MLP_H * sizeof(unsigned short)));
        // Glorot-uniform weights, zero biases; b3 = 0 gives s = 2 before training
        SplitMix rng{seed};
        std::vector<float> p(MLP_NPARAM, 0.f);
        const double a1 = std::sqrt(6.0 / (MLP_IN + MLP_H)), a2 = std::sqrt(6.0 / (2 * MLP_H)), a3 = std::sqrt(6.0 / (MLP_H + 1));
        for (int k = 0; k < MLP_IN * MLP_H; k++) p[MLP_OFF_W1 + k] = (float)((2.0 * rng.uniform() - 1.0) * a1);
        for (int k = 0; k < MLP_H * MLP_H; k++) p[MLP_OFF_W2 + k] = (float)((2.0 * rng.uniform() - 1.0) * a2);
        for (int k = 0; k < MLP_H; k++) p[MLP_OFF_W3 + k] = (float)((2.0 * rng.uniform() - 1.0) * a3);
        HIP_CHECK(hipMemcpy(M.p, p.data(), MLP_NPARAM * sizeof(float), hipMemcpyHostToDevice));
        HIP_CHECK(hipMemset(M.m, 0, MLP_NPARAM * sizeof(float)));
        HIP_CHECK(hipMemset(M.v, 0, MLP_NPARAM * sizeof(float)));
        mlp_refresh(M);
        HIP_CHECK(hipDeviceSynchronize());
    });
    if (st != LIMA_OK) { mlp_free(h->m); delete h; return st; }
    *out = h;
    return LIMA_OK;
}

lima_status lima_mlp_destroy(lima_mlp* h) {
    if (!h) return LIMA_OK;
    mlp_free(h->m);
    delete h;
    return LIMA_OK;
}

lima_status lima_mlp_get_params(lima_mlp* h, float* params) {
    if (!h || !params) { lima::g_err = "null argument"; return LIMA_ERR_BAD_INPUT; }
    return mguard([&] { HIP_CHECK(hipMemcpy(params, h->m.p, lima::MLP_NPARAM * sizeof(float), hipMemcpyDeviceToHost)); });
}

lima_status lima_mlp_set_params(lima_mlp* h, const float* params) {
    if (!h || !params) { lima::g_err = "null argument"; return LIMA_ERR_BAD_INPUT; }
    return mguard([&] {
        for (int k = 0; k < lima::MLP_NPARAM; k++)
            if (!std::isfinite(params[k])) throw lima::LimaFail(LIMA_ERR_BAD_INPUT, "non-finite parameter");
        HIP_CHECK(hipMemcpy(h->m.p, params, lima::MLP_NPARAM * sizeof(float), hipMemcpyHostToDevice));
        lima::mlp_refresh(h->m);
        HIP_CHECK(hipDeviceSynchronize());
    });
}

lima_status lima_mlp_forward(lima_mlp* h, const float* features, int64_t n, float* scale_out) {
    if (!h || !features || !scale_out || n < 0) { lima::g_err = "bad argument"; return LIMA_ERR_BAD_INPUT; }
    if (n == 0) return LIMA_OK;
    return mguard([&] {
        using namespace lima;
        float *dx = nullptr, *dy = nullptr;
        HIP_CHECK(hipMalloc(&dx, n * MLP_IN * sizeof(float)));
        HIP_CHECK(hipMalloc(&dy, n * sizeof(float)));
        HIP_CHECK(hipMemcpy(dx, features, n * MLP_IN * sizeof(float), hipMemcpyHostToDevice));
        hipLaunchKernelGGL(k_mlp_forward, dim3(nblk(n, 16)), dim3(64), 0, 0, h->m.view(), dx, n, dy);
        const hipError_t e = hipGetLastError();
        HIP_CHECK(hipMemcpy(scale_out, dy, n * sizeof(float), hipMemcpyDeviceToHost));
        (void)hipFree(dx);
        (void)hipFree(dy);
        HIP_CHECK(e);
    });
}

lima_status lima_mlp_forward_noalloc(lima_mlp* h, const float* features, int64_t n, float* scale_out) {
    if (!h || !features || !scale_out || n < 0) { lima::g_err = "bad argument"; return LIMA_ERR_BAD_INPUT; }
    if (n > LIMA_MLP_NOALLOC_ROWS) {
        lima::g_err = "lima_mlp_forward_noalloc: " + std::to_string(n) + " rows exceed the static staging array of " +
                      std::to_string(LIMA_MLP_NOALLOC_ROWS) + " rows; use lima_mlp_forward or split the call";
        return LIMA_ERR_BAD_INPUT;
    }
    if (n == 0) return LIMA_OK;
    return mguard([&] {
        using namespace lima;
        // Held until the results are on the host: calls from several host threads run one after
        // another instead of sharing the array at once.
        std::lock_guard<std::mutex> lock(g_scratch_mu);
        float* dx = scratch_address();
        float* dy = dx + (size_t)LIMA_MLP_NOALLOC_ROWS * MLP_IN;
        // every row the kernel reads (< n) is written here and every scale it writes is copied back:
        // nothing of an earlier call is read, so nothing needs clearing
        HIP_CHECK(hipMemcpy(dx, features, n * MLP_IN * sizeof(float), hipMemcpyHostToDevice));
        hipLaunchKernelGGL(k_mlp_forward, dim3(nblk(n, 16)), dim3(64), 0, 0, h->m.view(), dx, n, dy);
        HIP_CHECK(hipGetLastError());
        HIP_CHECK(hipMemcpy(scale_out, dy, n * sizeof(float), hipMemcpyDeviceToHost));
    });
}

lima_status lima_mlp_gradient(lima_mlp* h, const float* features, const float* targets, int64_t n, float* grad_out, double* loss) {
    if (!h || !features || !targets || !grad_out || n <= 0) { lima::g_err = "bad argument"; return LIMA_ERR_BAD_INPUT; }
    return mguard([&] {
        using namespace lima;
        const int blocks = (int)((n + 15) / 16);
        float *dx = nullptr, *dt = nullptr, *grad = nullptr, *dloss = nullptr;
        int* didx = nullptr;
        std::vector<int> idx(n);
        for (int64_t k = 0; k < n; k++) idx[k] = (int)k;
        HIP_CHECK(hipMalloc(&dx, n * MLP_IN * sizeof(float)));
        HIP_CHECK(hipMalloc(&dt, n * sizeof(float)));
        HIP_CHECK(hipMalloc(&didx, n * sizeof(int)));
        HIP_CHECK(hipMalloc(&grad, (size_t)blocks * MLP_NPARAM * sizeof(float)));
        HIP_CHECK(hipMalloc(&dloss, blocks * sizeof(float)));
        HIP_CHECK(hipMemcpy(dx, features, n * MLP_IN * sizeof(float), hipMemcpyHostToDevice));
        HIP_CHECK(hipMemcpy(dt, targets, n * sizeof(float), hipMemcpyHostToDevice));
        HIP_CHECK(hipMemcpy(didx, idx.data(), n * sizeof(int), hipMemcpyHostToDevice));
        hipLaunchKernelGGL(k_mlp_grad, dim3(blocks), dim3(64), 0, 0, h->m.view(), dx, dt, didx, n, grad, dloss);
        const hipError_t e = hipGetLastError();
        std::vector<float> g((size_t)blocks * MLP_NPARAM), l(blocks);
        HIP_CHECK(hipMemcpy(g.data(), grad, g.size() * sizeof(float), hipMemcpyDeviceToHost));
        HIP_CHECK(hipMemcpy(l.data(), dloss, blocks * sizeof(float), hipMemcpyDeviceToHost));
        (void)hipFree(dx); (void)hipFree(dt); (void)hipFree(didx); (void)hipFree(grad); (void)hipFree(dloss);
        HIP_CHECK(e);
        double ls = 0.0;
        for (int k = 0; k < MLP_NPARAM; k++) {
            double acc = 0.0;
            for (int b = 0; b < blocks; b++) acc += g[(size_t)b * MLP_NPARAM + k];
            grad_out[k] = (float)acc;
        }
        for (int b = 0; b < blocks; b++) ls += l[b];
        if (loss) *loss = ls;
    });
}

lima_status lima_mlp_train(lima_mlp* h, const float* features, const float* targets, int64_t n, int32_t epochs, int32_t batch,
                           float lr, uint64_t seed, float* loss_per_epoch, double* seconds) {
    if (!h || !features || !targets || n <= 0 || epochs < 0 || batch < 16 || batch % 16 != 0 || !(lr > 0.f)) {
        lima::g_err = "bad argument (n > 0, batch a positive multiple of 16, lr > 0)";
        return LIMA_ERR_BAD_INPUT;
    }
    return mguard([&] {
        using namespace lima;
        Mlp& M = h->m;
        const auto t0 = std::chrono::steady_clock::now();
        float *dx = nullptr, *dt = nullptr, *grad = nullptr, *dloss = nullptr;
        int* didx = nullptr;
        const int nb = batch / 16;
        HIP_CHECK(hipMalloc(&dx, n * MLP_IN * sizeof(float)));
        HIP_CHECK(hipMalloc(&dt, n * sizeof(float)));
        HIP_CHECK(hipMalloc(&didx, n * sizeof(int)));
        HIP_CHECK(hipMalloc(&grad, (size_t)nb * MLP_NPARAM * sizeof(float)));
        HIP_CHECK(hipMalloc(&dloss, nb * sizeof(float)));
        HIP_CHECK(hipMemcpy(dx, features, n * MLP_IN * sizeof(float), hipMemcpyHostToDevice));
        HIP_CHECK(hipMemcpy(dt, targets, n * sizeof(float), hipMemcpyHostToDevice));
        SplitMix rng{seed};
        std::vector<int> perm(n);
        std::vector<float> lbuf(nb);
        try {
            for (int ep = 0; ep < epochs; ep++) {
                for (int64_t k = 0; k < n; k++) perm[k] = (int)k;
                for (int64_t k = n - 1; k > 0; k--) std::swap(perm[k], perm[(int64_t)(rng.next() % (uint64_t)(k + 1))]);
                HIP_CHECK(hipMemcpy(didx, perm.data(), n * sizeof(int), hipMemcpyHostToDevice));
                double lsum = 0.0;
                for (int64_t s0 = 0; s0 < n; s0 += batch) {
                    const int64_t nbat = std::min<int64_t>(batch, n - s0);
                    const int blocks = (int)((nbat + 15) / 16);
                    hipLaunchKernelGGL(k_mlp_grad, dim3(blocks), dim3(64), 0, 0, M.view(), dx, dt, didx + s0, nbat, grad, dloss);
                    M.adam_t++;
                    const float bc1 = 1.f - std::pow(0.9f, (float)M.adam_t), bc2 = 1.f - std::pow(0.999f, (float)M.adam_t);
                    hipLaunchKernelGGL(k_mlp_adam, dim3(nblk(MLP_NPARAM, 256)), dim3(256), 0, 0, M.p, M.m, M.v, grad, blocks,
                                       1.f / (float)nbat, lr, bc1, bc2);
                    hipLaunchKernelGGL(k_mlp_pack, dim3(nblk(MLP_H * MLP_H, 256)), dim3(256), 0, 0, M.p, M.w1t, M.w2t, M.w2);
                    if (loss_per_epoch) {
                        HIP_CHECK(hipMemcpy(lbuf.data(), dloss, blocks * sizeof(float), hipMemcpyDeviceToHost));
                        for (int b = 0; b < blocks; b++) lsum += lbuf[b];
                    }
                }
                HIP_CHECK(hipGetLastError());
                if (loss_per_epoch) loss_per_epoch[ep] = (float)(lsum / (double)n);
            }
            HIP_CHECK(hipDeviceSynchronize());
        } catch (...) {
            (void)hipFree(dx); (void)hipFree(dt); (void)hipFree(didx); (void)hipFree(grad); (void)hipFree(dloss);
            throw;
        }
        (void)hipFree(dx); (void)hipFree(dt); (void)hipFree(didx); (void)hipFree(grad); (void)hipFree(dloss);
        if (seconds) *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    });
}

}  // extern "C"
