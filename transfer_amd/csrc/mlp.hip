// mlp.hip -- the learned part (extension; the reference has no learned
// component): a per-vertex MLP that predicts the scalar radiation as a
// fraction of its infinite-medium value, trained and run on the device.
//
//   xh = (x - mu) / sigma,  h1 = silu(W1 xh + b1),  h2 = silu(W2 h1 + b2),
//   y  = sigmoid(w3 . h2 + b3),          loss = mean (y - t)^2
//
// 8 features, width 64.  Weights live as FP32 masters (Adam) and BF16 copies
// that the kernels use, with FP32 accumulation.  One workgroup = 64 samples
// (one per lane); a sample's pre-activations, activations and deltas sit in
// LDS rows of 65 floats (the pad keeps column walks conflict-free), and the
// weight gradients are the workgroup's sums over its 64 samples, written as
// one partial per workgroup and reduced in a fixed order (deterministic).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "internal.h"
#include "transfer.h"

#define MF 8
#define MH 64
#define MROW 65
#define NPAR (MH * MF + MH + MH * MH + MH + MH + 1)
// parameter offsets
#define O_W1 0
#define O_B1 (MH * MF)
#define O_W2 (O_B1 + MH)
#define O_B2 (O_W2 + MH * MH)
#define O_W3 (O_B2 + MH)
#define O_B3 (O_W3 + MH)

struct tr_mlp {
    int device = 0;
    hipStream_t stream = nullptr;
    float *master = nullptr, *m1 = nullptr, *m2 = nullptr;  // FP32 parameters and Adam moments
    unsigned short *wbf = nullptr;                           // BF16 W1 | W2 | w3
    float *norm = nullptr;                                   // mu[8], sigma[8]
    float *part = nullptr;                                   // per-workgroup gradient partials
    float *grad = nullptr, *loss = nullptr;
    size_t part_cap = 0;
    int step = 0;
    std::vector<void *> owned;  // buffers of tr_mlp_alloc
};

namespace tr {

__device__ inline float bf2f(unsigned short h) { return __uint_as_float(((unsigned)h) << 16); }
__device__ inline unsigned short f2bf(float x) {
    unsigned u = __float_as_uint(x);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}
__device__ inline float sigm(float z) { return 1.0f / (1.0f + __expf(-z)); }

// BF16 copies of the weight matrices
__global__ void k_mlp_round(const float *__restrict__ p, unsigned short *__restrict__ w) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < MH * MF + MH * MH + MH; i += gridDim.x * blockDim.x) {
        const int src = i < MH * MF ? O_W1 + i : (i < MH * MF + MH * MH ? O_W2 + (i - MH * MF) : O_W3 + (i - MH * MF - MH * MH));
        w[i] = f2bf(p[src]);
    }
}

// sample index of slot i of a minibatch (all samples in order when step < 0)
__device__ inline uint32_t pick(uint32_t i, uint32_t n, long long step, unsigned long long seed) {
    if (step < 0) return i;
    unsigned long long z = seed + 0x9e3779b97f4a7c15ull * ((unsigned long long)step * 0x100000001ull + i + 1);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    z ^= z >> 31;
    return (uint32_t)(z % n);
}

// Forward for `cnt` samples; with t != null also the backward pass, each
// workgroup writing its gradient sums (NPAR floats) and loss sum to part.
template <bool GRAD>
__global__ __launch_bounds__(64) void k_mlp(const float *__restrict__ feat, const float *__restrict__ t,
                                            uint32_t n, uint32_t cnt, long long step, unsigned long long seed,
                                            const unsigned short *__restrict__ wbf, const float *__restrict__ p,
                                            const float *__restrict__ norm, float *__restrict__ y,
                                            float *__restrict__ part, float scale) {
    __shared__ unsigned short W[MH * MF + MH * MH + MH];
    __shared__ float XH[64][MF + 1];
    __shared__ float Z1[64][MROW], Z2[64][MROW], D1[GRAD ? 64 : 1][MROW], D2[GRAD ? 64 : 1][MROW];
    __shared__ float D3[64];
    const int s = threadIdx.x;
    for (int i = s; i < MH * MF + MH * MH + MH; i += 64) W[i] = wbf[i];
    const uint32_t i = blockIdx.x * 64u + (uint32_t)s;
    const bool live = i < cnt;
    const uint32_t k = live ? pick(i, n, step, seed) : 0;
    float xh[MF];
#pragma unroll
    for (int f = 0; f < MF; f++) {
        xh[f] = live ? (feat[(size_t)k * MF + f] - norm[f]) / norm[MF + f] : 0.0f;
        XH[s][f] = xh[f];
    }
    __syncthreads();
    const unsigned short *W1 = W, *W2 = W + MH * MF, *w3 = W + MH * MF + MH * MH;
    // h1 = SiLU(z1) once per unit, in registers (fully unrolled: static
    // indices), for the 64 x 64 layer
    float h1[MH];
#pragma unroll
    for (int a = 0; a < MH; a++) {
        float z = p[O_B1 + a];
#pragma unroll
        for (int f = 0; f < MF; f++) z = fmaf(bf2f(W1[a * MF + f]), xh[f], z);
        Z1[s][a] = z;
        h1[a] = z * sigm(z);
    }
    float z3 = p[O_B3];
    for (int a = 0; a < MH; a++) {
        float z = p[O_B2 + a];
#pragma unroll
        for (int b = 0; b < MH; b++) z = fmaf(bf2f(W2[a * MH + b]), h1[b], z);
        Z2[s][a] = z;
        z3 = fmaf(bf2f(w3[a]), z * sigm(z), z3);
    }
    const float yy = sigm(z3);
    if (!GRAD) {
        if (live) y[i] = yy;
        return;
    }
    const float e = live ? yy - t[k] : 0.0f;
    const float d3 = live ? 2.0f * e * scale * yy * (1.0f - yy) : 0.0f;
    D3[s] = d3;
    // delta of z2, then of z1; Z1/Z2 become the activations h1/h2
    for (int a = 0; a < MH; a++) {
        const float z = Z2[s][a], sg = sigm(z);
        D2[s][a] = d3 * bf2f(w3[a]) * sg * (1.0f + z * (1.0f - sg));
    }
    for (int b = 0; b < MH; b++) {
        float dh = 0;
        for (int a = 0; a < MH; a++) dh = fmaf(bf2f(W2[a * MH + b]), D2[s][a], dh);
        const float z = Z1[s][b], sg = sigm(z);
        D1[s][b] = dh * sg * (1.0f + z * (1.0f - sg));
    }
    for (int a = 0; a < MH; a++) {
        const float z1 = Z1[s][a], z2 = Z2[s][a];
        Z1[s][a] = z1 * sigm(z1);
        Z2[s][a] = z2 * sigm(z2);
    }
    __syncthreads();
    // workgroup sums over the 64 samples: lane s owns parameters s, s+64, ...
    float *out = part + (size_t)blockIdx.x * (NPAR + 1);
    for (int q = s; q < NPAR; q += 64) {
        float acc = 0;
        if (q < O_B1) {
            const int a = q / MF, f = q % MF;
            for (int r = 0; r < 64; r++) acc = fmaf(D1[r][a], XH[r][f], acc);
        } else if (q < O_W2) {
            const int a = q - O_B1;
            for (int r = 0; r < 64; r++) acc += D1[r][a];
        } else if (q < O_B2) {
            const int a = (q - O_W2) / MH, b = (q - O_W2) % MH;
            for (int r = 0; r < 64; r++) acc = fmaf(D2[r][a], Z1[r][b], acc);
        } else if (q < O_W3) {
            const int a = q - O_B2;
            for (int r = 0; r < 64; r++) acc += D2[r][a];
        } else if (q < O_B3) {
            const int a = q - O_W3;
            for (int r = 0; r < 64; r++) acc = fmaf(D3[r], Z2[r][a], acc);
        } else {
            for (int r = 0; r < 64; r++) acc += D3[r];
        }
        out[q] = acc;
    }
    // loss partial: sum of squared errors of this workgroup
    __shared__ float L[64];
    L[s] = e * e;
    __syncthreads();
    if (s == 0) {
        float l = 0;
        for (int r = 0; r < 64; r++) l += L[r];
        out[NPAR] = l;
    }
}

// grad[q] = sum over workgroups (fixed order)
__global__ void k_mlp_reduce(const float *__restrict__ part, uint32_t nblk, float *__restrict__ grad) {
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q <= NPAR; q += gridDim.x * blockDim.x) {
        float acc = 0;
        for (uint32_t b = 0; b < nblk; b++) acc += part[(size_t)b * (NPAR + 1) + q];
        grad[q] = acc;
    }
}

__global__ void k_mlp_adam(float *__restrict__ p, float *__restrict__ m1, float *__restrict__ m2,
                           const float *__restrict__ g, float lr, float c1, float c2) {
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < NPAR; q += gridDim.x * blockDim.x) {
        const float gq = g[q];
        m1[q] = b1 * m1[q] + (1 - b1) * gq;
        m2[q] = b2 * m2[q] + (1 - b2) * gq * gq;
        p[q] -= lr * (m1[q] * c1) / (sqrtf(m2[q] * c2) + eps);
    }
}

}  // namespace tr

using namespace tr;

#define MCHECK(x)                                                                        \
    do {                                                                                 \
        hipError_t e_ = (x);                                                             \
        if (e_ != hipSuccess) return fail(TR_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

static int mlp_parts(tr_mlp *m, size_t nblk) {
    if (nblk <= m->part_cap) return TR_OK;
    if (m->part) hipFree(m->part);
    m->part = nullptr;
    MCHECK(hipMalloc((void **)&m->part, sizeof(float) * nblk * (NPAR + 1)));
    m->part_cap = nblk;
    return TR_OK;
}

extern "C" {

int tr_mlp_create(int device, const float *params, uint64_t seed, tr_mlp **out) {
    if (!out) return fail(TR_ERR_ARG, "tr_mlp_create: null argument");
    *out = nullptr;
    std::vector<float> p(NPAR, 0.0f);
    if (params) {
        std::memcpy(p.data(), params, sizeof(float) * NPAR);
    } else {
        // scaled uniform initial weights from the SplitMix64 stream, zero biases
        for (int q = 0; q < MH * MF; q++) p[O_W1 + q] = (float)((2 * rng_uniform(seed, 1, q) - 1) * std::sqrt(3.0 / MF));
        for (int q = 0; q < MH * MH; q++) p[O_W2 + q] = (float)((2 * rng_uniform(seed, 2, q) - 1) * std::sqrt(3.0 / MH));
        for (int q = 0; q < MH; q++) p[O_W3 + q] = (float)((2 * rng_uniform(seed, 3, q) - 1) * std::sqrt(3.0 / MH));
    }
    tr_mlp *m = new tr_mlp();
    m->device = device;
    const float nrm[2 * MF] = {0, 0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 1, 1, 1, 1};
    hipError_t e = hipSetDevice(device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipMalloc((void **)&m->master, sizeof(float) * NPAR * 3);
    if (e == hipSuccess) e = hipMalloc((void **)&m->wbf, sizeof(unsigned short) * (MH * MF + MH * MH + MH));
    if (e == hipSuccess) e = hipMalloc((void **)&m->norm, sizeof(float) * 2 * MF);
    if (e == hipSuccess) e = hipMalloc((void **)&m->grad, sizeof(float) * (NPAR + 1));
    if (e == hipSuccess) {
        m->m1 = m->master + NPAR;
        m->m2 = m->master + 2 * NPAR;
        e = hipMemcpyAsync(m->master, p.data(), sizeof(float) * NPAR, hipMemcpyHostToDevice, m->stream);
    }
    if (e == hipSuccess) e = hipMemsetAsync(m->m1, 0, sizeof(float) * 2 * NPAR, m->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(m->norm, nrm, sizeof nrm, hipMemcpyHostToDevice, m->stream);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_mlp_round, dim3(16), dim3(256), 0, m->stream, m->master, m->wbf);
        e = hipStreamSynchronize(m->stream);
    }
    if (e != hipSuccess) {
        tr_mlp_destroy(m);
        return fail(TR_ERR_HIP, std::string("tr_mlp_create: ") + hipGetErrorString(e));
    }
    *out = m;
    return TR_OK;
}

int tr_mlp_destroy(tr_mlp *m) {
    if (!m) return TR_OK;
    if (m->stream) hipStreamSynchronize(m->stream);
    for (void *q : {(void *)m->master, (void *)m->wbf, (void *)m->norm, (void *)m->grad, (void *)m->part})
        if (q) hipFree(q);
    for (void *q : m->owned) hipFree(q);
    if (m->stream) hipStreamDestroy(m->stream);
    delete m;
    return TR_OK;
}

int tr_mlp_alloc(tr_mlp *m, size_t bytes, void **out) {
    if (!m || !out) return fail(TR_ERR_ARG, "tr_mlp_alloc: null argument");
    *out = nullptr;
    MCHECK(hipMalloc(out, std::max<size_t>(bytes, 1)));
    m->owned.push_back(*out);
    return TR_OK;
}

int tr_mlp_copy(tr_mlp *m, void *dst, const void *src, size_t bytes, int to_host) {
    if (!m || !dst || !src) return fail(TR_ERR_ARG, "tr_mlp_copy: null argument");
    MCHECK(hipMemcpyAsync(dst, src, bytes, to_host ? hipMemcpyDeviceToHost : hipMemcpyHostToDevice, m->stream));
    MCHECK(hipStreamSynchronize(m->stream));
    return TR_OK;
}

int tr_mlp_set_norm(tr_mlp *m, const float *mu, const float *sigma) {
    if (!m || !mu || !sigma) return fail(TR_ERR_ARG, "tr_mlp_set_norm: null argument");
    float h[2 * MF];
    for (int f = 0; f < MF; f++) {
        if (!(sigma[f] > 0)) return fail(TR_ERR_ARG, "tr_mlp_set_norm: sigma must be positive");
        h[f] = mu[f];
        h[MF + f] = sigma[f];
    }
    MCHECK(hipMemcpyAsync(m->norm, h, sizeof h, hipMemcpyHostToDevice, m->stream));
    MCHECK(hipStreamSynchronize(m->stream));
    return TR_OK;
}

int tr_mlp_get_params(tr_mlp *m, float *params) {
    if (!m || !params) return fail(TR_ERR_ARG, "tr_mlp_get_params: null argument");
    MCHECK(hipMemcpyAsync(params, m->master, sizeof(float) * NPAR, hipMemcpyDeviceToHost, m->stream));
    MCHECK(hipStreamSynchronize(m->stream));
    return TR_OK;
}

int tr_mlp_predict(tr_mlp *m, const float *feat, uint32_t n, float *y) {
    if (!m || !feat || !y) return fail(TR_ERR_ARG, "tr_mlp_predict: null argument");
    if (!n) return TR_OK;
    hipLaunchKernelGGL(k_mlp<false>, dim3((n + 63) / 64), dim3(64), 0, m->stream, feat, nullptr, n, n, -1LL, 0ULL,
                       m->wbf, m->master, m->norm, y, nullptr, 0.0f);
    MCHECK(hipGetLastError());
    MCHECK(hipStreamSynchronize(m->stream));
    return TR_OK;
}

int tr_mlp_grad(tr_mlp *m, const float *feat, const float *target, uint32_t n, float *loss, float *grad) {
    if (!m || !feat || !target || !n || !grad) return fail(TR_ERR_ARG, "tr_mlp_grad: bad argument");
    const uint32_t nblk = (n + 63) / 64;
    int rc = mlp_parts(m, nblk);
    if (rc) return rc;
    hipLaunchKernelGGL(k_mlp<true>, dim3(nblk), dim3(64), 0, m->stream, feat, target, n, n, -1LL, 0ULL, m->wbf,
                       m->master, m->norm, nullptr, m->part, 1.0f / n);
    hipLaunchKernelGGL(k_mlp_reduce, dim3((NPAR + 256) / 256), dim3(256), 0, m->stream, m->part, nblk, m->grad);
    MCHECK(hipGetLastError());
    std::vector<float> g(NPAR + 1);
    MCHECK(hipMemcpyAsync(g.data(), m->grad, sizeof(float) * (NPAR + 1), hipMemcpyDeviceToHost, m->stream));
    MCHECK(hipStreamSynchronize(m->stream));
    std::memcpy(grad, g.data(), sizeof(float) * NPAR);
    if (loss) *loss = g[NPAR] / n;
    return TR_OK;
}

// Adam on minibatches of `batch` samples drawn with replacement (a
// counter-based hash of seed, step, slot) from the n samples; loss_hist[k]
// = mean loss of step k's minibatch (may be NULL).
int tr_mlp_train(tr_mlp *m, const float *feat, const float *target, uint32_t n, int steps, uint32_t batch, float lr,
                 uint64_t seed, float *loss_hist) {
    if (!m || !feat || !target || !n || steps < 0 || !batch) return fail(TR_ERR_ARG, "tr_mlp_train: bad argument");
    const uint32_t nblk = (batch + 63) / 64;
    int rc = mlp_parts(m, nblk);
    if (rc) return rc;
    std::vector<float> lh(steps > 0 ? steps : 1, 0.0f);
    for (int k = 0; k < steps; k++) {
        m->step++;
        hipLaunchKernelGGL(k_mlp<true>, dim3(nblk), dim3(64), 0, m->stream, feat, target, n, batch, (long long)m->step,
                           (unsigned long long)seed, m->wbf, m->master, m->norm, nullptr, m->part, 1.0f / batch);
        hipLaunchKernelGGL(k_mlp_reduce, dim3((NPAR + 256) / 256), dim3(256), 0, m->stream, m->part, nblk, m->grad);
        const float c1 = 1.0f / (1.0f - std::pow(0.9f, (float)m->step)), c2 = 1.0f / (1.0f - std::pow(0.999f, (float)m->step));
        hipLaunchKernelGGL(k_mlp_adam, dim3((NPAR + 255) / 256), dim3(256), 0, m->stream, m->master, m->m1, m->m2,
                           m->grad, lr, c1, c2);
        hipLaunchKernelGGL(k_mlp_round, dim3(16), dim3(256), 0, m->stream, m->master, m->wbf);
        if (loss_hist) MCHECK(hipMemcpyAsync(&lh[k], m->grad + NPAR, sizeof(float), hipMemcpyDeviceToHost, m->stream));
    }
    MCHECK(hipGetLastError());
    MCHECK(hipStreamSynchronize(m->stream));
    if (loss_hist)
        for (int k = 0; k < steps; k++) loss_hist[k] = lh[k] / batch;
    return TR_OK;
}

}  // extern "C"
