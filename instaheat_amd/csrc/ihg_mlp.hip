// ihg_mlp.hip -- the learned DOPRI853 step proposer: a 20-64-64-1 tanh MLP
// in FP32 with hand-written inference, gradient and Adam kernels.
//
// It has no counterpart in the reference (whose controller is success(),
// src/dopri853.c:505-540); the host stepper consults it before success() on
// accepted steps, and the embedded error estimate still judges every step.
//
// Parameters, one flat float array:
//   W1[64][20] | b1[64] | W2[64][64] | b2[64] | W3[64] | b3
// Model: h1 = tanh(W1 x + b1), h2 = tanh(W2 h1 + b2), y = W3 . h2 + b3.
// Loss of a batch of n samples: mean of (y - t)^2.
//
// The gradient is deterministic: a work-group of 64 samples keeps every
// sample's activations and back-propagated errors in LDS and sums each
// parameter's contributions over its samples in a fixed order; the
// per-group partials are summed in a fixed order by a second kernel.
#include <cmath>
#include <cstdint>
#include <vector>

#include "ihg_internal.h"

namespace ihg {

constexpr int MI = IHG_MLP_IN, MH = IHG_MLP_HIDDEN;
constexpr int OFF_W1 = 0, OFF_B1 = MH * MI, OFF_W2 = OFF_B1 + MH,
              OFF_B2 = OFF_W2 + MH * MH, OFF_W3 = OFF_B2 + MH,
              OFF_B3 = OFF_W3 + MH, NPARAM = OFF_B3 + 1;
constexpr int MLP_WG = 64; // samples per work-group in the gradient kernel

__device__ __forceinline__ void load_params(float *sp, const float *p)
{
    for (int i = threadIdx.x; i < NPARAM; i += blockDim.x)
        sp[i] = p[i];
    __syncthreads();
}

// forward pass of one sample, keeping both hidden layers
__device__ __forceinline__ float mlp_sample(const float *sp, const float *x,
                                            float *h1, float *h2)
{
    for (int m = 0; m < MH; ++m) {
        float s = sp[OFF_B1 + m];
        for (int i = 0; i < MI; ++i)
            s += sp[OFF_W1 + m * MI + i] * x[i];
        h1[m] = tanhf(s);
    }
    float y = sp[OFF_B3];
    for (int k = 0; k < MH; ++k) {
        float s = sp[OFF_B2 + k];
        for (int m = 0; m < MH; ++m)
            s += sp[OFF_W2 + k * MH + m] * h1[m];
        h2[k] = tanhf(s);
        y += sp[OFF_W3 + k] * h2[k];
    }
    return y;
}

// inference: y[s] for x[s][20], s < n
__global__ __launch_bounds__(MLP_WG) void k_mlp_forward(
    const float *__restrict__ p, const float *__restrict__ x,
    float *__restrict__ y, int n)
{
    __shared__ float sp[NPARAM];
    load_params(sp, p);
    const int s = blockIdx.x * MLP_WG + threadIdx.x;
    if (s >= n) return;
    float xs[MI], h1[MH], h2[MH];
    for (int i = 0; i < MI; ++i) xs[i] = x[(long)s * MI + i];
    y[s] = mlp_sample(sp, xs, h1, h2);
}

// per-work-group partial gradients of sum_s (y_s - t_s)^2 over the samples
// idx[g * 64 + 0 .. 63] (g = blockIdx.x) of the first n entries of idx;
// partial[g][NPARAM], ploss[g]
__global__ __launch_bounds__(MLP_WG) void k_mlp_grad(
    const float *__restrict__ p, const float *__restrict__ x,
    const float *__restrict__ t, const int *__restrict__ idx, int n,
    float *__restrict__ partial, float *__restrict__ ploss)
{
    __shared__ float sp[NPARAM];
    __shared__ float sx[MLP_WG][MI + 1];
    __shared__ float sh1[MLP_WG][MH + 1], sh2[MLP_WG][MH + 1];
    __shared__ float sd1[MLP_WG][MH + 1], sd2[MLP_WG][MH + 1];
    __shared__ float sdy[MLP_WG], se2[MLP_WG];
    load_params(sp, p);
    const int tid = threadIdx.x;
    const int k = blockIdx.x * MLP_WG + tid;
    const bool act = k < n;
    {
        float xs[MI], h1[MH], h2[MH];
        const long s = act ? idx[k] : 0;
        for (int i = 0; i < MI; ++i) xs[i] = x[s * MI + i];
        const float y = mlp_sample(sp, xs, h1, h2);
        const float e = act ? y - t[s] : 0.0f;
        const float dy = 2.0f * e;
        sdy[tid] = dy;
        se2[tid] = e * e;
        for (int i = 0; i < MI; ++i) sx[tid][i] = xs[i];
        float d2[MH];
        for (int kk = 0; kk < MH; ++kk) {
            d2[kk] = dy * sp[OFF_W3 + kk] * (1.0f - h2[kk] * h2[kk]);
            sh2[tid][kk] = h2[kk];
            sd2[tid][kk] = d2[kk];
        }
        for (int m = 0; m < MH; ++m) {
            float s1 = 0.0f;
            for (int kk = 0; kk < MH; ++kk)
                s1 += sp[OFF_W2 + kk * MH + m] * d2[kk];
            sh1[tid][m] = h1[m];
            sd1[tid][m] = s1 * (1.0f - h1[m] * h1[m]);
        }
    }
    __syncthreads();
    float *out = partial + (long)blockIdx.x * NPARAM;
    for (int q = tid; q < NPARAM; q += MLP_WG) {
        float g = 0.0f;
        if (q < OFF_B1) {
            const int m = q / MI, i = q - m * MI;
            for (int s = 0; s < MLP_WG; ++s) g += sd1[s][m] * sx[s][i];
        } else if (q < OFF_W2) {
            const int m = q - OFF_B1;
            for (int s = 0; s < MLP_WG; ++s) g += sd1[s][m];
        } else if (q < OFF_B2) {
            const int kk = (q - OFF_W2) / MH, m = (q - OFF_W2) - kk * MH;
            for (int s = 0; s < MLP_WG; ++s) g += sd2[s][kk] * sh1[s][m];
        } else if (q < OFF_W3) {
            const int kk = q - OFF_B2;
            for (int s = 0; s < MLP_WG; ++s) g += sd2[s][kk];
        } else if (q < OFF_B3) {
            const int kk = q - OFF_W3;
            for (int s = 0; s < MLP_WG; ++s) g += sdy[s] * sh2[s][kk];
        } else {
            for (int s = 0; s < MLP_WG; ++s) g += sdy[s];
        }
        out[q] = g;
    }
    if (tid == 0) {
        float l = 0.0f;
        for (int s = 0; s < MLP_WG; ++s) l += se2[s];
        ploss[blockIdx.x] = l;
    }
}

// grad[q] = sum_g partial[g][q] / n (fixed order); loss_acc += sum ploss / n
__global__ __launch_bounds__(256) void k_mlp_reduce(
    const float *__restrict__ partial, const float *__restrict__ ploss,
    int ngroups, int n, float *__restrict__ grad, float *__restrict__ loss)
{
    const int q = blockIdx.x * 256 + threadIdx.x;
    if (q < NPARAM) {
        float g = 0.0f;
        for (int b = 0; b < ngroups; ++b) g += partial[(long)b * NPARAM + q];
        grad[q] = g / (float)n;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        float l = 0.0f;
        for (int b = 0; b < ngroups; ++b) l += ploss[b];
        loss[0] = l / (float)n;
        loss[1] += l; // running sum over an epoch
    }
}

// Adam (Kingma & Ba 2015) with bias correction; step >= 1
__global__ __launch_bounds__(256) void k_mlp_adam(
    float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m1,
    float *__restrict__ m2, float lr, float b1, float b2, float eps, float c1,
    float c2)
{
    const int q = blockIdx.x * 256 + threadIdx.x;
    if (q >= NPARAM) return;
    const float gq = g[q];
    const float a = b1 * m1[q] + (1.0f - b1) * gq;
    const float v = b2 * m2[q] + (1.0f - b2) * gq * gq;
    m1[q] = a;
    m2[q] = v;
    p[q] -= lr * (a / c1) / (sqrtf(v / c2) + eps);
}

} // namespace ihg

using namespace ihg;

struct ihg_mlp {
    hipStream_t stream = nullptr;
    float *p = nullptr, *g = nullptr, *m1 = nullptr, *m2 = nullptr;
    float *loss = nullptr; // [batch loss, epoch sum]
    float *x = nullptr, *t = nullptr, *y = nullptr, *partial = nullptr,
          *ploss = nullptr;
    int *idx = nullptr;
    long cap_n = 0, cap_groups = 0;
    long adam_step = 0;
};

namespace {

int mlp_reserve(ihg_mlp *m, long n)
{
    if (n <= m->cap_n) return 0;
    (void)hipFree(m->x);
    (void)hipFree(m->t);
    (void)hipFree(m->y);
    (void)hipFree(m->idx);
    (void)hipFree(m->partial);
    (void)hipFree(m->ploss);
    m->x = m->t = m->y = m->partial = m->ploss = nullptr;
    m->idx = nullptr;
    m->cap_n = m->cap_groups = 0;
    const long groups = (n + MLP_WG - 1) / MLP_WG;
    if (hipMalloc(&m->x, n * MI * sizeof(float)) != hipSuccess ||
        hipMalloc(&m->t, n * sizeof(float)) != hipSuccess ||
        hipMalloc(&m->y, n * sizeof(float)) != hipSuccess ||
        hipMalloc(&m->idx, n * sizeof(int)) != hipSuccess ||
        hipMalloc(&m->partial, groups * NPARAM * sizeof(float)) != hipSuccess ||
        hipMalloc(&m->ploss, groups * sizeof(float)) != hipSuccess) {
        (void)hipGetLastError();
        return set_error("ihg_mlp: device allocation of %ld samples failed", n);
    }
    m->cap_n = n;
    m->cap_groups = groups;
    return 0;
}

// splitmix64 (Steele, Lea & Flood 2014): initialisation and shuffles
uint64_t splitmix(uint64_t &s)
{
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
double uniform01(uint64_t &s) { return (splitmix(s) >> 11) * 0x1.0p-53; }

int grad_launch(ihg_mlp *m, int n)
{
    const int groups = (n + MLP_WG - 1) / MLP_WG;
    hipLaunchKernelGGL(k_mlp_grad, dim3(groups), dim3(MLP_WG), 0, m->stream,
                       m->p, m->x, m->t, m->idx, n, m->partial, m->ploss);
    IHG_TRY(hip_check(hipGetLastError(), "k_mlp_grad"));
    hipLaunchKernelGGL(k_mlp_reduce, dim3((NPARAM + 255) / 256), dim3(256), 0,
                       m->stream, m->partial, m->ploss, groups, n, m->g,
                       m->loss);
    return hip_check(hipGetLastError(), "k_mlp_reduce");
}

} // namespace

extern "C" int ihg_mlp_nparams(void) { return NPARAM; }

extern "C" int ihg_mlp_create(ihg_mlp **out, const float *params,
                              unsigned long seed)
{
    if (!out) return set_error("NULL argument");
    *out = nullptr;
    ihg_mlp *m = new ihg_mlp();
    std::vector<float> h(NPARAM, 0.0f);
    if (params) {
        for (int i = 0; i < NPARAM; ++i) h[i] = params[i];
    } else {
        // Glorot-uniform weights, zero biases
        uint64_t s = seed;
        auto fill = [&](int off, int fan_in, int fan_out, int cnt) {
            const double r = std::sqrt(6.0 / (fan_in + fan_out));
            for (int i = 0; i < cnt; ++i)
                h[off + i] = (float)((2.0 * uniform01(s) - 1.0) * r);
        };
        fill(OFF_W1, MI, MH, MH * MI);
        fill(OFF_W2, MH, MH, MH * MH);
        fill(OFF_W3, MH, 1, MH);
    }
    int rc = 0;
    if (hipStreamCreate(&m->stream) != hipSuccess ||
        hipMalloc(&m->p, NPARAM * sizeof(float)) != hipSuccess ||
        hipMalloc(&m->g, NPARAM * sizeof(float)) != hipSuccess ||
        hipMalloc(&m->m1, NPARAM * sizeof(float)) != hipSuccess ||
        hipMalloc(&m->m2, NPARAM * sizeof(float)) != hipSuccess ||
        hipMalloc(&m->loss, 2 * sizeof(float)) != hipSuccess) {
        (void)hipGetLastError();
        rc = set_error("ihg_mlp_create: device allocation failed");
    }
    if (!rc) rc = hip_check(hipMemcpy(m->p, h.data(), NPARAM * sizeof(float),
                                      hipMemcpyHostToDevice), "upload params");
    if (!rc) rc = hip_check(hipMemset(m->m1, 0, NPARAM * sizeof(float)), "memset");
    if (!rc) rc = hip_check(hipMemset(m->m2, 0, NPARAM * sizeof(float)), "memset");
    if (!rc) rc = hip_check(hipMemset(m->loss, 0, 2 * sizeof(float)), "memset");
    if (rc) {
        ihg_mlp_destroy(m);
        return rc;
    }
    *out = m;
    return 0;
}

extern "C" int ihg_mlp_destroy(ihg_mlp *m)
{
    if (!m) return 0;
    if (m->stream) (void)hipStreamSynchronize(m->stream);
    float *bufs[] = {m->p, m->g, m->m1, m->m2, m->loss, m->x,
                     m->t, m->y, m->partial, m->ploss};
    for (float *b : bufs) (void)hipFree(b);
    (void)hipFree(m->idx);
    if (m->stream) (void)hipStreamDestroy(m->stream);
    delete m;
    return 0;
}

extern "C" int ihg_mlp_get_params(ihg_mlp *m, float *params)
{
    if (!m || !params) return set_error("NULL argument");
    IHG_TRY(hip_check(hipMemcpyAsync(params, m->p, NPARAM * sizeof(float),
                                     hipMemcpyDeviceToHost, m->stream),
                      "download params"));
    return hip_check(hipStreamSynchronize(m->stream), "sync");
}

extern "C" int ihg_mlp_set_params(ihg_mlp *m, const float *params)
{
    if (!m || !params) return set_error("NULL argument");
    IHG_TRY(hip_check(hipMemcpyAsync(m->p, params, NPARAM * sizeof(float),
                                     hipMemcpyHostToDevice, m->stream),
                      "upload params"));
    return hip_check(hipStreamSynchronize(m->stream), "sync");
}

extern "C" int ihg_mlp_forward(ihg_mlp *m, const float *x, float *y, int n)
{
    if (!m || !x || !y) return set_error("NULL argument");
    if (n <= 0) return set_error("ihg_mlp_forward: n = %d", n);
    IHG_TRY(mlp_reserve(m, n));
    IHG_TRY(hip_check(hipMemcpyAsync(m->x, x, (size_t)n * MI * sizeof(float),
                                     hipMemcpyHostToDevice, m->stream), "upload x"));
    hipLaunchKernelGGL(k_mlp_forward, dim3((n + MLP_WG - 1) / MLP_WG),
                       dim3(MLP_WG), 0, m->stream, m->p, m->x, m->y, n);
    IHG_TRY(hip_check(hipGetLastError(), "k_mlp_forward"));
    IHG_TRY(hip_check(hipMemcpyAsync(y, m->y, (size_t)n * sizeof(float),
                                     hipMemcpyDeviceToHost, m->stream), "download y"));
    return hip_check(hipStreamSynchronize(m->stream), "sync");
}

extern "C" int ihg_mlp_grad(ihg_mlp *m, const float *x, const float *t, int n,
                            float *grad, float *loss)
{
    if (!m || !x || !t || !grad) return set_error("NULL argument");
    if (n <= 0) return set_error("ihg_mlp_grad: n = %d", n);
    IHG_TRY(mlp_reserve(m, n));
    std::vector<int> id(n);
    for (int i = 0; i < n; ++i) id[i] = i;
    IHG_TRY(hip_check(hipMemcpyAsync(m->x, x, (size_t)n * MI * sizeof(float),
                                     hipMemcpyHostToDevice, m->stream), "upload x"));
    IHG_TRY(hip_check(hipMemcpyAsync(m->t, t, (size_t)n * sizeof(float),
                                     hipMemcpyHostToDevice, m->stream), "upload t"));
    IHG_TRY(hip_check(hipMemcpyAsync(m->idx, id.data(), (size_t)n * sizeof(int),
                                     hipMemcpyHostToDevice, m->stream), "upload idx"));
    IHG_TRY(grad_launch(m, n));
    float l[2];
    IHG_TRY(hip_check(hipMemcpyAsync(grad, m->g, NPARAM * sizeof(float),
                                     hipMemcpyDeviceToHost, m->stream), "download grad"));
    IHG_TRY(hip_check(hipMemcpyAsync(l, m->loss, 2 * sizeof(float),
                                     hipMemcpyDeviceToHost, m->stream), "download loss"));
    IHG_TRY(hip_check(hipStreamSynchronize(m->stream), "sync"));
    if (loss) *loss = l[0];
    return 0;
}

extern "C" int ihg_mlp_train(ihg_mlp *m, const float *x, const float *t, int n,
                             int epochs, int batch, float lr,
                             unsigned long seed, float *epoch_loss)
{
    if (!m || !x || !t) return set_error("NULL argument");
    if (n <= 0 || epochs < 0 || batch <= 0)
        return set_error("ihg_mlp_train: n %d epochs %d batch %d", n, epochs,
                         batch);
    IHG_TRY(mlp_reserve(m, n));
    IHG_TRY(hip_check(hipMemcpyAsync(m->x, x, (size_t)n * MI * sizeof(float),
                                     hipMemcpyHostToDevice, m->stream), "upload x"));
    IHG_TRY(hip_check(hipMemcpyAsync(m->t, t, (size_t)n * sizeof(float),
                                     hipMemcpyHostToDevice, m->stream), "upload t"));
    std::vector<int> perm(n);
    uint64_t s = seed;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    for (int e = 0; e < epochs; ++e) {
        // Fisher-Yates shuffle of the sample order
        for (int i = 0; i < n; ++i) perm[i] = i;
        for (int i = n - 1; i > 0; --i) {
            const int j = (int)(splitmix(s) % (uint64_t)(i + 1));
            const int tmp = perm[i];
            perm[i] = perm[j];
            perm[j] = tmp;
        }
        IHG_TRY(hip_check(hipMemcpyAsync(m->idx, perm.data(),
                                         (size_t)n * sizeof(int),
                                         hipMemcpyHostToDevice, m->stream),
                          "upload idx"));
        IHG_TRY(hip_check(hipMemsetAsync(m->loss, 0, 2 * sizeof(float),
                                         m->stream), "memset"));
        for (int off = 0; off < n; off += batch) {
            const int nb = (n - off < batch) ? n - off : batch;
            const int groups = (nb + MLP_WG - 1) / MLP_WG;
            hipLaunchKernelGGL(k_mlp_grad, dim3(groups), dim3(MLP_WG), 0,
                               m->stream, m->p, m->x, m->t, m->idx + off, nb,
                               m->partial, m->ploss);
            IHG_TRY(hip_check(hipGetLastError(), "k_mlp_grad"));
            hipLaunchKernelGGL(k_mlp_reduce, dim3((NPARAM + 255) / 256),
                               dim3(256), 0, m->stream, m->partial, m->ploss,
                               groups, nb, m->g, m->loss);
            IHG_TRY(hip_check(hipGetLastError(), "k_mlp_reduce"));
            m->adam_step += 1;
            const float c1 = 1.0f - std::pow(b1, (float)m->adam_step);
            const float c2 = 1.0f - std::pow(b2, (float)m->adam_step);
            hipLaunchKernelGGL(k_mlp_adam, dim3((NPARAM + 255) / 256),
                               dim3(256), 0, m->stream, m->p, m->g, m->m1,
                               m->m2, lr, b1, b2, eps, c1, c2);
            IHG_TRY(hip_check(hipGetLastError(), "k_mlp_adam"));
        }
        if (epoch_loss) {
            float l[2];
            IHG_TRY(hip_check(hipMemcpyAsync(l, m->loss, 2 * sizeof(float),
                                             hipMemcpyDeviceToHost, m->stream),
                              "download loss"));
            IHG_TRY(hip_check(hipStreamSynchronize(m->stream), "sync"));
            epoch_loss[e] = l[1] / (float)n;
        }
    }
    return hip_check(hipStreamSynchronize(m->stream), "sync");
}
