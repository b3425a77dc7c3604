// Learned troubled-cell indicator: a 125-64-64-1 MLP (GELU hidden layers, 0.5*sigmoid output)
// trained on the GPU by hand-written forward / backward / Adam kernels. Weights are held in FP32
// (master copy, Adam moments) and rounded to BF16 for every forward pass; all sums are FP32.
// Inference for the solver runs in k_mlp (plens_gpu.hip, on the matrix cores) on the BF16-rounded
// copy exported by plens_mlp_device_bf16_params().
//
// There is no counterpart in the reference: it replaces BlenderCalculator::get_blender
// (dgsem/blender_calculator.cc:123-187) as the source of the cell blender value, with the plain
// value kept as fallback.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <new>
#include <random>
#include <vector>

#include "plens_mlp.h"
#include <algorithm>

namespace {

constexpr int IN = 125, H1 = 64, H2 = 64;
constexpr int OFF_W1 = 0, OFF_B1 = OFF_W1 + H1 * IN, OFF_W2 = OFF_B1 + H1, OFF_B2 = OFF_W2 + H2 * H1,
              OFF_W3 = OFF_B2 + H2, OFF_B3 = OFF_W3 + H2, NPARAM = OFF_B3 + 1;

__device__ __forceinline__ float bf16_round(float x) {
    unsigned int b = __float_as_uint(x);
    b = (b + 0x7FFFu + ((b >> 16) & 1u)) & 0xFFFF0000u;
    return __uint_as_float(b);
}
__device__ __forceinline__ float gelu(float x) {
    const float t = tanhf(0.7978845608028654f * (x + 0.044715f * x * x * x));
    return 0.5f * x * (1.0f + t);
}
__device__ __forceinline__ float gelu_grad(float x) {
    const float t = tanhf(0.7978845608028654f * (x + 0.044715f * x * x * x));
    const float du = 0.7978845608028654f * (1.0f + 3.0f * 0.044715f * x * x);
    return 0.5f * (1.0f + t) + 0.5f * x * (1.0f - t * t) * du;
}

// BF16-rounded weights (kept as FP32 values) from the FP32 master parameters
__global__ void k_round(const float* __restrict__ P, float* __restrict__ Pb) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < NPARAM) {
        const bool weight = (i < OFF_B1) || (i >= OFF_W2 && i < OFF_B2) || (i >= OFF_W3 && i < OFF_B3);
        Pb[i] = weight ? bf16_round(P[i]) : P[i];
    }
}

// layer 1: z1[s][h] = b1[h] + sum_k X[idx[s]][k] W1[h][k]; one thread per (sample, unit)
__global__ void k_l1(const float* __restrict__ X, const int* __restrict__ idx, int n, const float* __restrict__ Pb,
                     float* __restrict__ z1, float* __restrict__ a1) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * H1) return;
    const int s = t / H1, h = t - s * H1;
    const float* x = X + (long long)(idx ? idx[s] : s) * IN;
    const float* w = Pb + OFF_W1 + h * IN;
    float acc = Pb[OFF_B1 + h];
    for (int k = 0; k < IN; ++k) acc = fmaf(x[k], w[k], acc);
    z1[t] = acc;
    a1[t] = gelu(acc);
}
__global__ void k_l2(const float* __restrict__ a1, int n, const float* __restrict__ Pb, float* __restrict__ z2,
                     float* __restrict__ a2) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * H2) return;
    const int s = t / H2, h = t - s * H2;
    const float* x = a1 + s * H1;
    const float* w = Pb + OFF_W2 + h * H1;
    float acc = Pb[OFF_B2 + h];
    for (int k = 0; k < H1; ++k) acc = fmaf(x[k], w[k], acc);
    z2[t] = acc;
    a2[t] = gelu(acc);
}
__global__ void k_l3(const float* __restrict__ a2, int n, const float* __restrict__ Pb, float* __restrict__ y) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const float* x = a2 + s * H2;
    float acc = Pb[OFF_B3];
    for (int k = 0; k < H2; ++k) acc = fmaf(x[k], Pb[OFF_W3 + k], acc);
    y[s] = 0.5f / (1.0f + expf(-acc));
}
// output residual, dL/dz3 and the per-sample squared error (MSE loss)
__global__ void k_out_grad(const float* __restrict__ y, const float* __restrict__ tgt, const int* __restrict__ idx,
                           int n, float* __restrict__ dz3, float* __restrict__ sqerr) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const float r = y[s] - tgt[idx ? idx[s] : s];
    sqerr[s] = r * r;
    dz3[s] = (2.0f * r / (float)n) * y[s] * (1.0f - 2.0f * y[s]);
}
__global__ void k_back2(const float* __restrict__ dz3, const float* __restrict__ z2, int n, const float* __restrict__ Pb,
                        float* __restrict__ dz2) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * H2) return;
    const int s = t / H2, h = t - s * H2;
    dz2[t] = dz3[s] * Pb[OFF_W3 + h] * gelu_grad(z2[t]);
}
__global__ void k_back1(const float* __restrict__ dz2, const float* __restrict__ z1, int n, const float* __restrict__ Pb,
                        float* __restrict__ dz1) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * H1) return;
    const int s = t / H1, h = t - s * H1;
    float acc = 0.0f;
    for (int j = 0; j < H2; ++j) acc = fmaf(dz2[s * H2 + j], Pb[OFF_W2 + j * H1 + h], acc);
    dz1[t] = acc * gelu_grad(z1[t]);
}
// parameter gradients: one thread per parameter, sums over the batch in sample order
__global__ void k_wgrad(const float* __restrict__ X, const int* __restrict__ idx, const float* __restrict__ a1,
                        const float* __restrict__ a2, const float* __restrict__ dz1, const float* __restrict__ dz2,
                        const float* __restrict__ dz3, int n, float* __restrict__ G) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= NPARAM) return;
    float acc = 0.0f;
    if (i < OFF_B1) {
        const int h = i / IN, k = i - h * IN;
        for (int s = 0; s < n; ++s) acc = fmaf(dz1[s * H1 + h], X[(long long)(idx ? idx[s] : s) * IN + k], acc);
    } else if (i < OFF_W2) {
        const int h = i - OFF_B1;
        for (int s = 0; s < n; ++s) acc += dz1[s * H1 + h];
    } else if (i < OFF_B2) {
        const int h = (i - OFF_W2) / H1, k = (i - OFF_W2) - h * H1;
        for (int s = 0; s < n; ++s) acc = fmaf(dz2[s * H2 + h], a1[s * H1 + k], acc);
    } else if (i < OFF_W3) {
        const int h = i - OFF_B2;
        for (int s = 0; s < n; ++s) acc += dz2[s * H2 + h];
    } else if (i < OFF_B3) {
        const int h = i - OFF_W3;
        for (int s = 0; s < n; ++s) acc = fmaf(dz3[s], a2[s * H2 + h], acc);
    } else {
        for (int s = 0; s < n; ++s) acc += dz3[s];
    }
    G[i] = acc;
}
__global__ void k_adam(float* __restrict__ P, float* __restrict__ M, float* __restrict__ V, const float* __restrict__ G,
                       float lr, float c1, float c2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= NPARAM) return;
    const float g = G[i];
    const float m = 0.9f * M[i] + 0.1f * g;
    const float v = 0.999f * V[i] + 0.001f * g * g;
    M[i] = m;
    V[i] = v;
    P[i] = P[i] - lr * (m / c1) / (sqrtf(v / c2) + 1e-8f);
}
__global__ void k_sum(const float* __restrict__ x, int n, float* __restrict__ out) {
    __shared__ float red[256];
    float a = 0.0f;
    for (int i = threadIdx.x; i < n; i += blockDim.x) a += x[i];
    red[threadIdx.x] = a;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = red[0];
}

inline unsigned blocks(long long n, int t) { return (unsigned)((n + t - 1) / t); }

}  // namespace

struct plens_mlp {
    float *P = nullptr, *Pb = nullptr, *M = nullptr, *V = nullptr, *G = nullptr;
    float *X = nullptr, *Y = nullptr;  // training data on the device
    int *idx = nullptr;
    long long n_data = 0;
    int cap = 0;  // batch capacity of the work arrays
    float *z1 = nullptr, *a1 = nullptr, *z2 = nullptr, *a2 = nullptr, *y = nullptr, *dz1 = nullptr, *dz2 = nullptr,
          *dz3 = nullptr, *sq = nullptr, *sum = nullptr;
    int adam_t = 0;
    hipStream_t stream = nullptr;
};

namespace {

void free_work(plens_mlp* m) {
    float* f[] = {m->z1, m->a1, m->z2, m->a2, m->y, m->dz1, m->dz2, m->dz3, m->sq};
    for (float* p : f)
        if (p) (void)hipFree(p);
    m->z1 = m->a1 = m->z2 = m->a2 = m->y = m->dz1 = m->dz2 = m->dz3 = m->sq = nullptr;
    m->cap = 0;
}

int ensure_work(plens_mlp* m, int n) {
    if (n <= m->cap) return PLENS_OK;
    free_work(m);
    const size_t h = (size_t)n * H1 * sizeof(float), s = (size_t)n * sizeof(float);
    if (hipMalloc((void**)&m->z1, h) || hipMalloc((void**)&m->a1, h) || hipMalloc((void**)&m->z2, h) ||
        hipMalloc((void**)&m->a2, h) || hipMalloc((void**)&m->dz1, h) || hipMalloc((void**)&m->dz2, h) ||
        hipMalloc((void**)&m->y, s) || hipMalloc((void**)&m->dz3, s) || hipMalloc((void**)&m->sq, s))
        return PLENS_E_NOMEM;
    m->cap = n;
    return PLENS_OK;
}

void launch_forward(plens_mlp* m, const float* X, const int* idx, int n) {
    const int T = 256;
    hipLaunchKernelGGL(k_l1, dim3(blocks((long long)n * H1, T)), dim3(T), 0, m->stream, X, idx, n, m->Pb, m->z1, m->a1);
    hipLaunchKernelGGL(k_l2, dim3(blocks((long long)n * H2, T)), dim3(T), 0, m->stream, m->a1, n, m->Pb, m->z2, m->a2);
    hipLaunchKernelGGL(k_l3, dim3(blocks(n, T)), dim3(T), 0, m->stream, m->a2, n, m->Pb, m->y);
}

void launch_backward(plens_mlp* m, const float* X, const float* tgt, const int* idx, int n) {
    const int T = 256;
    hipLaunchKernelGGL(k_out_grad, dim3(blocks(n, T)), dim3(T), 0, m->stream, m->y, tgt, idx, n, m->dz3, m->sq);
    hipLaunchKernelGGL(k_back2, dim3(blocks((long long)n * H2, T)), dim3(T), 0, m->stream, m->dz3, m->z2, n, m->Pb,
                       m->dz2);
    hipLaunchKernelGGL(k_back1, dim3(blocks((long long)n * H1, T)), dim3(T), 0, m->stream, m->dz2, m->z1, n, m->Pb,
                       m->dz1);
    hipLaunchKernelGGL(k_wgrad, dim3(blocks(NPARAM, T)), dim3(T), 0, m->stream, X, idx, m->a1, m->a2, m->dz1, m->dz2,
                       m->dz3, n, m->G);
    hipLaunchKernelGGL(k_sum, dim3(1), dim3(256), 0, m->stream, m->sq, n, m->sum);
}

void launch_adam(plens_mlp* m, float lr) {
    ++m->adam_t;
    const float c1 = 1.0f - powf(0.9f, (float)m->adam_t), c2 = 1.0f - powf(0.999f, (float)m->adam_t);
    hipLaunchKernelGGL(k_adam, dim3(blocks(NPARAM, 256)), dim3(256), 0, m->stream, m->P, m->M, m->V, m->G, lr, c1, c2);
    hipLaunchKernelGGL(k_round, dim3(blocks(NPARAM, 256)), dim3(256), 0, m->stream, m->P, m->Pb);
}

}  // namespace

extern "C" {

int32_t plens_mlp_n_params(void) { return NPARAM; }

int plens_mlp_create(const float* init_params, plens_mlp** out) {
    if (!init_params || !out) return PLENS_E_INVALID;
    int nd = 0;
    if (hipGetDeviceCount(&nd) != hipSuccess || nd < 1) return PLENS_E_NODEVICE;
    plens_mlp* m = new (std::nothrow) plens_mlp;
    if (!m) return PLENS_E_NOMEM;
    const size_t b = NPARAM * sizeof(float);
    if (hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) || hipMalloc((void**)&m->P, b) ||
        hipMalloc((void**)&m->Pb, b) || hipMalloc((void**)&m->M, b) || hipMalloc((void**)&m->V, b) ||
        hipMalloc((void**)&m->G, b) || hipMalloc((void**)&m->sum, sizeof(float))) {
        plens_mlp_destroy(m);
        return PLENS_E_NOMEM;
    }
    if (hipMemcpy(m->P, init_params, b, hipMemcpyHostToDevice) || hipMemset(m->M, 0, b) || hipMemset(m->V, 0, b)) {
        plens_mlp_destroy(m);
        return PLENS_E_HIP;
    }
    hipLaunchKernelGGL(k_round, dim3(blocks(NPARAM, 256)), dim3(256), 0, m->stream, m->P, m->Pb);
    if (hipStreamSynchronize(m->stream) != hipSuccess) {
        plens_mlp_destroy(m);
        return PLENS_E_HIP;
    }
    *out = m;
    return PLENS_OK;
}

int plens_mlp_destroy(plens_mlp* m) {
    if (!m) return PLENS_OK;
    if (m->stream) (void)hipStreamSynchronize(m->stream);
    free_work(m);
    void* p[] = {m->P, m->Pb, m->M, m->V, m->G, m->X, m->Y, m->idx, m->sum};
    for (void* x : p)
        if (x) (void)hipFree(x);
    if (m->stream) (void)hipStreamDestroy(m->stream);
    delete m;
    return PLENS_OK;
}

int plens_mlp_get_params(const plens_mlp* m, float* params) {
    if (!m || !params) return PLENS_E_INVALID;
    return hipMemcpy(params, m->P, NPARAM * sizeof(float), hipMemcpyDeviceToHost) ? PLENS_E_HIP : PLENS_OK;
}

const float* plens_mlp_device_bf16_params(const plens_mlp* m) { return m ? m->Pb : nullptr; }

int plens_mlp_predict(plens_mlp* m, const float* X, int64_t n, float* y) {
    if (!m || !X || !y || n < 1) return PLENS_E_INVALID;
    const int chunk = 65536;
    float* dX = nullptr;
    if (ensure_work(m, (int)std::min<int64_t>(n, chunk)) ||
        hipMalloc((void**)&dX, (size_t)std::min<int64_t>(n, chunk) * IN * sizeof(float)))
        return PLENS_E_NOMEM;
    int rc = PLENS_OK;
    for (int64_t o = 0; o < n && !rc; o += chunk) {
        const int c = (int)std::min<int64_t>(chunk, n - o);
        if (hipMemcpyAsync(dX, X + o * IN, (size_t)c * IN * sizeof(float), hipMemcpyHostToDevice, m->stream)) {
            rc = PLENS_E_HIP;
            break;
        }
        launch_forward(m, dX, nullptr, c);
        if (hipMemcpyAsync(y + o, m->y, c * sizeof(float), hipMemcpyDeviceToHost, m->stream) ||
            hipStreamSynchronize(m->stream))
            rc = PLENS_E_HIP;
    }
    (void)hipFree(dX);
    return rc;
}

// one forward + backward on a batch (no update): gradients of the mean squared error
int plens_mlp_grad(plens_mlp* m, const float* X, const float* tgt, int32_t n, float* grads, float* loss) {
    if (!m || !X || !tgt || !grads || n < 1) return PLENS_E_INVALID;
    float *dX = nullptr, *dT = nullptr;
    if (ensure_work(m, n) || hipMalloc((void**)&dX, (size_t)n * IN * sizeof(float)) ||
        hipMalloc((void**)&dT, (size_t)n * sizeof(float)))
        return PLENS_E_NOMEM;
    int rc = PLENS_OK;
    if (hipMemcpy(dX, X, (size_t)n * IN * sizeof(float), hipMemcpyHostToDevice) ||
        hipMemcpy(dT, tgt, (size_t)n * sizeof(float), hipMemcpyHostToDevice)) {
        rc = PLENS_E_HIP;
    } else {
        launch_forward(m, dX, nullptr, n);
        launch_backward(m, dX, dT, nullptr, n);
        float s = 0;
        if (hipMemcpyAsync(grads, m->G, NPARAM * sizeof(float), hipMemcpyDeviceToHost, m->stream) ||
            hipMemcpyAsync(&s, m->sum, sizeof(float), hipMemcpyDeviceToHost, m->stream) ||
            hipStreamSynchronize(m->stream))
            rc = PLENS_E_HIP;
        if (loss) *loss = s / n;
    }
    (void)hipFree(dX);
    (void)hipFree(dT);
    return rc;
}

// mini-batch Adam training on (X, y); the data stay on the device between calls with the same
// pointers and size. Batches are drawn from a seeded permutation each epoch.
int plens_mlp_train(plens_mlp* m, const float* X, const float* tgt, int64_t n, int32_t epochs, int32_t batch,
                    float lr, uint64_t seed, float* epoch_loss) {
    if (!m || !X || !tgt || n < 1 || epochs < 1 || batch < 1) return PLENS_E_INVALID;
    if (n != m->n_data) {
        if (m->X) (void)hipFree(m->X);
        if (m->Y) (void)hipFree(m->Y);
        if (m->idx) (void)hipFree(m->idx);
        m->X = m->Y = nullptr;
        m->idx = nullptr;
        if (hipMalloc((void**)&m->X, (size_t)n * IN * sizeof(float)) || hipMalloc((void**)&m->Y, n * sizeof(float)) ||
            hipMalloc((void**)&m->idx, n * sizeof(int)))
            return PLENS_E_NOMEM;
        m->n_data = n;
    }
    if (hipMemcpy(m->X, X, (size_t)n * IN * sizeof(float), hipMemcpyHostToDevice) ||
        hipMemcpy(m->Y, tgt, n * sizeof(float), hipMemcpyHostToDevice))
        return PLENS_E_HIP;
    if (ensure_work(m, batch)) return PLENS_E_NOMEM;
    std::vector<int> perm(n);
    for (int64_t i = 0; i < n; ++i) perm[i] = (int)i;
    std::mt19937_64 rng(seed);
    for (int e = 0; e < epochs; ++e) {
        std::shuffle(perm.begin(), perm.end(), rng);
        if (hipMemcpyAsync(m->idx, perm.data(), n * sizeof(int), hipMemcpyHostToDevice, m->stream)) return PLENS_E_HIP;
        double tot = 0;
        int nb = 0;
        for (int64_t o = 0; o + batch <= n; o += batch, ++nb) {
            launch_forward(m, m->X, m->idx + o, batch);
            launch_backward(m, m->X, m->Y, m->idx + o, batch);
            launch_adam(m, lr);
            float s = 0;
            if (hipMemcpyAsync(&s, m->sum, sizeof(float), hipMemcpyDeviceToHost, m->stream) ||
                hipStreamSynchronize(m->stream))
                return PLENS_E_HIP;
            tot += s / batch;
        }
        if (epoch_loss) epoch_loss[e] = nb ? (float)(tot / nb) : 0.0f;
    }
    return hipStreamSynchronize(m->stream) ? PLENS_E_HIP : PLENS_OK;
}

// one Adam update from the gradients of a given batch (for the optimiser check)
int plens_mlp_step(plens_mlp* m, const float* X, const float* tgt, int32_t n, float lr) {
    float* g = new (std::nothrow) float[NPARAM];
    if (!g) return PLENS_E_NOMEM;
    int rc = plens_mlp_grad(m, X, tgt, n, g, nullptr);
    delete[] g;
    if (rc) return rc;
    launch_adam(m, lr);
    return hipStreamSynchronize(m->stream) ? PLENS_E_HIP : PLENS_OK;
}

}  // extern "C"
