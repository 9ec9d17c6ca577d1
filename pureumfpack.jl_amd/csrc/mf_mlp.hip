// Small MLP cost model for the gemm tile policy: forward, backward and Adam on the GPU.
//
// Network: d_in -> hidden (ReLU) -> hidden (ReLU) -> 1, FP32, trained full-batch on mean
// squared error of log-time samples that the autotuner measures (mf_time_gemm).  The model
// only ranks tile shapes; the shapes it picks never change a factorization result.
//
// Layout: samples row-major (n x d_in); a dense layer's weights W (d_out x d_in, row-major)
// followed by its bias (d_out); the three layers back to back in one parameter buffer, the
// gradient buffer and both Adam moment buffers in the same layout.  Every reduction over the
// samples runs in sample order in one thread, so results do not depend on scheduling.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <utility>
#include <vector>

#include "../../include/mfgpu.h"

namespace {

struct Layer {
    int din, dout;
    int64_t w, b;   // offsets of W and b in the parameter buffer
};

// Z = X W^T + b, Y = relu(Z) (or Z for the output layer); one thread per (sample, unit)
__global__ void k_dense_fwd(const float *__restrict__ X, const float *__restrict__ P, Layer L, int64_t n,
                            int relu, float *__restrict__ Z, float *__restrict__ Y)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n * L.dout) return;
    const int64_t s = e / L.dout;
    const int o = (int)(e % L.dout);
    const float *w = P + L.w + (int64_t)o * L.din;
    const float *x = X + s * L.din;
    float z = P[L.b + o];
    for (int i = 0; i < L.din; i++) z = fmaf(w[i], x[i], z);
    Z[e] = z;
    Y[e] = relu ? fmaxf(z, 0.0f) : z;
}

// loss = mean (y - t)^2; dY = 2 (y - t) / n; one thread (serial sum, deterministic)
__global__ void k_mse(const float *__restrict__ Y, const float *__restrict__ t, int64_t n, float *__restrict__ dY,
                      float *__restrict__ loss)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    float acc = 0.0f;
    const float sc = 2.0f / (float)n;
    for (int64_t s = 0; s < n; s++) {
        const float d = Y[s] - t[s];
        acc = fmaf(d, d, acc);
        dY[s] = sc * d;
    }
    *loss = acc / (float)n;
}

// dZ = dY * relu'(Z) in place (relu layers)
__global__ void k_relu_bwd(const float *__restrict__ Z, float *__restrict__ dY, int64_t cnt)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < cnt && Z[e] <= 0.0f) dY[e] = 0.0f;
}

// dW[o][i] = sum_s dZ[s][o] X[s][i], db[o] = sum_s dZ[s][o]; one thread per weight
__global__ void k_dense_wgrad(const float *__restrict__ X, const float *__restrict__ dZ, Layer L, int64_t n,
                              float *__restrict__ G)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    const int nw = L.dout * L.din;
    if (e < nw) {
        const int o = e / L.din, i = e % L.din;
        float acc = 0.0f;
        for (int64_t s = 0; s < n; s++) acc = fmaf(dZ[s * L.dout + o], X[s * L.din + i], acc);
        G[L.w + e] = acc;
    } else if (e < nw + L.dout) {
        const int o = e - nw;
        float acc = 0.0f;
        for (int64_t s = 0; s < n; s++) acc += dZ[s * L.dout + o];
        G[L.b + o] = acc;
    }
}

// dX[s][i] = sum_o dZ[s][o] W[o][i]; one thread per (sample, input)
__global__ void k_dense_xgrad(const float *__restrict__ P, const float *__restrict__ dZ, Layer L, int64_t n,
                              float *__restrict__ dX)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n * L.din) return;
    const int64_t s = e / L.din;
    const int i = (int)(e % L.din);
    float acc = 0.0f;
    for (int o = 0; o < L.dout; o++) acc = fmaf(dZ[s * L.dout + o], P[L.w + (int64_t)o * L.din + i], acc);
    dX[e] = acc;
}

// Adam (Kingma & Ba 2015, Algorithm 1), step t >= 1
__global__ void k_adam(float *__restrict__ P, const float *__restrict__ G, float *__restrict__ M1,
                       float *__restrict__ M2, int64_t np, float lr, float b1, float b2, float eps, int t)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= np) return;
    const float g = G[e];
    const float m = b1 * M1[e] + (1.0f - b1) * g;
    const float v = b2 * M2[e] + (1.0f - b2) * g * g;
    M1[e] = m;
    M2[e] = v;
    const float mh = m / (1.0f - powf(b1, (float)t));
    const float vh = v / (1.0f - powf(b2, (float)t));
    P[e] -= lr * mh / (sqrtf(vh) + eps);
}

inline unsigned blocks(int64_t cnt) { return (unsigned)((cnt + 255) / 256); }

thread_local std::string g_mlp_err;

}  // namespace

struct mf_mlp {
    int din, hidden;
    Layer L[3];
    int64_t np;
    float *P = nullptr, *G = nullptr, *M1 = nullptr, *M2 = nullptr;
    int step = 0;
    // work buffers for n samples
    int64_t cap = 0;
    float *X = nullptr, *t = nullptr, *Z[3] = {}, *Y[3] = {}, *dA = nullptr, *dB = nullptr, *loss = nullptr;
    hipStream_t st = nullptr;
};

#define MLPCHK(expr)                                                                 \
    do {                                                                             \
        hipError_t e_ = (expr);                                                      \
        if (e_ != hipSuccess) {                                                      \
            g_mlp_err = std::string("HIP error: ") + hipGetErrorString(e_);          \
            return MF_HIP;                                                           \
        }                                                                            \
    } while (0)

static void mlp_free_work(mf_mlp *m)
{
    float **bufs[] = {&m->X, &m->t, &m->Z[0], &m->Z[1], &m->Z[2], &m->Y[0], &m->Y[1], &m->Y[2], &m->dA, &m->dB,
                      &m->loss};
    for (float **b : bufs)
        if (*b) { (void)hipFree(*b); *b = nullptr; }
    m->cap = 0;
}

static int mlp_reserve(mf_mlp *m, int64_t n)
{
    if (n <= m->cap) return MF_OK;
    mlp_free_work(m);
    const int64_t w = std::max<int64_t>(m->din, m->hidden);
    MLPCHK(hipMalloc(&m->X, sizeof(float) * n * m->din));
    MLPCHK(hipMalloc(&m->t, sizeof(float) * n));
    for (int l = 0; l < 3; l++) {
        MLPCHK(hipMalloc(&m->Z[l], sizeof(float) * n * m->L[l].dout));
        MLPCHK(hipMalloc(&m->Y[l], sizeof(float) * n * m->L[l].dout));
    }
    MLPCHK(hipMalloc(&m->dA, sizeof(float) * n * w));
    MLPCHK(hipMalloc(&m->dB, sizeof(float) * n * w));
    MLPCHK(hipMalloc(&m->loss, sizeof(float)));
    m->cap = n;
    return MF_OK;
}

// forward over the samples already in m->X
static hipError_t mlp_fwd(mf_mlp *m, int64_t n)
{
    const float *in = m->X;
    for (int l = 0; l < 3; l++) {
        hipLaunchKernelGGL(k_dense_fwd, dim3(blocks(n * m->L[l].dout)), dim3(256), 0, m->st, in, m->P, m->L[l], n,
                           l < 2 ? 1 : 0, m->Z[l], m->Y[l]);
        in = m->Y[l];
    }
    return hipGetLastError();
}

// loss and gradient (into m->G) for the samples and targets in m->X, m->t
static hipError_t mlp_grad(mf_mlp *m, int64_t n)
{
    hipError_t e = mlp_fwd(m, n);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_mse, dim3(1), dim3(1), 0, m->st, m->Y[2], m->t, n, m->dA, m->loss);
    float *dZ = m->dA, *dX = m->dB;
    for (int l = 2; l >= 0; l--) {
        const Layer &L = m->L[l];
        if (l < 2) hipLaunchKernelGGL(k_relu_bwd, dim3(blocks(n * L.dout)), dim3(256), 0, m->st, m->Z[l], dZ, n * L.dout);
        const float *in = l == 0 ? m->X : m->Y[l - 1];
        hipLaunchKernelGGL(k_dense_wgrad, dim3(blocks((int64_t)L.dout * L.din + L.dout)), dim3(256), 0, m->st, in, dZ,
                           L, n, m->G);
        if (l > 0) {
            hipLaunchKernelGGL(k_dense_xgrad, dim3(blocks(n * L.din)), dim3(256), 0, m->st, m->P, dZ, L, n, dX);
            std::swap(dZ, dX);
        }
    }
    return hipGetLastError();
}

static int mlp_upload(mf_mlp *m, const float *X, const float *t, int64_t n)
{
    int rc = mlp_reserve(m, n);
    if (rc != MF_OK) return rc;
    MLPCHK(hipMemcpyAsync(m->X, X, sizeof(float) * n * m->din, hipMemcpyHostToDevice, m->st));
    if (t) MLPCHK(hipMemcpyAsync(m->t, t, sizeof(float) * n, hipMemcpyHostToDevice, m->st));
    return MF_OK;
}

extern "C" int mf_mlp_create(int32_t d_in, int32_t hidden, uint64_t seed, mf_mlp **out)
{
    if (!out || d_in < 1 || hidden < 1 || d_in > 4096 || hidden > 4096) {
        g_mlp_err = "bad MLP dimensions";
        return MF_BADARG;
    }
    mf_mlp *m = new mf_mlp();
    m->din = d_in;
    m->hidden = hidden;
    const int dims[4] = {d_in, hidden, hidden, 1};
    int64_t off = 0;
    for (int l = 0; l < 3; l++) {
        m->L[l].din = dims[l];
        m->L[l].dout = dims[l + 1];
        m->L[l].w = off;
        off += (int64_t)dims[l] * dims[l + 1];
        m->L[l].b = off;
        off += dims[l + 1];
    }
    m->np = off;
    // Glorot-uniform weights, zero biases, from a 64-bit LCG (deterministic per seed)
    std::vector<float> p(m->np, 0.0f);
    uint64_t x = seed * 6364136223846793005ULL + 1442695040888963407ULL;
    for (int l = 0; l < 3; l++) {
        const float a = std::sqrt(6.0f / (float)(m->L[l].din + m->L[l].dout));
        for (int64_t i = 0; i < (int64_t)m->L[l].din * m->L[l].dout; i++) {
            x = x * 6364136223846793005ULL + 1442695040888963407ULL;
            const float u = (float)((x >> 40) & 0xFFFFFF) / 16777216.0f;
            p[m->L[l].w + i] = (2.0f * u - 1.0f) * a;
        }
    }
    hipError_t e = hipStreamCreate(&m->st);
    if (e == hipSuccess) e = hipMalloc(&m->P, sizeof(float) * m->np);
    if (e == hipSuccess) e = hipMalloc(&m->G, sizeof(float) * m->np);
    if (e == hipSuccess) e = hipMalloc(&m->M1, sizeof(float) * m->np);
    if (e == hipSuccess) e = hipMalloc(&m->M2, sizeof(float) * m->np);
    if (e == hipSuccess) e = hipMemcpy(m->P, p.data(), sizeof(float) * m->np, hipMemcpyHostToDevice);
    std::vector<float> zero(m->np, 0.0f);
    if (e == hipSuccess) e = hipMemcpy(m->M1, zero.data(), sizeof(float) * m->np, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(m->M2, zero.data(), sizeof(float) * m->np, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        g_mlp_err = std::string("HIP error: ") + hipGetErrorString(e);
        mf_mlp_destroy(m);
        return MF_HIP;
    }
    *out = m;
    return MF_OK;
}

extern "C" void mf_mlp_destroy(mf_mlp *m)
{
    if (!m) return;
    mlp_free_work(m);
    float **bufs[] = {&m->P, &m->G, &m->M1, &m->M2};
    for (float **b : bufs)
        if (*b) (void)hipFree(*b);
    if (m->st) (void)hipStreamDestroy(m->st);
    delete m;
}

extern "C" int64_t mf_mlp_num_params(const mf_mlp *m) { return m ? m->np : 0; }

extern "C" int mf_mlp_get_params(mf_mlp *m, float *params)
{
    if (!m || !params) return MF_BADARG;
    MLPCHK(hipMemcpy(params, m->P, sizeof(float) * m->np, hipMemcpyDeviceToHost));
    return MF_OK;
}

extern "C" int mf_mlp_set_params(mf_mlp *m, const float *params)
{
    if (!m || !params) return MF_BADARG;
    MLPCHK(hipMemcpy(m->P, params, sizeof(float) * m->np, hipMemcpyHostToDevice));
    return MF_OK;
}

extern "C" int mf_mlp_forward(mf_mlp *m, const float *X, int64_t n, float *Y)
{
    if (!m || !X || !Y || n < 1) return MF_BADARG;
    int rc = mlp_upload(m, X, nullptr, n);
    if (rc != MF_OK) return rc;
    MLPCHK(mlp_fwd(m, n));
    MLPCHK(hipMemcpyAsync(Y, m->Y[2], sizeof(float) * n, hipMemcpyDeviceToHost, m->st));
    MLPCHK(hipStreamSynchronize(m->st));
    return MF_OK;
}

extern "C" int mf_mlp_gradient(mf_mlp *m, const float *X, const float *t, int64_t n, float *loss, float *grad)
{
    if (!m || !X || !t || n < 1) return MF_BADARG;
    int rc = mlp_upload(m, X, t, n);
    if (rc != MF_OK) return rc;
    MLPCHK(mlp_grad(m, n));
    if (loss) MLPCHK(hipMemcpyAsync(loss, m->loss, sizeof(float), hipMemcpyDeviceToHost, m->st));
    if (grad) MLPCHK(hipMemcpyAsync(grad, m->G, sizeof(float) * m->np, hipMemcpyDeviceToHost, m->st));
    MLPCHK(hipStreamSynchronize(m->st));
    return MF_OK;
}

extern "C" int mf_mlp_train(mf_mlp *m, const float *X, const float *t, int64_t n, int32_t epochs, float lr,
                            float *final_loss)
{
    if (!m || !X || !t || n < 1 || epochs < 0) return MF_BADARG;
    int rc = mlp_upload(m, X, t, n);
    if (rc != MF_OK) return rc;
    for (int ep = 0; ep < epochs; ep++) {
        MLPCHK(mlp_grad(m, n));
        m->step++;
        hipLaunchKernelGGL(k_adam, dim3(blocks(m->np)), dim3(256), 0, m->st, m->P, m->G, m->M1, m->M2, m->np, lr, 0.9f,
                           0.999f, 1e-8f, m->step);
        MLPCHK(hipGetLastError());
    }
    MLPCHK(mlp_grad(m, n));   // loss at the final parameters
    if (final_loss) MLPCHK(hipMemcpyAsync(final_loss, m->loss, sizeof(float), hipMemcpyDeviceToHost, m->st));
    MLPCHK(hipStreamSynchronize(m->st));
    return MF_OK;
}

extern "C" const char *mf_mlp_last_error(void) { return g_mlp_err.c_str(); }
