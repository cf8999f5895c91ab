// Forward and backward kernels of the autotuner's cost model, a small fully connected network
// (fp32) that predicts the time per cell update of a traverser configuration. Sizes are
// small (batch of a few thousand configurations, widths <= 128), so each kernel maps one
// thread to one output element; the point is that training runs on the GPU with the
// project's own kernels, not throughput.
#include <hip/hip_runtime.h>

#include <string>

#include "cellato_hip.h"
#include "plan.hpp"

namespace cellato::hip {
namespace {

// Y[n, o] = act(sum_i X[n, i] W[i, o] + b[o])
__global__ void linear_fwd(const float* X, const float* W, const float* b, float* Y, int N, int Din, int Dout, int relu) {
    const long long t = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= static_cast<long long>(N) * Dout) return;
    const int n = static_cast<int>(t / Dout), o = static_cast<int>(t % Dout);
    float acc = b ? b[o] : 0.0f;
    for (int i = 0; i < Din; ++i) acc = fmaf(X[static_cast<size_t>(n) * Din + i], W[static_cast<size_t>(i) * Dout + o], acc);
    Y[t] = relu && acc < 0.0f ? 0.0f : acc;
}

// G = dY * act'(Y) (ReLU: Y > 0), kept in dYg
__global__ void act_bwd(const float* Y, const float* dY, float* G, long long n, int relu) {
    const long long t = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= n) return;
    G[t] = relu ? (Y[t] > 0.0f ? dY[t] : 0.0f) : dY[t];
}

// dW[i, o] = sum_n X[n, i] G[n, o]
__global__ void weight_grad(const float* X, const float* G, float* dW, int N, int Din, int Dout) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= Din * Dout) return;
    const int i = t / Dout, o = t % Dout;
    float acc = 0.0f;
    for (int n = 0; n < N; ++n) acc = fmaf(X[static_cast<size_t>(n) * Din + i], G[static_cast<size_t>(n) * Dout + o], acc);
    dW[t] = acc;
}

// db[o] = sum_n G[n, o]
__global__ void bias_grad(const float* G, float* db, int N, int Dout) {
    const int o = blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= Dout) return;
    float acc = 0.0f;
    for (int n = 0; n < N; ++n) acc += G[static_cast<size_t>(n) * Dout + o];
    db[o] = acc;
}

// dX[n, i] = sum_o G[n, o] W[i, o]
__global__ void input_grad(const float* G, const float* W, float* dX, int N, int Din, int Dout) {
    const long long t = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= static_cast<long long>(N) * Din) return;
    const int n = static_cast<int>(t / Din), i = static_cast<int>(t % Din);
    float acc = 0.0f;
    for (int o = 0; o < Dout; ++o) acc = fmaf(G[static_cast<size_t>(n) * Dout + o], W[static_cast<size_t>(i) * Dout + o], acc);
    dX[t] = acc;
}

// mean squared error: loss = mean((p - y)^2), dP = 2 (p - y) / N; loss accumulated atomically
__global__ void mse(const float* P, const float* Yt, float* dP, float* loss, int N) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    float l = 0.0f;
    if (t < N) {
        const float d = P[t] - Yt[t];
        dP[t] = 2.0f * d / static_cast<float>(N);
        l = d * d / static_cast<float>(N);
    }
    for (int off = 32; off > 0; off >>= 1) l += __shfl_xor(l, off);
    if ((threadIdx.x & 63) == 0 && loss) atomicAdd(loss, l);
}

// Adam (Kingma & Ba 2015) with bias correction, step t >= 1
__global__ void adam(float* p, const float* g, float* m, float* v, int n, float lr, float b1, float b2, float eps,
                     float c1, float c2) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const float gt = g[t];
    const float mt = b1 * m[t] + (1.0f - b1) * gt;
    const float vt = b2 * v[t] + (1.0f - b2) * gt * gt;
    m[t] = mt;
    v[t] = vt;
    p[t] -= lr * (mt / c1) / (sqrtf(vt / c2) + eps);
}

inline int nblocks(long long n) { return static_cast<int>((n + 255) / 256 > 0 ? (n + 255) / 256 : 1); }

int mfail(int code, const std::string& m) {
    last_error() = m;
    return code;
}

int done(const char* what) {
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return mfail(CELLATO_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
    return CELLATO_OK;
}

} // namespace
} // namespace cellato::hip

using namespace cellato::hip;

extern "C" {

int cellato_mlp_linear_forward(const float* X, const float* W, const float* b, float* Y, int N, int Din, int Dout,
                               int relu, void* stream) {
    if (!X || !W || !Y || N <= 0 || Din <= 0 || Dout <= 0) return mfail(CELLATO_E_ARG, "mlp forward: bad arguments");
    hipLaunchKernelGGL(linear_fwd, dim3(nblocks(static_cast<long long>(N) * Dout)), dim3(256), 0,
                       static_cast<hipStream_t>(stream), X, W, b, Y, N, Din, Dout, relu);
    return done("linear_fwd");
}

int cellato_mlp_linear_backward(const float* X, const float* W, const float* Y, const float* dY, float* G, float* dX,
                                float* dW, float* db, int N, int Din, int Dout, int relu, void* stream) {
    if (!X || !W || !Y || !dY || !G || !dW || N <= 0 || Din <= 0 || Dout <= 0)
        return mfail(CELLATO_E_ARG, "mlp backward: bad arguments");
    hipStream_t s = static_cast<hipStream_t>(stream);
    const long long ny = static_cast<long long>(N) * Dout;
    hipLaunchKernelGGL(act_bwd, dim3(nblocks(ny)), dim3(256), 0, s, Y, dY, G, ny, relu);
    hipLaunchKernelGGL(weight_grad, dim3(nblocks(static_cast<long long>(Din) * Dout)), dim3(256), 0, s, X, G, dW, N, Din, Dout);
    if (db) hipLaunchKernelGGL(bias_grad, dim3(nblocks(Dout)), dim3(256), 0, s, G, db, N, Dout);
    if (dX) hipLaunchKernelGGL(input_grad, dim3(nblocks(static_cast<long long>(N) * Din)), dim3(256), 0, s, G, W, dX, N, Din, Dout);
    return done("linear backward");
}

int cellato_mlp_mse(const float* pred, const float* target, float* dpred, float* loss, int N, void* stream) {
    if (!pred || !target || !dpred || N <= 0) return mfail(CELLATO_E_ARG, "mse: bad arguments");
    hipLaunchKernelGGL(mse, dim3(nblocks(N)), dim3(256), 0, static_cast<hipStream_t>(stream), pred, target, dpred, loss, N);
    return done("mse");
}

int cellato_mlp_adam(float* param, const float* grad, float* m, float* v, int n, float lr, float beta1, float beta2,
                     float eps, int step, void* stream) {
    if (!param || !grad || !m || !v || n <= 0 || step < 1) return mfail(CELLATO_E_ARG, "adam: bad arguments");
    float c1 = 1.0f, c2 = 1.0f, p1 = 1.0f, p2 = 1.0f;
    for (int i = 0; i < step; ++i) { p1 *= beta1; p2 *= beta2; }
    c1 = 1.0f - p1;
    c2 = 1.0f - p2;
    hipLaunchKernelGGL(adam, dim3(nblocks(n)), dim3(256), 0, static_cast<hipStream_t>(stream), param, grad, m, v, n, lr,
                       beta1, beta2, eps, c1, c2);
    return done("adam");
}

} // extern "C"
