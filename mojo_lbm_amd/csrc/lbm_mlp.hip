// lbm_mlp.hip -- training kernels of the learned LES closure (include/lbm_hip.h,
// "Learned closure"): batched forward, mean-squared-error gradient and Adam on the
// 1,409 parameters of the MLP 9 -> 32 -> 32 -> 1 that collide evaluates per node
// (lbm_physics.hpp, mlp_nu). No reference counterpart.
//
// Gradient: one wave per workgroup takes 64 samples at a time; each lane runs the
// forward and backward pass of one sample and leaves the per-sample factors in
// LDS; then each lane accumulates its 22 of the 1,409 gradient entries over the 64
// samples (the weight gradients are sample-sum outer products). Workgroup partials
// are summed in FP64 in a fixed order by a second kernel, so the result is
// deterministic.
#include <algorithm>
#include <cmath>
#include <mutex>
#include <string>

#include "lbm_launch.hpp"

namespace lbm {

constexpr int MLP_CHUNK = 64;
constexpr int MLP_OWN = (MLP_NPARAM + 63) / 64;  // entries per lane
constexpr int MLP_MAX_BLOCKS = 1024;             // workgroups of mlp_grad_kernel (grid-stride beyond)

// Scratch of lbm_mlp_loss_grad_noalloc: the per-workgroup gradient partials and loss partials of
// the largest gradient launch (MLP_MAX_BLOCKS workgroups), 5,779,456 bytes per GPU, held until the
// process ends. Host code never uses this name as an address: it reads each device's address once
// with hipGetSymbolAddress (scratch_address below) and passes that.
constexpr size_t MLP_SCRATCH_BYTES = (size_t)MLP_MAX_BLOCKS * MLP_NPARAM * sizeof(float) +
                                     (size_t)MLP_MAX_BLOCKS * sizeof(double);
static __device__ __align__(256) unsigned char scratch[MLP_SCRATCH_BYTES];

__device__ __forceinline__ float sigmoidf_(float x) { return 1.0f / (1.0f + expf(-x)); }

__global__ void __launch_bounds__(256) mlp_forward_kernel(const float* __restrict__ P, const float* __restrict__ X,
                                                          int64_t n, float* __restrict__ nu) {
    const int64_t s = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (s >= n) return;
    float z[MLP_IN];
#pragma unroll
    for (int k = 0; k < MLP_IN; ++k) z[k] = (X[s * MLP_IN + k] - P[MLP_MEAN + k]) * P[MLP_ISTD + k];
    float h1[MLP_H];
#pragma unroll
    for (int j = 0; j < MLP_H; ++j) {
        float a = P[MLP_B1 + j];
#pragma unroll
        for (int k = 0; k < MLP_IN; ++k) a = a + P[MLP_W1 + k * MLP_H + j] * z[k];
        h1[j] = silu(a);
    }
    float out = P[MLP_B3];
    for (int j = 0; j < MLP_H; ++j) {
        float a = P[MLP_B2 + j];
#pragma unroll
        for (int k = 0; k < MLP_H; ++k) a = a + P[MLP_W2 + k * MLP_H + j] * h1[k];
        out = out + P[MLP_W3 + j] * silu(a);
    }
    nu[s] = softplus(out) * P[MLP_NUSCALE];
}

// Loss = mean over samples of (softplus(out) - y / nu_scale)^2.
__global__ void __launch_bounds__(64) mlp_grad_kernel(const float* __restrict__ P, const float* __restrict__ X,
                                                      const float* __restrict__ Y, int64_t n,
                                                      float* __restrict__ partial, double* __restrict__ loss_part) {
    __shared__ float s_z[MLP_CHUNK][MLP_IN];
    __shared__ float s_h1[MLP_CHUNK][MLP_H];
    __shared__ float s_g1[MLP_CHUNK][MLP_H];
    __shared__ float s_h2[MLP_CHUNK][MLP_H];
    __shared__ float s_g2[MLP_CHUNK][MLP_H];
    __shared__ float s_go[MLP_CHUNK];
    const int lane = threadIdx.x;
    const float inv_n = 1.0f / (float)n;
    const float inv_scale = 1.0f / P[MLP_NUSCALE];
    float acc[MLP_OWN];
#pragma unroll
    for (int e = 0; e < MLP_OWN; ++e) acc[e] = 0.0f;
    double loss = 0.0;
    for (int64_t base = (int64_t)blockIdx.x * MLP_CHUNK; base < n; base += (int64_t)gridDim.x * MLP_CHUNK) {
        const int64_t s = base + lane;
        const bool live = s < n;
        const int64_t sc = live ? s : 0;  // every load stays in bounds; dead lanes are masked below
        float z[MLP_IN], a1[MLP_H], h1[MLP_H], a2[MLP_H];
#pragma unroll
        for (int k = 0; k < MLP_IN; ++k) {
            const float x = X[sc * MLP_IN + k];
            z[k] = live ? (x - P[MLP_MEAN + k]) * P[MLP_ISTD + k] : 0.0f;
        }
#pragma unroll
        for (int j = 0; j < MLP_H; ++j) {
            float a = P[MLP_B1 + j];
#pragma unroll
            for (int k = 0; k < MLP_IN; ++k) a = a + P[MLP_W1 + k * MLP_H + j] * z[k];
            a1[j] = a;
            h1[j] = silu(a);
        }
        float out = P[MLP_B3];
#pragma unroll
        for (int j = 0; j < MLP_H; ++j) {
            float a = P[MLP_B2 + j];
#pragma unroll
            for (int k = 0; k < MLP_H; ++k) a = a + P[MLP_W2 + k * MLP_H + j] * h1[k];
            a2[j] = a;
            out = out + P[MLP_W3 + j] * silu(a);
        }
        const float sp = softplus(out);
        const float yv = Y[sc];
        const float r = live ? sp - yv * inv_scale : 0.0f;
        loss += (double)r * (double)r;
        const float go = 2.0f * r * inv_n * sigmoidf_(out);  // dL/dout
        float gh1[MLP_H];
#pragma unroll
        for (int k = 0; k < MLP_H; ++k) gh1[k] = 0.0f;
#pragma unroll
        for (int j = 0; j < MLP_H; ++j) {
            const float sg = sigmoidf_(a2[j]);
            const float g2 = go * P[MLP_W3 + j] * (sg * (1.0f + a2[j] * (1.0f - sg)));  // silu'
            s_g2[lane][j] = g2;
            s_h2[lane][j] = a2[j] * sg;
#pragma unroll
            for (int k = 0; k < MLP_H; ++k) gh1[k] = gh1[k] + P[MLP_W2 + k * MLP_H + j] * g2;
        }
#pragma unroll
        for (int k = 0; k < MLP_H; ++k) {
            const float sg = sigmoidf_(a1[k]);
            s_g1[lane][k] = gh1[k] * (sg * (1.0f + a1[k] * (1.0f - sg)));
            s_h1[lane][k] = h1[k];
        }
#pragma unroll
        for (int k = 0; k < MLP_IN; ++k) s_z[lane][k] = z[k];
        s_go[lane] = go;
        __syncthreads();
#pragma unroll
        for (int e = 0; e < MLP_OWN; ++e) {
            const int id = lane + 64 * e;
            float t = 0.0f;
            if (id < MLP_B1) {
                const int k = id / MLP_H, j = id % MLP_H;
                for (int q = 0; q < MLP_CHUNK; ++q) t = t + s_z[q][k] * s_g1[q][j];
            } else if (id < MLP_W2) {
                for (int q = 0; q < MLP_CHUNK; ++q) t = t + s_g1[q][id - MLP_B1];
            } else if (id < MLP_B2) {
                const int k = (id - MLP_W2) / MLP_H, j = (id - MLP_W2) % MLP_H;
                for (int q = 0; q < MLP_CHUNK; ++q) t = t + s_h1[q][k] * s_g2[q][j];
            } else if (id < MLP_W3) {
                for (int q = 0; q < MLP_CHUNK; ++q) t = t + s_g2[q][id - MLP_B2];
            } else if (id < MLP_B3) {
                for (int q = 0; q < MLP_CHUNK; ++q) t = t + s_go[q] * s_h2[q][id - MLP_W3];
            } else if (id == MLP_B3) {
                for (int q = 0; q < MLP_CHUNK; ++q) t = t + s_go[q];
            }
            acc[e] = acc[e] + t;
        }
        __syncthreads();
    }
#pragma unroll
    for (int e = 0; e < MLP_OWN; ++e) {
        const int id = lane + 64 * e;
        if (id < MLP_NPARAM) partial[(size_t)blockIdx.x * MLP_NPARAM + id] = acc[e];
    }
    for (int off = 32; off > 0; off >>= 1) loss += __shfl_down(loss, off, 64);
    if (lane == 0) loss_part[blockIdx.x] = loss;
}

// grad[i] = sum over workgroups (fixed order, FP64); loss = sum / n.
__global__ void __launch_bounds__(256) mlp_grad_finish(const float* __restrict__ partial,
                                                       const double* __restrict__ loss_part, int blocks, int64_t n,
                                                       float* __restrict__ grad, double* __restrict__ loss) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < MLP_NPARAM) {
        double s = 0.0;
        for (int b = 0; b < blocks; ++b) s += (double)partial[(size_t)b * MLP_NPARAM + i];
        grad[i] = (float)s;
    }
    if (i == 0 && loss) {
        double s = 0.0;
        for (int b = 0; b < blocks; ++b) s += loss_part[b];
        *loss = s / (double)n;
    }
}

// Adam (Kingma & Ba 2015) with bias correction at step t >= 1.
__global__ void __launch_bounds__(256) mlp_adam_kernel(float* __restrict__ P, float* __restrict__ m,
                                                       float* __restrict__ v, const float* __restrict__ g, int t,
                                                       float lr, float b1, float b2, float eps) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= MLP_NPARAM) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float mh = mi / (1.0f - powf(b1, (float)t));
    const float vh = vi / (1.0f - powf(b2, (float)t));
    P[i] = P[i] - lr * mh / (sqrtf(vh) + eps);
}

// The closure exactly as collide evaluates it (wave-level, matrix cores): nu[s] for raw features
// X[s], or -1 where a feature is outside the training range. Lanes past n repeat sample n-1 so
// every wave stays whole.
__global__ void __launch_bounds__(256) mlp_closure_kernel(const float* __restrict__ blob, const float* __restrict__ X,
                                                          int64_t n, float* __restrict__ nu) {
    const int64_t s = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t sc = s < n ? s : n - 1;
    float x[MLP_IN];
#pragma unroll
    for (int k = 0; k < MLP_IN; ++k) x[k] = X[sc * MLP_IN + k];
    bool inside;
    const float v = mlp_nu_wave(blob, x, inside);
    if (s < n) nu[s] = inside ? v : -1.0f;
}

}  // namespace lbm

using namespace lbm;

extern "C" {

int lbm_mlp_forward(const float* params_dev, const float* X_dev, int64_t n, float* nu_dev, void* stream) {
    if (!params_dev || (!X_dev && n) || (!nu_dev && n) || n < 0) return set_error(LBM_ERR_INVALID, "bad closure forward arguments");
    if (!n) return LBM_OK;
    hipLaunchKernelGGL(mlp_forward_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream,
                       params_dev, X_dev, n, nu_dev);
    return check_launch("mlp_forward_kernel");
}

size_t lbm_mlp_workspace_bytes(int64_t n) {
    if (n <= 0) return 0;
    const int64_t blocks = std::min<int64_t>((n + MLP_CHUNK - 1) / MLP_CHUNK, MLP_MAX_BLOCKS);
    return (size_t)blocks * MLP_NPARAM * sizeof(float) + (size_t)blocks * sizeof(double);
}

int lbm_mlp_loss_grad(const float* params_dev, const float* X_dev, const float* y_dev, int64_t n, float* grad_dev,
                      double* loss_dev, void* workspace_dev, void* stream) {
    if (!params_dev || !X_dev || !y_dev || !grad_dev || !workspace_dev || n <= 0)
        return set_error(LBM_ERR_INVALID, "bad closure gradient arguments");
    const int blocks = (int)std::min<int64_t>((n + MLP_CHUNK - 1) / MLP_CHUNK, MLP_MAX_BLOCKS);
    float* partial = (float*)workspace_dev;
    double* loss_part = (double*)(partial + (size_t)blocks * MLP_NPARAM);
    hipLaunchKernelGGL(mlp_grad_kernel, dim3(blocks), dim3(64), 0, (hipStream_t)stream, params_dev, X_dev, y_dev, n,
                       partial, loss_part);
    if (int rc = check_launch("mlp_grad_kernel")) return rc;
    hipLaunchKernelGGL(mlp_grad_finish, dim3((MLP_NPARAM + 255) / 256), dim3(256), 0, (hipStream_t)stream, partial,
                       loss_part, blocks, n, grad_dev, loss_dev);
    return check_launch("mlp_grad_finish");
}

}  // extern "C"

namespace {

constexpr int SCRATCH_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void* g_scratch_addr[SCRATCH_MAX_DEVICES] = {};

// Device address of `scratch` on the current device, looked up once per device (with that device
// current) on the first call there.
int scratch_address(void** out) {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0)
        return set_error(LBM_ERR_NO_DEVICE, "lbm_mlp_loss_grad_noalloc: no current HIP device");
    if (dev >= SCRATCH_MAX_DEVICES)
        return set_error(LBM_ERR_UNSUPPORTED, "lbm_mlp_loss_grad_noalloc: device index above 63");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_addr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return set_error(LBM_ERR_HIP, std::string("hipGetSymbolAddress of the gradient scratch failed: ") +
                                              hipGetErrorString(e));
        g_scratch_addr[dev] = p;
    }
    *out = g_scratch_addr[dev];
    return LBM_OK;
}

}  // namespace

extern "C" {

int lbm_mlp_loss_grad_noalloc(const float* params_dev, const float* X_dev, const float* y_dev, int64_t n,
                              float* grad_dev, double* loss_dev, void* stream) {
    if (!params_dev || !X_dev || !y_dev || !grad_dev || n <= 0)
        return set_error(LBM_ERR_INVALID, "bad closure gradient arguments");
    const size_t need = lbm_mlp_workspace_bytes(n);
    if (need > MLP_SCRATCH_BYTES)
        return set_error(LBM_ERR_INVALID, "lbm_mlp_loss_grad_noalloc: " + std::to_string(need) +
                                              " bytes of scratch needed, the static scratch holds " +
                                              std::to_string(MLP_SCRATCH_BYTES));
    void* ws = nullptr;
    if (int rc = scratch_address(&ws)) return rc;
    // Every workgroup of the gradient kernel overwrites all of its partials before the finishing
    // kernel reads them, so nothing in the scratch has to start at zero.
    return lbm_mlp_loss_grad(params_dev, X_dev, y_dev, n, grad_dev, loss_dev, ws, stream);
}

int lbm_mlp_closure(lbm_handle* hh, const float* X_dev, int64_t n, float* nu_dev, void* stream) {
    if (!hh || !X_dev || !nu_dev || n <= 0) return set_error(LBM_ERR_INVALID, "bad closure arguments");
    const Handle& h = *reinterpret_cast<Handle*>(hh);
    if (!h.mlp) return set_error(LBM_ERR_INVALID, "lbm_set_mlp first");
    hipLaunchKernelGGL(mlp_closure_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream,
                       (const float*)h.mlp, X_dev, n, nu_dev);
    return check_launch("mlp_closure_kernel");
}

int lbm_mlp_adam(float* params_dev, float* m_dev, float* v_dev, const float* grad_dev, int32_t t, float lr,
                 float beta1, float beta2, float eps, void* stream) {
    if (!params_dev || !m_dev || !v_dev || !grad_dev || t < 1) return set_error(LBM_ERR_INVALID, "bad Adam arguments");
    hipLaunchKernelGGL(mlp_adam_kernel, dim3((MLP_NPARAM + 255) / 256), dim3(256), 0, (hipStream_t)stream, params_dev,
                       m_dev, v_dev, grad_dev, (int)t, lr, beta1, beta2, eps);
    return check_launch("mlp_adam_kernel");
}

}  // extern "C"
