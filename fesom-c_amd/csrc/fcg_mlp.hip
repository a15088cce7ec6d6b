// Learned vertical closure: a per-(column, level) network 5 -> 128 -> 128 ->
// 128 -> 2 with ReLU hidden layers, standing in for d3_end
// (src/fv_mixing.F90:73-347); shapes and the matrix-core fragment layout are
// in fcg_mlp.h.
//
// Parameters are FP32 masters in one flat array; a cast kernel rebuilds the
// BF16 fragment arrays (forward F1..F4, back-propagation G4, G3, G2) and a
// contiguous bias array after each update.  Training: each wave takes
// 16-sample tiles in turn, runs the forward and the backward pass on the
// matrix cores, stages the activations and errors in LDS as [row][sample]
// BF16 tiles and forms the weight gradients from them with 16x16x16 products
// over the tile's samples, added into its own partial gradient in global
// memory (one lane per entry).  The partials are summed in wave order
// (deterministic) and applied by Adam.
#include <algorithm>
#include <cmath>
#include <mutex>
#include <random>
#include <vector>

#include "fcg_internal.h"
#include "fcg_mlp.h"

namespace fcgm {

// ---- fragments and biases from the FP32 masters --------------------------
__global__ __launch_bounds__(256) void k_mlp_cast(const float *__restrict__ w, bf16x8 *__restrict__ frag,
                                                  float *__restrict__ bias) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < kMlpBias) {  // b1 b2 b3 b4 contiguous
        const int src = t < kMlpH ? kMlpB1 + t
                        : t < 2 * kMlpH ? kMlpB2 + t - kMlpH
                        : t < 3 * kMlpH ? kMlpB3 + t - 2 * kMlpH
                                        : kMlpB4 + t - 3 * kMlpH;
        bias[t] = w[src];
    }
    if (t >= kMlpFragCount * 64) return;
    const int f = t / 64, l = t % 64, q = l / 16, c = l % 16;
    const float *W1 = w + kMlpW1, *W2 = w + kMlpW2, *W3 = w + kMlpW3, *W4 = w + kMlpW4;
    float v[8];
    for (int j = 0; j < 8; ++j) {
        float x = 0.0f;
        if (f < kFragF2) {  // F1[mt]: W1 rows 16mt+c, inputs 8q+j
            x = W1[(16 * f + c) * kMlpIn + 8 * q + j];
        } else if (f < kFragF3) {  // F2[mt][k]: W2 rows 16mt+c, hidden kappa
            const int mt = (f - kFragF2) / 4, k = (f - kFragF2) % 4;
            x = W2[(16 * mt + c) * kMlpH + kappa(k, q, j)];
        } else if (f < kFragF4) {  // F3[mt][k]
            const int mt = (f - kFragF3) / 4, k = (f - kFragF3) % 4;
            x = W3[(16 * mt + c) * kMlpH + kappa(k, q, j)];
        } else if (f < kFragG4) {  // F4[k]: W4 rows c (outputs), hidden kappa
            x = W4[c * kMlpH + kappa(f - kFragF4, q, j)];
        } else if (f < kFragG3) {  // G4[mt] (16x16x16): W4^T rows 16mt+c, outputs 4q+j
            const int mt = f - kFragG4;
            x = j < 4 ? W4[(4 * q + j) * kMlpH + 16 * mt + c] : 0.0f;
        } else if (f < kFragG2) {  // G3[mt][k]: W3^T rows 16mt+c, hidden kappa
            const int mt = (f - kFragG3) / 4, k = (f - kFragG3) % 4;
            x = W3[kappa(k, q, j) * kMlpH + 16 * mt + c];
        } else {  // G2[mt][k]
            const int mt = (f - kFragG2) / 4, k = (f - kFragG2) % 4;
            x = W2[kappa(k, q, j) * kMlpH + 16 * mt + c];
        }
        v[j] = x;
    }
    frag[t] = slab(f32x4{v[0], v[1], v[2], v[3]}, f32x4{v[4], v[5], v[6], v[7]});
}

// ---- inference on normalized features: X[n][kMlpF] -> Y[n][2] -----------
// one wavefront per block, 64 samples (four tiles) per iteration
__global__ __launch_bounds__(64) void k_mlp_forward(const bf16x8 *__restrict__ frag, const float *__restrict__ bias,
                                                    const float *__restrict__ X, float *__restrict__ Y, long n) {
    __shared__ float xs[64 * kMlpF];
    const int lane = threadIdx.x;
    const long ngrp = (n + 63) / 64;
#pragma unroll 1
    for (long gi = blockIdx.x; gi < ngrp; gi += gridDim.x) {
        const long s0 = gi * 64;
        for (int j = lane; j < 64 * kMlpF; j += 64) {
            const long s = s0 + j / kMlpF;
            xs[j] = s < n ? X[s * kMlpF + j % kMlpF] : 0.0f;
        }
        __syncthreads();
        bf16x8 x[4];
        for (int t = 0; t < 4; ++t) x[t] = load_bx(&xs[t * 16 * kMlpF], lane);
        f32x4 y[4];
        infer_tiles<4>(frag, bias, x, lane, y);
        for (int t = 0; t < 4; ++t) {
            const long s = s0 + 16 * t + lane;
            if (lane < 16 && s < n) {
                Y[s * 2] = y[t][0];
                Y[s * 2 + 1] = y[t][1];
            }
        }
        __syncthreads();
    }
}

// ---- training: gradient of 0.5*mean over the batch of |y - t|^2 ------------
// BF16 [row][16 samples] staging tiles of one wave
struct GradLds {
    unsigned short x[16 * 16], dy[16 * 16];
    unsigned short a1[kMlpH * 16], a2[kMlpH * 16], a3[kMlpH * 16];
    unsigned short d1[kMlpH * 16], d2[kMlpH * 16], d3[kMlpH * 16];
};

// lane (q, c): the 4 samples 4q..4q+3 of row `row` as a 16x16x16 operand
__device__ __forceinline__ bf16x4 row4(const unsigned short *L, int row, int q) {
    const unsigned short *p = L + row * 16 + 4 * q;
    return bf16x4{(short)p[0], (short)p[1], (short)p[2], (short)p[3]};
}

// a B slab (K order kappa, slab p) of one lane into a [row][sample] tile
__device__ __forceinline__ void stage_slab(unsigned short *L, bf16x8 v, int p, int q, int c) {
    for (int j = 0; j < 8; ++j) L[kappa(p, q, j) * 16 + c] = (unsigned short)v[j];
}

// D tile mt of W^T * delta (rows 16mt+4q+i), times the ReLU derivative of the
// layer's activations (slab mt/2, half mt%2 of a)
__device__ __forceinline__ f32x4 relu_mask(f32x4 d, const bf16x8 (&a)[4], int mt) {
    const bf16x8 v = a[mt >> 1];
    const int h = (mt & 1) * 4;
    f32x4 r;
    for (int i = 0; i < 4; ++i) r[i] = v[h + i] > 0 ? d[i] : 0.0f;
    return r;
}

__global__ __launch_bounds__(64) void k_mlp_grad(const bf16x8 *__restrict__ frag, const float *__restrict__ bias,
                                                 const float *__restrict__ X, const float *__restrict__ T, long n,
                                                 float *__restrict__ part) {
    __shared__ float xs[16 * kMlpF];
    __shared__ GradLds L;
    const int lane = threadIdx.x, q = lane / 16, c = lane % 16;
    const float inv_n = 1.0f / (float)n;
    float *P = part + (size_t)blockIdx.x * kMlpParams;
    for (int p = lane; p < kMlpParams; p += 64) P[p] = 0.0f;
    __syncthreads();
    const long ntile = (n + 15) / 16;
#pragma unroll 1
    for (long t = blockIdx.x; t < ntile; t += gridDim.x) {
        const long s0 = t * 16;
        for (int j = lane; j < 16 * kMlpF; j += 64) {
            const long s = s0 + j / kMlpF;
            xs[j] = s < n ? X[s * kMlpF + j % kMlpF] : 0.0f;
        }
        __syncthreads();
        bf16x8 x[1] = {load_bx(xs, lane)};
        bf16x8 a1[1][4], a2[1][4], a3[1][4];
        f32x4 y[1];
        fwd_tiles<1>(frag, bias, x, lane, a1, a2, a3, y);
        // output error (lanes 0-15 hold outputs 0..3 of sample c)
        f32x4 dy = {0, 0, 0, 0};
        if (q == 0 && s0 + c < n) {
            dy[0] = (y[0][0] - T[(s0 + c) * 2]) * inv_n;
            dy[1] = (y[0][1] - T[(s0 + c) * 2 + 1]) * inv_n;
        }
        const bf16x4 dyB = pack4(dy);
        // back-propagation: delta3 = (W4^T dy) . [a3 > 0], delta2, delta1 likewise
        bf16x8 d3[4], d2[4], d1[4];
        f32x4 e[8];
        for (int mt = 0; mt < 8; ++mt) {
            const bf16x8 g = frag[(kFragG4 + mt) * 64 + lane];
            e[mt] = relu_mask(mfma16(g.s0123, dyB, f32x4{0, 0, 0, 0}), a3[0], mt);
        }
        for (int p = 0; p < 4; ++p) d3[p] = slab(e[2 * p], e[2 * p + 1]);
#pragma unroll 2
        for (int mt = 0; mt < 8; ++mt) {
            f32x4 acc = {0, 0, 0, 0};
            for (int k = 0; k < 4; ++k) acc = mfma32(frag[(kFragG3 + mt * 4 + k) * 64 + lane], d3[k], acc);
            e[mt] = relu_mask(acc, a2[0], mt);
        }
        for (int p = 0; p < 4; ++p) d2[p] = slab(e[2 * p], e[2 * p + 1]);
#pragma unroll 2
        for (int mt = 0; mt < 8; ++mt) {
            f32x4 acc = {0, 0, 0, 0};
            for (int k = 0; k < 4; ++k) acc = mfma32(frag[(kFragG2 + mt * 4 + k) * 64 + lane], d2[k], acc);
            e[mt] = relu_mask(acc, a1[0], mt);
        }
        for (int p = 0; p < 4; ++p) d1[p] = slab(e[2 * p], e[2 * p + 1]);
        // stage [row][sample] tiles
        if (q < 2)
            for (int j = 0; j < 8; ++j) L.x[(8 * q + j) * 16 + c] = (unsigned short)x[0][j];
        for (int i = 0; i < 4; ++i) L.dy[(4 * q + i) * 16 + c] = (unsigned short)dyB[i];
        for (int p = 0; p < 4; ++p) {
            stage_slab(L.a1, a1[0][p], p, q, c);
            stage_slab(L.a2, a2[0][p], p, q, c);
            stage_slab(L.a3, a3[0][p], p, q, c);
            stage_slab(L.d1, d1[p], p, q, c);
            stage_slab(L.d2, d2[p], p, q, c);
            stage_slab(L.d3, d3[p], p, q, c);
        }
        __syncthreads();
        // weight gradients dW[o][i] = sum_s delta[o][s] act[i][s] (D: rows 16mo+4q+i, col c)
#pragma unroll 1
        for (int mo = 0; mo < 8; ++mo) {
            const f32x4 g1 = mfma16(row4(L.d1, 16 * mo + c, q), row4(L.x, c, q), f32x4{0, 0, 0, 0});
            for (int i = 0; i < 4; ++i) P[kMlpW1 + (16 * mo + 4 * q + i) * kMlpIn + c] += g1[i];
            const bf16x4 a2r = row4(L.d2, 16 * mo + c, q), a3r = row4(L.d3, 16 * mo + c, q);
#pragma unroll 2
            for (int nt = 0; nt < 8; ++nt) {
                const f32x4 g2 = mfma16(a2r, row4(L.a1, 16 * nt + c, q), f32x4{0, 0, 0, 0});
                const f32x4 g3 = mfma16(a3r, row4(L.a2, 16 * nt + c, q), f32x4{0, 0, 0, 0});
                for (int i = 0; i < 4; ++i) {
                    P[kMlpW2 + (16 * mo + 4 * q + i) * kMlpH + 16 * nt + c] += g2[i];
                    P[kMlpW3 + (16 * mo + 4 * q + i) * kMlpH + 16 * nt + c] += g3[i];
                }
            }
        }
        {
            const bf16x4 ar = row4(L.dy, c, q);
            for (int nt = 0; nt < 8; ++nt) {
                const f32x4 g4 = mfma16(ar, row4(L.a3, 16 * nt + c, q), f32x4{0, 0, 0, 0});
                for (int i = 0; i < 4; ++i) P[kMlpW4 + (4 * q + i) * kMlpH + 16 * nt + c] += g4[i];
            }
        }
        // bias gradients: the BF16 errors summed over the tile's samples in order
#pragma unroll 1
        for (int r = lane; r < kMlpBias; r += 64) {
            const unsigned short *row = r < kMlpH ? L.d1 + r * 16
                                        : r < 2 * kMlpH ? L.d2 + (r - kMlpH) * 16
                                        : r < 3 * kMlpH ? L.d3 + (r - 2 * kMlpH) * 16
                                                        : L.dy + (r - 3 * kMlpH) * 16;
            float sum = 0.0f;
            for (int sI = 0; sI < 16; ++sI) sum += bf2f(row[sI]);
            const int dst = r < kMlpH ? kMlpB1 + r
                            : r < 2 * kMlpH ? kMlpB2 + r - kMlpH
                            : r < 3 * kMlpH ? kMlpB3 + r - 2 * kMlpH
                                            : kMlpB4 + r - 3 * kMlpH;
            P[dst] += sum;
        }
        __syncthreads();
    }
}

// Sum the per-wave partials in wave order (deterministic) into g.
__global__ __launch_bounds__(256) void k_mlp_reduce(const float *__restrict__ part, int nparts, float *__restrict__ g) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= kMlpParams) return;
    float s = 0.0f;
    for (int b = 0; b < nparts; ++b) s += part[(size_t)b * kMlpParams + p];
    g[p] = s;
}

// Adam (Kingma & Ba 2015) on the masters; padded weight entries stay zero.
__global__ __launch_bounds__(256) void k_mlp_adam(float *__restrict__ w, float *__restrict__ m, float *__restrict__ v,
                                                  const float *__restrict__ g, float lr, float b1, float b2,
                                                  float eps, int step) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= kMlpParams || !mlp_param_used(p)) return;
    const float gi = g[p];
    const float mi = b1 * m[p] + (1.0f - b1) * gi;
    const float vi = b2 * v[p] + (1.0f - b2) * gi * gi;
    m[p] = mi;
    v[p] = vi;
    const float mh = mi / (1.0f - powf(b1, (float)step));
    const float vh = vi / (1.0f - powf(b2, (float)step));
    w[p] -= lr * mh / (sqrtf(vh) + eps);
}

}  // namespace fcgm

// ---------------------------------------------------------------------------
// C-ABI of the network object (training and stand-alone inference)
// ---------------------------------------------------------------------------
namespace {

constexpr int kMaxParts = 128;  // waves per gradient launch (each sums several 16-sample tiles)

int mlp_fail(hipError_t e) { return e == hipSuccess ? FCG_OK : FCG_E_HIP; }
#define MLP_CHECK(x)                         \
    do {                                     \
        hipError_t e__ = (x);                \
        if (e__ != hipSuccess) return FCG_E_HIP; \
    } while (0)

int mlp_cast(fcg_mlp *net) {
    hipLaunchKernelGGL(fcgm::k_mlp_cast, dim3((kMlpFragCount * 64 + 255) / 256), dim3(256), 0, net->stream, net->w,
                       (fcgm::bf16x8 *)net->frag, net->bias);
    return mlp_fail(hipGetLastError());
}

// gradient of the batch into g, with the per-wave partials in part
// (kMaxParts x kMlpParams floats), on stream st
int mlp_grad_into(fcg_mlp *net, const float *X, const float *T, long n, float *part, float *g, hipStream_t st) {
    const long ntile = (n + 15) / 16;
    const int parts = (int)std::min<long>(kMaxParts, ntile);
    hipLaunchKernelGGL(fcgm::k_mlp_grad, dim3(parts), dim3(64), 0, st, (const fcgm::bf16x8 *)net->frag,
                       (const float *)net->bias, X, T, n, part);
    hipLaunchKernelGGL(fcgm::k_mlp_reduce, dim3((kMlpParams + 255) / 256), dim3(256), 0, st, (const float *)part,
                       parts, g);
    return mlp_fail(hipGetLastError());
}

int mlp_grad(fcg_mlp *net, const float *X, const float *T, long n) {
    return mlp_grad_into(net, X, T, n, net->part, net->g, net->stream);
}

// Partial sums of fcg_mlp_gradient_static: one array per GPU for the whole
// process (128 x 39312 floats, 20 MB), addressed through the device pointer
// looked up once per device in fcg_mlp_create.  Each wave clears its own
// partial before adding its tiles into it, and the reduction reads only the
// partials of the waves of the same launch, so nothing needs clearing between
// calls; calls that overlap on one GPU race on it.
constexpr int kMaxDevices = 64;
static __device__ __align__(256) float scratch[kMaxParts * kMlpParams];
float *g_scratch_ptr[kMaxDevices] = {};
std::mutex g_scratch_mu;

int scratch_lookup(int dev) {
    if (dev < 0 || dev >= kMaxDevices) return FCG_E_ARG;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[dev]) return FCG_OK;
    void *ptr = nullptr;
    if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) return FCG_E_HIP;
    g_scratch_ptr[dev] = (float *)ptr;
    return FCG_OK;
}

}  // namespace

extern "C" int fcg_mlp_create(unsigned seed, fcg_mlp **out) {
    if (!out) return FCG_E_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return FCG_E_NOGPU;
    auto *net = new fcg_mlp();
    *out = net;
    MLP_CHECK(hipGetDevice(&net->device));
    if (int rc = scratch_lookup(net->device)) return rc;  // device address of the static partials
    MLP_CHECK(hipStreamCreateWithFlags(&net->stream, hipStreamNonBlocking));
    const size_t pb = sizeof(float) * kMlpParams;
    MLP_CHECK(hipMalloc((void **)&net->w, pb));
    MLP_CHECK(hipMalloc((void **)&net->m, pb));
    MLP_CHECK(hipMalloc((void **)&net->v, pb));
    MLP_CHECK(hipMalloc((void **)&net->g, pb));
    MLP_CHECK(hipMalloc((void **)&net->part, pb * kMaxParts));
    MLP_CHECK(hipMalloc(&net->frag, sizeof(fcgm::bf16x8) * 64 * kMlpFragCount));
    MLP_CHECK(hipMalloc((void **)&net->bias, sizeof(float) * kMlpBias));
    MLP_CHECK(hipMemset(net->m, 0, pb));
    MLP_CHECK(hipMemset(net->v, 0, pb));
    // He-uniform weights (ReLU layers; Glorot for the linear output), zero biases and padding
    std::vector<float> w(kMlpParams, 0.0f);
    std::mt19937 rng(seed);
    auto fill = [&](int off, int rows, int cols, int stride, int fan_in, int fan_out) {
        const float a = std::sqrt(6.0f / (float)(fan_in + fan_out));
        std::uniform_real_distribution<float> u(-a, a);
        for (int r = 0; r < rows; ++r)
            for (int c = 0; c < cols; ++c) w[off + r * stride + c] = u(rng);
    };
    fill(kMlpW1, kMlpH, kMlpF, kMlpIn, kMlpF, 0);       // bound sqrt(6/fan_in)
    fill(kMlpW2, kMlpH, kMlpH, kMlpH, kMlpH, 0);
    fill(kMlpW3, kMlpH, kMlpH, kMlpH, kMlpH, 0);
    fill(kMlpW4, kMlpOut, kMlpH, kMlpH, kMlpH, kMlpOut);  // sqrt(6/(fan_in+fan_out))
    MLP_CHECK(hipMemcpy(net->w, w.data(), pb, hipMemcpyHostToDevice));
    for (int f = 0; f < kMlpF; ++f) {
        net->norm.x_mean[f] = 0.0f;
        net->norm.x_std[f] = 1.0f;
        net->norm.x_lo[f] = -1e30f;
        net->norm.x_hi[f] = 1e30f;
    }
    for (int o = 0; o < kMlpOut; ++o) {
        net->norm.y_mean[o] = 0.0f;
        net->norm.y_std[o] = 1.0f;
    }
    if (int rc = mlp_cast(net)) return rc;
    MLP_CHECK(hipStreamSynchronize(net->stream));
    return FCG_OK;
}

extern "C" void fcg_mlp_destroy(fcg_mlp *net) {
    if (!net) return;
    if (net->stream) (void)hipStreamSynchronize(net->stream);
    for (void *p : {(void *)net->w, (void *)net->m, (void *)net->v, (void *)net->g, (void *)net->part, net->frag,
                    (void *)net->bias})
        if (p) (void)hipFree(p);
    if (net->stream) (void)hipStreamDestroy(net->stream);
    delete net;
}

extern "C" int fcg_mlp_get_params(fcg_mlp *net, float *host, long count) {
    if (!net || !host || count != kMlpParams) return FCG_E_ARG;
    MLP_CHECK(hipStreamSynchronize(net->stream));
    MLP_CHECK(hipMemcpy(host, net->w, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost));
    return FCG_OK;
}

extern "C" int fcg_mlp_set_params(fcg_mlp *net, const float *host, long count) {
    if (!net || !host || count != kMlpParams) return FCG_E_ARG;
    std::vector<float> w(host, host + kMlpParams);
    for (int p = 0; p < kMlpParams; ++p)
        if (!mlp_param_used(p)) w[p] = 0.0f;
    MLP_CHECK(hipStreamSynchronize(net->stream));
    MLP_CHECK(hipMemcpy(net->w, w.data(), sizeof(float) * kMlpParams, hipMemcpyHostToDevice));
    if (int rc = mlp_cast(net)) return rc;
    MLP_CHECK(hipStreamSynchronize(net->stream));
    return FCG_OK;
}

extern "C" int fcg_mlp_set_norm(fcg_mlp *net, const float *x_mean, const float *x_std, const float *y_mean,
                                const float *y_std, const float *x_lo, const float *x_hi) {
    if (!net || !x_mean || !x_std || !y_mean || !y_std || !x_lo || !x_hi) return FCG_E_ARG;
    for (int f = 0; f < kMlpF; ++f) {
        if (!(x_std[f] > 0.0f)) return FCG_E_ARG;
        net->norm.x_mean[f] = x_mean[f];
        net->norm.x_std[f] = x_std[f];
        net->norm.x_lo[f] = x_lo[f];
        net->norm.x_hi[f] = x_hi[f];
    }
    for (int o = 0; o < kMlpOut; ++o) {
        if (!(y_std[o] > 0.0f)) return FCG_E_ARG;
        net->norm.y_mean[o] = y_mean[o];
        net->norm.y_std[o] = y_std[o];
    }
    return FCG_OK;
}

extern "C" int fcg_mlp_forward(fcg_mlp *net, const float *X, float *Y, long n) {
    if (!net || !X || !Y || n <= 0) return FCG_E_ARG;
    const long ntile = (n + 15) / 16;
    hipLaunchKernelGGL(fcgm::k_mlp_forward, dim3((unsigned)std::min<long>((ntile + 3) / 4, 8192)), dim3(64), 0,
                       net->stream, (const fcgm::bf16x8 *)net->frag, (const float *)net->bias, X, Y, n);
    MLP_CHECK(hipGetLastError());
    MLP_CHECK(hipStreamSynchronize(net->stream));
    return FCG_OK;
}

extern "C" int fcg_mlp_gradient(fcg_mlp *net, const float *X, const float *T, long n, float *grad) {
    if (!net || !X || !T || !grad || n <= 0) return FCG_E_ARG;
    if (int rc = mlp_grad(net, X, T, n)) return rc;
    MLP_CHECK(hipStreamSynchronize(net->stream));
    MLP_CHECK(hipMemcpy(grad, net->g, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost));
    return FCG_OK;
}

extern "C" int fcg_mlp_gradient_static(fcg_mlp *net, const float *X, const float *T, long n, float *grad_dev,
                                       void *stream) {
    if (!net || !X || !T || !grad_dev) return FCG_E_ARG;
    if (n <= 0 || n > FCG_MLP_STATIC_MAX_SAMPLES) return FCG_E_ARG;  // batch outside 1..2^31-1
    if (net->device < 0 || net->device >= kMaxDevices || !g_scratch_ptr[net->device]) return FCG_E_ARG;
    hipStream_t st = stream ? (hipStream_t)stream : net->stream;
    return mlp_grad_into(net, X, T, n, g_scratch_ptr[net->device], grad_dev, st);
}

extern "C" int fcg_mlp_train_step(fcg_mlp *net, const float *X, const float *T, long n, float lr) {
    if (!net || !X || !T || n <= 0 || !(lr > 0.0f)) return FCG_E_ARG;
    if (int rc = mlp_grad(net, X, T, n)) return rc;
    net->step += 1;
    hipLaunchKernelGGL(fcgm::k_mlp_adam, dim3((kMlpParams + 255) / 256), dim3(256), 0, net->stream, net->w, net->m,
                       net->v, (const float *)net->g, lr, 0.9f, 0.999f, 1e-8f, net->step);
    MLP_CHECK(hipGetLastError());
    if (int rc = mlp_cast(net)) return rc;
    MLP_CHECK(hipStreamSynchronize(net->stream));
    return FCG_OK;
}
