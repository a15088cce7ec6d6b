// Learned vertical closure: a per-(column, level) MLP 5 -> 64 -> 64 -> 2 with
// tanh hidden layers, standing in for d3_end (src/fv_mixing.F90:73-347).
//
// All products run on the matrix cores: BF16 operands, FP32 accumulation,
// v_mfma_f32_16x16x16_bf16.  Layers are evaluated transposed, D = W * A with
// the weights as the A operand and activations (features x samples) as the B
// operand, so the D fragment of one layer is already the B fragment of the
// next (lane l holds rows 4*(l/16)+i of sample column l%16 in both).  The
// backward pass uses W^T fragments the same way; the weight gradients
// (sum over samples) stage activations through LDS as [feature][sample] so
// that both MFMA operands read 4 consecutive samples per lane.
//
// Parameters are FP32 masters in one flat array (layout below); a cast kernel
// rebuilds the BF16 fragment arrays after each update.  Gradients are summed
// per wave, then over waves in a fixed order (deterministic), and applied by
// Adam.
#include <algorithm>
#include <cmath>
#include <mutex>
#include <random>
#include <vector>

#include "fcg_internal.h"
#include "fcg_mlp.h"

namespace fcgm {

using fcgm::bf16x4;

// ---- fragment construction (master FP32 -> BF16 A operands) -------------
// F1[mt][lane]      W1 rows 16mt+l%16, cols 4(l/16)+i
// F2[mt][kt][lane]  W2 rows 16mt+l%16, cols 16kt+4(l/16)+i
// F3[kt][lane]      W3 rows l%16,      cols 16kt+4(l/16)+i
// G3[mt][lane]      W3^T rows 16mt+l%16 (hidden), cols 4(l/16)+i (output)
// G2[mt][kt][lane]  W2^T rows 16mt+l%16, cols 16kt+4(l/16)+i
__global__ __launch_bounds__(256) void k_mlp_cast(const float *__restrict__ w, bf16x4 *__restrict__ frag) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= kMlpFragCount * 64) return;
    const int f = t / 64, l = t % 64, r = l % 16, c0 = 4 * (l / 16);
    const float *W1 = w + kMlpW1, *W2 = w + kMlpW2, *W3 = w + kMlpW3;
    f32x4 v;
    if (f < 4) {  // F1
        const int mt = f;
        for (int i = 0; i < 4; ++i) v[i] = W1[(16 * mt + r) * kMlpIn + c0 + i];
    } else if (f < 20) {  // F2
        const int mt = (f - 4) / 4, kt = (f - 4) % 4;
        for (int i = 0; i < 4; ++i) v[i] = W2[(16 * mt + r) * kMlpH + 16 * kt + c0 + i];
    } else if (f < 24) {  // F3
        const int kt = f - 20;
        for (int i = 0; i < 4; ++i) v[i] = W3[r * kMlpH + 16 * kt + c0 + i];
    } else if (f < 28) {  // G3
        const int mt = f - 24;
        for (int i = 0; i < 4; ++i) v[i] = W3[(c0 + i) * kMlpH + 16 * mt + r];
    } else {  // G2
        const int mt = (f - 28) / 4, kt = (f - 28) % 4;
        for (int i = 0; i < 4; ++i) v[i] = W2[(16 * kt + c0 + i) * kMlpH + 16 * mt + r];
    }
    frag[t] = pack4(v);
}

// ---- inference on normalized features: X[n][kMlpF] -> Y[n][2] -----------
__global__ __launch_bounds__(64) void k_mlp_forward(const float *__restrict__ w, const bf16x4 *__restrict__ frag,
                                                    const float *__restrict__ X, float *__restrict__ Y, long n) {
    __shared__ float xs[16 * kMlpF];
    const int lane = threadIdx.x;
    const long ntile = (n + 15) / 16;
    for (long t = blockIdx.x; t < ntile; t += gridDim.x) {
        const long s0 = t * 16;
        for (int j = lane; j < 16 * kMlpF; j += 64) {
            const long s = s0 + j / kMlpF;
            xs[j] = s < n ? X[s * kMlpF + j % kMlpF] : 0.0f;
        }
        __syncthreads();
        const f32x4 y = fwd_infer(frag, w, load_bx(xs, lane), lane);
        if (lane < 16 && s0 + lane < n) {
            Y[(s0 + lane) * 2] = y[0];
            Y[(s0 + lane) * 2 + 1] = y[1];
        }
        __syncthreads();
    }
}

// ---- training: per-wave gradients of 0.5*mean over the batch of |y - t|^2 -----
// (the factor 1/n and sample weights are folded into dy).  Each block (one
// wave) takes tiles t = blockIdx.x, blockIdx.x + gridDim.x, ... and writes
// its partial gradient to part[blockIdx.x][kMlpParams].
__global__ __launch_bounds__(64) void k_mlp_grad(const float *__restrict__ w, const bf16x4 *__restrict__ frag,
                                                 const float *__restrict__ X, const float *__restrict__ T, long n,
                                                 float *__restrict__ part) {
    __shared__ float xs[16 * kMlpF];
    __shared__ unsigned short xL[16 * 16], dyL[16 * 16], t1L[64 * 16], t2L[64 * 16], d2L[64 * 16], d1L[64 * 16];
    const int lane = threadIdx.x, q = lane / 16, smp = lane % 16;
    const bf16x4 *G3 = frag + 24 * 64, *G2 = frag + 28 * 64;
    const float inv_n = 1.0f / (float)n;
    f32x4 gW1[4], gW2[4][4], gW3[4], gb1[4], gb2[4], gb3;
    for (int a = 0; a < 4; ++a) {
        gW1[a] = f32x4{0, 0, 0, 0};
        gW3[a] = f32x4{0, 0, 0, 0};
        gb1[a] = f32x4{0, 0, 0, 0};
        gb2[a] = f32x4{0, 0, 0, 0};
        for (int b = 0; b < 4; ++b) gW2[a][b] = f32x4{0, 0, 0, 0};
    }
    gb3 = f32x4{0, 0, 0, 0};
    const long ntile = (n + 15) / 16;
    for (long t = blockIdx.x; t < ntile; t += gridDim.x) {
        const long s0 = t * 16;
        for (int j = lane; j < 16 * kMlpF; j += 64) {
            const long s = s0 + j / kMlpF;
            xs[j] = s < n ? X[s * kMlpF + j % kMlpF] : 0.0f;
        }
        __syncthreads();
        const bf16x4 bx = load_bx(xs, lane);
        Fwd o;
        fwd_tile(frag, w, bx, lane, o);
        // output error (D layout: lanes 0-15 hold outputs 0..3 of sample lane)
        const bool valid = s0 + smp < n;
        f32x4 dy = {0, 0, 0, 0};
        if (q == 0 && valid) {
            dy[0] = (o.y[0] - T[(s0 + smp) * 2]) * inv_n;
            dy[1] = (o.y[1] - T[(s0 + smp) * 2 + 1]) * inv_n;
        }
        const bf16x4 dyB = pack4(dy);
        // dh2 = (W3^T dy) * (1 - t2^2); dh1 = (W2^T dh2) * (1 - t1^2)
        f32x4 d2[4], d1[4];
        bf16x4 d2B[4];
        for (int mt = 0; mt < 4; ++mt) {
            f32x4 c = mfma(G3[mt * 64 + lane], dyB, f32x4{0, 0, 0, 0});
            for (int i = 0; i < 4; ++i) c[i] = valid ? c[i] * (1.0f - o.t2[mt][i] * o.t2[mt][i]) : 0.0f;
            d2[mt] = c;
            d2B[mt] = pack4(c);
        }
        for (int mt = 0; mt < 4; ++mt) {
            f32x4 c = {0, 0, 0, 0};
            for (int kt = 0; kt < 4; ++kt) c = mfma(G2[(mt * 4 + kt) * 64 + lane], d2B[kt], c);
            for (int i = 0; i < 4; ++i) c[i] = valid ? c[i] * (1.0f - o.t1[mt][i] * o.t1[mt][i]) : 0.0f;
            d1[mt] = c;
        }
        // bias gradients: per-lane partial sums, reduced over samples at the end
        for (int mt = 0; mt < 4; ++mt) {
            gb1[mt] += d1[mt];
            gb2[mt] += d2[mt];
        }
        gb3 += dy;
        // stage [feature][sample] BF16 tiles for the weight-gradient products
        for (int i = 0; i < 4; ++i) {
            const int f = 4 * q + i;
            xL[f * 16 + smp] = (unsigned short)bx[i];
            dyL[f * 16 + smp] = (unsigned short)dyB[i];
            for (int mt = 0; mt < 4; ++mt) {
                const int h = 16 * mt + f;
                t1L[h * 16 + smp] = (unsigned short)o.b2[mt][i];
                t2L[h * 16 + smp] = (unsigned short)o.b3[mt][i];
                d2L[h * 16 + smp] = (unsigned short)d2B[mt][i];
                d1L[h * 16 + smp] = f2bf(d1[mt][i]);
            }
        }
        __syncthreads();
        auto frag_of = [&](const unsigned short *L, int row) {
            bf16x4 r;
            for (int i = 0; i < 4; ++i) r[i] = (short)L[row * 16 + 4 * q + i];
            return r;
        };
        // dW3[o][h] = sum_s dy[o][s] t2[h][s]
        const bf16x4 aDy = frag_of(dyL, smp);
        for (int nt = 0; nt < 4; ++nt) gW3[nt] = mfma(aDy, frag_of(t2L, 16 * nt + smp), gW3[nt]);
        // dW2[o][h] = sum_s dh2[o][s] t1[h][s]
        for (int mo = 0; mo < 4; ++mo) {
            const bf16x4 a = frag_of(d2L, 16 * mo + smp);
            for (int nt = 0; nt < 4; ++nt) gW2[mo][nt] = mfma(a, frag_of(t1L, 16 * nt + smp), gW2[mo][nt]);
        }
        // dW1[o][in] = sum_s dh1[o][s] x[in][s]
        const bf16x4 bX = frag_of(xL, smp);
        for (int mo = 0; mo < 4; ++mo) gW1[mo] = mfma(frag_of(d1L, 16 * mo + smp), bX, gW1[mo]);
        __syncthreads();
    }
    // reduce bias partials over the 16 sample lanes of each lane group
    for (int off = 1; off < 16; off <<= 1) {
        for (int mt = 0; mt < 4; ++mt)
            for (int i = 0; i < 4; ++i) {
                gb1[mt][i] += __shfl_xor(gb1[mt][i], off);
                gb2[mt][i] += __shfl_xor(gb2[mt][i], off);
            }
        for (int i = 0; i < 4; ++i) gb3[i] += __shfl_xor(gb3[i], off);
    }
    float *P = part + (size_t)blockIdx.x * kMlpParams;
    // D layout: lane holds rows 4q+i (+16*tile) of column smp (+16*tile)
    for (int i = 0; i < 4; ++i) {
        const int r = 4 * q + i;
        for (int mo = 0; mo < 4; ++mo) {
            P[kMlpW1 + (16 * mo + r) * kMlpIn + smp] = gW1[mo][i];
            for (int nt = 0; nt < 4; ++nt) P[kMlpW2 + (16 * mo + r) * kMlpH + 16 * nt + smp] = gW2[mo][nt][i];
        }
        for (int nt = 0; nt < 4; ++nt) P[kMlpW3 + r * kMlpH + 16 * nt + smp] = gW3[nt][i];
        if (smp == 0) {
            for (int mt = 0; mt < 4; ++mt) {
                P[kMlpB1 + 16 * mt + r] = gb1[mt][i];
                P[kMlpB2 + 16 * mt + r] = gb2[mt][i];
            }
            P[kMlpB3 + r] = gb3[i];
        }
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

constexpr int kMaxParts = 256;  // waves per gradient launch

int mlp_fail(hipError_t e) { return e == hipSuccess ? FCG_OK : FCG_E_HIP; }
#define MLP_CHECK(x)                         \
    do {                                     \
        hipError_t e__ = (x);                \
        if (e__ != hipSuccess) return FCG_E_HIP; \
    } while (0)

int mlp_cast(fcg_mlp *net) {
    hipLaunchKernelGGL(fcgm::k_mlp_cast, dim3((kMlpFragCount * 64 + 255) / 256), dim3(256), 0, net->stream, net->w,
                       (fcgm::bf16x4 *)net->frag);
    return mlp_fail(hipGetLastError());
}

// gradient of the batch into g, with the per-wave partials in part
// (kMaxParts x kMlpParams floats), on stream st
int mlp_grad_into(fcg_mlp *net, const float *X, const float *T, long n, float *part, float *g, hipStream_t st) {
    const long ntile = (n + 15) / 16;
    const int parts = (int)std::min<long>(kMaxParts, ntile);
    hipLaunchKernelGGL(fcgm::k_mlp_grad, dim3(parts), dim3(64), 0, st, net->w, (const fcgm::bf16x4 *)net->frag, X,
                       T, n, part);
    hipLaunchKernelGGL(fcgm::k_mlp_reduce, dim3((kMlpParams + 255) / 256), dim3(256), 0, st, (const float *)part,
                       parts, g);
    return mlp_fail(hipGetLastError());
}

int mlp_grad(fcg_mlp *net, const float *X, const float *T, long n) {
    return mlp_grad_into(net, X, T, n, net->part, net->g, net->stream);
}

// Partial sums of fcg_mlp_gradient_static: one array per GPU for the whole
// process (6.4 MB), addressed through the device pointer looked up once per
// device in fcg_mlp_create.  Every wave writes all kMlpParams entries of its
// partial before the reduction reads them, so nothing needs clearing between
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
    MLP_CHECK(hipMalloc(&net->frag, sizeof(fcgm::bf16x4) * 64 * kMlpFragCount));
    MLP_CHECK(hipMemset(net->m, 0, pb));
    MLP_CHECK(hipMemset(net->v, 0, pb));
    // Glorot-uniform weights, zero biases and padding
    std::vector<float> w(kMlpParams, 0.0f);
    std::mt19937 rng(seed);
    auto fill = [&](int off, int rows, int cols, int stride, int fan_in, int fan_out) {
        const float a = std::sqrt(6.0f / (float)(fan_in + fan_out));
        std::uniform_real_distribution<float> u(-a, a);
        for (int r = 0; r < rows; ++r)
            for (int c = 0; c < cols; ++c) w[off + r * stride + c] = u(rng);
    };
    fill(kMlpW1, kMlpH, kMlpF, kMlpIn, kMlpF, kMlpH);
    fill(kMlpW2, kMlpH, kMlpH, kMlpH, kMlpH, kMlpH);
    fill(kMlpW3, kMlpOut, kMlpH, kMlpH, kMlpH, kMlpOut);
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
    for (void *p : {(void *)net->w, (void *)net->m, (void *)net->v, (void *)net->g, (void *)net->part, net->frag})
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
    hipLaunchKernelGGL(fcgm::k_mlp_forward, dim3((unsigned)std::min<long>(ntile, 8192)), dim3(64), 0, net->stream,
                       net->w, (const fcgm::bf16x4 *)net->frag, X, Y, n);
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
