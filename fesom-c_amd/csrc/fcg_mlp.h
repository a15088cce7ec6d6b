// Shapes and parameter layout of the learned column closure (internal).
#pragma once

#include <hip/hip_runtime.h>

constexpr int kMlpF = 5;     // input features per (column, level) sample
constexpr int kMlpIn = 16;   // padded input width (row stride of W1)
constexpr int kMlpH = 64;    // hidden width
constexpr int kMlpOut = 2;   // log10 snu, log10 Kv
// flat FP32 parameter array: W1[64][16] b1[64] W2[64][64] b2[64] W3[16][64] b3[16]
constexpr int kMlpW1 = 0;
constexpr int kMlpB1 = kMlpW1 + kMlpH * kMlpIn;
constexpr int kMlpW2 = kMlpB1 + kMlpH;
constexpr int kMlpB2 = kMlpW2 + kMlpH * kMlpH;
constexpr int kMlpW3 = kMlpB2 + kMlpH;
constexpr int kMlpB3 = kMlpW3 + 16 * kMlpH;
constexpr int kMlpParams = kMlpB3 + 16;
// BF16 A-operand fragments (64 lanes x 4 values each): F1 4, F2 16, F3 4, G3 4, G2 16
constexpr int kMlpFragCount = 44;

// Entries that are real weights (the rest is zero padding and stays zero).
__host__ __device__ inline bool mlp_param_used(int p) {
    if (p < kMlpB1) return (p - kMlpW1) % kMlpIn < kMlpF;
    if (p >= kMlpW3 && p < kMlpB3) return (p - kMlpW3) / kMlpH < kMlpOut;
    if (p >= kMlpB3) return p - kMlpB3 < kMlpOut;
    return true;
}

// Feature / target normalisation and the out-of-distribution box (normalised
// units), passed by value to the closure kernel.
struct MlpNorm {
    float x_mean[kMlpF], x_std[kMlpF], x_lo[kMlpF], x_hi[kMlpF];
    float y_mean[kMlpOut], y_std[kMlpOut];
};

struct fcg_mlp {
    float *w = nullptr, *m = nullptr, *v = nullptr, *g = nullptr, *part = nullptr;
    void *frag = nullptr;
    int nparts = 0;   // waves per gradient launch
    int step = 0;     // Adam step count
    int device = 0;   // the HIP device the network lives on
    MlpNorm norm{};
    hipStream_t stream = nullptr;
};

// ---- device building blocks shared by the training kernels and the fused
// closure kernel ------------------------------------------------------------
namespace fcgm {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ unsigned short f2bf(float f) {  // round to nearest even
    unsigned u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return 0x7fc0;
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf2f(unsigned short h) { return __uint_as_float((unsigned)h << 16); }

// Four floats -> BF16 (round to nearest even); the compiler emits the gfx950
// packed conversion v_cvt_pk_bf16_f32 for the vector convert.
typedef __bf16 bf16x4v __attribute__((ext_vector_type(4)));
__device__ __forceinline__ bf16x4 pack4(f32x4 v) {
    return __builtin_bit_cast(bf16x4, __builtin_convertvector(v, bf16x4v));
}

__device__ __forceinline__ f32x4 mfma(bf16x4 a, bf16x4 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// Bias fragment: lane l holds bias[16mt + 4(l/16) + i]
__device__ __forceinline__ f32x4 bias4(const float *b, int mt, int lane) {
    f32x4 c;
    for (int i = 0; i < 4; ++i) c[i] = b[16 * mt + 4 * (lane / 16) + i];
    return c;
}

// tanh(x) = 1 - 2/(exp(2x) + 1) on the hardware exp2 and reciprocal: a few
// instructions instead of a library call, ~1e-7 absolute error, far below the
// BF16 rounding of the activations that follows.
__device__ __forceinline__ float fast_tanh(float x) {
    return 1.0f - 2.0f * __builtin_amdgcn_rcpf(__expf(2.0f * x) + 1.0f);
}

__device__ __forceinline__ f32x4 tanh4(f32x4 v) {
    f32x4 r;
    for (int i = 0; i < 4; ++i) r[i] = fast_tanh(v[i]);
    return r;
}

struct Fwd {
    f32x4 t1[4], t2[4], y;  // activations (FP32) in D layout
    bf16x4 b2[4], b3[4];    // BF16 copies feeding the next layer
};

// Forward of one 16-sample tile.  bx: features as the B operand (lane l:
// features 4(l/16)+i of sample l%16, zero beyond the 5 inputs).
__device__ __forceinline__ void fwd_tile(const bf16x4 *__restrict__ frag, const float *__restrict__ w, bf16x4 bx,
                                         int lane, Fwd &o) {
    const bf16x4 *F1 = frag, *F2 = frag + 4 * 64, *F3 = frag + 20 * 64;
    for (int mt = 0; mt < 4; ++mt) {
        o.t1[mt] = tanh4(mfma(F1[mt * 64 + lane], bx, bias4(w + kMlpB1, mt, lane)));
        o.b2[mt] = pack4(o.t1[mt]);
    }
    for (int mt = 0; mt < 4; ++mt) {
        f32x4 c = bias4(w + kMlpB2, mt, lane);
        for (int kt = 0; kt < 4; ++kt) c = mfma(F2[(mt * 4 + kt) * 64 + lane], o.b2[kt], c);
        o.t2[mt] = tanh4(c);
        o.b3[mt] = pack4(o.t2[mt]);
    }
    f32x4 c = bias4(w + kMlpB3, 0, lane);
    for (int kt = 0; kt < 4; ++kt) c = mfma(F3[kt * 64 + lane], o.b3[kt], c);
    o.y = c;  // lanes 0-15: outputs 0..3 (0 and 1 used) of sample lane%16
}

// The same forward without keeping activations for a backward pass (fewer
// live registers); its outputs equal fwd_tile's bit for bit.
__device__ __forceinline__ f32x4 fwd_infer(const bf16x4 *__restrict__ frag, const float *__restrict__ w, bf16x4 bx,
                                           int lane) {
    const bf16x4 *F1 = frag, *F2 = frag + 4 * 64, *F3 = frag + 20 * 64;
    bf16x4 b2[4], b3[4];
    for (int mt = 0; mt < 4; ++mt) b2[mt] = pack4(tanh4(mfma(F1[mt * 64 + lane], bx, bias4(w + kMlpB1, mt, lane))));
    for (int mt = 0; mt < 4; ++mt) {
        f32x4 c = bias4(w + kMlpB2, mt, lane);
        for (int kt = 0; kt < 4; ++kt) c = mfma(F2[(mt * 4 + kt) * 64 + lane], b2[kt], c);
        b3[mt] = pack4(tanh4(c));
    }
    f32x4 c = bias4(w + kMlpB3, 0, lane);
    for (int kt = 0; kt < 4; ++kt) c = mfma(F3[kt * 64 + lane], b3[kt], c);
    return c;
}

// Features of a sample tile -> B fragment, from an LDS tile xs[16][kMlpF]
__device__ __forceinline__ bf16x4 load_bx(const float *xs, int lane) {
    f32x4 v;
    const int q = lane / 16, smp = lane % 16;
    for (int i = 0; i < 4; ++i) {
        const int f = 4 * q + i;
        v[i] = f < kMlpF ? xs[smp * kMlpF + f] : 0.0f;
    }
    return pack4(v);
}


}  // namespace fcgm
