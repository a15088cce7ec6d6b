// Shapes, parameter layout and device building blocks of the learned column
// closure (internal): a per-(column, level) network 5 -> 128 -> 128 -> 128 ->
// 2 with ReLU hidden layers, BF16 weights and activations, FP32 accumulation.
//
// Matrix-core layout (gfx950 v_mfma_f32_16x16x32_bf16: lane l = 16q + c holds
// A[row c][k 8q..8q+7], B[k 8q..8q+7][col c], D[rows 4q..4q+3][col c]).
// Layers are evaluated transposed, D = W * A, weights as the A operand and
// activations (features x 16 samples) as B.  A layer's 8 D tiles (rows 16mt
// + 4q + i) become the next layer's B operand without moving data: slab p
// (hidden rows 32p..32p+31) takes tiles 2p and 2p+1, so lane group q, slot j
// of the B operand holds hidden row kappa_p(q, j) = 32p + 4q + j (j < 4) or
// 32p + 16 + 4q + j - 4 (j >= 4); the weight fragments are built with the
// same K order.  Back-propagation uses the transposed weights the same way.
#pragma once

#include <hip/hip_runtime.h>

constexpr int kMlpF = 5;      // input features per (column, level) sample
constexpr int kMlpIn = 32;    // padded input width (row stride of W1): one K slab
constexpr int kMlpH = 128;    // hidden width
constexpr int kMlpOut = 2;    // log10 snu, log10 Kv
constexpr int kMlpOutP = 16;  // padded output rows
// flat FP32 parameters: W1[128][32] b1[128] W2[128][128] b2[128] W3[128][128]
// b3[128] W4[16][128] b4[16]
constexpr int kMlpW1 = 0;
constexpr int kMlpB1 = kMlpW1 + kMlpH * kMlpIn;
constexpr int kMlpW2 = kMlpB1 + kMlpH;
constexpr int kMlpB2 = kMlpW2 + kMlpH * kMlpH;
constexpr int kMlpW3 = kMlpB2 + kMlpH;
constexpr int kMlpB3 = kMlpW3 + kMlpH * kMlpH;
constexpr int kMlpW4 = kMlpB3 + kMlpH;
constexpr int kMlpB4 = kMlpW4 + kMlpOutP * kMlpH;
constexpr int kMlpParams = kMlpB4 + kMlpOutP;  // 39312
constexpr int kMlpBias = 3 * kMlpH + kMlpOutP;  // biases b1..b4, 400 floats
// BF16 operand fragments, 64 lanes x 8 values (16 bytes) each: forward
// F1[8] F2[8][4] F3[8][4] F4[4]; backward G4[8] (4 values used, 16x16x16),
// G3[8][4] G2[8][4]
constexpr int kFragF1 = 0, kFragF2 = 8, kFragF3 = 40, kFragF4 = 72, kFragFwd = 76;
constexpr int kFragG4 = 76, kFragG3 = 84, kFragG2 = 116;
constexpr int kMlpFragCount = 148;

// The fused closure kernel: threads per (persistent) block and its LDS: the
// forward fragments, biases and the sigma levels, shared by the block, then one record per
// wavefront: the features and outputs of its 64 samples, per sample the four
// density coefficients and the two velocity components (doubles), and three
// doubles and a flag for each of its at most 16 columns (4 <= nsigma <= 64:
// a wavefront holds 64/nsigma whole columns).
constexpr int kMlpT = 512;
constexpr int kMlpMinK = 4, kMlpMaxK = 64;
constexpr size_t kMlpWaveLds = sizeof(float) * 64 * (kMlpF + 2) + sizeof(double) * 6 * 64 +
                               (sizeof(double) * 5 + sizeof(int)) * (64 / kMlpMinK);
// (the block's sigma tables: sigma, 1-sigma, and per interface 1/d, 1/d^2)
constexpr size_t kMlpClosureLds = 16 * kFragFwd * 64 + sizeof(float) * kMlpBias + 4 * sizeof(double) * kMlpMaxK +
                                  (kMlpT / 64) * kMlpWaveLds;

// Entries that are real weights (the rest is zero padding and stays zero).
__host__ __device__ inline bool mlp_param_used(int p) {
    if (p < kMlpB1) return (p - kMlpW1) % kMlpIn < kMlpF;
    if (p >= kMlpW4 && p < kMlpB4) return (p - kMlpW4) / kMlpH < kMlpOut;
    if (p >= kMlpB4) return p - kMlpB4 < kMlpOut;
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
    void *frag = nullptr;    // BF16 operand fragments (kMlpFragCount x 64 lanes x 16 B)
    float *bias = nullptr;   // b1 b2 b3 b4 contiguous (kMlpBias)
    int nparts = 0;   // waves per gradient launch
    int step = 0;     // Adam step count
    int device = 0;   // the HIP device the network lives on
    MlpNorm norm{};
    hipStream_t stream = nullptr;
};

namespace fcgm {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef short bf16x8 __attribute__((ext_vector_type(8)));
typedef short short2v __attribute__((ext_vector_type(2)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x4v __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8v __attribute__((ext_vector_type(8)));

__host__ __device__ inline int kappa(int p, int q, int j) {  // hidden row of slab p, lane group q, slot j
    return 32 * p + (j < 4 ? 4 * q + j : 16 + 4 * q + (j - 4));
}

__device__ __forceinline__ unsigned short f2bf(float f) {  // round to nearest even
    unsigned u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return 0x7fc0;
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf2f(unsigned short h) { return __uint_as_float((unsigned)h << 16); }

typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2v __attribute__((ext_vector_type(2)));

// floats -> BF16 (round to nearest even): the gfx950 packed conversion, one
// instruction per pair (a 4-wide conversion is split into single lanes and
// re-packed by the compiler)
__device__ __forceinline__ unsigned pack2(float a, float b) {
    return __builtin_bit_cast(unsigned, __builtin_convertvector(f32x2{a, b}, bf16x2v));
}
__device__ __forceinline__ bf16x4 pack4(f32x4 v) {
    typedef unsigned u32x2 __attribute__((ext_vector_type(2)));
    return __builtin_bit_cast(bf16x4, u32x2{pack2(v[0], v[1]), pack2(v[2], v[3])});
}
// two D tiles -> one B slab (K order kappa), ReLU applied to the BF16 values:
// a BF16 value is negative exactly when its sign bit is set, so a signed
// 16-bit maximum with 0 is the ReLU of the rounded value (= the rounded ReLU)
__device__ __forceinline__ bf16x8 slab_relu(f32x4 lo, f32x4 hi) {
    const bf16x4 a = pack4(lo), b = pack4(hi);
    bf16x8 r;
    r.s0123 = a;
    r.s4567 = b;
    const short2v z = {0, 0};
    short2v *p = reinterpret_cast<short2v *>(&r);
#pragma unroll
    for (int i = 0; i < 4; ++i) p[i] = __builtin_elementwise_max(p[i], z);
    return r;
}
__device__ __forceinline__ bf16x8 slab(f32x4 lo, f32x4 hi) {
    bf16x8 r;
    r.s0123 = pack4(lo);
    r.s4567 = pack4(hi);
    return r;
}

__device__ __forceinline__ f32x4 mfma32(bf16x8 a, bf16x8 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8v, a), __builtin_bit_cast(bf16x8v, b), c,
                                                   0, 0, 0);
}
__device__ __forceinline__ f32x4 mfma16(bf16x4 a, bf16x4 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// bias fragment of D tile mt: lane l holds bias[16mt + 4(l/16) + i]
__device__ __forceinline__ f32x4 bias4(const float *b, int mt, int lane) {
    const int r = 16 * mt + 4 * (lane >> 4);
    return f32x4{b[r], b[r + 1], b[r + 2], b[r + 3]};
}

// Features of a 16-sample tile from xs[16][kMlpF] (normalised) -> the B
// operand of layer 1 (natural K order: lane group q, slot j = feature 8q+j).
__device__ __forceinline__ bf16x8 load_bx(const float *xs, int lane) {
    const int q = lane >> 4, smp = lane & 15;
    f32x4 lo = {0, 0, 0, 0}, hi = {0, 0, 0, 0};
    if (q == 0) {
        lo = f32x4{xs[smp * kMlpF], xs[smp * kMlpF + 1], xs[smp * kMlpF + 2], xs[smp * kMlpF + 3]};
        hi[0] = xs[smp * kMlpF + 4];
    }
    return slab(lo, hi);
}

// One hidden layer (128 -> 128) on T tiles: output slab p = D tiles 2p, 2p+1
// over the 4 input K slabs, bias in the accumulator, ReLU on the BF16 values.
// The (p, k) steps run in one sequence whose next pair of weight fragments is
// read while the current step's products issue, so the LDS latency hides
// behind the matrix cores.
template <int T>
__device__ __forceinline__ void hidden_layer(const bf16x8 *__restrict__ frag, int F, const float *__restrict__ bl,
                                             const bf16x8 (&in)[T][4], bf16x8 (&out)[T][4], int lane) {
    bf16x8 w0 = frag[(F + 0) * 64 + lane], w1 = frag[(F + 4) * 64 + lane];
    f32x4 c0[T], c1[T];
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        const int p = i >> 2, k = i & 3;
        bf16x8 n0 = w0, n1 = w1;
        if (i < 15) {  // fragments of the next step: (p, k+1) or (p+1, 0)
            const int pn = (i + 1) >> 2, kn = (i + 1) & 3;
            n0 = frag[(F + (2 * pn) * 4 + kn) * 64 + lane];
            n1 = frag[(F + (2 * pn + 1) * 4 + kn) * 64 + lane];
        }
        __builtin_amdgcn_sched_barrier(0);  // keep the reads ahead of this step's products
        if (k == 0) {
            const f32x4 z0 = bias4(bl, 2 * p, lane), z1 = bias4(bl, 2 * p + 1, lane);
#pragma unroll
            for (int t = 0; t < T; ++t) {
                c0[t] = z0;
                c1[t] = z1;
            }
        }
#pragma unroll
        for (int t = 0; t < T; ++t) {
            c0[t] = mfma32(w0, in[t][k], c0[t]);
            c1[t] = mfma32(w1, in[t][k], c1[t]);
        }
        if (k == 3) {
#pragma unroll
            for (int t = 0; t < T; ++t) out[t][p] = slab_relu(c0[t], c1[t]);
        }
        w0 = n0;
        w1 = n1;
    }
}

// Forward of T tiles of 16 samples held by one wavefront.  frag: F1..F4 (in
// LDS or global memory); bias: b1 b2 b3 b4 (400 floats).  a1..a3 receive the
// hidden activations (BF16 B slabs, after ReLU); y[t] holds, in lanes 0-15,
// outputs 0..3 (0 and 1 used) of sample lane of tile t.
template <int T>
__device__ __forceinline__ void fwd_tiles(const bf16x8 *__restrict__ frag, const float *__restrict__ bias,
                                          const bf16x8 (&x)[T], int lane, bf16x8 (&a1)[T][4], bf16x8 (&a2)[T][4],
                                          bf16x8 (&a3)[T][4], f32x4 (&y)[T]) {
#pragma unroll
    for (int p = 0; p < 4; ++p) {  // layer 1: one K slab of inputs
        const bf16x8 w0 = frag[(kFragF1 + 2 * p) * 64 + lane], w1 = frag[(kFragF1 + 2 * p + 1) * 64 + lane];
        const f32x4 c0 = bias4(bias, 2 * p, lane), c1 = bias4(bias, 2 * p + 1, lane);
#pragma unroll
        for (int t = 0; t < T; ++t) a1[t][p] = slab_relu(mfma32(w0, x[t], c0), mfma32(w1, x[t], c1));
    }
    hidden_layer<T>(frag, kFragF2, bias + kMlpH, a1, a2, lane);
    hidden_layer<T>(frag, kFragF3, bias + 2 * kMlpH, a2, a3, lane);
    const f32x4 z = bias4(bias + 3 * kMlpH, 0, lane);  // output layer: rows 0..15 (0, 1 used)
#pragma unroll
    for (int t = 0; t < T; ++t) y[t] = z;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const bf16x8 w = frag[(kFragF4 + k) * 64 + lane];
#pragma unroll
        for (int t = 0; t < T; ++t) y[t] = mfma32(w, a3[t][k], y[t]);
    }
}

// Inference only: the same arithmetic, two activation buffers in turn.
template <int T>
__device__ __forceinline__ void infer_tiles(const bf16x8 *__restrict__ frag, const float *__restrict__ bias,
                                            const bf16x8 (&x)[T], int lane, f32x4 (&y)[T]) {
    bf16x8 a[T][4], b[T][4];
#pragma unroll
    for (int p = 0; p < 4; ++p) {
        const bf16x8 w0 = frag[(kFragF1 + 2 * p) * 64 + lane], w1 = frag[(kFragF1 + 2 * p + 1) * 64 + lane];
        const f32x4 c0 = bias4(bias, 2 * p, lane), c1 = bias4(bias, 2 * p + 1, lane);
#pragma unroll
        for (int t = 0; t < T; ++t) a[t][p] = slab_relu(mfma32(w0, x[t], c0), mfma32(w1, x[t], c1));
    }
    hidden_layer<T>(frag, kFragF2, bias + kMlpH, a, b, lane);
    hidden_layer<T>(frag, kFragF3, bias + 2 * kMlpH, b, a, lane);
    const f32x4 z = bias4(bias + 3 * kMlpH, 0, lane);
#pragma unroll
    for (int t = 0; t < T; ++t) y[t] = z;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const bf16x8 w = frag[(kFragF4 + k) * 64 + lane];
#pragma unroll
        for (int t = 0; t < T; ++t) y[t] = mfma32(w, a[t][k], y[t]);
    }
}

}  // namespace fcgm
