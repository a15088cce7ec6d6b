// HIP kernels of the FESOM-C internal time step for gfx950 (FP64 drop-in).
//
// Every reference edge loop that scatters into nodes or elements
// (ssh_rhs, Wvel, tracer rhs, U_rhsAB, U_rhs_2D, U_rhs_2D_3D) is a
// deterministic gather here: one thread per receiving node/element walks its
// incident edges in ascending edge id, recomputing each edge flux from the
// same operands, so the summation order -- and in FP64 every bit -- equals
// the reference's serial scatter (no atomics).
//
// The file is compiled twice, once per storage type R of the state
// (FCG_STORAGE_DOUBLE = 1: the FP64 drop-in, with -ffp-contract=off to keep
// the reference's rounding; = 0: the mixed-precision path, where the
// compiler may fuse a multiplication and an addition written in one
// expression into one FMA, which rounds once instead of twice).  Each build
// instantiates the kernels of its own storage type; the kernels that do not
// depend on it, and the learned closure's, are built with the FP64 one.
//
// Vertical Jacobians are not stored: Jc = a*(sigma_k - sigma_k+1),
// zbar = (1-sigma_k)*a, Z = mid(zbar), Jd = (Jc1+Jc2)/2 are recomputed from
// the node scalar a = max(depth+eta_n, Dmin) (src/fv_main.F90:1228-1263),
// which gives the same bits as the reference's stored arrays.
#include "fcg_internal.h"
#include "fcg_logcr.h"
#include "fcg_mlp.h"

namespace fcgk {

// Loads through pointers of a named address space.  Where one branch reads a
// neighbour from the block's LDS columns and the other from global memory,
// these keep the two loads apart; through plain pointers the compiler merges
// them into one generic (flat) load behind a pointer select.
template <typename T>
__device__ __forceinline__ T lds_ld(const T *p) {
    return *(const __attribute__((address_space(3))) T *)p;
}
template <typename T>
__device__ __forceinline__ T glb_ld(const T *p) {
    return *(const __attribute__((address_space(1))) T *)p;
}

// Level vectors: W = 16/sizeof(R) consecutive levels of one column (4 in
// FP32, 2 in FP64), moved as one 16-byte access.  The vector kernels that use
// them need nsigma to be a multiple of W, so that every column starts on a
// 16-byte boundary; a lane then owns levels [W*j, W*j+W) of its column.
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
template <typename R>
struct LV {
    static constexpr int W = 16 / (int)sizeof(R);
    R x[W];
};
template <typename R>
__device__ __forceinline__ LV<R> ldv(const R *p) {
    const u32x4 u = *(const u32x4 *)p;
    LV<R> r;
    __builtin_memcpy(r.x, &u, 16);
    return r;
}
template <typename R>
__device__ __forceinline__ LV<R> ldv_lds(const R *p) {
    const u32x4 u = *(const __attribute__((address_space(3))) u32x4 *)p;
    LV<R> r;
    __builtin_memcpy(r.x, &u, 16);
    return r;
}
template <typename R>
__device__ __forceinline__ LV<R> ldv_glb(const R *p) {
    const u32x4 u = *(const __attribute__((address_space(1))) u32x4 *)p;
    LV<R> r;
    __builtin_memcpy(r.x, &u, 16);
    return r;
}
template <typename R>
__device__ __forceinline__ void stv(R *p, const LV<R> &v) {
    u32x4 u;
    __builtin_memcpy(&u, v.x, 16);
    *(u32x4 *)p = u;
}
// the first n (<= W) levels of v: a 16-byte store when n == W; the lane that
// holds a column's unused last slot stores its layers one by one, so that
// slot keeps what it held
template <typename R>
__device__ __forceinline__ void stv_n(R *p, const LV<R> &v, int n) {
    if (n >= LV<R>::W) {
        stv(p, v);
    } else {
#pragma unroll
        for (int i = 0; i < LV<R>::W; ++i)
            if (i < n) p[i] = v.x[i];
    }
}

// (2,E)/(4,E) interleaved records of the barotropic momentum path in the
// storage type, widened to double on load and rounded on store
__device__ __forceinline__ double2 ld2(const double *p, size_t i) { return ((const double2 *)p)[i]; }
__device__ __forceinline__ double2 ld2(const float *p, size_t i) {
    const float2 v = ((const float2 *)p)[i];
    return make_double2(v.x, v.y);
}
__device__ __forceinline__ void st2(double *p, size_t i, double a, double b) { ((double2 *)p)[i] = make_double2(a, b); }
__device__ __forceinline__ void st2(float *p, size_t i, double a, double b) {
    ((float2 *)p)[i] = make_float2((float)a, (float)b);
}
__device__ __forceinline__ double4 ld4(const double *p, size_t i) { return ((const double4 *)p)[i]; }
__device__ __forceinline__ double4 ld4(const float *p, size_t i) {
    const float4 v = ((const float4 *)p)[i];
    return make_double4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ void st4(double *p, size_t i, double a, double b, double c, double d) {
    ((double4 *)p)[i] = make_double4(a, b, c, d);
}
__device__ __forceinline__ void st4(float *p, size_t i, double a, double b, double c, double d) {
    ((float4 *)p)[i] = make_float4((float)a, (float)b, (float)c, (float)d);
}

__device__ __forceinline__ double sum4(const double4 w, double x0, double x1, double x2, double x3) {
    double s = w.x * x0;
    s = s + w.y * x1;
    s = s + w.z * x2;
    s = s + w.w * x3;
    return s;
}

__device__ __forceinline__ double sum4a(const double *w, double x0, double x1, double x2, double x3) {
    double r = w[0] * x0;
    r = r + w[1] * x1;
    r = r + w[2] * x2;
    r = r + w[3] * x3;
    return r;
}

__device__ __forceinline__ double dmaxr(double a, double b) { return a > b ? a : b; }

// densityJM: src/fv_pressure_new.F90:340-414 (un-suffixed literals are REAL(4))
__device__ double densityJM(double t, double s, double pz, double rho0) {
    double ss3 = sqrt(s * s * s);
    double bulk = 19092.56 + t * (209.8925 - t * (3.041638 - t * (-(double)1.852732e-3f - t * ((double)1.361629e-5f))))
                + s * (104.4077 - t * (6.500517 - t * (.1553190 - t * (-(double)2.326469e-4f))))
                + ss3 * (-5.587545 + t * (0.7390729 - t * ((double)1.909078e-2f)))
                - pz * ((double)4.721788e-1f + t * ((double)1.028859e-2f + t * (-(double)2.512549e-4f - t * ((double)5.939910e-7f))))
                - pz * s * (-(double)1.571896e-2f - t * ((double)2.598241e-4f + t * (-(double)7.267926e-6f)))
                - pz * ss3 * (double)2.042967e-3f
                + pz * pz * ((double)1.045941e-5f - t * ((double)5.782165e-10f - t * ((double)1.296821e-7f)))
                + pz * pz * s * (-(double)2.595994e-7f + t * (-(double)1.248266e-9f + t * (-(double)3.508914e-9f)));
    double rhopot = 999.842594
                  + t * ((double)6.793952e-2f + t * (-(double)9.095290e-3f + t * ((double)1.001685e-4f
                    + t * (-(double)1.120083e-6f + t * (double)6.536332e-9f))))
                  + s * (0.824493 + t * (-(double)4.08990e-3f + t * ((double)7.64380e-5f
                    + t * (-(double)8.24670e-7f + t * (double)5.38750e-9f))))
                  + ss3 * (-(double)5.72466e-3f + t * ((double)1.02270e-4f + t * (-(double)1.65460e-6f)))
                  + (double)4.8314e-4f * s * s;
    return rhopot * bulk / (bulk + 0.1 * pz) - rho0;
}

// densityJM split by level, for the learned closure's features: the
// pressure-free parts of one (T, S) -- rhopot and the bulk modulus as
// c0 - pz*c1 + pz^2*c2 -- formed once, then the density at any pressure from
// them (the same polynomial, summed in another order).
struct EosLevel {
    double rp, c0, c1, c2;
};
__device__ EosLevel eos_level(double t, double s) {
    const double ss3 = sqrt(s * s * s);
    EosLevel e;
    e.c0 = 19092.56 + t * (209.8925 - t * (3.041638 - t * (-(double)1.852732e-3f - t * ((double)1.361629e-5f))))
         + s * (104.4077 - t * (6.500517 - t * (.1553190 - t * (-(double)2.326469e-4f))))
         + ss3 * (-5.587545 + t * (0.7390729 - t * ((double)1.909078e-2f)));
    e.c1 = ((double)4.721788e-1f + t * ((double)1.028859e-2f + t * (-(double)2.512549e-4f - t * ((double)5.939910e-7f))))
         + s * (-(double)1.571896e-2f - t * ((double)2.598241e-4f + t * (-(double)7.267926e-6f)))
         + ss3 * (double)2.042967e-3f;
    e.c2 = ((double)1.045941e-5f - t * ((double)5.782165e-10f - t * ((double)1.296821e-7f)))
         + s * (-(double)2.595994e-7f + t * (-(double)1.248266e-9f + t * (-(double)3.508914e-9f)));
    e.rp = 999.842594
         + t * ((double)6.793952e-2f + t * (-(double)9.095290e-3f + t * ((double)1.001685e-4f
           + t * (-(double)1.120083e-6f + t * (double)6.536332e-9f))))
         + s * (0.824493 + t * (-(double)4.08990e-3f + t * ((double)7.64380e-5f
           + t * (-(double)8.24670e-7f + t * (double)5.38750e-9f))))
         + ss3 * (-(double)5.72466e-3f + t * ((double)1.02270e-4f + t * (-(double)1.65460e-6f)))
         + (double)4.8314e-4f * s * s;
    return e;
}
__device__ __forceinline__ double eos_rho(const EosLevel &e, double pz, double rho0) {
    const double bulk = e.c0 - pz * e.c1 + pz * pz * e.c2;
    return e.rp * bulk / (bulk + 0.1 * pz) - rho0;
}

// XCD-aware block order: the hardware deals consecutive workgroups round-robin
// to the 8 XCDs; remap so each XCD (own L2) walks one contiguous range of the
// locality-ordered entities instead of every 8th block.
__device__ __forceinline__ unsigned xcd_block() {
    const unsigned G = gridDim.x, b = blockIdx.x;
    const unsigned q = G / 8u, r = G % 8u, x = b % 8u;
    return x * q + (x < r ? x : r) + b / 8u;
}
#define TID() (xcd_block() * blockDim.x + threadIdx.x)
// 3-D fields are stored by column: level k of entity n at n*KC + k, KC = the
// number of levels (nsigma); layer fields leave their last slot unused.
#define PL(arr, k, n, kc) (arr)[(size_t)(n) * (kc) + (k)]

// (entity, layer) of this thread in the layer-parallel kernels: consecutive
// threads take consecutive layers of one column, then the next column, so a
// wavefront covers about 64/L whole columns and every gather of a neighbour
// column reads one contiguous run of its levels.  Launch with
// ceil(ents*L/block) blocks.
__device__ __forceinline__ bool ent_layer(int ents, int L, int &n, int &k) {
    const unsigned i = xcd_block() * blockDim.x + threadIdx.x;
    n = (int)(i / (unsigned)L);
    k = (int)(i - (unsigned)n * (unsigned)L);
    return n < ents;
}

// Column tiles.  Kernels with a recurrence along the column take 64
// consecutive columns per block (kColB), which are one contiguous range of
// every column-stored field.  Data move between HBM and LDS tiles
// t[c*KP + k] (KP = K + 1) in flat loops over the block's (column, slot)
// items, so global accesses stay contiguous; per-level work runs four
// threads per column (slots sub, sub+4, ...), which load the column's
// constants once; the recurrence itself runs one lane per column on the
// tiles, in the block's first wavefront.
constexpr int kColB = 64;   // columns per tile block
constexpr int kColT = 256;  // threads per tile block

// block-strided walk over the (column, slot) items of nc columns x S slots
struct Items {
    int c, k, S, dc, dk;
    __device__ __forceinline__ explicit Items(int S_) : S(S_) {
        c = (int)threadIdx.x / S;
        k = (int)threadIdx.x % S;
        dc = (int)blockDim.x / S;
        dk = (int)blockDim.x % S;
    }
    __device__ __forceinline__ void next() {
        c += dc;
        k += dk;
        if (k >= S) {
            k -= S;
            ++c;
        }
    }
};

__device__ __forceinline__ int col_first() { return (int)(xcd_block() * kColB); }
__device__ __forceinline__ int tcol() { return (int)threadIdx.x >> 2; }  // column of this thread (per-level phases)
__device__ __forceinline__ int tsub() { return (int)threadIdx.x & 3; }   // its first slot

// the K slots of columns [c0, c0+nc) of g -> tile t; loads are issued in
// batches of eight per thread before their LDS stores, so their latencies
// overlap
template <typename T, typename G>
__device__ __forceinline__ void tile_load(T *t, const G *__restrict__ g, int c0, int nc, int K) {
    const size_t base = (size_t)c0 * K;
    Items it(K);
    while (it.c < nc) {
        G v[8];
        int d[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            d[j] = -1;
            if (it.c < nc) {
                v[j] = g[base + (size_t)it.c * K + it.k];
                d[j] = it.c * (K + 1) + it.k;
            }
            it.next();
        }
#pragma unroll
        for (int j = 0; j < 8; ++j)
            if (d[j] >= 0) t[d[j]] = (T)v[j];
    }
}

// the first S slots of each tile column -> columns [c0, c0+nc) of g
template <typename T, typename G>
__device__ __forceinline__ void tile_store(G *__restrict__ g, const T *t, int c0, int nc, int K, int S) {
    const size_t base = (size_t)c0 * K;
    for (Items it(S); it.c < nc; it.next()) g[base + (size_t)it.c * K + it.k] = (G)t[it.c * (K + 1) + it.k];
}

// slots [k0, k1) of tile t -> level planes x[k*N + n] of columns [c0, c0+nc):
// each wavefront writes one 64-column row of a plane
template <typename T, typename G>
__device__ __forceinline__ void tile_to_planes(G *__restrict__ x, const T *t, int c0, int nc, int K, int N, int k0,
                                               int k1) {
    const int j = threadIdx.x & 63;
    if (j >= nc) return;
    for (int k = k0 + (int)(threadIdx.x >> 6); k < k1; k += (int)(blockDim.x >> 6))
        x[(size_t)k * N + c0 + j] = (G)t[j * (K + 1) + k];
}

template <typename T, typename G>
__device__ __forceinline__ void planes_to_tile(T *t, const G *__restrict__ x, int c0, int nc, int K, int N) {
    const int j = threadIdx.x & 63;
    if (j >= nc) return;
    for (int k = (int)(threadIdx.x >> 6); k < K; k += (int)(blockDim.x >> 6))
        t[j * (K + 1) + k] = (T)x[(size_t)k * N + c0 + j];
}

// Column-aligned flat blocks: cpb = blockDim.x / S whole columns per block,
// one thread per (column, slot); a column recurrence then runs on threads
// 0..nc-1 over LDS arrays x[c*S + k].  Every thread has a single item, so the
// block's loads are all in flight together.
struct ColFlat {
    int cpb, c0, nc, c, k;
    __device__ __forceinline__ bool on() const { return c < nc; }
};
__device__ __forceinline__ ColFlat col_flat(int ents, int S) {
    ColFlat f;
    f.cpb = (int)blockDim.x / S;
    f.c0 = (int)xcd_block() * f.cpb;
    f.nc = min(f.cpb, ents - f.c0);
    f.c = (int)threadIdx.x / S;
    f.k = (int)threadIdx.x - f.c * S;
    return f;
}

// Node-gather blocks: the column-aligned flat blocks of the node kernels
// first copy the per-edge (or per-element) operands of all their nodes' CSR
// slots -- one contiguous range of the gather lists -- into LDS, one load per
// slot, so the per-level threads of a node read them from LDS instead of each
// loading the same records through the vector memory path.
__device__ __forceinline__ int2 csr_block(const int *ptr, const ColFlat &f) {
    return make_int2(ptr[f.c0], ptr[f.c0 + f.nc]);
}

// ---------------------------------------------------------------------------
// pressure: hydrostatic column scan, src/fv_pressure_new.F90:91-107
// ---------------------------------------------------------------------------
template <typename R, bool ATM>
__device__ __forceinline__ void pressure_body(DevMesh m, DevStateT<R> s, double gd0, double rho0) {
    extern __shared__ __align__(16) unsigned char lds_[];
    double *th = (double *)lds_;  // per column: the layer increments, then their running sums
    const int L = m.L, KC = m.K;
    const ColFlat f = col_flat(m.N, L);
    if (f.nc <= 0) return;
    const int n = f.c0 + f.c, k = f.k;
    if (f.on()) {
        const double a = s.a[n];
        const double Z = 0.5 * (m.oms[k] * a + m.oms[k + 1] * a);
        const double r = densityJM(PL(s.TF, k, n, KC), PL(s.SF, k, n, KC), -Z, rho0);
        if (ATM) s.rho_c[(size_t)n * L + k] = r;  // rho_c, read by the surface heat source
        th[f.c * L + k] = r * (a * m.dsig[k]) * gd0;
    }
    __syncthreads();
    if ((int)threadIdx.x < f.nc) {
        double *col = th + threadIdx.x * L;
        double hp = 0.0;
        for (int j = 0; j < L; ++j) {
            hp = hp + col[j];
            col[j] = hp;
        }
    }
    __syncthreads();
    if (f.on()) {
        PL(s.hp, k + 1, n, KC) = th[f.c * L + k];
        if (k == 0) PL(s.hp, 0, n, KC) = 0.0;
    }
}
template <typename R>
__global__ __launch_bounds__(256) void k_pressure(DevMesh m, DevStateT<R> s, double gd0, double rho0) {
    pressure_body<R, false>(m, s, gd0, rho0);
}
// the same, also storing rho_c for the surface heat source (key_atm)
template <typename R>
__global__ __launch_bounds__(256) void k_pressure_atm(DevMesh m, DevStateT<R> s, double gd0, double rho0) {
    pressure_body<R, true>(m, s, gd0, rho0);
}

// Per layer (four threads per element column) the gradient and its
// Bpu/Bpv; the vertical integral Bar_pr_2D then sums the layers in order, one
// lane per column.
template <typename R>
__global__ __launch_bounds__(kColT) void k_bpg(DevMesh m, DevStateT<R> s) {
    extern __shared__ __align__(16) unsigned char lds_[];
    const int K = m.K, L = m.L, KP = K + 1, KC = K;
    double *tu = (double *)lds_, *tv = tu + kColB * KP;  // bu*Je, bv*Je
    const int c0 = col_first(), nc = min(kColB, m.E - c0);
    if (nc <= 0) return;
    if (tcol() < nc) {
        const int c = tcol(), e = c0 + c;
        const int4 q = m.en[e];
        const double4 w = m.wcv[e], gx = m.gsx[e], gy = m.gsy[e];
        const double d0 = m.depth[q.x], d1 = m.depth[q.y], d2 = m.depth[q.z], d3 = m.depth[q.w];
        const double h0 = d0 + s.eta_n[q.x], h1 = d1 + s.eta_n[q.y], h2 = d2 + s.eta_n[q.z], h3 = d3 + s.eta_n[q.w];
        const double Ax = sum4(gx, h0, h1, h2, h3), Ay = sum4(gy, h0, h1, h2, h3);
        const double Bx = sum4(gx, d0, d1, d2, d3), By = sum4(gy, d0, d1, d2, d3);
        const double a0 = s.a[q.x], a1 = s.a[q.y], a2 = s.a[q.z], a3 = s.a[q.w];
        const double accB = m.accB[e];
        for (int k = tsub(); k < L; k += 4) {
            const double ds = m.dsig[k];
            const double je = sum4(w, a0 * ds, a1 * ds, a2 * ds, a3 * ds);
            const double aJ = -accB / je;
            const double p0 = PL(s.hp, k, q.x, KC), p1 = PL(s.hp, k, q.y, KC), p2 = PL(s.hp, k, q.z, KC),
                         p3 = PL(s.hp, k, q.w, KC);
            const double r0 = PL(s.hp, k + 1, q.x, KC), r1 = PL(s.hp, k + 1, q.y, KC), r2 = PL(s.hp, k + 1, q.z, KC),
                         r3 = PL(s.hp, k + 1, q.w, KC);
            const double qd_up = sum4(w, p0, p1, p2, p3), qd_low = sum4(w, r0, r1, r2, r3);
            const double pr0 = 0.5 * (p0 + r0) * (a0 * ds), pr1 = 0.5 * (p1 + r1) * (a1 * ds);
            const double pr2 = 0.5 * (p2 + r2) * (a2 * ds), pr3 = 0.5 * (p3 + r3) * (a3 * ds);
            const double bu = aJ * (sum4(gx, pr0, pr1, pr2, pr3) + qd_low * (m.sigma[k + 1] * Ax - Bx) - qd_up * (m.sigma[k] * Ax - Bx));
            const double bv = aJ * (sum4(gy, pr0, pr1, pr2, pr3) + qd_low * (m.sigma[k + 1] * Ay - By) - qd_up * (m.sigma[k] * Ay - By));
            PL(s.Bpu, k, e, KC) = bu;
            PL(s.Bpv, k, e, KC) = bv;
            tu[c * KP + k] = bu * je;
            tv[c * KP + k] = bv * je;
        }
    }
    __syncthreads();
    if ((int)threadIdx.x < nc) {
        const int c = threadIdx.x, e = c0 + c;
        double su = 0.0, sv = 0.0;
        for (int k = 0; k < L; ++k) {
            su = su + tu[c * KP + k];
            sv = sv + tv[c * KP + k];
        }
        st2(s.Bar_pr_2D, e, su, sv);
    }
}

// ---------------------------------------------------------------------------
// d3_end (src/fv_mixing.F90:73-347) in three kernels: the per-level
// right-hand side (flat over the column slots), the iterated tridiagonal
// solve (one thread per node column), and the per-level closure values
// (flat).  Its scratch is private to these kernels and stored by level,
// PP(x, k, n) = x[k*N + n], so the column-sequential solve reads it
// contiguously across the wavefront: p = -b1-b2, q = b_t/dt, c2, the snu
// column, xx1/yy1, and the btr ping-pong (mx_bA, mx_bB).  With the learned
// closure, `only` restricts all three to the columns it sent back.
// ---------------------------------------------------------------------------
#define PP(arr, k, n, N) (arr)[(size_t)(k) * (N) + (n)]
#define SG(nz) m.sigma[(nz) - 1]

// level-plane rows of the block's nc columns: x[k*N + c0 + j] = t[j*S + k]
// for k in [k0, k1), each row written by consecutive threads
template <typename T, typename G>
__device__ __forceinline__ void rows_store(G *__restrict__ x, const T *t, int c0, int nc, int S, int N, int k0, int k1,
                                           const int *__restrict__ only) {
    for (int r = (int)threadIdx.x; r < (k1 - k0) * nc; r += (int)blockDim.x) {
        const int k = k0 + r / nc, j = r % nc;
        if (only && only[c0 + j] == 0) continue;
        x[(size_t)k * N + c0 + j] = (G)t[j * S + k];
    }
}

template <typename R>
__global__ __launch_bounds__(256) void k_d3end_prep(DevMesh m, DevStateT<R> s, double dt, double Dmin, double rho0,
                                                    const int *__restrict__ only) {
    extern __shared__ __align__(16) unsigned char lds_[];
    const int N = m.N, K = m.K, KC = K;
    const ColFlat f = col_flat(N, K);
    const int T = f.cpb * K;
    R *tp = (R *)lds_, *tq = tp + T, *tc = tq + T, *ts = tc + T, *tb = ts + T;
    if (f.nc <= 0) return;
    const int n = f.c0 + f.c, lev = f.k;
    if (only && !__syncthreads_or(f.on() && only[n] != 0)) return;  // no flagged column in this block
    if (f.on() && !(only && only[n] == 0)) {
#define SNU(z) PL(s.snu, (z) - 1, n, KC)
#define BT(z) PL(s.bt, (z) - 1, n, KC)
        const int i = f.c * K + lev;
        ts[i] = SNU(lev + 1);
        tb[i] = BT(lev + 1);
        const int nz = lev + 1;
        if (nz >= 2 && nz <= K - 1) {
            const double t_inv = 1.0 / dt;
            const double a = s.a[n];
            const double dmean = dmaxr(Dmin, m.depth[n] + s.eta_n[n]);
            const double dmean2 = dmean * dmean;
            const double d1 = SG(nz - 1) - SG(nz);
            const double d2 = SG(nz) - SG(nz + 1);
            const double d = 0.5 * (d1 + d2);
            // node-averaged velocities of compute_el2nodes_3D (src/fv_main.F90:2006)
            const double uprev = PL(s.Unode, nz - 2, n, KC), vprev = PL(s.Vnode, nz - 2, n, KC);
            const double ucur = PL(s.Unode, nz - 1, n, KC), vcur = PL(s.Vnode, nz - 1, n, KC);
            double ap1 = uprev - ucur;
            const double ap2 = vprev - vcur;
            double b1 = SNU(nz) * (ap1 * ap1 + ap2 * ap2) / (d * d);
            b1 = b1 / dmean2;
            const double zb = m.oms[nz - 1] * a;
            const double rho_up = densityJM(PL(s.TF, nz - 2, n, KC), PL(s.SF, nz - 2, n, KC), -zb, rho0);
            const double rho_dn = densityJM(PL(s.TF, nz - 1, n, KC), PL(s.SF, nz - 1, n, KC), -zb, rho0);
            ap1 = (rho_up - rho_dn) / d;
            const double b2 = -0.5 * kG * PL(s.Kv, nz - 1, n, KC) * ap1 / (rho0 * dmean);
            const double snu_p = 0.5 * (SNU(nz - 1) + SNU(nz));
            const double snu_m = 0.5 * (SNU(nz) + SNU(nz + 1));
            double c2 = 0.5 * 0.73 * (snu_p * (BT(nz - 1) - BT(nz)) / d1 - snu_m * (BT(nz) - BT(nz + 1)) / d2) / d;
            c2 = c2 / dmean2;
            tp[i] = -b1 - b2;
            tq[i] = t_inv * BT(nz);
            tc[i] = c2;
        }
#undef SNU
#undef BT
    }
    __syncthreads();
    rows_store(s.mx_s, ts, f.c0, f.nc, K, N, 0, K, only);
    rows_store(s.mx_bA, tb, f.c0, f.nc, K, N, 0, K, only);
    rows_store(s.mx_p, tp, f.c0, f.nc, K, N, 1, K - 1, only);
    rows_store(s.mx_q, tq, f.c0, f.nc, K, N, 1, K - 1, only);
    rows_store(s.mx_c2, tc, f.c0, f.nc, K, N, 1, K - 1, only);
}

// One thread per node column, kD3endT columns per block.  The iterated
// column -- the btr ping-pong and the sweep's xx1/yy1 -- stays in the block's
// LDS, slot [k*kD3endT + thread], for all iterations; the per-level operands
// of k_d3end_prep (p, q, c2, the snu column) are read from their level planes.
template <typename R>
__global__ __launch_bounds__(kD3endT) void k_d3end(DevMesh m, DevStateT<R> s, double dt, double Dmin, double z0b_min,
                                                   const int *__restrict__ only) {
    extern __shared__ __align__(16) unsigned char lds_[];
    const int KB = m.K * kD3endT, th = threadIdx.x;
    R *lA = (R *)lds_, *lB = lA + KB, *lxx = lB + KB, *lyy = lxx + KB;
#define LX(arr, k) (arr)[(k) * kD3endT + th]
    int n = TID();
    if (n >= m.N) return;
    if (only && only[n] == 0) return;  // learned closure active: only the columns it sent back
    const int N = m.N, K = m.K, L = m.L, KC = K;
    const double t_inv = 1.0 / dt;
    const double B1cka2 = 16.6 * (kCka * kCka);
    const double a = s.a[n];
    const double dmean = dmaxr(Dmin, m.depth[n] + s.eta_n[n]);
    const double dmean2 = dmean * dmean;
#define SNU(z) PP(s.mx_s, (z) - 1, n, N)
    // bottom: the deepest layer (nsigma-1) of the node-averaged velocity
    const double ub = PL(s.Unode, L - 1, n, KC), vb = PL(s.Vnode, L - 1, n, KC);
    const double d1w = 0.5 * (SG(K - 1) - SG(K));
    const double wx = s.windx[n], wy = s.windy[n];
    const double uds = 0.05 * (wx * wx + wy * wy);
    const double snu_mw = 0.5 * (SNU(K) + SNU(K - 1));
    const double yy2 = (0.4e-3 * (uds * uds * uds) * dmean * d1w) / (snu_mw * 0.76);
    const double delta_b = a * m.dsig[L - 1] + z0b_min;
    const double lg = fcglog::log_cr(delta_b / z0b_min);  // correctly rounded, as the host libm's
    const double bott_R = B1cka2 / (lg * lg);
    for (int k = 0; k < K; ++k) LX(lA, k) = PP(s.mx_bA, k, n, N);
    R *bc = lA, *bn = lB;
    for (int it = 1; it <= 200; ++it) {
        double xxp = 1.0, yyp = yy2;
        for (int nz = 2; nz <= K - 1; ++nz) {
            double d1 = SG(nz - 1) - SG(nz);
            double d2 = SG(nz) - SG(nz + 1);
            double d = 0.5 * (d1 + d2);
            double btr = LX(bc, nz - 1);
            double sn = SNU(nz);
            double c1 = 0.046 * (btr * btr) / sn;
            double fr = PP(s.mx_p, nz - 1, n, N) - c1 - PP(s.mx_q, nz - 1, n, N) - PP(s.mx_c2, nz - 1, n, N);
            double snu_p = 0.5 * (SNU(nz - 1) + sn);
            double snu_m = 0.5 * (sn + SNU(nz + 1));
            double am = 0.5 * (snu_p * 0.73) / (dmean2 * d1 * d);
            double cm = 0.5 * (snu_m * 0.73) / (dmean2 * d2 * d);
            double bm = am + cm + (2.0 * 0.046 * btr) / sn + t_inv;
            double aa = bm - am * xxp;
            double xxn = cm / aa;
            double yyn = (am * yyp - fr) / aa;
            LX(lxx, nz) = xxn;   // xx1(nz+1)
            LX(lyy, nz) = yyn;
            xxp = xxn;
            yyp = yyn;
        }
        // back substitution from the bottom boundary value, with convergence measure
        double b1v = bott_R * (ub * ub + vb * vb);
        LX(bn, K - 1) = b1v;
        double bbb = 0.0;
        {
            double bo = LX(bc, K - 1);
            double aa = bo - b1v;
            double bb = dmaxr(b1v, bo) + 1.0e-8;
            double cc = fabs(aa / bb);
            if (cc > bbb) bbb = cc;
        }
        for (int nz = K; nz >= 2; --nz) {
            double xx = (nz == 2) ? 1.0 : (double)LX(lxx, nz - 1);
            double yy = (nz == 2) ? yy2 : (double)LX(lyy, nz - 1);
            if (nz == K) { xx = xxp; yy = yyp; }
            b1v = b1v * xx + yy;
            LX(bn, nz - 2) = b1v;
            double bo = LX(bc, nz - 2);
            double aa = bo - b1v;
            double bb = dmaxr(b1v, bo) + 1.0e-8;
            double cc = fabs(aa / bb);
            if (cc > bbb) bbb = cc;
        }
        if (bbb < 1.0e-6) break;
        R *t = bc; bc = bn; bn = t;
    }
    // the converged column, for the per-level closure values
    for (int k = 0; k < K; ++k) PP(s.mx_xx, k, n, N) = LX(bc, k);
#undef SNU
#undef LX
}

template <typename R>
__global__ __launch_bounds__(256) void k_d3end_fin(DevMesh m, DevStateT<R> s, double Dmin, double z0b_min,
                                                   double z0s_min, double snul, const int *__restrict__ only) {
    extern __shared__ __align__(16) unsigned char lds_[];
    const int N = m.N, K = m.K, KC = K;
    const ColFlat f = col_flat(N, K);
    R *tx = (R *)lds_;
    if (f.nc <= 0) return;
    if (only && !__syncthreads_or(f.on() && only[f.c0 + f.c] != 0)) return;  // no flagged column in this block
    for (int r = (int)threadIdx.x; r < K * f.nc; r += (int)blockDim.x) {
        const int k = r / f.nc, j = r % f.nc;
        if (only && only[f.c0 + j] == 0) continue;
        tx[j * K + k] = PP(s.mx_xx, k, f.c0 + j, N);
    }
    __syncthreads();
    if (!f.on()) return;
    const int n = f.c0 + f.c, lev = f.k;
    if (only && only[n] == 0) return;
    const int nz = lev + 1;
    const double dmean = dmaxr(Dmin, m.depth[n] + s.eta_n[n]);
    const double btr = tx[f.c * K + lev];
    const double btv = dmaxr(btr, 1.0e-8);
    PL(s.bt, lev, n, KC) = btv;
    const double zt = SG(nz) * dmean;
    const double Zh = zt + z0b_min;
    const double Zd = fabs(dmean - zt) + z0s_min;
    const double Z0 = 1.0 - kBetaScale * Zh * Zd / (dmean * dmean);
    const double scale_t = kCka * Zh * Zd * Z0 / dmean;
    double sv = scale_t * sqrt(btv) + snul;
    if (sv >= 0.2) sv = 0.2;
    PL(s.snu, lev, n, KC) = sv;
    PL(s.Kv, lev, n, KC) = kPrNum * (sv);
}
#undef SG

// ---------------------------------------------------------------------------
// vel_gradients (3-D, free slip): src/fv_gradients.F90:2-93
// ---------------------------------------------------------------------------
// vel_gradients (3-D, free slip) on column-aligned blocks: the element's
// gradient_vec, neighbours and boundary-side metric are read once per column
// into LDS.
template <typename R>
__global__ __launch_bounds__(256) void k_velgrad3d(DevMesh m, DevStateT<R> s) {
    extern __shared__ __align__(16) unsigned char lds_[];
    const int KC = m.K;
    const ColFlat f = col_flat(m.E, m.L);
    VelGradRec *rec = (VelGradRec *)lds_;
    if (f.nc <= 0) return;
    if ((int)threadIdx.x < f.nc) {
        const int e = f.c0 + threadIdx.x;
        const int4 q = m.en[e];
        const int4 nb = m.elnb[e];
        const int4 ed = m.eledges[e];
        const double4 gx = m.gvx[e], gy = m.gvy[e];
        const int nbs[4] = {nb.x, nb.y, nb.z, nb.w};
        const int eds[4] = {ed.x, ed.y, ed.z, ed.w};
        const double gxs[4] = {gx.x, gx.y, gx.z, gx.w}, gys[4] = {gy.x, gy.y, gy.z, gy.w};
        const double ec = m.elem_cos[e];
        VelGradRec r;
        r.jend = (q.x == q.w) ? 3 : 4;
        for (int j = 0; j < 4; ++j) {
            r.gx[j] = gxs[j];
            r.gy[j] = gys[j];
            r.nbK[j] = nbs[j] >= 0 ? nbs[j] * KC : -1;
            r.zc0[j] = 0.0;
            r.zc1[j] = 0.0;
            if (j < r.jend && nbs[j] < 0) {
                const double2 dd = m.edxdy[eds[j]];
                r.zc0[j] = dd.x * ec;
                r.zc1[j] = dd.y;
            }
        }
        rec[threadIdx.x] = r;
    }
    __syncthreads();
    if (!f.on()) return;
    const int e = f.c0 + f.c, k = f.k;
    const VelGradRec &r = rec[f.c];
    const double u = PL(s.U_n, k, e, KC), v = PL(s.V_n, k, e, KC);
    double ux = 0.0, uy = 0.0, vx = 0.0, vy = 0.0;
    for (int j = 0; j < r.jend; ++j) {
        if (r.nbK[j] >= 0) {
            double du = s.U_n[r.nbK[j] + k] - u;
            double dv = s.V_n[r.nbK[j] + k] - v;
            ux = ux + r.gx[j] * du;
            uy = uy + r.gy[j] * du;
            vx = vx + r.gx[j] * dv;
            vy = vy + r.gy[j] * dv;
        } else {
            const double zc0 = r.zc0[j], zc1 = r.zc1[j];
            double un = -2 * (u * zc1 - v * zc0) / (zc0 * zc0 + zc1 * zc1);
            ux = ux + r.gx[j] * un * zc1;
            uy = uy + r.gy[j] * un * zc1;
            vx = vx - r.gx[j] * un * zc0;
            vy = vy - r.gy[j] * un * zc0;
        }
    }
    PL(s.gux, k, e, KC) = ux;
    PL(s.guy, k, e, KC) = uy;
    PL(s.gvx, k, e, KC) = vx;
    PL(s.gvy, k, e, KC) = vy;
}

// momentum_adv_upwind edge flux at one level: src/fv_advection.F90:120-165
template <typename R>
__device__ __forceinline__ void edge_uu3d(const DevMesh &m, const DevStateT<R> &s, int ed, int k, R &uu, R &vv) {
    const int KC = m.K;
    // in the storage type: FP32 advective fluxes in the mixed-precision path
    const int E = m.E;
    const int2 nd = m.ednodes[ed];
    const int2 et = m.edtri[ed];
    const R acc = (R)(0.5 * (m.ac[nd.x] + m.ac[nd.y]));
    const double4 cd = m.ecross[ed];
    const double2 dd = m.edxdy[ed];
    const R cx = (R)cd.x, cy = (R)cd.y, cz = (R)cd.z, cw = (R)cd.w, dx = (R)dd.x, dy = (R)dd.y;
    const int e1 = et.x, e2 = et.y;
    const R u1 = PL(s.U_n, k, e1, KC) - PL(s.gux, k, e1, KC) * cx - PL(s.guy, k, e1, KC) * cy;
    const R v1 = PL(s.V_n, k, e1, KC) - PL(s.gvx, k, e1, KC) * cx - PL(s.gvy, k, e1, KC) * cy;
    const R u2 = PL(s.U_n, k, e2, KC) - PL(s.gux, k, e2, KC) * cz - PL(s.guy, k, e2, KC) * cw;
    const R v2 = PL(s.V_n, k, e2, KC) - PL(s.gvx, k, e2, KC) * cz - PL(s.gvy, k, e2, KC) * cw;
    const double ds = m.dsig[k];
    const R jd = (R)(0.5 * (s.a[nd.x] * ds + s.a[nd.y] * ds));
    const R un = R(0.5) * R(kREarthD) * jd * ((u1 + u2) * dy - (v1 * (R)m.elem_cos[e1] + v2 * (R)m.elem_cos[e2]) * dx);
    if (un >= R(0)) {
        uu = u1 * un * acc;
        vv = v1 * un * acc;
    } else {
        uu = u2 * un * acc;
        vv = v2 * un * acc;
    }
}

// Element quantities shared by the two compute_vel_rhs kernels.
struct ElemCol {
    int4 q;
    double4 w;
    double a0, a1, a2, a3;
    __device__ double Je(const DevMesh &m, int k) const {
        const double ds = m.dsig[k];
        return sum4(w, a0 * ds, a1 * ds, a2 * ds, a3 * ds);
    }
    template <typename R>
    __device__ double Av(const DevMesh &m, const DevStateT<R> &s, int k) const {
    const int KC = m.K;
        const int N = m.N;
        return sum4(w, PL(s.snu, k, q.x, KC), PL(s.snu, k, q.y, KC), PL(s.snu, k, q.z, KC), PL(s.snu, k, q.w, KC));
    }
    __device__ double zbw(const DevMesh &m, int lev) const {  // sum w_cv*zbar at 0-based level
        const double o = m.oms[lev];
        return sum4(w, o * a0, o * a1, o * a2, o * a3);
    }
    __device__ double Zw(const DevMesh &m, int lay) const {  // sum w_cv*Z at 0-based layer
        const double o0 = m.oms[lay], o1 = m.oms[lay + 1];
        return sum4(w, 0.5 * (o0 * a0 + o1 * a0), 0.5 * (o0 * a1 + o1 * a1), 0.5 * (o0 * a2 + o1 * a2),
                    0.5 * (o0 * a3 + o1 * a3));
    }
};

template <typename R>
__device__ __forceinline__ ElemCol elem_col(const DevMesh &m, const DevStateT<R> &s, int e) {
    ElemCol c;
    c.q = m.en[e];
    c.w = m.wcv[e];
    c.a0 = s.a[c.q.x];
    c.a1 = s.a[c.q.y];
    c.a2 = s.a[c.q.z];
    c.a3 = s.a[c.q.w];
    return c;
}

// momentum_vert_adv_upwind flux through 1-based interface nz (2..K-1) of an
// element column (fv_advection.F90:389-433)
template <typename R>
__device__ __forceinline__ void uvert(const DevMesh &m, const DevStateT<R> &s, const ElemCol &c, int e, double area,
                                      int nz, double &fu, double &fv) {
    const int KC = m.K;
    const int N = m.N, E = m.E, K = m.K;
    const int4 q = c.q;
    auto U = [&](int k) { return PL(s.U_n, k, e, KC); };
    auto V = [&](int k) { return PL(s.V_n, k, e, KC); };
    const double wv = sum4(c.w, PL(s.W, nz - 1, q.x, KC), PL(s.W, nz - 1, q.y, KC), PL(s.W, nz - 1, q.z, KC),
                           PL(s.W, nz - 1, q.w, KC)) * area;
    double um = 0.0, vm = 0.0;
    if (wv >= 0.0) {
        if (nz == K - 1) {
            um = 0.5 * (U(nz - 2) + U(nz - 1));
            vm = 0.5 * (V(nz - 2) + V(nz - 1));
        } else {
            double Z1 = c.zbw(m, nz - 1);
            double aa = Z1 - c.Zw(m, nz - 2);
            double bb = c.Zw(m, nz - 1) - Z1;
            double cc = c.Zw(m, nz) - Z1;
            double d = (cc + aa) * (bb * bb - aa * aa) - (cc * cc - aa * aa) * (bb + aa);
            double dg = aa * bb * (aa + bb) / d;
            double db = -aa * cc * (aa + cc) / d;
            double da = 1.0 - dg - db;
            um = U(nz - 2) * da + U(nz - 1) * db + U(nz) * dg;
            vm = V(nz - 2) * da + V(nz - 1) * db + V(nz) * dg;
        }
    } else {
        if (nz == 2) {
            um = 0.5 * (U(nz - 2) + U(nz - 1));
            vm = 0.5 * (V(nz - 2) + V(nz - 1));
        } else {
            double Z1 = c.zbw(m, nz - 1);
            double aa = c.Zw(m, nz - 1) - Z1;
            double bb = Z1 - c.Zw(m, nz - 2);
            double cc = Z1 - c.Zw(m, nz - 3);
            double d = (cc + aa) * (bb * bb - aa * aa) - (cc * cc - aa * aa) * (bb + aa);
            double dg = aa * bb * (aa + bb) / d;
            double db = -aa * cc * (aa + cc) / d;
            double da = 1.0 - dg - db;
            um = U(nz - 1) * da + U(nz - 2) * db + U(nz - 3) * dg;
            vm = V(nz - 1) * da + V(nz - 2) * db + V(nz - 3) * dg;
        }
    }
    fu = -um * wv;
    fv = -vm * wv;
}

// uvert on the column arrays a block of compute_vel_rhs shares in LDS: ws =
// sum w_cv*W at each level, u/v the element's velocity, Zw/zbw the
// element-averaged layer and level depths (the same operands uvert forms).
template <typename R, typename C>
__device__ __forceinline__ void uvert_lds(const C *ws, const R *u, const R *v, const C *Zw, const C *zbw,
                                          int K, double area, int nz, double &fu, double &fv) {
    const double wv = ws[nz - 1] * area;
    double um = 0.0, vm = 0.0;
    if (wv >= 0.0) {
        if (nz == K - 1) {
            um = 0.5 * (u[nz - 2] + u[nz - 1]);
            vm = 0.5 * (v[nz - 2] + v[nz - 1]);
        } else {
            double Z1 = zbw[nz - 1];
            double aa = Z1 - Zw[nz - 2];
            double bb = Zw[nz - 1] - Z1;
            double cc = Zw[nz] - Z1;
            double d = (cc + aa) * (bb * bb - aa * aa) - (cc * cc - aa * aa) * (bb + aa);
            double dg = aa * bb * (aa + bb) / d;
            double db = -aa * cc * (aa + cc) / d;
            double da = 1.0 - dg - db;
            um = u[nz - 2] * da + u[nz - 1] * db + u[nz] * dg;
            vm = v[nz - 2] * da + v[nz - 1] * db + v[nz] * dg;
        }
    } else {
        if (nz == 2) {
            um = 0.5 * (u[nz - 2] + u[nz - 1]);
            vm = 0.5 * (v[nz - 2] + v[nz - 1]);
        } else {
            double Z1 = zbw[nz - 1];
            double aa = Zw[nz - 1] - Z1;
            double bb = Z1 - Zw[nz - 2];
            double cc = Z1 - Zw[nz - 3];
            double d = (cc + aa) * (bb * bb - aa * aa) - (cc * cc - aa * aa) * (bb + aa);
            double dg = aa * bb * (aa + bb) / d;
            double db = -aa * cc * (aa + cc) / d;
            double da = 1.0 - dg - db;
            um = u[nz - 1] * da + u[nz - 2] * db + u[nz - 3] * dg;
            vm = v[nz - 1] * da + v[nz - 2] * db + v[nz - 3] * dg;
        }
    }
    fu = -um * wv;
    fv = -vm * wv;
}

// uvert_lds as one instruction stream for both signs of W, so lanes of a
// wavefront whose interfaces carry opposite W do not diverge: the downward
// branch is the upward one on the mirrored operands (layers nz-1, nz-2,
// nz-3 instead of nz-2, nz-1, nz) with the depth differences negated, and
// x - y = -(y - x) exactly, so every value equals uvert_lds's bit for bit.
template <typename R, typename C>
__device__ __forceinline__ void uvert_uni(const C *ws, const R *u, const R *v, const C *Zw, const C *zbw,
                                          int K, double area, int nz, double &fu, double &fv) {
    const double wv = ws[nz - 1] * area;
    const bool up = wv >= 0.0;
    double um, vm;
    if (up ? nz == K - 1 : nz == 2) {
        um = 0.5 * (u[nz - 2] + u[nz - 1]);
        vm = 0.5 * (v[nz - 2] + v[nz - 1]);
    } else {
        const int iP = up ? nz - 2 : nz - 1, iQ = up ? nz - 1 : nz - 2, iR = up ? nz : nz - 3;
        const double sg = up ? 1.0 : -1.0;
        const double Z1 = zbw[nz - 1];
        const double aa = sg * (Z1 - Zw[iP]);
        const double bb = sg * (Zw[iQ] - Z1);
        const double cc = sg * (Zw[iR] - Z1);
        const double d = (cc + aa) * (bb * bb - aa * aa) - (cc * cc - aa * aa) * (bb + aa);
        const double dg = aa * bb * (aa + bb) / d;
        const double db = -aa * cc * (aa + cc) / d;
        const double da = 1.0 - dg - db;
        um = u[iP] * da + u[iQ] * db + u[iR] * dg;
        vm = v[iP] * da + v[iQ] * db + v[iR] * dg;
    }
    fu = -um * wv;
    fv = -vm * wv;
}

// Per-edge constants of momentum_adv_upwind (fv_advection.F90:120-165), as
// the flux uses them (storage type R for the geometry, FP64 node depths).
template <typename R>
struct EdgeRec {
    int cc, oK, oc;      // signed edge id (kEmpty: none; < 0: the element is edge_tri(1)), the other element's
                         // column offset, and its column in this block (-1: outside)
    R acc, cx, cy, cz, cw, dx, dy, cos1, cos2;
    double an1, an2;     // a at the edge's nodes
};

// Element constants of compute_vel_rhs
struct ElemRec {
    ElemCol col;
    double Fx, Fy, acc, area, cor, mask;
};

static_assert(sizeof(ElemRec) <= kVelrhsElemBytes && alignof(ElemRec) <= 16 + 16, "element record size");
static_assert(sizeof(EdgeRec<double>) == velrhs_edge_bytes(8) && sizeof(EdgeRec<float>) == velrhs_edge_bytes(4),
              "edge record size");

// Biharmonic viscosity helpers (the kernels are with the barotropic filters below).
// corrected normal flux across an internal edge (fv_viscosity.F90:158-169):
// the mean gradient with its centre-line component replaced by the difference
// quotient, dotted with the edge normal
__device__ __forceinline__ double bh_flux(const BhEdge &g, double gx1, double gx2, double gy1, double gy2, double f1,
                                          double f2) {
    double g1 = 0.5 * (gx1 + gx2), g2 = 0.5 * (gy1 + gy2);
    const double tt = g1 * g.tx + g2 * g.ty - (f2 - f1) / g.tl;
    g1 = g1 - g.tx * tt;
    g2 = g2 - g.ty * tt;
    return g1 * g.ye - g2 * g.xe;
}

// the element's sum of a 3-D pass's edge fluxes at layer k (internal edges
// +/- by side, then boundary edges +)
template <typename R>
__device__ __forceinline__ void bh_gather(const DevMesh &m, const DevStateT<R> &s, int e, int k, double &u, double &v) {
    const int K = m.K;
    const int4 ied = m.el_ied[e];
    const int slots[4] = {ied.x, ied.y, ied.z, ied.w};
    for (int j = 0; j < 4; ++j) {
        const int c = slots[j];
        if (c == kEmpty) break;
        const size_t q = (size_t)(c >= 0 ? c : ~c) * K + k;
        if (c >= 0) {  // element is edge_tri(2)
            u = u - (double)s.bhfu[q];
            v = v - (double)s.bhfv[q];
        } else {
            u = u + (double)s.bhfu[q];
            v = v + (double)s.bhfv[q];
        }
    }
    const int4 bed = m.el_bed[e];
    const int bs[4] = {bed.x, bed.y, bed.z, bed.w};
    for (int j = 0; j < 4; ++j) {
        if (bs[j] == kEmpty) break;
        const size_t q = (size_t)bs[j] * K + k;
        u = u + (double)s.bhfu[q];
        v = v + (double)s.bhfv[q];
    }
}

// the element's sum of the 2-D pass's edge fluxes, in ascending edge id
template <typename R>
__device__ __forceinline__ void bh2_gather(const DevMesh &m, const DevStateT<R> &s, int e, double &u, double &v) {
    const double2 *fl = (const double2 *)s.bh_fl2;
    const int4 ied = m.el_ied[e];
    const int slots[4] = {ied.x, ied.y, ied.z, ied.w};
    for (int j = 0; j < 4; ++j) {
        const int c = slots[j];
        if (c == kEmpty) break;
        const double2 f = fl[c >= 0 ? c : ~c];
        if (c >= 0) {
            u = u - f.x;
            v = v - f.y;
        } else {
            u = u + f.x;
            v = v + f.y;
        }
    }
    const int4 bed = m.el_bed[e];
    const int bs[4] = {bed.x, bed.y, bed.z, bed.w};
    for (int j = 0; j < 4; ++j) {
        if (bs[j] == kEmpty) break;
        const double2 f = fl[bs[j]];
        u = u + f.x;
        v = v + f.y;
    }
}

// compute_vel_rhs (src/fv_3D_vel.F90:1-197), explicit part, on column-aligned
// blocks (cpb = 256/L element columns, one thread per element and layer):
// the element's and its internal edges' constants are read once per column
// into LDS; each thread then forms the elemental pressure/Coriolis/AB2 terms
// (35-78), the momentum_adv_upwind gather over the element's internal edges
// in ascending id (fv_advection.F90:64-197), momentum_vert_adv_upwind at the
// layer's two interfaces (354-469) and the explicit update (124-143).  Writes
// the new AB history and the provisional velocity into U_rhs/V_rhs;
// k_velvisc then applies the implicit vertical viscosity column by column.
template <typename R, bool ATM, bool FILT>
__device__ __forceinline__ void velrhs_body(DevMesh m, DevStateT<R> s, double dt, double ff, int hadv, int vadv,
                                            double d0i, double tau) {
    extern __shared__ __align__(16) unsigned char lds_[];
    const int L = m.L, K = m.K, KC = K;
    const ColFlat f = col_flat(m.E, L);
    ElemRec *er = (ElemRec *)lds_;                                        // [cpb]
    EdgeRec<R> *edr = (EdgeRec<R> *)(lds_ + f.cpb * kVelrhsElemBytes);   // [cpb][4]
    // per column and layer: sum w_cv*W, Zw, zbw, U_n, V_n, and the velocity
    // gradients (read by the block's neighbours across internal edges)
    const int T = f.cpb * L;
    // the vertical advection's column arrays in the storage type (as the
    // level-vector kernel keeps them, so both variants round alike)
    R *cws = (R *)(edr + 4 * f.cpb), *cZw = cws + T, *czb = cZw + T;
    R *cu = czb + T, *cv = cu + T, *cgux = cv + T, *cguy = cgux + T, *cgvx = cguy + T, *cgvy = cgvx + T;
    if (f.nc <= 0) return;
    if ((int)threadIdx.x < f.nc) {
        const int c = threadIdx.x, e = f.c0 + c;
        ElemRec r;
        r.col = elem_col(m, s, e);
        const int4 q = r.col.q;
        const double4 gx = m.gsx[e], gy = m.gsy[e];
        r.area = m.elem_area[e];
        r.cor = m.coriolis[e];
        r.mask = s.mask_wd[e];
        double pre0 = -kG * (m.atp1[q.x] * s.eta_n[q.x] - m.atp2[q.x] * 0.0);
        double pre1 = -kG * (m.atp1[q.y] * s.eta_n[q.y] - m.atp2[q.y] * 0.0);
        double pre2 = -kG * (m.atp1[q.z] * s.eta_n[q.z] - m.atp2[q.z] * 0.0);
        double pre3 = -kG * (m.atp1[q.w] * s.eta_n[q.w] - m.atp2[q.w] * 0.0);
        if (ATM) {  // sea-level pressure and freshwater load (fv_3D_vel.F90:40-53)
            const double4 w = r.col.w;
            const double am = w.x * s.mslp[q.x] + w.y * s.mslp[q.y] + w.z * s.mslp[q.z] + w.w * s.mslp[q.w];
            const double em = w.x * s.emp[q.x] + w.y * s.emp[q.y] + w.z * s.emp[q.z] + w.w * s.emp[q.w];
            pre0 = pre0 - (kPref - am) * d0i + kG * em * d0i;
            pre1 = pre1 - (kPref - am) * d0i + kG * em * d0i;
            pre2 = pre2 - (kPref - am) * d0i + kG * em * d0i;
            pre3 = pre3 - (kPref - am) * d0i + kG * em * d0i;
        }
        r.Fx = sum4(gx, pre0, pre1, pre2, pre3) * r.area;
        r.Fy = sum4(gy, pre0, pre1, pre2, pre3) * r.area;
        r.acc = sum4(r.col.w, m.ac[q.x], m.ac[q.y], m.ac[q.z], m.ac[q.w]);
        er[c] = r;
        const int4 ied = m.el_ied[e];
        for (int j = 0; j < 4; ++j) {
            const int cc = j == 0 ? ied.x : (j == 1 ? ied.y : (j == 2 ? ied.z : ied.w));
            EdgeRec<R> g;
            g.cc = cc;
            if (cc != kEmpty && hadv) {
                const int ed = cc >= 0 ? cc : ~cc;
                const int2 nd = m.ednodes[ed];
                const int2 et = m.edtri[ed];
                const double4 cd = m.ecross[ed];
                const double2 dd = m.edxdy[ed];
                const int oe = cc >= 0 ? et.x : et.y;
                g.oK = oe * K;
                g.oc = (oe >= f.c0 && oe < f.c0 + f.nc) ? oe - f.c0 : -1;
                g.acc = (R)(0.5 * (m.ac[nd.x] + m.ac[nd.y]));
                g.cx = (R)cd.x;
                g.cy = (R)cd.y;
                g.cz = (R)cd.z;
                g.cw = (R)cd.w;
                g.dx = (R)dd.x;
                g.dy = (R)dd.y;
                g.cos1 = (R)m.elem_cos[et.x];
                g.cos2 = (R)m.elem_cos[et.y];
                g.an1 = s.a[nd.x];
                g.an2 = s.a[nd.y];
            }
            edr[c * 4 + j] = g;
        }
    }
    __syncthreads();
    const int c = f.c, e = f.c0 + c, k = f.k;
    if (f.on()) {  // the column arrays of the vertical advection
        const ElemCol &col = er[c].col;
        const int4 q = col.q;
        cws[c * L + k] = (R)sum4(col.w, PL(s.W, k, q.x, KC), PL(s.W, k, q.y, KC), PL(s.W, k, q.z, KC), PL(s.W, k, q.w, KC));
        cZw[c * L + k] = (R)col.Zw(m, k);
        czb[c * L + k] = (R)col.zbw(m, k);
        cu[c * L + k] = PL(s.U_n, k, e, KC);
        cv[c * L + k] = PL(s.V_n, k, e, KC);
        if (hadv) {
            cgux[c * L + k] = PL(s.gux, k, e, KC);
            cguy[c * L + k] = PL(s.guy, k, e, KC);
            cgvx[c * L + k] = PL(s.gvx, k, e, KC);
            cgvy[c * L + k] = PL(s.gvy, k, e, KC);
        }
    }
    __syncthreads();
    if (!f.on()) return;
    const ElemRec &r = er[c];
    const ElemCol &col = r.col;
    const double area = r.area;
    const double area_inv = 1.0 / area;
    const double je = col.Je(m, k);
    const double u = cu[c * L + k], v = cv[c * L + k];
    double ur = -(0.5 + kEps) * PL(s.U_rhsAB, k, e, KC) + r.Fx * je;
    double vr = -(0.5 + kEps) * PL(s.V_rhsAB, k, e, KC) + r.Fy * je;
    if (FILT) {  // viscosity_filt_3D (fv_viscosity.F90:822-868): neighbours minus own U_n, elem_neighbors order
        const int4 q4 = m.en[e];
        const int4 nb4 = m.elnb[e];
        const int nbs[4] = {nb4.x, nb4.y, nb4.z, nb4.w};
        const int jend = (q4.x == q4.w) ? 3 : 4;
        double fu = 0.0, fv = 0.0;
        for (int j = 0; j < jend; ++j) {
            if (nbs[j] >= 0) {
                fu = fu + (double)PL(s.U_n, k, nbs[j], KC) - u;
                fv = fv + (double)PL(s.V_n, k, nbs[j], KC) - v;
            }
        }
        ur = ur + tau * fu * area * je;
        vr = vr + tau * fv * area * je;
    }
    if (s.bhfu) bh_gather(m, s, e, k, ur, vr);  // biharmonic_viscosity, second pass (fv_viscosity.F90:262-333)
    R abu = (R)(v * r.cor * area * je);
    R abv = (R)(-u * r.cor * area * je);
    if (hadv) {
        const double ds = m.dsig[k];
        const int ck = c * L + k;
        // the element's own velocity and gradients at this layer (one end of every edge)
        const R uo = cu[ck], vo = cv[ck], guxo = cgux[ck], guyo = cguy[ck], gvxo = cgvx[ck], gvyo = cgvy[ck];
        for (int j = 0; j < 4; ++j) {
            const EdgeRec<R> &g = edr[c * 4 + j];
            if (g.cc == kEmpty) break;
            const bool own1 = g.cc < 0;  // the element is edge_tri(1)
            R ux, vx, guxx, guyx, gvxx, gvyx;
            if (g.oc >= 0) {  // the other element is a column of this block
                const int i = g.oc * L + k;
                ux = lds_ld(cu + i);
                vx = lds_ld(cv + i);
                guxx = lds_ld(cgux + i);
                guyx = lds_ld(cguy + i);
                gvxx = lds_ld(cgvx + i);
                gvyx = lds_ld(cgvy + i);
            } else {
                const int io = g.oK + k;
                ux = glb_ld(s.U_n + io);
                vx = glb_ld(s.V_n + io);
                guxx = glb_ld(s.gux + io);
                guyx = glb_ld(s.guy + io);
                gvxx = glb_ld(s.gvx + io);
                gvyx = glb_ld(s.gvy + io);
            }
            const R U1 = own1 ? uo : ux, V1 = own1 ? vo : vx, U2 = own1 ? ux : uo, V2 = own1 ? vx : vo;
            const R gux1 = own1 ? guxo : guxx, guy1 = own1 ? guyo : guyx, gvx1 = own1 ? gvxo : gvxx, gvy1 = own1 ? gvyo : gvyx;
            const R gux2 = own1 ? guxx : guxo, guy2 = own1 ? guyx : guyo, gvx2 = own1 ? gvxx : gvxo, gvy2 = own1 ? gvyx : gvyo;
            // in the storage type: FP32 advective fluxes in the mixed-precision path
            const R u1 = U1 - gux1 * g.cx - guy1 * g.cy;
            const R v1 = V1 - gvx1 * g.cx - gvy1 * g.cy;
            const R u2 = U2 - gux2 * g.cz - guy2 * g.cw;
            const R v2 = V2 - gvx2 * g.cz - gvy2 * g.cw;
            const R jd = (R)(0.5 * (g.an1 * ds + g.an2 * ds));
            const R un = R(0.5) * R(kREarthD) * jd * ((u1 + u2) * g.dy - (v1 * g.cos1 + v2 * g.cos2) * g.dx);
            R uu, vv;
            if (un >= R(0)) {
                uu = u1 * un * g.acc;
                vv = v1 * un * g.acc;
            } else {
                uu = u2 * un * g.acc;
                vv = v2 * un * g.acc;
            }
            if (g.cc >= 0) {
                abu = abu + uu;
                abv = abv + vv;
            } else {
                abu = abu - uu;
                abv = abv - vv;
            }
        }
    }
    if (vadv) {
        if (r.mask != 0.0) {
            // fluxes through the layer's top (t = 0) and bottom (t = 1) interface
            double ftu = 0.0, ftv = 0.0, fbu = 0.0, fbv = 0.0;
#pragma nounroll
            for (int t = 0; t < 2; ++t) {
                const int nz = k + 1 + t;
                double fu = 0.0, fv = 0.0;
                if (nz == 1) {
                    const double w1 = cws[c * L] * area;
                    fu = -w1 * cu[c * L];
                    fv = -w1 * cv[c * L];
                } else if (nz <= K - 1) {
                    uvert_lds(cws + c * L, cu + c * L, cv + c * L, cZw + c * L, czb + c * L, K, area, nz, fu, fv);
                }
                if (t == 0) {
                    ftu = fu;
                    ftv = fv;
                } else {
                    fbu = fu;
                    fbv = fv;
                }
            }
            abu = abu + (R)((ftu - fbu) * r.acc);
            abv = abv + (R)((ftv - fbv) * r.acc);
        } else {
            abu = R(0);
            abv = R(0);
        }
    }
    const double mask = r.mask;
    PL(s.U_rhsAB, k, e, KC) = abu;
    PL(s.V_rhsAB, k, e, KC) = abv;
    // UV_rhs: the filt_bi_3D tendency (zero when that filter is off)
    const double uvx = s.uvu ? (double)PL(s.uvu, k, e, KC) : 0.0, uvy = s.uvu ? (double)PL(s.uvv, k, e, KC) : 0.0;
    PL(s.U_rhs, k, e, KC) = mask * (u + PL(s.Bpu, k, e, KC) * dt + (dt * (ur + abu * ff) * area_inv + uvx) / je);
    PL(s.V_rhs, k, e, KC) = mask * (v + PL(s.Bpv, k, e, KC) * dt + (dt * (vr + abv * ff) * area_inv + uvy) / je);
}

template <typename R>
__global__ __launch_bounds__(1024) void k_velrhs(DevMesh m, DevStateT<R> s, double dt, double ff, int hadv, int vadv,
                                                  double d0i) {
    velrhs_body<R, false, false>(m, s, dt, ff, hadv, vadv, d0i, 0.0);
}

// compute_vel_rhs as velrhs_body, with each lane owning W consecutive layers
// of its element column (LV<R>; nsigma a multiple of W): every column field
// and every neighbour column is read with one 16-byte load per W layers, so
// the kernel issues a quarter (FP32) or half (FP64) of the vector-memory
// instructions.  LDS holds, per column at a stride of K slots (each lane's
// run 16-byte aligned), the vertical-advection operands (sum w_cv*W, Zw,
// zbw, formed once per slot) and the velocity, which neighbours in the block
// read from there; velocity gradients come through 16-byte global loads.
// Each edge is oriented once (own/other element instead of edge_tri 1/2:
// the sums of the flux formula commute exactly), and the vertical fluxes are
// formed once per interface of the lane's run (the bottom flux of one layer
// is the top flux of the next) without divergence (uvert_uni).  Per layer
// the arithmetic and the order of the edge sum are those of velrhs_body:
// results are identical bit for bit.
template <typename R, bool ATM, bool FILT>
__device__ __forceinline__ void velrhs_vec_body(DevMesh m, DevStateT<R> s, double dt, double ff, int hadv, int vadv,
                                                double d0i, double tau) {
    constexpr int W = LV<R>::W;
    extern __shared__ __align__(16) unsigned char lds_[];
    const int L = m.L, K = m.K, KC = K;
    const ColFlat f = col_flat(m.E, K / W);  // f.k: the lane's run of W slots in its column
    ElemRec *er = (ElemRec *)lds_;                                        // [cpb]
    EdgeRec<R> *edr = (EdgeRec<R> *)(lds_ + f.cpb * kVelrhsElemBytes);   // [cpb][4]
    const int T = f.cpb * K;
    // the vertical advection's column arrays in the storage type: FP64 in the
    // drop-in; FP32 in the mixed path, which keeps the block's LDS within a
    // quarter of the compute unit's
    R *cws = (R *)(edr + 4 * f.cpb), *cZw = cws + T, *czb = cZw + T;
    R *cu = czb + T, *cv = cu + T;
    if (f.nc <= 0) return;
    // element and edge records, one thread per record (the element's four
    // edge slots, then the element), while every lane stages its run of the
    // column arrays: all of the block's loads are in flight before the one
    // barrier
    for (int t = threadIdx.x; t < 5 * f.nc; t += blockDim.x) {
        if (t >= 4 * f.nc) {  // element record: as velrhs_body
            const int c = t - 4 * f.nc, e = f.c0 + c;
            ElemRec r;
            r.col = elem_col(m, s, e);
            const int4 q = r.col.q;
            const double4 gx = m.gsx[e], gy = m.gsy[e];
            r.area = m.elem_area[e];
            r.cor = m.coriolis[e];
            r.mask = s.mask_wd[e];
            double pre0 = -kG * (m.atp1[q.x] * s.eta_n[q.x] - m.atp2[q.x] * 0.0);
            double pre1 = -kG * (m.atp1[q.y] * s.eta_n[q.y] - m.atp2[q.y] * 0.0);
            double pre2 = -kG * (m.atp1[q.z] * s.eta_n[q.z] - m.atp2[q.z] * 0.0);
            double pre3 = -kG * (m.atp1[q.w] * s.eta_n[q.w] - m.atp2[q.w] * 0.0);
            if (ATM) {
                const double4 w = r.col.w;
                const double am = w.x * s.mslp[q.x] + w.y * s.mslp[q.y] + w.z * s.mslp[q.z] + w.w * s.mslp[q.w];
                const double em = w.x * s.emp[q.x] + w.y * s.emp[q.y] + w.z * s.emp[q.z] + w.w * s.emp[q.w];
                pre0 = pre0 - (kPref - am) * d0i + kG * em * d0i;
                pre1 = pre1 - (kPref - am) * d0i + kG * em * d0i;
                pre2 = pre2 - (kPref - am) * d0i + kG * em * d0i;
                pre3 = pre3 - (kPref - am) * d0i + kG * em * d0i;
            }
            r.Fx = sum4(gx, pre0, pre1, pre2, pre3) * r.area;
            r.Fy = sum4(gy, pre0, pre1, pre2, pre3) * r.area;
            r.acc = sum4(r.col.w, m.ac[q.x], m.ac[q.y], m.ac[q.z], m.ac[q.w]);
            er[c] = r;
        } else {  // edge slot j of element column c
            const int c = t >> 2, j = t & 3, e = f.c0 + c;
            const int4 ied = m.el_ied[e];
            const int cc = j == 0 ? ied.x : (j == 1 ? ied.y : (j == 2 ? ied.z : ied.w));
            EdgeRec<R> g;
            g.cc = cc;
            if (cc != kEmpty && hadv) {
                const int ed = cc >= 0 ? cc : ~cc;
                const int2 nd = m.ednodes[ed];
                const int2 et = m.edtri[ed];
                const double4 cd = m.ecross[ed];
                const double2 dd = m.edxdy[ed];
                const bool own1 = cc < 0;   // this element is edge_tri(1)
                const int oe = own1 ? et.y : et.x;
                g.oK = oe * K;
                g.oc = (oe >= f.c0 && oe < f.c0 + f.nc) ? oe - f.c0 : -1;
                g.acc = (R)(0.5 * (m.ac[nd.x] + m.ac[nd.y]));
                // cross distances and cosines of the own (cx, cy, cos1) and
                // the other element (cz, cw, cos2)
                g.cx = (R)(own1 ? cd.x : cd.z);
                g.cy = (R)(own1 ? cd.y : cd.w);
                g.cz = (R)(own1 ? cd.z : cd.x);
                g.cw = (R)(own1 ? cd.w : cd.y);
                g.dx = (R)dd.x;
                g.dy = (R)dd.y;
                g.cos1 = (R)m.elem_cos[own1 ? et.x : et.y];
                g.cos2 = (R)m.elem_cos[own1 ? et.y : et.x];
                g.an1 = s.a[nd.x];
                g.an2 = s.a[nd.y];
            }
            edr[c * 4 + j] = g;
        }
    }
    const int c = f.c, e = f.c0 + c, k0 = f.k * W, ck0 = c * K + k0;
    const size_t ek0 = (size_t)e * KC + k0;
    ElemCol col;
    if (f.on()) {  // the column arrays of the vertical advection and the velocity
        col = elem_col(m, s, e);
        const int4 q = col.q;
        const LV<R> w0 = ldv_glb(s.W + (size_t)q.x * KC + k0), w1 = ldv_glb(s.W + (size_t)q.y * KC + k0);
        const LV<R> w2 = ldv_glb(s.W + (size_t)q.z * KC + k0), w3 = ldv_glb(s.W + (size_t)q.w * KC + k0);
#pragma unroll
        for (int i = 0; i < W; ++i) {
            const int k = k0 + i;
            cws[ck0 + i] = (R)sum4(col.w, w0.x[i], w1.x[i], w2.x[i], w3.x[i]);
            cZw[ck0 + i] = (R)(k < L ? col.Zw(m, k) : 0.0);
            czb[ck0 + i] = (R)col.zbw(m, k);
        }
        stv(cu + ck0, ldv_glb(s.U_n + ek0));
        stv(cv + ck0, ldv_glb(s.V_n + ek0));
    }
    __syncthreads();
    if (!f.on()) return;
    const int nl = min(W, L - k0);  // layers of this lane (the column's last lane: W-1)
    const ElemRec &r = er[c];
    const double area = r.area;
    const double area_inv = 1.0 / area;
    const LV<R> ua = ldv_glb(s.U_rhsAB + ek0), va = ldv_glb(s.V_rhsAB + ek0);
    double je[W], ur[W], vr[W];
    R abu[W], abv[W];
    const LV<R> uo = ldv_lds(cu + ck0), vo = ldv_lds(cv + ck0);
#pragma unroll
    for (int i = 0; i < W; ++i) {
        je[i] = col.Je(m, k0 + i < L ? k0 + i : L - 1);
        ur[i] = -(0.5 + kEps) * ua.x[i] + r.Fx * je[i];
        vr[i] = -(0.5 + kEps) * va.x[i] + r.Fy * je[i];
    }
    if (FILT) {  // viscosity_filt_3D: neighbours minus own U_n, elem_neighbors order
        const int4 q4 = m.en[e];
        const int4 nb4 = m.elnb[e];
        const int nbs[4] = {nb4.x, nb4.y, nb4.z, nb4.w};
        const int jend = (q4.x == q4.w) ? 3 : 4;
        double fu[W], fv[W];
#pragma unroll
        for (int i = 0; i < W; ++i) fu[i] = fv[i] = 0.0;
        for (int j = 0; j < jend; ++j) {
            if (nbs[j] >= 0) {
                const LV<R> un = ldv_glb(s.U_n + (size_t)nbs[j] * KC + k0), vn = ldv_glb(s.V_n + (size_t)nbs[j] * KC + k0);
#pragma unroll
                for (int i = 0; i < W; ++i) {
                    fu[i] = fu[i] + (double)un.x[i] - uo.x[i];
                    fv[i] = fv[i] + (double)vn.x[i] - vo.x[i];
                }
            }
        }
#pragma unroll
        for (int i = 0; i < W; ++i) {
            ur[i] = ur[i] + tau * fu[i] * area * je[i];
            vr[i] = vr[i] + tau * fv[i] * area * je[i];
        }
    }
    if (s.bhfu) {  // biharmonic_viscosity, second pass (fv_viscosity.F90:262-333)
#pragma unroll
        for (int i = 0; i < W; ++i)
            if (k0 + i < L) bh_gather(m, s, e, k0 + i, ur[i], vr[i]);
    }
#pragma unroll
    for (int i = 0; i < W; ++i) {
        abu[i] = (R)(vo.x[i] * r.cor * area * je[i]);
        abv[i] = (R)(-uo.x[i] * r.cor * area * je[i]);
    }
    if (hadv) {
        const LV<R> gxo = ldv_glb(s.gux + ek0), gyo = ldv_glb(s.guy + ek0);
        const LV<R> hxo = ldv_glb(s.gvx + ek0), hyo = ldv_glb(s.gvy + ek0);
        for (int j = 0; j < 4; ++j) {
            const EdgeRec<R> &g = edr[c * 4 + j];
            if (g.cc == kEmpty) break;
            const bool own1 = g.cc < 0;  // the element is edge_tri(1)
            const size_t io = (size_t)g.oK + k0;
            LV<R> ux, vx;
            if (g.oc >= 0) {  // the other element is a column of this block
                ux = ldv_lds(cu + g.oc * K + k0);
                vx = ldv_lds(cv + g.oc * K + k0);
            } else {
                ux = ldv_glb(s.U_n + io);
                vx = ldv_glb(s.V_n + io);
            }
            const LV<R> gxx = ldv_glb(s.gux + io), gyx = ldv_glb(s.guy + io);
            const LV<R> hxx = ldv_glb(s.gvx + io), hyx = ldv_glb(s.gvy + io);
            // g: own element's cross distances (cx, cy) and cosine (cos1),
            // the other's (cz, cw, cos2)
#pragma unroll
            for (int i = 0; i < W; ++i) {
                const double ds = m.dsig[k0 + i < L ? k0 + i : L - 1];
                const R uor = uo.x[i] - gxo.x[i] * g.cx - gyo.x[i] * g.cy;
                const R vor = vo.x[i] - hxo.x[i] * g.cx - hyo.x[i] * g.cy;
                const R uxr = ux.x[i] - gxx.x[i] * g.cz - gyx.x[i] * g.cw;
                const R vxr = vx.x[i] - hxx.x[i] * g.cz - hyx.x[i] * g.cw;
                const R jd = (R)(0.5 * (g.an1 * ds + g.an2 * ds));
                const R un = R(0.5) * R(kREarthD) * jd * ((uor + uxr) * g.dy - (vor * g.cos1 + vxr * g.cos2) * g.dx);
                const bool po = (un >= R(0)) == own1;   // the upwind element is this one
                const R uu = (po ? uor : uxr) * un * g.acc;
                const R vv = (po ? vor : vxr) * un * g.acc;
                if (g.cc >= 0) {
                    abu[i] = abu[i] + uu;
                    abv[i] = abv[i] + vv;
                } else {
                    abu[i] = abu[i] - uu;
                    abv[i] = abv[i] - vv;
                }
            }
        }
    }
    if (vadv) {
        if (r.mask != 0.0) {
            // flux through 1-based interface nz (0 at the bottom, nz = K);
            // the bottom flux of one layer is the top flux of the next
            const R *ws = cws + c * K, *zw = cZw + c * K, *zb = czb + c * K;
            const R *uc = cu + c * K, *vc = cv + c * K;
            auto flux = [&](int nz, double &fu, double &fv) {
                fu = 0.0;
                fv = 0.0;
                if (nz == 1) {
                    const double w1 = ws[0] * area;
                    fu = -w1 * uc[0];
                    fv = -w1 * vc[0];
                } else if (nz <= K - 1) {
                    uvert_uni(ws, uc, vc, zw, zb, K, area, nz, fu, fv);
                }
            };
            double ftu, ftv;
            flux(k0 + 1, ftu, ftv);
#pragma unroll
            for (int i = 0; i < W; ++i) {
                double fbu = 0.0, fbv = 0.0;
                if (i < nl) flux(k0 + i + 2, fbu, fbv);
                abu[i] = abu[i] + (R)((ftu - fbu) * r.acc);
                abv[i] = abv[i] + (R)((ftv - fbv) * r.acc);
                ftu = fbu;
                ftv = fbv;
            }
        } else {
#pragma unroll
            for (int i = 0; i < W; ++i) abu[i] = abv[i] = R(0);
        }
    }
    const double mask = r.mask;
    const LV<R> bu = ldv_glb(s.Bpu + ek0), bv = ldv_glb(s.Bpv + ek0);
    // the layer thicknesses and the own velocity again (from the LDS record and
    // columns), rather than held in registers through the flux loops
    const LV<R> ue = ldv_lds(cu + ck0), ve = ldv_lds(cv + ck0);
    LV<R> oa, ob, oc, od;
#pragma unroll
    for (int i = 0; i < W; ++i) {
        oa.x[i] = abu[i];
        ob.x[i] = abv[i];
        const double jei = r.col.Je(m, k0 + i < L ? k0 + i : L - 1);
        // UV_rhs: the filt_bi_3D tendency (zero when that filter is off)
        const double uvx = (s.uvu && i < nl) ? (double)s.uvu[ek0 + i] : 0.0;
        const double uvy = (s.uvu && i < nl) ? (double)s.uvv[ek0 + i] : 0.0;
        oc.x[i] = mask * (ue.x[i] + bu.x[i] * dt + (dt * (ur[i] + abu[i] * ff) * area_inv + uvx) / jei);
        od.x[i] = mask * (ve.x[i] + bv.x[i] * dt + (dt * (vr[i] + abv[i] * ff) * area_inv + uvy) / jei);
    }
    stv_n(s.U_rhsAB + ek0, oa, nl);
    stv_n(s.V_rhsAB + ek0, ob, nl);
    stv_n(s.U_rhs + ek0, oc, nl);
    stv_n(s.V_rhs + ek0, od, nl);
}

template <typename R>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_velrhs_v(DevMesh m, DevStateT<R> s, double dt, double ff, int hadv, int vadv,
                                                   double d0i) {
    velrhs_vec_body<R, false, false>(m, s, dt, ff, hadv, vadv, d0i, 0.0);
}
template <typename R>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_velrhs_v_atm(DevMesh m, DevStateT<R> s, double dt, double ff, int hadv,
                                                       int vadv, double d0i) {
    velrhs_vec_body<R, true, false>(m, s, dt, ff, hadv, vadv, d0i, 0.0);
}
template <typename R>
__global__ __launch_bounds__(256) void k_velrhs_v_filt(DevMesh m, DevStateT<R> s, double dt, double ff, int hadv,
                                                        int vadv, double d0i, double tau) {
    velrhs_vec_body<R, false, true>(m, s, dt, ff, hadv, vadv, d0i, tau);
}
template <typename R>
__global__ __launch_bounds__(256) void k_velrhs_v_atm_filt(DevMesh m, DevStateT<R> s, double dt, double ff, int hadv,
                                                            int vadv, double d0i, double tau) {
    velrhs_vec_body<R, true, true>(m, s, dt, ff, hadv, vadv, d0i, tau);
}
// the same with surface forcing (key_atm)
template <typename R>
__global__ __launch_bounds__(1024) void k_velrhs_atm(DevMesh m, DevStateT<R> s, double dt, double ff, int hadv,
                                                      int vadv, double d0i) {
    velrhs_body<R, true, false>(m, s, dt, ff, hadv, vadv, d0i, 0.0);
}
// the same with the neighbour-difference filter (filt_3D), without and with surface forcing
template <typename R>
__global__ __launch_bounds__(1024) void k_velrhs_filt(DevMesh m, DevStateT<R> s, double dt, double ff, int hadv,
                                                       int vadv, double d0i, double tau) {
    velrhs_body<R, false, true>(m, s, dt, ff, hadv, vadv, d0i, tau);
}
template <typename R>
__global__ __launch_bounds__(1024) void k_velrhs_atm_filt(DevMesh m, DevStateT<R> s, double dt, double ff, int hadv,
                                                           int vadv, double d0i, double tau) {
    velrhs_body<R, true, true>(m, s, dt, ff, hadv, vadv, d0i, tau);
}

// momentum_vert_impl_visc_tmp (fv_viscosity.F90:1374-1486) on tiles of 64
// element columns: the provisional velocity and Av (formed from snu,
// fv_mixing.F90:308-313) are staged in LDS; one lane per column
// then runs the surface stress, the forward sweep, the bottom condition by
// fic_point and the back substitution.  The sweep keeps its coefficients in
// the tile slots it has consumed (xx1(k+1) in Av's slot k, yu/yv in the
// velocity's slot k), rounded to the storage type as the Thomas scratch is.
// Dry elements get zero velocity.
template <typename R>
__global__ __launch_bounds__(kColT) void k_velvisc(DevMesh m, DevStateT<R> s, double dt, double rho0, double Cd_bz, int fic_point) {
    extern __shared__ __align__(16) unsigned char lds_[];
    const int K = m.K, L = m.L, KP = K + 1, KC = K;
    double *tA = (double *)lds_;
    R *tU = (R *)(tA + kColB * KP), *tV = tU + kColB * KP;
    const int c0 = col_first(), nc = min(kColB, m.E - c0);
    if (nc <= 0) return;
    tile_load(tU, s.U_rhs, c0, nc, K);
    tile_load(tV, s.V_rhs, c0, nc, K);
    if (tcol() < nc) {
        const int c = tcol(), e = c0 + c;
        const int4 q = m.en[e];
        const double4 w = m.wcv[e];
        for (int k0 = tsub(); k0 < K; k0 += 32) {  // eight levels per batch, loads first
            R v[8][4];
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const int k = k0 + 4 * j;
                if (k < K) {
                    v[j][0] = PL(s.snu, k, q.x, KC);
                    v[j][1] = PL(s.snu, k, q.y, KC);
                    v[j][2] = PL(s.snu, k, q.z, KC);
                    v[j][3] = PL(s.snu, k, q.w, KC);
                }
            }
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const int k = k0 + 4 * j;
                if (k < K) tA[c * KP + k] = sum4(w, v[j][0], v[j][1], v[j][2], v[j][3]);
            }
        }
    }
    __syncthreads();
    if ((int)threadIdx.x < nc) {
        const int c = threadIdx.x, e = c0 + c;
        double *A = tA + c * KP;
        R *U = tU + c * KP, *V = tV + c * KP;
        const double mask = s.mask_wd[e];
        if (mask == 0.0) {
            for (int k = 0; k < L; ++k) {
                U[k] = R(0);
                V[k] = R(0);
            }
        } else {
            const ElemCol col = elem_col(m, s, e);
            const int4 q = col.q;
            const double acc = sum4(col.w, m.ac[q.x], m.ac[q.y], m.ac[q.z], m.ac[q.w]);
            const double avL1 = A[L - 1], avK1 = A[K - 1];
            double xxp = 1.0, yup, yvp;
            {
                const double de = col.Je(m, 0) / (A[0] + A[1]);
                const double rw = de / rho0;
                yup = s.taux[e] * rw;
                yvp = s.tauy[e] * rw;
            }
            double je_prev = 0.0, je_cur = col.Je(m, 0);
            double av_cur = A[0];
            for (int k = 0; k < L; ++k) {
                const double je = je_cur;
                const double je_next = (k + 1 < L) ? col.Je(m, k + 1) : 0.0;
                const double ur = U[k], vr = V[k];
                const double av_next = A[k + 1];
                const double delta_1 = (k == 0) ? 0.5 * je : 0.5 * (je + je_prev);
                const double delta_2 = (k == L - 1) ? 0.5 * je : 0.5 * (je + je_next);
                const double zt = dt / je;
                const double aa = zt * av_cur / delta_1;
                const double cc = zt * av_next / delta_2;
                const double bb = aa + cc + 1.0;
                const double d = bb - aa * xxp;
                xxp = cc / d;
                yup = (aa * yup + ur) / d;
                yvp = (aa * yvp + vr) / d;
                A[k] = (double)(R)xxp;  // vxx(k+1)
                U[k] = (R)yup;          // vyu(k+1)
                V[k] = (R)yvp;          // vyv(k+1)
                av_cur = av_next;
                je_prev = je;
                je_cur = je_next;
            }
            const double a_Cd = s.C_d_el[e] * ((1.0 - acc) * Cd_bz + 1.0);
            const double de = col.Je(m, L - 1) / (avK1 + avL1);
            const double ub = PL(s.U_n, L - 1, e, KC), vb = PL(s.V_n, L - 1, e, KC);
            const double friction = de * a_Cd * sqrt(ub * ub + vb * vb);
            double ut, vt;
            if (fic_point == 0) {
                ut = 0.0;
                vt = 0.0;
            } else if (fic_point == 1) {
                const double aa = friction + 1.0 - xxp;
                ut = yup / aa;
                vt = yvp / aa;
            } else {
                const double aa = 1.0 / (1.0 - friction) - xxp;
                ut = yup / aa;
                vt = yvp / aa;
            }
            for (int k = L - 1; k >= 0; --k) {
                const double xx = A[k];
                ut = ut * xx + (double)U[k];
                vt = vt * xx + (double)V[k];
                U[k] = (R)(ut * mask);
                V[k] = (R)(vt * mask);
            }
        }
    }
    __syncthreads();
    tile_store(s.U_rhs, tU, c0, nc, K, L);
    tile_store(s.V_rhs, tV, c0, nc, K, L);
}

// momentum_vert_impl_visc (fv_viscosity.F90:1266-1371, the im_vert_visc
// option): a Thomas solve per element column whose off-diagonals come from Av
// over the spacing of the w_cv-weighted layer depths Z, with the wind stress
// added to the surface row and the bottom drag -a_Cd*|u_b|*u_b (explicit in
// the bottom layer's velocity U_n) to the bottom row.  Same tiles as
// k_velvisc; the forward sweep keeps c' in Av's consumed slot and the primed
// right-hand sides in the velocity slots.  Every element is solved and the
// result multiplied by mask_wd, as the reference does.
template <typename R>
__global__ __launch_bounds__(kColT) void k_velvisc_im(DevMesh m, DevStateT<R> s, double dt, double rho0, double Cd_bz) {
    extern __shared__ __align__(16) unsigned char lds_[];
    const int K = m.K, L = m.L, KP = K + 1, KC = K;
    double *tA = (double *)lds_;
    R *tU = (R *)(tA + kColB * KP), *tV = tU + kColB * KP;
    const int c0 = col_first(), nc = min(kColB, m.E - c0);
    if (nc <= 0) return;
    tile_load(tU, s.U_rhs, c0, nc, K);
    tile_load(tV, s.V_rhs, c0, nc, K);
    if (tcol() < nc) {
        const int c = tcol(), e = c0 + c;
        const int4 q = m.en[e];
        const double4 w = m.wcv[e];
        for (int k = tsub(); k < K; k += 4)
            tA[c * KP + k] = sum4(w, PL(s.snu, k, q.x, KC), PL(s.snu, k, q.y, KC), PL(s.snu, k, q.z, KC),
                                  PL(s.snu, k, q.w, KC));
    }
    __syncthreads();
    if ((int)threadIdx.x < nc) {
        const int c = threadIdx.x, e = c0 + c;
        double *A = tA + c * KP;
        R *U = tU + c * KP, *V = tV + c * KP;
        const double mask = s.mask_wd[e];
        const ElemCol col = elem_col(m, s, e);
        const int4 q = col.q;
        const double acc = sum4(col.w, m.ac[q.x], m.ac[q.y], m.ac[q.z], m.ac[q.w]);
        const double a_Cd = s.C_d_el[e] * ((1.0 - acc) * Cd_bz + 1.0);
        const double ub = PL(s.U_n, L - 1, e, KC), vb = PL(s.V_n, L - 1, e, KC);
        const double friction = -a_Cd * sqrt(ub * ub + vb * vb);
        const double zinvB = dt / col.Je(m, L - 1);
        double Zu = 0.0, Zc = col.Zw(m, 0);
        double cpp = 0.0, upp = 0.0, vpp = 0.0;  // c', u', v' of the row above
        for (int k = 0; k < L; ++k) {
            const double Zb = (k + 1 < L) ? col.Zw(m, k + 1) : 0.0;
            const double zinv = dt / col.Je(m, k);
            double aa, bb, cc;
            if (k == 0) {  // first row (fv_viscosity.F90:1315-1320)
                cc = -A[1] / (Zb - Zc) * zinv;
                aa = 0.0;
                bb = -cc + 1.0;
            } else if (k == L - 1) {  // last row (1308-1313)
                aa = -A[k] / (Zc - Zu) * zinv;
                bb = -aa + 1.0;
                cc = 0.0;
            } else {  // regular rows (1299-1306)
                aa = -A[k] / (Zc - Zu) * zinv;
                cc = -A[k + 1] / (Zb - Zc) * zinv;
                bb = -aa - cc + 1.0;
            }
            double ur = U[k], vr = V[k];
            if (k == 0) {  // surface stress (1327-1328)
                ur = ur + zinv * s.taux[e] / rho0;
                vr = vr + zinv * s.tauy[e] / rho0;
            }
            if (k == L - 1) {  // bottom friction (1330-1334)
                ur = ur + zinvB * friction * ub;
                vr = vr + zinvB * friction * vb;
            }
            if (k == 0) {
                cpp = cc / bb;
                upp = ur / bb;
                vpp = vr / bb;
            } else {
                const double mm = bb - cpp * aa;
                cpp = cc / mm;
                upp = (ur - upp * aa) / mm;
                vpp = (vr - vpp * aa) / mm;
            }
            A[k] = cpp;  // Av(k) is not read again
            U[k] = (R)upp;
            V[k] = (R)vpp;
            Zu = Zc;
            Zc = Zb;
        }
        double ut = U[L - 1], vt = V[L - 1];
        U[L - 1] = (R)(ut * mask);
        V[L - 1] = (R)(vt * mask);
        for (int k = L - 2; k >= 0; --k) {
            ut = (double)U[k] - A[k] * ut;
            vt = (double)V[k] - A[k] * vt;
            U[k] = (R)(ut * mask);
            V[k] = (R)(vt * mask);
        }
    }
    __syncthreads();
    tile_store(s.U_rhs, tU, c0, nc, K, L);
    tile_store(s.V_rhs, tV, c0, nc, K, L);
}

// ---------------------------------------------------------------------------
// barotropic substep, timestep_AB_2D: src/fv_main.F90:1372-1634
// ---------------------------------------------------------------------------
constexpr double kAB0 = 1.5 + kBeta, kAB1 = 0.5 + 2.0 * kBeta;

__device__ __forceinline__ double ab3(double a, double b, double c) { return kAB0 * a - kAB1 * b + kBeta * c; }

template <typename R>
__device__ __forceinline__ double etaAB3(const DevStateT<R> &s, int n) {
    return ab3(s.eta_n[n], s.eta_n_1[n], s.eta_n_2[n]);
}

// Geometry of the 2-D momentum path: the FP64 arrays for the drop-in, their
// FP32 copies in the mixed-precision path (half the bytes per element/edge).
template <typename R>
__device__ __forceinline__ double4 mom_geo4(const double4 *d, const float4 *f, int i) {
    if constexpr (sizeof(R) == sizeof(float)) {
        const float4 v = f[i];
        return make_double4(v.x, v.y, v.z, v.w);
    } else {
        return d[i];
    }
}
template <typename R>
__device__ __forceinline__ double2 mom_geo2(const double2 *d, const float2 *f, int i) {
    if constexpr (sizeof(R) == sizeof(float)) {
        const float2 v = f[i];
        return make_double2(v.x, v.y);
    } else {
        return d[i];
    }
}


// AB3 velocity + element depth of compute_ssh_rhs_elem (fv_average_dynamic.F90:66-86),
// first substep of a step; k_bt_elem2 forms both for the substeps after it.
template <typename R>
__global__ __launch_bounds__(256) void k_bt_elem1(DevMesh m, DevStateT<R> s, double Dmin) {
    int e = TID();
    if (e >= m.E) return;
    const int4 q = m.en[e];
    const double4 w = m.wcv[e];
    const double2 u0 = ld2(s.U2, e), u1 = ld2(s.U2_1, e), u2 = ld2(s.U2_2, e);
    st2(s.UAB, e, ab3(u0.x, u1.x, u2.x), ab3(u0.y, u1.y, u2.y));
    const int ns[4] = {q.x, q.y, q.z, q.w};
    double ad[4];
    for (int i = 0; i < 4; ++i) {
        ad[i] = Dmin;
        double t = etaAB3(s, ns[i]) + m.depth[ns[i]];
        if (t > Dmin) ad[i] = t;
    }
    s.aux_el[e] = sum4(w, ad[0], ad[1], ad[2], ad[3]);
}

// Per edge, the transport through it of compute_ssh_rhs_elem
// (fv_average_dynamic.F90:106-169): both elements' AB3 velocity times their
// depth, computed once per edge for the gather of its two nodes.
template <typename R>
__global__ __launch_bounds__(256) void k_bt_edge_ssh(DevMesh m, DevStateT<R> s) {
    const int ed = TID();
    if (ed >= m.D) return;
    const int2 et = m.edtri[ed];
    const double4 cd = m.ecross[ed];
    const double2 ua = ld2(s.UAB, et.x);
    double c1 = (ua.y * cd.x - ua.x * cd.y) * s.aux_el[et.x];
    if (et.y >= 0) {
        const double2 ub = ld2(s.UAB, et.y);
        c1 = c1 + (-ub.y * cd.z + ub.x * cd.w) * s.aux_el[et.y];
    }
    s.fl_ssh[ed] = c1;
}

// ssh_rhs gather (fv_average_dynamic.F90:106-169) with the river discharge of
// the mouths (177-183; runoff null without rivers), eta update (fv_main.F90:1432-1435),
// tidal OB value (fv_average_dynamic.F90:190-216, host-evaluated), the AM4
// elevation of update_2D_vel (fv_average_dynamic.F90:288-296) and the time-level
// shift (fv_main.F90:1547-1551).  Also the clamped AB3 depth of the coming
// substep's compute_ssh_rhs_elem (its elevation levels are the shifted ones),
// which k_bt_elem2 averages to the element.
template <typename R>
__global__ __launch_bounds__(256) void k_bt_node(DevMesh m, DevStateT<R> s, double dt_2D, const double *__restrict__ tide,
                                                 int atm, double d0i, const double *__restrict__ runoff, double Dmin) {
    int n = TID();
    if (n >= m.N) return;
    double ssh = 0.0;
    #pragma unroll 4
    for (int j = m.n2e_ptr[n]; j < m.n2e_ptr[n + 1]; ++j) {
        const int c = m.n2e[j];
        const double c1 = s.fl_ssh[c >= 0 ? c : ~c];
        ssh = (c >= 0) ? ssh + c1 : ssh - c1;
    }
    if (runoff) {
        const double q = runoff[n];
        if (q != 0.0) ssh = ssh + q;
    }
    s.ssh_rhs[n] = ssh;
    const double e0 = s.eta_n[n], e1 = s.eta_n_1[n], e2 = s.eta_n_2[n];
    double en;
    if (m.index[n] < 2) en = e0 + dt_2D * ssh / m.area[n];
    else en = tide ? tide[m.obidx[n]] : ssh;
    const double de = 0.614, ga = 0.088, ep = 0.013;
    const double fe = 1.0 - de - ga - ep;
    double ea = de * en + fe * e0 + ga * e1 + ep * e2;
    // with surface forcing: evaporation minus precipitation and sea-level
    // pressure (fv_average_dynamic.F90:278-285)
    const double eg = atm ? kG * (m.atp1[n] * ea - s.emp[n] * d0i - m.atp2[n] * 0.0) + (kPref - s.mslp[n]) * d0i
                          : kG * (m.atp1[n] * ea - m.atp2[n] * 0.0);
    if (!isfinite(en)) atomicOr(s.err + 2, 1);  // non-finite state flag (FCG_E_NAN)
    s.etaAB[n] = ea;
    s.eta_g[n] = eg;
    s.eta_n_2[n] = e1;
    s.eta_n_1[n] = e0;
    s.eta_n[n] = en;
    // the four node values update_2D_vel gathers, as one record per node
    ((double4 *)s.node2d)[n] = make_double4(ea, eg, e0, en);
    const double t = ab3(en, e0, e1) + m.depth[n];
    s.adn[n] = t > Dmin ? t : Dmin;
}

// vel_gradients_2D(UAB): src/fv_gradients.F90:97-176
template <typename R>
__global__ __launch_bounds__(256) void k_bt_velgrad(DevMesh m, DevStateT<R> s) {
    int e = TID();
    if (e >= m.E) return;
    const int4 q = m.en[e];
    const int jend = (q.x == q.w) ? 3 : 4;
    const int4 nb = m.elnb[e];
    const int4 ed = m.eledges[e];
    const double4 gx = mom_geo4<R>(m.gvx, m.gvx_f, e), gy = mom_geo4<R>(m.gvy, m.gvy_f, e);
    const int nbs[4] = {nb.x, nb.y, nb.z, nb.w};
    const int eds[4] = {ed.x, ed.y, ed.z, ed.w};
    const double gxs[4] = {gx.x, gx.y, gx.z, gx.w}, gys[4] = {gy.x, gy.y, gy.z, gy.w};
    const double2 uv = ld2(s.UAB, e);
    const double u = uv.x, v = uv.y;
    double g0 = 0.0, g1 = 0.0, g2 = 0.0, g3 = 0.0;
    for (int j = 0; j < jend; ++j) {
        int e2 = nbs[j];
        if (e2 >= 0) {
            const double2 w2 = ld2(s.UAB, e2);
            g0 = g0 + gxs[j] * (w2.x - u);
            g1 = g1 + gys[j] * (w2.x - u);
            g2 = g2 + gxs[j] * (w2.y - v);
            g3 = g3 + gys[j] * (w2.y - v);
        } else {
            double2 dd = mom_geo2<R>(m.edxdy, m.edxdy_f, eds[j]);
            double zc0 = dd.x * m.elem_cos[e], zc1 = dd.y;
            double un = -2.0 * (u * zc1 - v * zc0) / (zc0 * zc0 + zc1 * zc1);
            g0 = g0 + gxs[j] * un * zc1;
            g1 = g1 + gys[j] * un * zc1;
            g2 = g2 - gxs[j] * un * zc0;
            g3 = g3 - gys[j] * un * zc0;
        }
    }
    st4(s.vel_grad, e, g0, g1, g2, g3);
}

// first substep only: dmean_n (fv_main.F90:1467-1473) and compute_puls_vel
// (fv_3D_vel.F90:257-307), per element: the factor a = dmean_n/D^{n+1} of
// U_puls = U_n - a*U_n_2D.  The pulsation velocity itself is formed where it
// is consumed, in compute_el2nodes_3D (k_el2nodes with puls = 1): on the
// supported path nothing else reads it.  After k_bt_node's shift eta_n holds
// eta^{n+1} (= ssh_rhs of the reference) and eta_n_1 holds eta^n.  Also caches
// the bottom-layer velocity U_n(nsigma-1), which update_2D_vel reads in every
// substep, as a 2-D field.
template <typename R>
__global__ __launch_bounds__(256) void k_bt_puls(DevMesh m, DevStateT<R> s, double Dmin) {
    const int KC = m.K;
    const int e = TID();
    if (e >= m.E) return;
    const int4 q = m.en[e];
    const double4 w = m.wcv[e];
    const double dm = dmaxr(Dmin, w.x * (s.eta_n_1[q.x] + m.depth[q.x]) + w.y * (s.eta_n_1[q.y] + m.depth[q.y])
                                  + w.z * (s.eta_n_1[q.z] + m.depth[q.z]) + w.w * (s.eta_n_1[q.w] + m.depth[q.w]));
    s.dmean_n[e] = dm;
    const double snew = sum4(w, s.eta_n[q.x] + m.depth[q.x], s.eta_n[q.y] + m.depth[q.y],
                             s.eta_n[q.z] + m.depth[q.z], s.eta_n[q.w] + m.depth[q.w]);
    s.puls_ap[e] = dm / dmaxr(Dmin, snew);
    st2(s.ubot, e, PL(s.U_n, m.L - 1, e, KC), PL(s.V_n, m.L - 1, e, KC));
}

// compute_el2nodes_3D (src/fv_utilit.F90:669-689), one thread per node and
// layer on node-gather blocks: area-weighted mean over the node's elements
// in ascending id.  P: the element values are the pulsation velocity
// U_puls = U_n - a*U_n_2D (compute_puls_vel, fv_3D_vel.F90:286-295), formed
// here in the storage type from U_n, a (k_bt_puls) and U_n_2D.
template <typename R, bool P>
__global__ __launch_bounds__(1024) void k_el2nodes(DevMesh m, const R *__restrict__ Ue, const R *__restrict__ Ve,
                           R *__restrict__ Un, R *__restrict__ Vn, const double *__restrict__ ap,
                           const double *__restrict__ U2) {
    extern __shared__ __align__(16) unsigned char lds_[];
    const int KC = m.K;
    const ColFlat f = col_flat(m.N, m.L);
    if (f.nc <= 0) return;
    const int2 jr = csr_block(m.n2el_ptr, f);
    PulsRec *rec = (PulsRec *)lds_;
    for (int t = threadIdx.x; t < jr.y - jr.x; t += blockDim.x) {
        PulsRec r;
        const int e = m.n2el[jr.x + t];
        r.elK = m.n2elK[jr.x + t];
        r.ar = m.elem_area[e];
        if (P) {
            r.ap = ap[e];
            r.u2 = U2[2 * e];
            r.v2 = U2[2 * e + 1];
        }
        rec[t] = r;
    }
    __syncthreads();
    if (!f.on()) return;
    const int n = f.c0 + f.c, k = f.k;
    const int p0 = m.n2el_ptr[n] - jr.x, p1 = m.n2el_ptr[n + 1] - jr.x;
    double sa = 0.0;
    for (int j = p0; j < p1; ++j) sa = sa + rec[j].ar;
    const double tvi = 1.0 / sa;
    double su = 0.0, sv = 0.0;
    for (int j = p0; j < p1; ++j) {
        const double ar = rec[j].ar;
        R ue = Ue[rec[j].elK + k], ve = Ve[rec[j].elK + k];
        if (P) {
            const double un = ue, vn = ve;
            ue = (R)(un - rec[j].ap * rec[j].u2);
            ve = (R)(vn - rec[j].ap * rec[j].v2);
        }
        su = su + ue * ar;
        sv = sv + ve * ar;
    }
    PL(Un, k, n, KC) = su * tvi;
    PL(Vn, k, n, KC) = sv * tvi;
}

// momentum_adv_P1_3D_to_2D (fv_advection.F90:676-793), edge pass: one thread
// per (edge, layer) forms uu*un*Je for both elements of an internal edge at
// its layer (each element weighs the edge term with its own Je); one thread
// per (edge, element side, component) then sums them over the layers in order
// into fl_p1[4*ed + 2*side + comp] (side 0: edge_tri(1), side 1: edge_tri(2)).
// Device edge ids follow the locality order, so internal and boundary edges
// are interleaved: a boundary edge (one element) gets zeros and is never read.
template <typename R>
__global__ __launch_bounds__(256) void k_bt_p1edge(DevMesh m, DevStateT<R> s) {
    extern __shared__ __align__(16) unsigned char lds_[];
    const int L = m.L, KC = m.K;
    const ColFlat f = col_flat(m.D, L);
    const int T = f.cpb * L;
    double *t = (double *)lds_;             // [side*2 + comp][column*L + layer]
    P1Edge *rec = (P1Edge *)(t + 4 * T);    // [column]
    if (f.nc <= 0) return;
    if ((int)threadIdx.x < f.nc) {
        const int ed = f.c0 + threadIdx.x;
        const int2 nd = m.ednodes[ed];
        const int2 et = m.edtri[ed];
        const double2 dd = m.edxdy[ed];
        P1Edge r;
        r.n1K = nd.x * KC;
        r.n2K = nd.y * KC;
        r.e1K = -1;
        r.e2K = -1;
        r.xe = r.ye = r.an1 = r.an2 = 0.0;
        if (et.x >= 0 && et.y >= 0) {  // internal edge
            r.e1K = et.x * KC;
            r.e2K = et.y * KC;
            r.xe = dd.x * kREarthD * 0.5 * (m.elem_cos[et.x] + m.elem_cos[et.y]);
            r.ye = dd.y * kREarthD;
            r.an1 = s.a[nd.x];
            r.an2 = s.a[nd.y];
        }
        rec[threadIdx.x] = r;
    }
    __syncthreads();
    if (f.on() && rec[f.c].e1K < 0) {
        const int i = f.c * L + f.k;
        t[i] = t[T + i] = t[2 * T + i] = t[3 * T + i] = 0.0;
    } else if (f.on()) {
        const P1Edge &r = rec[f.c];
        const int k = f.k;
        const double ds = m.dsig[k];
        double uu = 0.5 * (s.Unode_p[r.n1K + k] + s.Unode_p[r.n2K + k]);
        double vv = 0.5 * (s.Vnode_p[r.n1K + k] + s.Vnode_p[r.n2K + k]);
        double jd = 0.5 * (r.an1 * ds + r.an2 * ds);
        double un = (uu * r.ye - vv * r.xe) * jd;
        const double pu = uu * un, pv = vv * un;
        const double je1 = s.Je[r.e1K + k], je2 = s.Je[r.e2K + k];
        const int i = f.c * L + k;
        t[i] = pu * je1;
        t[T + i] = pv * je1;
        t[2 * T + i] = pu * je2;
        t[3 * T + i] = pv * je2;
    }
    __syncthreads();
    for (int q = threadIdx.x; q < 4 * f.nc; q += blockDim.x) {
        const int c = q >> 2, j = q & 3;
        const double *p = t + j * T + c * L;
        double sum = 0.0;
        for (int k = 0; k < L; ++k) sum = sum + p[k];
        s.fl_p1[4 * (size_t)(f.c0 + c) + j] = sum;
    }
}

// element pass: the element's internal edges in ascending id, each term with
// the edge's ac weight, then U_rhs_2D_3D /= dmean_n (fv_main.F90:1504-1513)
template <typename R>
__global__ __launch_bounds__(256) void k_bt_p1elem(DevMesh m, DevStateT<R> s, double tau) {
    const int KC = m.K;
    const int e = TID();
    if (e >= m.E) return;
    const int4 ied = m.el_ied[e];
    const int slots[4] = {ied.x, ied.y, ied.z, ied.w};
    const double2 *fl = (const double2 *)s.fl_p1;
    double ru = 0.0, rv = 0.0;
    for (int j = 0; j < 4; ++j) {
        const int cc = slots[j];
        if (cc == kEmpty) break;
        const int ed = cc >= 0 ? cc : ~cc;
        const int2 nd = m.ednodes[ed];
        const double acc = (m.ac[nd.x] + m.ac[nd.y]) * .5;
        if (cc >= 0) {  // element is edge_tri(2)
            const double2 v = fl[2 * (size_t)ed + 1];
            ru = ru + acc * v.x;
            rv = rv + acc * v.y;
        } else {
            const double2 v = fl[2 * (size_t)ed];
            ru = ru - acc * v.x;
            rv = rv - acc * v.y;
        }
    }
    if (tau != 0.0) {  // filt_3D: viscosity_filt_3D_to_2D (fv_viscosity.F90:870-939, fv_main.F90:1499)
        // on the pulsation transports U_puls*Je, U_puls = U_n - puls_ap*U_n_2D
        // (compute_puls_vel), layer by layer, added to the column sum in order
        const int4 q4 = m.en[e];
        const int4 nb4 = m.elnb[e];
        const int nbs[4] = {nb4.x, nb4.y, nb4.z, nb4.w};
        const int jend = (q4.x == q4.w) ? 3 : 4;
        const double ao = s.puls_ap[e], uo2 = s.U2[2 * e], vo2 = s.U2[2 * e + 1];
        const double area = m.elem_area[e];
        for (int k = 0; k < m.L; ++k) {
            const double jo = (double)PL(s.Je, k, e, KC);
            const double po = ((double)PL(s.U_n, k, e, KC) - ao * uo2) * jo;
            const double qo = ((double)PL(s.V_n, k, e, KC) - ao * vo2) * jo;
            double fu = 0.0, fv = 0.0;
            for (int j = 0; j < jend; ++j) {
                const int nb = nbs[j];
                if (nb >= 0) {
                    const double an = s.puls_ap[nb], jn = (double)PL(s.Je, k, nb, KC);
                    fu = fu + ((double)PL(s.U_n, k, nb, KC) - an * s.U2[2 * nb]) * jn - po;
                    fv = fv + ((double)PL(s.V_n, k, nb, KC) - an * s.U2[2 * nb + 1]) * jn - qo;
                }
            }
            ru = ru + tau * fu * area * jo;
            rv = rv + tau * fv * area * jo;
        }
    }
    if (s.bhub) {  // biharmonic_viscosity_3D_to_2D: the column sum times Je (fv_viscosity.F90:1200-1206)
        for (int k = 0; k < m.L; ++k) {
            const double jo = (double)PL(s.Je, k, e, KC);
            ru = ru + (double)PL(s.bhub, k, e, KC) * jo;
            rv = rv + (double)PL(s.bhvb, k, e, KC) * jo;
        }
    }
    st2(s.U_rhs_2D_3D, e, ru / s.dmean_n[e], rv / s.dmean_n[e]);
}

// momentum_adv_upwind_2D edge flux: fv_advection.F90:589-643
template <typename R>
__device__ __forceinline__ void edge_uu2d(const DevMesh &m, const DevStateT<R> &s, int ed, double Dmin, double &uu,
                                          double &vv) {
    const int2 nd = m.ednodes[ed];
    const int2 et = m.edtri[ed];
    const double dm = dmaxr(Dmin, 0.5 * ((s.etaAB[nd.x] + m.depth[nd.x]) + (s.etaAB[nd.y] + m.depth[nd.y])));
    const double acc = (m.ac[nd.x] + m.ac[nd.y]) * 0.5;
    const double4 cd = mom_geo4<R>(m.ecross, m.ecross_f, ed);
    const double2 dd = mom_geo2<R>(m.edxdy, m.edxdy_f, ed);
    const int e1 = et.x, e2 = et.y;
    const double4 g1 = ld4(s.vel_grad, e1), g2 = ld4(s.vel_grad, e2);
    const double2 w1 = ld2(s.UAB, e1), w2 = ld2(s.UAB, e2);
    double u1 = w1.x - g1.x * cd.x - g1.y * cd.y;
    double v1 = w1.y - g1.z * cd.x - g1.w * cd.y;
    double u2 = w2.x - g2.x * cd.z - g2.y * cd.w;
    double v2 = w2.y - g2.z * cd.z - g2.w * cd.w;
    double un = 0.5 * kREarthD * dm * ((u1 + u2) * dd.y - (v1 * m.elem_cos[e1] + v2 * m.elem_cos[e2]) * dd.x);
    if (un >= 0.0) {
        uu = un * w1.x * acc;
        vv = un * w1.y * acc;
    } else {
        uu = un * w2.x * acc;
        vv = un * w2.y * acc;
    }
}

// momentum_adv_upwind_2D flux of every internal edge, once per edge (each is
// gathered by its two elements in k_bt_elem2)
template <typename R>
__global__ __launch_bounds__(256) void k_bt_edge_mom(DevMesh m, DevStateT<R> s, double Dmin) {
    const int ed = TID();
    if (ed >= m.D || m.edtri[ed].y < 0) return;
    double uu, vv;
    edge_uu2d(m, s, ed, Dmin, uu, vv);
    st2(s.fl_uv, ed, uu, vv);
}

// update_2D_vel (fv_average_dynamic.F90:228-443, 3-D branch) for one element,
// followed by the velocity shift and U_filt accumulation (fv_main.F90:1569-1606).
// With uab_next, also the AB3 velocity and element depth of the following
// substep's compute_ssh_rhs_elem (k_bt_elem1's values, from the shifted
// velocities and k_bt_node's clamped node depths); the velocity goes to the
// other of the two AB3 buffers, since the 2-D filter reads the neighbours'
// current one.  C_d_el = C_d is set on the first substep only: nothing
// changes it between substeps.
template <typename R>
__global__ __launch_bounds__(256, 5) void k_bt_elem2(DevMesh m, DevStateT<R> s, double dt_2D, double Dmin, double rho0, double C_d, double Cd_bz,
                           int wet_dry, int adv2d, int first, double tau2d, R *__restrict__ uab_next) {
    const int KC = m.K;
    int e = TID();
    if (e >= m.E) return;
    const int E = m.E, L = m.L;
    const int4 q = m.en[e];
    const double4 w = m.wcv[e], gx = mom_geo4<R>(m.gsx, m.gsx_f, e), gy = mom_geo4<R>(m.gsy, m.gsy_f, e);
    const double area = m.elem_area[e], cor = m.coriolis[e];
    const double d0 = m.depth[q.x], d1 = m.depth[q.y], d2 = m.depth[q.z], d3 = m.depth[q.w];
    // per node: etaAB, eta_g, eta^n (eta_n_1), eta^{n+1} (eta_n), written by k_bt_node this substep
    const double4 r0 = ((const double4 *)s.node2d)[q.x], r1 = ((const double4 *)s.node2d)[q.y];
    const double4 r2 = ((const double4 *)s.node2d)[q.z], r3 = ((const double4 *)s.node2d)[q.w];
    const double ea0 = r0.x, ea1 = r1.x, ea2 = r2.x, ea3 = r3.x;
    double mask;
    if (wet_dry) {  // wad_mask: fv_main.F90:1284-1294
        double mn = fmin(fmin(fmin(d0, d1), d2), d3);
        double mx = fmax(fmax(fmax(ea0, ea1), ea2), ea3);
        mask = (mn + mx > Dmin) ? 1.0 : 0.0;
        s.mask_wd[e] = mask;
    } else {
        mask = s.mask_wd[e];
    }
    if (first) s.C_d_el[e] = C_d;
    const double rho_inv = 1.0 / (rho0);
    const double dm = dmaxr(Dmin, w.x * (ea0 + d0) + w.y * (ea1 + d1) + w.z * (ea2 + d2) + w.w * (ea3 + d3));
    const double eg0 = r0.y, eg1 = r1.y, eg2 = r2.y, eg3 = r3.y;
    const double2 wab = ld2(s.UAB, e);
    const double uab = wab.x, vab = wab.y;
    double ru = area * (-dm * (gx.x * eg0 + gx.y * eg1 + gx.z * eg2 + gx.w * eg3) + vab * dm * cor + s.taux[e] * rho_inv);
    double rv = area * (-dm * (gy.x * eg0 + gy.y * eg1 + gy.z * eg2 + gy.w * eg3) - uab * dm * cor + s.tauy[e] * rho_inv);
    if (adv2d) {  // edge fluxes of k_bt_edge_mom, in ascending edge id
        const int4 ied = m.el_ied[e];
        const int slots[4] = {ied.x, ied.y, ied.z, ied.w};
        for (int j = 0; j < 4; ++j) {
            int c = slots[j];
            if (c == kEmpty) break;
            const double2 f = ld2(s.fl_uv, c >= 0 ? c : ~c);
            if (c >= 0) {
                ru = ru + f.x;
                rv = rv + f.y;
            } else {
                ru = ru - f.x;
                rv = rv - f.y;
            }
        }
    }
    if (tau2d != 0.0) {  // filt_2D: viscosity_filt_2D (fv_viscosity.F90:769-808, fv_average_dynamic.F90:346)
        const int4 nb4 = m.elnb[e];
        const int nbs[4] = {nb4.x, nb4.y, nb4.z, nb4.w};
        const int jend = (q.x == q.w) ? 3 : 4;
        double fu = 0.0, fv = 0.0;
        for (int j = 0; j < jend; ++j) {
            if (nbs[j] >= 0) {
                const double2 wn = ld2(s.UAB, nbs[j]);
                fu = fu + wn.x - uab;
                fv = fv + wn.y - vab;
            }
        }
        const double dsum = sum4(w, d0, d1, d2, d3);
        ru = ru + tau2d * fu * area * dsum;
        rv = rv + tau2d * fv * area * dsum;
    }
    if (s.bh_fl2) bh2_gather(m, s, e, ru, rv);  // biharmonic_2D, second pass (fv_viscosity.F90:456-541)
    const double2 r23 = ld2(s.U_rhs_2D_3D, e);
    ru = ru + r23.x;
    rv = rv + r23.y;
    // eta_n_1 holds eta^n, eta_n holds eta^{n+1} (the reference's ssh_rhs)
    const double dmo = dmaxr(Dmin, sum4(w, r0.z + d0, r1.z + d1, r2.z + d2, r3.z + d3));
    const double acc = m.acc_el[e];  // sum w_cv*ac, formed at set-up in the same order
    const double a_Cd = C_d * ((1.0 - acc) * Cd_bz + 1.0);
    const double2 bot = ld2(s.ubot, e);  // U_n, V_n at the bottom layer
    const double ub = bot.x, vb = bot.y;
    const double friction = a_Cd * sqrt(ub * ub + vb * vb);
    const double snew = sum4(w, r0.w + d0, r1.w + d1, r2.w + d2, r3.w + d3);
    const double ffc = 1.0 / (dmaxr(Dmin, snew) + friction * dt_2D);
    const double2 u0 = ld2(s.U2, e), h1 = ld2(s.U2_1, e), bp = ld2(s.Bar_pr_2D, e);
    const double U0 = u0.x, V0 = u0.y;
    double uv2x = 0.0, uv2y = 0.0;
    if (s.fb2c) {  // UV2_rhs of viscosity_filt2x_2D, second pass (fv_viscosity.F90:735-758)
        const int4 nb4 = m.elnb[e];
        const int nbs[4] = {nb4.x, nb4.y, nb4.z, nb4.w};
        const double2 c0 = ld2(s.fb2c, e);
        for (int j = 0; j < 4; ++j) {
            if (nbs[j] < 0 || (j == 3 && nbs[3] == nbs[0])) continue;
            const double2 cn = ld2(s.fb2c, nbs[j]);
            if (j == 0) {
                uv2x = -c0.x + cn.x;
                uv2y = -c0.y + cn.y;
            } else {
                uv2x = uv2x - c0.x + cn.x;
                uv2y = uv2y - c0.y + cn.y;
            }
        }
    }
    const double un = mask * ((dmo * U0 - dt_2D * friction * (ub - U0) + dt_2D * bp.x
                               + dt_2D * ru / area) * ffc + uv2x);
    const double vn = mask * ((dmo * V0 - dt_2D * friction * (vb - V0) + dt_2D * bp.y
                               + dt_2D * rv / area) * ffc + uv2y);
    st2(s.U2_2, e, h1.x, h1.y);
    st2(s.U2_1, e, U0, V0);
    st2(s.U2, e, un, vn);
    if (uab_next) {
        st2(uab_next, e, ab3(un, U0, h1.x), ab3(vn, V0, h1.y));
        s.aux_el[e] = sum4(w, s.adn[q.x], s.adn[q.y], s.adn[q.z], s.adn[q.w]);
    }
    const double dmf = dmaxr(Dmin, w.x * (r0.w + d0) + w.y * (r1.w + d1) + w.z * (r2.w + d2) + w.w * (r3.w + d3));
    const double uf0 = first ? 0.0 : s.U_filt[e];
    const double vf0 = first ? 0.0 : s.V_filt[e];
    s.U_filt[e] = uf0 + un * dmf;
    s.V_filt[e] = vf0 + vn * dmf;
}

// ---------------------------------------------------------------------------
// Twice-applied neighbour-difference filters (filt_bi_2D, filt_bi_3D):
// viscosity_filt2x_2D / _3D, src/fv_viscosity.F90:570-767.  The first pass
// forms U_c = K_bi * sum over the neighbours of (own - neighbour), K_bi =
// tau_inv*sqrt(scale_area/elem_area); the second pass sums (neighbour's U_c -
// own U_c) into the tendency UV_rhs / UV2_rhs that the velocity updates add.
// Neighbours are taken as the routines take them: slots 1-3 when present,
// slot 4 when present and not a repeat of slot 1 (a triangle's slot 4); each
// pass forms its first term apart and then accumulates, as they do.
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool f2x_take(const int (&nb)[4], int j) {
    return nb[j] >= 0 && (j < 3 || nb[3] != nb[0]);
}

// first pass of the 3-D filter on U_n*Je (fv_viscosity.F90:593-630), one
// thread per element layer
template <typename R>
__global__ __launch_bounds__(256) void k_filt2x_3d_c(DevMesh m, DevStateT<R> s, double tau_inv) {
    int e, k;
    if (!ent_layer(m.E, m.L, e, k)) return;
    const int K = m.K;
    const int4 nb4 = m.elnb[e];
    const int nb[4] = {nb4.x, nb4.y, nb4.z, nb4.w};
    const size_t i = (size_t)e * K + k;
    const double po = (double)s.U_n[i] * (double)s.Je[i], qo = (double)s.V_n[i] * (double)s.Je[i];
    double cu = 0.0, cv = 0.0;
    for (int j = 0; j < 4; ++j) {
        if (!f2x_take(nb, j)) continue;
        const size_t q = (size_t)nb[j] * K + k;
        const double pn = (double)s.U_n[q] * (double)s.Je[q], qn = (double)s.V_n[q] * (double)s.Je[q];
        if (j == 0) {
            cu = po - pn;
            cv = qo - qn;
        } else {
            cu = cu + po - pn;
            cv = cv + qo - qn;
        }
    }
    const double K_bi = tau_inv * m.kbi[e];
    s.fbu[i] = (R)(cu * K_bi);
    s.fbv[i] = (R)(cv * K_bi);
}

// second pass of the 3-D filter into UV_rhs (fv_viscosity.F90:638-660)
template <typename R>
__global__ __launch_bounds__(256) void k_filt2x_3d_uv(DevMesh m, DevStateT<R> s) {
    int e, k;
    if (!ent_layer(m.E, m.L, e, k)) return;
    const int K = m.K;
    const int4 nb4 = m.elnb[e];
    const int nb[4] = {nb4.x, nb4.y, nb4.z, nb4.w};
    const size_t i = (size_t)e * K + k;
    const double co = (double)s.fbu[i], do_ = (double)s.fbv[i];
    double u = 0.0, v = 0.0;
    for (int j = 0; j < 4; ++j) {
        if (!f2x_take(nb, j)) continue;
        const size_t q = (size_t)nb[j] * K + k;
        if (j == 0) {
            u = -co + (double)s.fbu[q];
            v = -do_ + (double)s.fbv[q];
        } else {
            u = u - co + (double)s.fbu[q];
            v = v - do_ + (double)s.fbv[q];
        }
    }
    s.uvu[i] = (R)u;
    s.uvv[i] = (R)v;
}

// first pass of the 2-D filter on U_n_2D (fv_viscosity.F90:700-727); the
// second pass runs inside k_bt_elem2, which adds UV2_rhs in the update
template <typename R>
__global__ __launch_bounds__(256) void k_bt_filt2x_c(DevMesh m, DevStateT<R> s, double tau_inv) {
    const int e = TID();
    if (e >= m.E) return;
    const int4 nb4 = m.elnb[e];
    const int nb[4] = {nb4.x, nb4.y, nb4.z, nb4.w};
    const double2 uo = ld2(s.U2, e);
    double cu = 0.0, cv = 0.0;
    for (int j = 0; j < 4; ++j) {
        if (!f2x_take(nb, j)) continue;
        const double2 un = ld2(s.U2, nb[j]);
        if (j == 0) {
            cu = uo.x - un.x;
            cv = uo.y - un.y;
        } else {
            cu = cu + uo.x - un.x;
            cv = cv + uo.y - un.y;
        }
    }
    const double K_bi = tau_inv * m.kbi[e];
    st2(s.fb2c, e, cu * K_bi, cv * K_bi);
}

// ---------------------------------------------------------------------------
// Biharmonic viscosity (bih_2D, bih_3D): biharmonic_2D, biharmonic_viscosity
// and biharmonic_viscosity_3D_to_2D, src/fv_viscosity.F90:104-568, 964-1220:
// a corrected edge Laplacian applied twice, with the gradients of the first
// result in between.  Edge kernels form each edge's flux once; element kernels
// gather them in ascending reference edge id (the internal edges, el_ied,
// then the boundary edges, el_bed), which is the order of the serial loops'
// scatters, so FP64 sums keep their bits.  free_slip = .true.
// ---------------------------------------------------------------------------


// one pass over (edge, layer) of the 3-D operator: the fluxes of (F, G) with
// the gradients (gux, guy, gvx, gvy) into bhfu/bhfv, times Jd (with_jd) and
// the edge's sponge weight (with_acc); boundary edges take the element's own
// gradient (fv_viscosity.F90:148-184, 208-218, 262-333, 998-1100, 1130-1196)
template <typename R>
__global__ __launch_bounds__(256) void k_bh3_edge(DevMesh m, DevStateT<R> s, const R *__restrict__ F,
                                                   const R *__restrict__ G, const R *__restrict__ gux,
                                                   const R *__restrict__ guy, const R *__restrict__ gvx,
                                                   const R *__restrict__ gvy, int with_jd, int with_acc) {
    int ed, k;
    if (!ent_layer(m.D, m.L, ed, k)) return;
    const int K = m.K;
    const int2 et = m.edtri[ed];
    const BhEdge g = m.bhe[ed];
    const size_t i1 = (size_t)et.x * K + k;
    double fu, fv;
    if (et.y >= 0) {
        const size_t i2 = (size_t)et.y * K + k;
        fu = bh_flux(g, gux[i1], gux[i2], guy[i1], guy[i2], F[i1], F[i2]);
        fv = bh_flux(g, gvx[i1], gvx[i2], gvy[i1], gvy[i2], G[i1], G[i2]);
    } else {
        fu = (double)gux[i1] * g.ye - (double)guy[i1] * g.xb;
        fv = (double)gvx[i1] * g.ye - (double)gvy[i1] * g.xb;
    }
    if (with_jd) {
        const int2 nd = m.ednodes[ed];
        const double ds = m.dsig[k];
        const double jd = 0.5 * (s.a[nd.x] * ds + s.a[nd.y] * ds);
        fu = fu * jd;
        fv = fv * jd;
    }
    if (with_acc) {
        fu = fu * g.acc;
        fv = fv * g.acc;
    }
    s.bhfu[(size_t)ed * K + k] = (R)fu;
    s.bhfv[(size_t)ed * K + k] = (R)fv;
}


// element sums of a pass into (oU, oV): mode 1 times Ah (biharmonic_viscosity,
// fv_viscosity.F90:224-235), mode 2 times (Ah_acc - Visc) with the Smagorinsky
// Visc = 0 (1084-1095; only the sediment module sets it), mode 0 as summed
template <typename R>
__global__ __launch_bounds__(256) void k_bh3_elem(DevMesh m, DevStateT<R> s, R *__restrict__ oU, R *__restrict__ oV,
                                                   int mode) {
    int e, k;
    if (!ent_layer(m.E, m.L, e, k)) return;
    double u = 0.0, v = 0.0;
    bh_gather(m, s, e, k, u, v);
    if (mode == 1) {
        u = u * m.bh_ah[e];
        v = v * m.bh_ah[e];
    } else if (mode == 2) {
        u = u * (m.bh_ah_acc[e] - 0.0);
        v = v * (m.bh_ah_acc[e] - 0.0);
    }
    const size_t i = (size_t)e * m.K + k;
    oU[i] = (R)u;
    oV[i] = (R)v;
}

// vel_lapl_gradients (fv_gradients.F90:327-402, mirror value -U at the
// boundary; puls = 0) or vel_gradients_puls (404-496, free-slip reflection of
// the pulsation velocity; puls = 1) of (Uc, Vc) into bhgux..bhgvy
template <typename R>
__global__ __launch_bounds__(256) void k_bh3_grad(DevMesh m, DevStateT<R> s, const R *__restrict__ Uc,
                                                   const R *__restrict__ Vc, int puls) {
    int e, k;
    if (!ent_layer(m.E, m.L, e, k)) return;
    const int K = m.K;
    const int4 q = m.en[e];
    const int jend = (q.x == q.w) ? 3 : 4;
    const int4 nb4 = m.elnb[e], ed4 = m.eledges[e];
    const int nbs[4] = {nb4.x, nb4.y, nb4.z, nb4.w}, eds[4] = {ed4.x, ed4.y, ed4.z, ed4.w};
    const double4 gx = m.gvx[e], gy = m.gvy[e];
    const double gxs[4] = {gx.x, gx.y, gx.z, gx.w}, gys[4] = {gy.x, gy.y, gy.z, gy.w};
    const size_t i = (size_t)e * K + k;
    const double U = (double)Uc[i], V = (double)Vc[i];
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    for (int j = 0; j < jend; ++j) {
        double u, v;
        if (nbs[j] >= 0) {
            u = (double)Uc[(size_t)nbs[j] * K + k] - U;
            v = (double)Vc[(size_t)nbs[j] * K + k] - V;
        } else if (puls) {
            const double2 dd = m.edxdy[eds[j]];
            const double zc0 = dd.x * m.elem_cos[e], zc1 = dd.y;
            const double un = -2.0 * (U * zc1 - V * zc0) / (zc0 * zc0 + zc1 * zc1);
            u = (U + un * zc1) - U;
            v = (V - un * zc0) - V;
        } else {
            u = -2.0 * U;
            v = -2.0 * V;
        }
        a0 = a0 + gxs[j] * u;
        a1 = a1 + gys[j] * u;
        a2 = a2 + gxs[j] * v;
        a3 = a3 + gys[j] * v;
    }
    s.bhgux[i] = (R)a0;
    s.bhguy[i] = (R)a1;
    s.bhgvx[i] = (R)a2;
    s.bhgvy[i] = (R)a3;
}

// the pulsation velocity U_puls = U_n - a*U_n_2D of compute_puls_vel
// (fv_3D_vel.F90:285-299) into bhpu/bhpv, formed as k_bt_p1elem forms it
template <typename R>
__global__ __launch_bounds__(256) void k_bh_puls(DevMesh m, DevStateT<R> s) {
    int e, k;
    if (!ent_layer(m.E, m.L, e, k)) return;
    const size_t i = (size_t)e * m.K + k;
    const double ao = s.puls_ap[e];
    s.bhpu[i] = (R)((double)s.U_n[i] - ao * s.U2[2 * e]);
    s.bhpv[i] = (R)((double)s.V_n[i] - ao * s.U2[2 * e + 1]);
}

// one pass of the 2-D operator (fv_viscosity.F90:381-432, 456-541): fluxes of
// Uin (2E) with the gradients grad (4E) into bh_fl2; boundary edges with the
// free-slip projection, whose v line takes the already projected u
template <typename R>
__global__ __launch_bounds__(256) void k_bh2_edge(DevMesh m, DevStateT<R> s, const R *__restrict__ Uin,
                                                   const R *__restrict__ grad) {
    const int ed = TID();
    if (ed >= m.D) return;
    const int2 et = m.edtri[ed];
    const BhEdge g = m.bhe[ed];
    const double4 a = ld4(grad, et.x);
    double u1, v1;
    if (et.y >= 0) {
        const double4 b = ld4(grad, et.y);
        const double2 w1 = ld2(Uin, et.x), w2 = ld2(Uin, et.y);
        u1 = bh_flux(g, a.x, b.x, a.y, b.y, w1.x, w2.x);
        v1 = bh_flux(g, a.z, b.z, a.w, b.w, w1.y, w2.y);
    } else {
        u1 = a.x * g.ye - a.y * g.xbd;
        v1 = a.z * g.ye - a.w * g.xbd;
        u1 = g.ye * (u1 * g.ye - v1 * g.xe) / (g.xe * g.xe + g.ye * g.ye);
        v1 = -g.xe * (u1 * g.ye - v1 * g.xe) / (g.xe * g.xe + g.ye * g.ye);
    }
    ((double2 *)s.bh_fl2)[ed] = make_double2(u1, v1);
}


// first 2-D pass summed per element and scaled: Ul = Ul*Ah*acc
// (fv_viscosity.F90:438-451)
template <typename R>
__global__ __launch_bounds__(256) void k_bh2_elem(DevMesh m, DevStateT<R> s) {
    const int e = TID();
    if (e >= m.E) return;
    double u = 0.0, v = 0.0;
    bh2_gather(m, s, e, u, v);
    const double ah = m.bh_ah[e], acc = m.acc_el[e];
    st2(s.bh_ul, e, u * (ah) * acc, v * (ah) * acc);
}

// ---------------------------------------------------------------------------
// jacobian: src/fv_main.F90:1228-1252 (node scalar + stored Je)
// ---------------------------------------------------------------------------
template <typename R>
__global__ __launch_bounds__(256) void k_jac_node(DevMesh m, DevStateT<R> s, double Dmin) {
    int n = TID();
    if (n >= m.N) return;
    s.a_old[n] = s.a[n];
    s.a[n] = fmax(m.depth[n] + s.eta_n[n], Dmin);
}

// the set-up jacobian after a restart has replaced eta_n (src/fv_main.F90:362
// after read_restart_separate at 336): Jc from the restored elevation; Jc_old
// keeps the value of the array set-up jacobian (src/fv_main.F90:162)
template <typename R>
__global__ __launch_bounds__(256) void k_jac_restart(DevMesh m, DevStateT<R> s, double Dmin) {
    int n = TID();
    if (n >= m.N) return;
    s.a[n] = fmax(m.depth[n] + s.eta_n[n], Dmin);
}

template <typename R>
__global__ __launch_bounds__(256) void k_je(DevMesh m, DevStateT<R> s) {
    extern __shared__ __align__(16) unsigned char lds_[];
    const int KC = m.K;
    const ColFlat f = col_flat(m.E, m.L);
    double *cs = (double *)lds_;  // per column: w_cv(4), a at the nodes(4)
    if (f.nc <= 0) return;
    if ((int)threadIdx.x < f.nc) {
        const int e = f.c0 + threadIdx.x;
        const int4 q = m.en[e];
        const double4 w = m.wcv[e];
        double *c = cs + 8 * threadIdx.x;
        c[0] = w.x;
        c[1] = w.y;
        c[2] = w.z;
        c[3] = w.w;
        c[4] = s.a[q.x];
        c[5] = s.a[q.y];
        c[6] = s.a[q.z];
        c[7] = s.a[q.w];
    }
    __syncthreads();
    if (!f.on()) return;
    const int e = f.c0 + f.c, k = f.k;
    const double *c = cs + 8 * f.c;
    const double ds = m.dsig[k];
    PL(s.Je, k, e, KC) = sum4a(c, c[4] * ds, c[5] * ds, c[6] * ds, c[7] * ds);
}

// update_3D_vel: src/fv_3D_vel.F90:201-255, with U_filt_2D *= 1/Mt from
// oce_barotropic_timestep (fv_main.F90:1670-1677); one thread per (element,
// layer) forms U_rhs*Je_old, the vertical integral runs one thread per
// column over those products in layer order.
template <typename R>
__global__ __launch_bounds__(256) void k_update3d(DevMesh m, DevStateT<R> s, double Dmin, double Mt_inv) {
    extern __shared__ __align__(16) unsigned char lds_[];
    const int L = m.L, KC = m.K;
    const ColFlat f = col_flat(m.E, L);
    double *pu = (double *)lds_, *pv = pu + f.cpb * L;
    double *sc = pv + f.cpb * L;  // per column: U2p, V2p, uf, vf, U2, V2, w_cv(4), a_old at the nodes(4)
    if (f.nc <= 0) return;
    if ((int)threadIdx.x < f.nc) {
        const int e = f.c0 + threadIdx.x;
        const int4 q = m.en[e];
        const double4 w = m.wcv[e];
        double *c = sc + 14 * threadIdx.x;
        c[4] = s.U2[2 * e];
        c[5] = s.U2[2 * e + 1];
        c[6] = w.x;
        c[7] = w.y;
        c[8] = w.z;
        c[9] = w.w;
        c[10] = s.a_old[q.x];
        c[11] = s.a_old[q.y];
        c[12] = s.a_old[q.z];
        c[13] = s.a_old[q.w];
    }
    __syncthreads();
    const int e = f.c0 + f.c, k = f.k;
    double ak = 0.0, ur = 0.0, vr = 0.0, je = 1.0;
    if (f.on()) {
        const double *c = sc + 14 * f.c;
        const double ds = m.dsig[k];
        ak = sum4a(c + 6, c[10] * ds, c[11] * ds, c[12] * ds, c[13] * ds);
        ur = PL(s.U_rhs, k, e, KC);
        vr = PL(s.V_rhs, k, e, KC);
        je = PL(s.Je, k, e, KC);
        pu[f.c * L + k] = ur * ak;
        pv[f.c * L + k] = vr * ak;
    }
    __syncthreads();
    if ((int)threadIdx.x < f.nc) {
        const int ci = threadIdx.x, ec = f.c0 + ci;
        const int4 q = m.en[ec];
        const double4 w = m.wcv[ec];
        const double dmi = 1.0 / dmaxr(Dmin, sum4(w, m.depth[q.x] + s.eta_n[q.x], m.depth[q.y] + s.eta_n[q.y],
                                                  m.depth[q.z] + s.eta_n[q.z], m.depth[q.w] + s.eta_n[q.w]));
        double su = 0.0, sv = 0.0;
        for (int j = 0; j < L; ++j) {
            su = su + pu[ci * L + j];
            sv = sv + pv[ci * L + j];
        }
        const double uf = (s.U_filt[ec] * Mt_inv) * dmi;
        const double vf = (s.V_filt[ec] * Mt_inv) * dmi;
        s.U_filt[ec] = uf;
        s.V_filt[ec] = vf;
        double *c = sc + 14 * ci;
        c[0] = su * dmi;
        c[1] = sv * dmi;
        c[2] = uf;
        c[3] = vf;
    }
    __syncthreads();
    if (!f.on()) return;
    const double *c = sc + 14 * f.c;
    const double U2p = c[0], V2p = c[1], uf = c[2], vf = c[3];
    ak = ak / je;
    PL(s.U_nf, k, e, KC) = ak * ur + uf - U2p;
    PL(s.V_nf, k, e, KC) = ak * vr + vf - V2p;
    PL(s.U_n, k, e, KC) = ak * ur + c[4] - U2p;
    PL(s.V_n, k, e, KC) = ak * vr + c[5] - V2p;
}

// vert_vel_sigma (src/fv_3D_vel.F90:309-456), first part: per node and layer
// the divergence of the baroclinic transport (U_n_filt - U_filt_2D)*Je,
// gathered over the node's edges in ascending id (354-437), stored in W.
template <typename R>
__global__ __launch_bounds__(1024) void k_vertvel_div(DevMesh m, DevStateT<R> s) {
    extern __shared__ __align__(16) unsigned char lds_[];
    const int KC = m.K;
    const ColFlat f = col_flat(m.N, m.L);
    if (f.nc <= 0) return;
    const int2 jr = csr_block(m.n2e_ptr, f);
    DivRec<R> *rec = (DivRec<R> *)lds_;
    {
        const EdgeGeo<R> *eg = (const EdgeGeo<R> *)m.egeo;
        for (int t = threadIdx.x; t < jr.y - jr.x; t += blockDim.x) {
            const int cc = m.n2e[jr.x + t];
            const EdgeGeo<R> g = eg[cc >= 0 ? cc : ~cc];
            DivRec<R> r;
            r.e1K = g.e1K;
            r.e2K = g.e2K;
            r.pos = cc >= 0;
            r.cx = g.cx;
            r.cy = g.cy;
            r.cz = g.cz;
            r.cw = g.cw;
            r.vf1 = (R)s.V_filt[g.e1];
            r.uf1 = (R)s.U_filt[g.e1];
            r.vf2 = g.e2K >= 0 ? (R)s.V_filt[g.e2] : R(0);
            r.uf2 = g.e2K >= 0 ? (R)s.U_filt[g.e2] : R(0);
            rec[t] = r;
        }
    }
    __syncthreads();
    if (!f.on()) return;
    const int n = f.c0 + f.c, k = f.k;
    const int p0 = m.n2e_ptr[n] - jr.x, p1 = m.n2e_ptr[n + 1] - jr.x;
    R c = R(0);  // storage type: FP32 transport fluxes in the mixed-precision path
    for (int j = p0; j < p1; ++j) {
        const DivRec<R> &g = rec[j];
        const int i1 = g.e1K + k;
        R fl = s.Je[i1] * ((s.V_nf[i1] - g.vf1) * g.cx - (s.U_nf[i1] - g.uf1) * g.cy);
        if (g.e2K >= 0) {
            const int i2 = g.e2K + k;
            fl = fl + s.Je[i2] * (-(s.V_nf[i2] - g.vf2) * g.cz + (s.U_nf[i2] - g.uf2) * g.cw);
        }
        c = g.pos ? c + fl : c - fl;
    }
    PL(s.W, k, n, KC) = c;
}

// second part: bottom-up cumulative sum and /area (444-450), one thread per
// column over the block's staged layers; Wvel(nsigma) stays zero.
template <typename R>
__global__ __launch_bounds__(256) void k_vertvel_scan(DevMesh m, DevStateT<R> s) {
    extern __shared__ __align__(16) unsigned char lds_[];
    const int L = m.L, KC = m.K;
    const ColFlat f = col_flat(m.N, L);
    R *tw = (R *)lds_;
    if (f.nc <= 0) return;
    const int n = f.c0 + f.c, k = f.k;
    if (f.on()) tw[f.c * L + k] = PL(s.W, k, n, KC);
    __syncthreads();
    if ((int)threadIdx.x < f.nc) {
        R *W = tw + threadIdx.x * L;
        const double area_inv = 1.0 / m.area[f.c0 + threadIdx.x];
        double run = 0.0;
        for (int j = L - 1; j >= 0; --j) {
            run = W[j] + run;
            W[j] = run * area_inv;
        }
    }
    __syncthreads();
    if (f.on()) PL(s.W, k, n, KC) = tw[f.c * L + k];
}

// ---------------------------------------------------------------------------
// tracers: solve_tracers (fv_main.F90:1707-1836)
// ---------------------------------------------------------------------------
// global min/max of TF and SF (fv_main.F90:1737-1777), per ensemble member:
// blockIdx.y = member, blockIdx.x = chunk of the member's (level, node)
// values; exact and order-free.  part[(m*gridDim.x + chunk)*4 + c]
template <typename R>
__global__ __launch_bounds__(256) void k_minmax_partial(DevMesh m, const R *__restrict__ T, const R *__restrict__ S,
                                                        double *part) {
    __shared__ double sm[4][256];
    const int mb = blockIdx.y;
    const int n0 = m.memb_noff[mb], cnt = m.memb_noff[mb + 1] - n0;
    const size_t tot = (size_t)cnt * m.L;
    double tmn = INFINITY, tmx = -INFINITY, smn = INFINITY, smx = -INFINITY;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < tot; i += (size_t)gridDim.x * blockDim.x) {
        const size_t n = n0 + i / m.L, k = i % m.L;
        const double t = T[n * m.K + k], v = S[n * m.K + k];
        tmn = fmin(tmn, t);
        tmx = fmax(tmx, t);
        smn = fmin(smn, v);
        smx = fmax(smx, v);
    }
    sm[0][threadIdx.x] = tmn;
    sm[1][threadIdx.x] = tmx;
    sm[2][threadIdx.x] = smn;
    sm[3][threadIdx.x] = smx;
    __syncthreads();
    for (int st = blockDim.x / 2; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) {
            sm[0][threadIdx.x] = fmin(sm[0][threadIdx.x], sm[0][threadIdx.x + st]);
            sm[1][threadIdx.x] = fmax(sm[1][threadIdx.x], sm[1][threadIdx.x + st]);
            sm[2][threadIdx.x] = fmin(sm[2][threadIdx.x], sm[2][threadIdx.x + st]);
            sm[3][threadIdx.x] = fmax(sm[3][threadIdx.x], sm[3][threadIdx.x + st]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int c = 0; c < 4; ++c) part[((size_t)mb * gridDim.x + blockIdx.x) * 4 + c] = sm[c][0];
}

// T_aver, S_aver of each member: midpoints of its extrema (one block per
// member, order-free min/max over its chunks)
#if FCG_STORAGE_DOUBLE
__global__ __launch_bounds__(256) void k_minmax_final(const double *part, int nb, int nmemb, double *tsa) {
    __shared__ double sm[4][256];
    const int mb = blockIdx.x;
    const double *p = part + (size_t)mb * nb * 4;
    double tmn = INFINITY, tmx = -INFINITY, smn = INFINITY, smx = -INFINITY;
    for (int b = threadIdx.x; b < nb; b += blockDim.x) {
        tmn = fmin(tmn, p[4 * b]);
        tmx = fmax(tmx, p[4 * b + 1]);
        smn = fmin(smn, p[4 * b + 2]);
        smx = fmax(smx, p[4 * b + 3]);
    }
    sm[0][threadIdx.x] = tmn;
    sm[1][threadIdx.x] = tmx;
    sm[2][threadIdx.x] = smn;
    sm[3][threadIdx.x] = smx;
    __syncthreads();
    for (int st = blockDim.x / 2; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) {
            sm[0][threadIdx.x] = fmin(sm[0][threadIdx.x], sm[0][threadIdx.x + st]);
            sm[1][threadIdx.x] = fmax(sm[1][threadIdx.x], sm[1][threadIdx.x + st]);
            sm[2][threadIdx.x] = fmin(sm[2][threadIdx.x], sm[2][threadIdx.x + st]);
            sm[3][threadIdx.x] = fmax(sm[3][threadIdx.x], sm[3][threadIdx.x + st]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        tsa[2 * mb] = 0.5 * (sm[0][0] + sm[1][0]);
        tsa[2 * mb + 1] = 0.5 * (sm[2][0] + sm[3][0]);
    }
}
#endif

// AB2 interpolation to n+1/2, de-meaned (fv_tracer.F90:432-444), over every
// slot of the tracer columns (the unused last slot stays finite)
template <typename R>
__global__ __launch_bounds__(256) void k_tr_ab(DevMesh m, DevStateT<R> s, size_t n) {
    size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int mb = m.memb[i / (size_t)m.K];
    const double Ta = s.tsa[2 * mb], Sa = s.tsa[2 * mb + 1];
    s.abT[i] = -(0.5 + kEps) * s.T_old[i] + (1.5 + kEps) * s.TF[i] - Ta;
    s.abS[i] = -(0.5 + kEps) * s.S_old[i] + (1.5 + kEps) * s.SF[i] - Sa;
}

// river water of salinity 0.2 diluting the salinity of its mouth's column
// (fv_main.F90:1805-1822), after the advection and before the implicit
// vertical diffusion: SF += (0.2 - S_old) * Q dsig dt / (area Jc), Jc =
// a dsig with a = max(depth+eta, Dmin) of this step; one thread per node,
// nodes without discharge return at once
template <typename R>
__global__ __launch_bounds__(256) void k_tr_rivers(DevMesh m, DevStateT<R> s, double dt) {
    const int n = TID();
    if (n >= m.N) return;
    const double q = s.runoff[n];
    if (q == 0.0) return;
    const double a = s.a[n], ar = m.area[n];
    for (int k = 0; k < m.L; ++k) {
        const size_t i = (size_t)n * m.K + k;
        const double coeff = q * m.dsig[k] * dt / (ar * (a * m.dsig[k]));
        const double delta = (double)0.2f - (double)s.S_old[i];
        s.SF[i] = (R)((double)s.SF[i] + delta * coeff);
    }
}

// QUICK vertical flux through 1-based interface nz of one node column
// (fv_tracer.F90:740-815), for tracer column X (de-meaned AB values)
template <typename R>
__device__ __forceinline__ double upw_vert(const DevMesh &m, const DevStateT<R> &s, const R *X, int n, int nz,
                                           double a) {
    const int KC = m.K;
    const int N = m.N, K = m.K;
    auto Wv = [&](int z) { return PL(s.W, z - 1, n, KC); };
    auto T = [&](int z) { return PL(X, z - 1, n, KC); };
    auto zb = [&](int z) { return m.oms[z - 1] * a; };
    auto Zc = [&](int z) { return 0.5 * (m.oms[z - 1] * a + m.oms[z] * a); };
    if (nz == 1) return -Wv(1) * T(1);
    if (nz == K) return 0.0;
    if (nz == 2) {
        if (Wv(2) > 0.0) {
            double dg = (zb(2) - Zc(1)) * (Zc(2) - zb(2)) / ((Zc(3) - Zc(1)) * (Zc(2) - Zc(3)));
            double db = -(zb(2) - Zc(1)) * (Zc(3) - zb(2)) / ((Zc(2) - Zc(1)) * (Zc(2) - Zc(3)));
            return -(T(1) * (1.0 - dg - db) + T(2) * db + T(3) * dg) * Wv(2);
        }
        return -T(1) * Wv(2);
    }
    if (nz == K - 1) {
        if (Wv(nz) >= 0.0) return -T(nz) * Wv(nz);
        double dg = (Zc(nz) - zb(nz)) * (zb(nz) - Zc(nz - 1)) / ((Zc(nz) - Zc(nz - 2)) * (Zc(nz - 2) - Zc(nz - 1)));
        double db = -(Zc(nz) - zb(nz)) * (zb(nz) - Zc(nz - 2)) / ((Zc(nz) - Zc(nz - 1)) * (Zc(nz - 2) - Zc(nz - 1)));
        return -(T(nz) * (1.0 - dg - db) + T(nz - 1) * db + T(nz - 2) * dg) * Wv(nz);
    }
    if (Wv(nz) >= 0.0) {
        double dg = (zb(nz) - Zc(nz - 1)) * (Zc(nz) - zb(nz)) / ((Zc(nz + 1) - Zc(nz - 1)) * (Zc(nz) - Zc(nz + 1)));
        double db = -(zb(nz) - Zc(nz - 1)) * (Zc(nz + 1) - zb(nz)) / ((Zc(nz) - Zc(nz - 1)) * (Zc(nz) - Zc(nz + 1)));
        return -(T(nz - 1) * (1.0 - dg - db) + T(nz) * db + T(nz + 1) * dg) * Wv(nz);
    }
    double dg = (Zc(nz) - zb(nz)) * (zb(nz) - Zc(nz - 1)) / ((Zc(nz) - Zc(nz - 2)) * (Zc(nz - 2) - Zc(nz - 1)));
    double db = -(Zc(nz) - zb(nz)) * (zb(nz) - Zc(nz - 2)) / ((Zc(nz) - Zc(nz - 1)) * (Zc(nz - 2) - Zc(nz - 1)));
    return -(T(nz) * (1.0 - dg - db) + T(nz - 1) * db + T(nz - 2) * dg) * Wv(nz);
}

// tracer_impl_vert_diff (fv_tracer.F90:1319-1494, key_atm off) of one node
// column held in LDS (T, S, Kv columns), applied to T/S in place, followed by
// the STOP checks of fv_tracer.F90:1466-1485 (reported through the device
// error flag).  The sweep keeps cp in the Kv slots it has consumed (cp(nz) in
// slot nz-1: the sweep reads Kv(2) at nz = 1 and Kv(nz) at nz >= 2, and never
// Kv(1)) and tp/sp in the columns Tp/Sp, each rounded to the storage type as
// the reference-order scratch arrays are.
// sw_cp: src/fv_tracer.F90:1775-1910, the UNESCO 1983 heat capacity of sea
// water (Fofonoff and Millard 1983), P in bar; literals the reference writes
// without _WP are REAL(4)
__device__ double sw_cp(double S, double Ti, double P) {
    const double T = Ti < 0 ? 0.0 : Ti;
    double S0, sS;
    if (S < 0) {
        S0 = 0.0;
        sS = 0.0;
    } else {
        S0 = S;
        sS = S * sqrt(S);
    }
    const double T68 = T * 1.00024;
    double c0 = 4217.4, c1 = -3.720283, c2 = 0.1412855, c3 = -(double)2.654387e-3f, c4 = (double)2.093236e-5f;
    double a0 = -7.64357, a1 = 0.1072763, a2 = -(double)1.38385e-3f;
    double b0 = 0.1770383, b1 = -(double)4.07718e-3f, b2 = (double)5.148e-5f;
    const double Cpst0 = (((c4 * T68 + c3) * T68 + c2) * T68 + c1) * T68 + c0 + (a0 + a1 * T68 + a2 * (T68 * T68)) * S0
                       + (b0 + b1 * T68 + b2 * (T68 * T68)) * sS;
    a0 = -(double)4.9592e-1f;
    a1 = (double)1.45747e-2f;
    a2 = -(double)3.13885e-4f;
    const double a3 = (double)2.0357e-6f, a4 = (double)1.7168e-8f;
    b0 = (double)2.4931e-4f;
    b1 = -(double)1.08645e-5f;
    b2 = (double)2.87533e-7f;
    const double b3 = -(double)4.0027e-9f, b4 = (double)2.2956e-11f;
    c0 = -(double)5.422e-8f;
    c1 = (double)2.6380e-9f;
    c2 = -(double)6.5637e-11f;
    c3 = (double)6.136e-13f;
    const double del_Cp0t0 = (((((c3 * T68 + c2) * T68 + c1) * T68 + c0) * P
                             + ((((b4 * T68 + b3) * T68 + b2) * T68 + b1) * T68 + b0)) * P
                             + ((((a4 * T68 + a3) * T68 + a2) * T68 + a1) * T68 + a0)) * P;
    const double d0 = (double)4.9247e-3f, d1 = -(double)1.28315e-4f, d2 = (double)9.802e-7f, d3 = (double)2.5941e-8f,
                 d4 = -(double)2.9179e-10f;
    const double e0 = -(double)1.2331e-4f, e1 = -(double)1.517e-6f, e2 = (double)3.122e-8f;
    const double f0 = -(double)2.9558e-6f, f1 = (double)1.17054e-7f, f2 = -(double)2.3905e-9f, f3 = (double)1.8448e-11f;
    const double g0 = (double)9.971e-8f;
    const double h0 = (double)5.540e-10f, h1 = -(double)1.7682e-11f, h2 = (double)3.513e-13f;
    const double j1 = -(double)1.4300e-12f;
    const double S3_2 = sS;
    const double del_Cpstp =
        (((((d4 * T68 + d3) * T68 + d2) * T68 + d1) * T68 + d0) * S0 + ((e2 * T68 + e1) * T68 + e0) * S3_2) * P
        + ((((f3 * T68 + f2) * T68 + f1) * T68 + f0) * S0 + g0 * S3_2) * (P * P)
        + (((h2 * T68 + h1) * T68 + h0) * S0 + j1 * T68 * S3_2) * (P * P * P);
    return Cpst0 + del_Cp0t0 + del_Cpstp;
}

// calc_heat_flux_sw (src/fv_tracer.F90:1635-1700): heat source of layer nz
// (1-based) of a column of depth a.  As the reference does: To/So (T_old,
// S_old) already hold full values and the tracer means are added again, and
// sw_cp(S, T, P) receives (T, S, P).  Shortwave absorption: Paulson and
// Simpson with constant R, a, b (swr_paul_sim, 1590-1610) or Soloviev
// (swr_soloviev, 1565-1588: exp(10 z) + 0 + 0).
__device__ double heat_src(const DevMesh &m, const AtmPar &ap, int nz, double a, double To, double So, double rc,
                           double Ta, double Sa, double qsr, double qns, double mslp) {
    const int L = m.L;
    const double d1 = m.oms[nz - 1] * a, d2 = m.oms[nz] * a;
    const double T = To + Ta, S = So + Sa;
    const double aux = (d1 + d2) * 0.5;
    const double rho = rc + ap.rho0;
    const double P = mslp == 0.0 ? (rho * kG * aux + kPref) * 10.0e-5 : (rho * kG * aux + mslp) * 10.0e-5;
    const double c_p = sw_cp(T, S, P);
    double swr1, swr2;
    if (ap.type_swr == 1) {
        swr1 = ap.Rjc * exp(-d1 / ap.ajc) + (1.0 - ap.Rjc) * exp(-d1 / ap.bjc);
        swr2 = ap.Rjc * exp(-d2 / ap.ajc) + (1.0 - ap.Rjc) * exp(-d2 / ap.bjc);
    } else {
        swr1 = exp(10.0 * -d1);
        swr2 = exp(10.0 * -d2);
    }
    const double jc = a * m.dsig[nz - 1];
    double src = nz == 1 ? (qsr * (swr1 - swr2) + qns) / (jc * c_p * rho) : qsr * (swr1 - swr2) / (jc * c_p * rho);
    if (nz == L) src = src + (double)0.2f * qsr * (swr2) / (jc * c_p * rho);
    return src;
}

// tracer_impl_vert_diff of one column (fv_tracer.F90:1319-1494), the
// coefficients and right-hand sides formed on the fly, the Thomas sweep
// keeping its coefficients in the slots it has consumed.  With surface
// forcing on a wet node (1376-1413) the heat source of every layer and the
// surface freshwater source (calc_evap_precip, 1702-1752) join the
// right-hand sides; TF/SF are anomalies there, T_old/S_old full values.
template <typename R, bool ATM>
__device__ __forceinline__ void vdiff_column(const DevMesh &m, const DevStateT<R> &s, R *Tc, R *Sc, R *Kc, R *Tp,
                                             R *Sp, double a, double Ta, double Sa, double dt, const AtmPar &ap,
                                             int n) {
    const int L = m.L, K = m.K;
    auto zb = [&](int z) { return m.oms[z - 1] * a; };
    auto Zc = [&](int z) { return 0.5 * (m.oms[z - 1] * a + m.oms[z] * a); };
    auto KV = [&](int z) { return Kc[z - 1]; };
    auto Td = [&](int z) { return Tc[z - 1] - Ta; };
    auto Sd = [&](int z) { return Sc[z - 1] - Sa; };
    const int KC = K;
    const bool atm = ATM && (ap.wet_dry ? (m.depth[n] + s.etaAB[n] > ap.Dmin ? 1.0 : 0.0) : 1.0) != 0.0;
    double frw = 0.0, qsr = 0.0, qns = 0.0, mslp = 0.0;
    if (atm) {
        qsr = s.qsr[n];
        qns = s.qns[n];
        mslp = s.mslp[n];
        const double T = Td(1) + Ta, S = Sd(1) + Sa;
        const double rho = densityJM(T, 0.0, -(m.oms[0] * a), ap.rho0);
        const double jc1 = a * m.dsig[0];
        if ((rho + ap.rho0) * jc1 != 0) {
            frw = S * s.emp[n] / ((rho + ap.rho0) * jc1);
            if (S <= 0.2 && frw < 0) frw = 0.0;
        }
    }
    auto src = [&](int z) {
        return heat_src(m, ap, z, a, (double)PL(s.T_old, z - 1, n, KC), (double)PL(s.S_old, z - 1, n, KC),
                        s.rho_c[(size_t)n * L + z - 1], Ta, Sa, qsr, qns, mslp);
    };
    double cp = 0, tp = 0, sp = 0;
    for (int nz = 1; nz <= L; ++nz) {
        double aa, bb, cc;
        if (nz == 1) {
            cc = -dt * KV(2) / ((Zc(1) - Zc(2)) * (zb(1) - zb(2)));
            aa = 0.0;
            bb = -cc + 1.0;
        } else if (nz == L) {
            aa = -dt * KV(L) / ((Zc(L - 1) - Zc(L)) * (zb(L) - zb(K)));
            bb = -aa + 1.0;
            cc = 0.0;
        } else {
            aa = -dt * KV(nz) / ((Zc(nz - 1) - Zc(nz)) * (zb(nz) - zb(nz + 1)));
            cc = -dt * KV(nz) / ((Zc(nz) - Zc(nz + 1)) * (zb(nz) - zb(nz + 1)));
            bb = -aa - cc + 1.0;
        }
        double tr, sr;
        if (nz == 1) {
            tr = cc * (-Td(2) + Td(1));
            sr = cc * (-Sd(2) + Sd(1));
            if (atm) {
                tr = tr + dt * src(1);
                sr = sr + dt * frw;
            }
        } else if (nz == L) {
            tr = aa * (Td(L) - Td(L - 1));
            sr = aa * (Sd(L) - Sd(L - 1));
            if (atm) tr = tr + dt * src(L);
        } else {
            tr = -aa * Td(nz - 1) - cc * Td(nz + 1) + (aa + cc) * Td(nz);
            sr = -aa * Sd(nz - 1) - cc * Sd(nz + 1) + (aa + cc) * Sd(nz);
            if (atm) tr = tr + dt * src(nz);
        }
        if (nz == 1) {
            cp = cc / bb;
            tp = tr / bb;
            sp = sr / bb;
        } else {
            double mm = 1.0 / (bb - cp * aa);
            cp = cc * mm;
            tp = (tr - tp * aa) * mm;
            sp = (sr - sp * aa) * mm;
        }
        Kc[nz - 1] = (R)cp;
        Tp[nz - 1] = (R)tp;
        Sp[nz - 1] = (R)sp;
    }
    double trn = tp, srn = sp;
    int bad = 0, nonfinite = 0;
    for (int nz = L; nz >= 1; --nz) {
        if (nz < L) {
            double c = Kc[nz - 1];
            trn = Tp[nz - 1] - c * trn;
            srn = Sp[nz - 1] - c * srn;
        }
        double tnew = Td(nz) + trn + Ta;
        double snew = Sd(nz) + srn + Sa;
        if (fabs(snew) < 1.e-12) snew = fabs(snew);
        if (snew < 0.0) bad = bad ? bad : FCG_E_SALINITY;
        if (tnew < -3.0) bad = bad ? bad : FCG_E_TEMP;
        if (!isfinite(tnew) || !isfinite(snew)) nonfinite = 1;  // NaN passes both tests above
        Tc[nz - 1] = tnew;
        Sc[nz - 1] = snew;
    }
    if (bad) atomicCAS(s.err, 0, bad);
    if (nonfinite) atomicOr(s.err + 2, 1);
}


// Net normal barotropic transport into an open-boundary node through its
// boundary edges (fv_tracer.F90:862-873 and the same block of the other
// schemes), summed in ascending edge id; decides inflow relaxation vs.
// replacement by climatology.
template <typename R>
__device__ __forceinline__ double ob_inflow(const DevMesh &m, const DevStateT<R> &s, int p0, int p1, double Dmin) {
    double vel_nor = 0.0;
    for (int j = p0; j < p1; ++j) {
        const int cc = m.n2e[j];
        const int ed = cc >= 0 ? cc : ~cc;
        if (m.edtri[ed].y >= 0) continue;  // internal edge
        const int2 nd = m.ednodes[ed];
        const int e1 = m.edtri[ed].x;
        const double4 cd = m.ecross[ed];
        double dmn = dmaxr(Dmin, 0.5 * (s.eta_n[nd.x] + m.depth[nd.x] + s.eta_n[nd.y] + m.depth[nd.y]));
        double un1 = (s.V_filt[e1] * cd.x - s.U_filt[e1] * cd.y) * dmn;
        vel_nor = (cc >= 0) ? vel_nor + un1 / m.area[nd.x] : vel_nor - un1 / m.area[nd.y];
    }
    return vel_nor;
}

// Vertical flux through 1-based interface nz for the MUSCL and Miura schemes
// (fv_tracer.F90:236-272, 1172-1208): QUICK from the upwind side, first
// order next to the surface/bottom.  The column value is PL(X) - sub (sub =
// 0 for the AB-interpolated column, the tracer mean for the current one).
// Tm keeps its previous value when W is exactly zero, as in the reference;
// the product -Tm*W is then a zero either way.  The two schemes spell the
// QUICK weights with one or two divisions, hence kChained.
template <bool kChained, typename R>
__device__ __forceinline__ double hi_vert(const DevMesh &m, const DevStateT<R> &s, const R *X, double sub, int n,
                                          int nz, double a, double &Tm) {
    const int KC = m.K;
    const int N = m.N, K = m.K;
    auto Wv = [&](int z) { return PL(s.W, z - 1, n, KC); };
    auto T = [&](int z) { return PL(X, z - 1, n, KC) - sub; };
    auto zb = [&](int z) { return m.oms[z - 1] * a; };
    auto Zc = [&](int z) { return 0.5 * (m.oms[z - 1] * a + m.oms[z] * a); };
    if (nz == 1) return -Wv(1) * T(1);
    if (nz == K) return 0.0;
    const double w = Wv(nz);
    if (w > 0.0) {
        if (nz == K - 1) {
            Tm = T(nz);
        } else {
            const double aa = zb(nz) - Zc(nz - 1), bb = Zc(nz) - zb(nz), cc = Zc(nz + 1) - zb(nz);
            double dg, db;
            if (kChained) {
                dg = aa * bb / (cc + aa) / (bb - cc);
                db = -aa * cc / (bb + aa) / (bb - cc);
            } else {
                dg = aa * bb / ((cc + aa) * (bb - cc));
                db = -aa * cc / ((bb + aa) * (bb - cc));
            }
            const double da = 1.0 - dg - db;
            Tm = T(nz - 1) * da + T(nz) * db + T(nz + 1) * dg;
        }
    }
    if (w < 0.0) {
        if (nz == 2) {
            Tm = T(nz - 1);
        } else {
            const double aa = Zc(nz) - zb(nz), bb = zb(nz) - Zc(nz - 1), cc = zb(nz) - Zc(nz - 2);
            double dg, db;
            if (kChained) {
                dg = aa * bb / (cc + aa) / (bb - cc);
                db = -aa * cc / (bb + aa) / (bb - cc);
            } else {
                dg = aa * bb / ((cc + aa) * (bb - cc));
                db = -aa * cc / ((bb + aa) * (bb - cc));
            }
            const double da = 1.0 - dg - db;
            Tm = T(nz) * da + T(nz - 1) * db + T(nz - 2) * dg;
        }
    }
    return -Tm * w;
}

// hi_vert for two tracer columns X and Y at once: the QUICK weights depend
// only on the geometry and on W, so they are formed once for both (the same
// operations on the same operands as two hi_vert calls, hence the same bits).
template <bool kChained, typename R>
__device__ __forceinline__ void hi_vert2(const DevMesh &m, const DevStateT<R> &s, const R *X, const R *Y, int n,
                                         int nz, double a, double &TmX, double &TmY, double &fx, double &fy) {
    const int K = m.K, KC = K;
    auto Wv = [&](int z) { return PL(s.W, z - 1, n, KC); };
    auto Xv = [&](int z) { return PL(X, z - 1, n, KC) - 0.0; };
    auto Yv = [&](int z) { return PL(Y, z - 1, n, KC) - 0.0; };
    auto zb = [&](int z) { return m.oms[z - 1] * a; };
    auto Zc = [&](int z) { return 0.5 * (m.oms[z - 1] * a + m.oms[z] * a); };
    if (nz == 1) {
        const double w1 = Wv(1);
        fx = -w1 * Xv(1);
        fy = -w1 * Yv(1);
        return;
    }
    if (nz == K) {
        fx = 0.0;
        fy = 0.0;
        return;
    }
    const double w = Wv(nz);
    if (w > 0.0) {
        if (nz == K - 1) {
            TmX = Xv(nz);
            TmY = Yv(nz);
        } else {
            const double aa = zb(nz) - Zc(nz - 1), bb = Zc(nz) - zb(nz), cc = Zc(nz + 1) - zb(nz);
            double dg, db;
            if (kChained) {
                dg = aa * bb / (cc + aa) / (bb - cc);
                db = -aa * cc / (bb + aa) / (bb - cc);
            } else {
                dg = aa * bb / ((cc + aa) * (bb - cc));
                db = -aa * cc / ((bb + aa) * (bb - cc));
            }
            const double da = 1.0 - dg - db;
            TmX = Xv(nz - 1) * da + Xv(nz) * db + Xv(nz + 1) * dg;
            TmY = Yv(nz - 1) * da + Yv(nz) * db + Yv(nz + 1) * dg;
        }
    }
    if (w < 0.0) {
        if (nz == 2) {
            TmX = Xv(nz - 1);
            TmY = Yv(nz - 1);
        } else {
            const double aa = Zc(nz) - zb(nz), bb = zb(nz) - Zc(nz - 1), cc = zb(nz) - Zc(nz - 2);
            double dg, db;
            if (kChained) {
                dg = aa * bb / (cc + aa) / (bb - cc);
                db = -aa * cc / (bb + aa) / (bb - cc);
            } else {
                dg = aa * bb / ((cc + aa) * (bb - cc));
                db = -aa * cc / ((bb + aa) * (bb - cc));
            }
            const double da = 1.0 - dg - db;
            TmX = Xv(nz) * da + Xv(nz - 1) * db + Xv(nz - 2) * dg;
            TmY = Yv(nz) * da + Yv(nz - 1) * db + Yv(nz - 2) * dg;
        }
    }
    fx = -TmX * w;
    fy = -TmY * w;
}

// Element gradient of a node column at level k (tracer_gradient_elements,
// fv_gradients.F90:180-205), the column value being PL(X) - sub.
template <typename R>
__device__ __forceinline__ double2 elem_grad(const DevMesh &m, const R *X, double sub, int el, int k) {
    const int KC = m.K;
    const int N = m.N;
    const int4 q = m.enK[el];  // column offsets of the element's nodes
    const double v0 = X[q.x + k] - sub, v1 = X[q.y + k] - sub;
    const double v2 = X[q.z + k] - sub, v3 = X[q.w + k] - sub;
    return make_double2(sum4(m.gsx[el], v0, v1, v2, v3), sum4(m.gsy[el], v0, v1, v2, v3));
}

// solve_tracer_upwind (fv_tracer.F90:390-923), one thread per node and layer:
// edge fluxes gathered in ascending edge id, QUICK vertical fluxes at the
// layer's two interfaces, OB relaxation, AB update.  The implicit vertical
// diffusion follows in k_tr_vdiff.
template <typename R>
__global__ __launch_bounds__(256) void k_tr_upwind(DevMesh m, DevStateT<R> s, double dt, double Dmin, int cl_relax, double clim_relax,
                            double relax2clim_ac) {
    const int KC = m.K;
    int n, k;
    if (!ent_layer(m.N, m.L, n, k)) return;
    const int N = m.N, E = m.E;
    const int mb = m.memb[n];
    const double Ta = s.tsa[2 * mb], Sa = s.tsa[2 * mb + 1];
    const double a = s.a[n];
    const int p0 = m.n2e_ptr[n], p1 = m.n2e_ptr[n + 1];
    const bool ob = cl_relax && m.index[n] == 2;
    double tc = 0.0, tt = 0.0, ts = 0.0;
    for (int j = p0; j < p1; ++j) {
        const int cc = m.n2e[j];
        const int ed = cc >= 0 ? cc : ~cc;
        const int2 nd = m.ednodes[ed];
        const int2 et = m.edtri[ed];
        const double4 cd = m.ecross[ed];
        const int e1 = et.x;
        const double c1 = PL(s.V_nf, k, e1, KC) * cd.x - PL(s.U_nf, k, e1, KC) * cd.y;
        const double T1 = PL(s.abT, k, nd.x, KC), T2 = PL(s.abT, k, nd.y, KC);
        const double S1 = PL(s.abS, k, nd.x, KC), S2 = PL(s.abS, k, nd.y, KC);
        double auxc, taux_, saux_;
        if (et.y >= 0) {
            const int e2 = et.y;
            const double c2 = -PL(s.V_nf, k, e2, KC) * cd.z + PL(s.U_nf, k, e2, KC) * cd.w;
            const double J1 = PL(s.Je, k, e1, KC), J2 = PL(s.Je, k, e2, KC);
            auxc = J1 * c1 + J2 * c2;
            if (c1 > 0.0 && c2 > 0.0) {
                taux_ = J1 * c1 * T2 + J2 * c2 * T2;
                saux_ = J1 * c1 * S2 + J2 * c2 * S2;
            } else if (c1 > 0.0) {
                taux_ = J1 * c1 * T2 + J2 * c2 * T1;
                saux_ = J1 * c1 * S2 + J2 * c2 * S1;
            } else if (c2 > 0.0) {
                taux_ = J1 * c1 * T1 + J2 * c2 * T2;
                saux_ = J1 * c1 * S1 + J2 * c2 * S2;
            } else {
                taux_ = J1 * c1 * T1 + J2 * c2 * T1;
                saux_ = J1 * c1 * S1 + J2 * c2 * S1;
            }
        } else {
            auxc = PL(s.Je, k, e1, KC) * c1;
            if (c1 > 0.0) {
                taux_ = auxc * T2;
                saux_ = auxc * S2;
            } else {
                taux_ = auxc * T1;
                saux_ = auxc * S1;
            }
        }
        if (cc >= 0) {
            tc = tc + auxc;
            tt = tt + taux_;
            ts = ts + saux_;
        } else {
            tc = tc - auxc;
            tt = tt - taux_;
            ts = ts - saux_;
        }
    }
    const double tvt = upw_vert(m, s, s.abT, n, k + 1, a), tvb = upw_vert(m, s, s.abT, n, k + 2, a);
    const double svt = upw_vert(m, s, s.abS, n, k + 1, a), svb = upw_vert(m, s, s.abS, n, k + 2, a);
    const double tab = PL(s.abT, k, n, KC), sab = PL(s.abS, k, n, KC);
    const double wdiff = PL(s.W, k, n, KC) - PL(s.W, k + 1, n, KC);
    const double jc = a * m.dsig[k];
    const double area_inv = 1.0 / m.area[n];
    double rt = ((tt - tab * tc) * area_inv + tab * wdiff + (tvt - tvb)) / jc;
    double rs = ((ts - sab * tc) * area_inv + sab * wdiff + (svt - svb)) / jc;
    if (ob) {  // fv_tracer.F90:832-895, on the AB-interpolated tracer
        if (ob_inflow(m, s, p0, p1, Dmin) > 0.0) {
            rt = rt + relax2clim_ac * (PL(s.Tclim, k, n, KC) - (tab + Ta));
            rs = rs + relax2clim_ac * (PL(s.Sclim, k, n, KC) - (sab + Sa));
        } else {
            rt = clim_relax * (PL(s.Tclim, k, n, KC) - (tab + Ta));
            rs = clim_relax * (PL(s.Sclim, k, n, KC) - (sab + Sa));
        }
    }
    const double tdm = PL(s.TF, k, n, KC) - Ta, sdm = PL(s.SF, k, n, KC) - Sa;
    PL(s.T_old, k, n, KC) = tdm + Ta;
    PL(s.S_old, k, n, KC) = sdm + Sa;
    PL(s.TF, k, n, KC) = tdm + rt * dt + Ta;
    PL(s.SF, k, n, KC) = sdm + rs * dt + Sa;
}

// Node-averaged gradients of the AB-interpolated T and S used by the MUSCL
// reconstruction (fill_up_dn_grad node-average branch,
// fv_gradients.F90:263-323): area-weighted mean over the node's elements in
// ascending element id, divided by the summed area.
template <typename R>
__global__ __launch_bounds__(1024) void k_tr_gradn_muscl(DevMesh m, DevStateT<R> s) {
    extern __shared__ __align__(16) unsigned char lds_[];
    const int KC = m.K;
    const ColFlat f = col_flat(m.N, m.L);
    if (f.nc <= 0) return;
    const int2 jr = csr_block(m.n2el_ptr, f);
    GradRec *rec = (GradRec *)lds_;
    for (int t = threadIdx.x; t < jr.y - jr.x; t += blockDim.x) {
        const int el = m.n2el[jr.x + t];
        const int4 q = m.enK[el];
        const double4 gx = m.gsx[el], gy = m.gsy[el];
        GradRec r;
        r.qK[0] = q.x;
        r.qK[1] = q.y;
        r.qK[2] = q.z;
        r.qK[3] = q.w;
        r.ea = m.elem_area[el];
        r.gx[0] = gx.x;
        r.gx[1] = gx.y;
        r.gx[2] = gx.z;
        r.gx[3] = gx.w;
        r.gy[0] = gy.x;
        r.gy[1] = gy.y;
        r.gy[2] = gy.z;
        r.gy[3] = gy.w;
        rec[t] = r;
    }
    __syncthreads();
    if (!f.on()) return;
    const int n = f.c0 + f.c, k = f.k;
    const int p0 = m.n2el_ptr[n] - jr.x, p1 = m.n2el_ptr[n + 1] - jr.x;
    double tvol = 0.0, tx = 0.0, ty = 0.0, sx = 0.0, sy = 0.0;
    for (int j = p0; j < p1; ++j) {
        const GradRec &r = rec[j];
        const double ea = r.ea;
        const double t0 = s.abT[r.qK[0] + k] - 0.0, t1 = s.abT[r.qK[1] + k] - 0.0;
        const double t2 = s.abT[r.qK[2] + k] - 0.0, t3 = s.abT[r.qK[3] + k] - 0.0;
        const double s0 = s.abS[r.qK[0] + k] - 0.0, s1 = s.abS[r.qK[1] + k] - 0.0;
        const double s2 = s.abS[r.qK[2] + k] - 0.0, s3 = s.abS[r.qK[3] + k] - 0.0;
        tvol = tvol + ea;
        tx = tx + sum4a(r.gx, t0, t1, t2, t3) * ea;
        ty = ty + sum4a(r.gy, t0, t1, t2, t3) * ea;
        sx = sx + sum4a(r.gx, s0, s1, s2, s3) * ea;
        sy = sy + sum4a(r.gy, s0, s1, s2, s3) * ea;
    }
    PL(s.gTx, k, n, KC) = tx / tvol;
    PL(s.gTy, k, n, KC) = ty / tvol;
    PL(s.gSx, k, n, KC) = sx / tvol;
    PL(s.gSy, k, n, KC) = sy / tvol;
}

// Node-averaged gradient of the current de-meaned T for the Miura scheme
// (fv_tracer.F90:984-998): the same mean, with 1/(summed area) formed first.
// Also stages the de-meaned value itself in tcur: the Miura update
// overwrites TF while neighbouring nodes still read it.
template <typename R>
__global__ __launch_bounds__(256) void k_tr_gradn_miura(DevMesh m, DevStateT<R> s) {
    const int KC = m.K;
    int n, k;
    if (!ent_layer(m.N, m.L, n, k)) return;
    const int N = m.N;
    const double Ta = s.tsa[2 * m.memb[n]];
    const int p0 = m.n2el_ptr[n], p1 = m.n2el_ptr[n + 1];
    double ssum = 0.0;
    for (int j = p0; j < p1; ++j) ssum = ssum + m.elem_area[m.n2el[j]];
    const double tvol = 1.0 / ssum;
    double tx = 0.0, ty = 0.0;
    for (int j = p0; j < p1; ++j) {
        const int el = m.n2el[j];
        const double ea = m.elem_area[el];
        const double2 g = elem_grad(m, s.TF, Ta, el, k);
        tx = tx + g.x * ea;
        ty = ty + g.y * ea;
    }
    PL(s.gTx, k, n, KC) = tx * tvol;
    PL(s.gTy, k, n, KC) = ty * tvol;
    PL(s.tcur, k, n, KC) = PL(s.TF, k, n, KC) - Ta;
}

// MUSCL edge value: linear reconstructions from both end nodes, 15 %
// upwind-biased blend (fv_tracer.F90:104-122).
// x/6 in the storage type: the reference's division in FP64; in the FP32
// path a multiplication by the rounded reciprocal (within an ulp of the
// quotient, at a fraction of the cost of an IEEE division)
__device__ __forceinline__ double div6(double x) { return x / 6.0; }
__device__ __forceinline__ float div6(float x) { return x * (1.0f / 6.0f); }

template <typename R>
__device__ __forceinline__ R muscl_mean(R T1, R T2, R dT, R dx, R dy, R g1x, R g1y, R g2x, R g2y, bool from2) {
    const R Tm2 = T2 - div6(dT + dx * g2x + dy * R(kREarthD) * g2y);
    const R Tm1 = T1 + div6(dT + dx * g1x + dy * R(kREarthD) * g1y);
    return from2 ? R(0.15) * Tm2 + R(0.85) * (Tm1 + Tm2) / R(2.0) : R(0.15) * Tm1 + R(0.85) * (Tm1 + Tm2) / R(2.0);
}

// solve_tracer_muscl for T and S (fv_tracer.F90:8-379), one thread per node
// and layer on node-gather blocks.  Per edge the first segment uses the
// filtered velocity of element 1; the second segment's flux uses the
// unfiltered U_n/V_n of element 2 (its upwind test and the volume flux use
// the filtered one), as the reference does.  The node's own T, S and
// gradients are loaded once; per edge only the other end's.  The implicit
// vertical diffusion follows in k_tr_vdiff.
template <typename R>
__global__ __launch_bounds__(1024) void k_tr_muscl(DevMesh m, DevStateT<R> s, double dt, double Dmin, int cl_relax, double clim_relax,
                           double relax2clim_ac) {
    extern __shared__ __align__(16) unsigned char lds_[];
    const int KC = m.K;
    const ColFlat f = col_flat(m.N, m.L);
    if (f.nc <= 0) return;
    const int2 jr = csr_block(m.n2e_ptr, f);
    const int L = m.L, T = f.cpb * L;
    EdgeGeo<R> *rec = (EdgeGeo<R> *)lds_;
    int *oc = (int *)(rec + (jr.y - jr.x));  // the other end's column in this block (-1: outside), sign: node is edge_nodes(1)
    R *cT = (R *)(oc + ((jr.y - jr.x + 1) & ~1)), *cS = cT + T, *cTx = cS + T, *cTy = cTx + T, *cSx = cTy + T, *cSy = cSx + T;
    {
        const EdgeGeo<R> *eg = (const EdgeGeo<R> *)m.egeo;
        for (int t = threadIdx.x; t < jr.y - jr.x; t += blockDim.x) {
            const int cc = m.n2e[jr.x + t];
            const EdgeGeo<R> g = eg[cc >= 0 ? cc : ~cc];
            rec[t] = g;
            const int on = (cc >= 0 ? g.n2K : g.n1K) / KC - f.c0;
            const int col = (on >= 0 && on < f.nc) ? on : -1;
            oc[t] = cc >= 0 ? col + 2 : -(col + 2);  // |oc| - 2 = column, sign = orientation
        }
    }
    const int n = f.c0 + f.c, k = f.k, nk = n * KC + k, ck = f.c * L + k;
    if (f.on()) {
        cT[ck] = s.abT[nk];
        cS[ck] = s.abS[nk];
        cTx[ck] = s.gTx[nk];
        cTy[ck] = s.gTy[nk];
        cSx[ck] = s.gSx[nk];
        cSy[ck] = s.gSy[nk];
    }
    __syncthreads();
    if (!f.on()) return;
    const int mb = m.memb[n];
    const double Ta = s.tsa[2 * mb], Sa = s.tsa[2 * mb + 1];
    const double a = s.a[n];
    const int p0 = m.n2e_ptr[n], p1 = m.n2e_ptr[n + 1];
    const R To = cT[ck], So = cS[ck], gTxo = cTx[ck], gTyo = cTy[ck], gSxo = cSx[ck], gSyo = cSy[ck];
    // horizontal fluxes in the storage type (FP32 in the mixed-precision path)
    R rt = R(0), rs = R(0), rc = R(0);
    for (int j = p0 - jr.x; j < p1 - jr.x; ++j) {
        const EdgeGeo<R> &g = rec[j];
        const int o = oc[j];
        const bool own1 = o > 0;
        const int col = (own1 ? o : -o) - 2;
        const R cdx = g.cx, cdy = g.cy, cdz = g.cz, cdw = g.cw, ddy = g.dy;
        const int i1 = g.e1K + k;
        R Tx, Sx, gTxx, gTyx, gSxx, gSyx;
        if (col >= 0) {  // the other end is a column of this block
            const int i = col * L + k;
            Tx = lds_ld(cT + i);
            Sx = lds_ld(cS + i);
            gTxx = lds_ld(cTx + i);
            gTyx = lds_ld(cTy + i);
            gSxx = lds_ld(cSx + i);
            gSyx = lds_ld(cSy + i);
        } else {
            const int ao = (own1 ? g.n2K : g.n1K) + k;
            Tx = glb_ld(s.abT + ao);
            Sx = glb_ld(s.abS + ao);
            gTxx = glb_ld(s.gTx + ao);
            gTyx = glb_ld(s.gTy + ao);
            gSxx = glb_ld(s.gSx + ao);
            gSyx = glb_ld(s.gSy + ao);
        }
        const R T1 = own1 ? To : Tx, T2 = own1 ? Tx : To;
        const R S1 = own1 ? So : Sx, S2 = own1 ? Sx : So;
        const R gT1x = own1 ? gTxo : gTxx, gT2x = own1 ? gTxx : gTxo;
        const R gT1y = own1 ? gTyo : gTyx, gT2y = own1 ? gTyx : gTyo;
        const R gS1x = own1 ? gSxo : gSxx, gS2x = own1 ? gSxx : gSxo;
        const R gS1y = own1 ? gSyo : gSyx, gS2y = own1 ? gSyx : gSyo;
        const R dTt = R(2.0) * (T2 - T1), dTs = R(2.0) * (S2 - S1);
        // first segment
        const R dxa = g.dxa;
        const R u1 = s.U_nf[i1], v1 = s.V_nf[i1], J1 = s.Je[i1];
        const bool f2 = v1 * cdx - u1 * cdy > R(0);
        const R mT = muscl_mean(T1, T2, dTt, dxa, ddy, gT1x, gT1y, gT2x, gT2y, f2);
        const R mS = muscl_mean(S1, S2, dTs, dxa, ddy, gS1x, gS1y, gS2x, gS2y, f2);
        const R cT = (J1 * v1 * mT) * cdx - (J1 * u1 * mT) * cdy;
        const R cS = (J1 * v1 * mS) * cdx - (J1 * u1 * mS) * cdy;
        const R cV = J1 * v1 * cdx - J1 * u1 * cdy;
        if (own1) {
            rt = rt + cT;
            rs = rs + cS;
        } else {
            rt = rt - cT;
            rs = rs - cS;
        }
        if (g.e2K >= 0) {  // second segment
            const int i2 = g.e2K + k;
            const R dxb = g.dxb;
            const R u2f = s.U_nf[i2], v2f = s.V_nf[i2];
            const R u2 = s.U_n[i2], v2 = s.V_n[i2], J2 = s.Je[i2];
            const bool g2 = v2f * cdz - u2f * cdw < R(0);
            const R nT = muscl_mean(T1, T2, dTt, dxb, ddy, gT1x, gT1y, gT2x, gT2y, g2);
            const R nS = muscl_mean(S1, S2, dTs, dxb, ddy, gS1x, gS1y, gS2x, gS2y, g2);
            const R c2T = -(J2 * v2 * nT) * cdz + (J2 * u2 * nT) * cdw;
            const R c2S = -(J2 * v2 * nS) * cdz + (J2 * u2 * nS) * cdw;
            const R c2V = -J2 * v2f * cdz + J2 * u2f * cdw;
            if (own1) {
                rt = rt + c2T;
                rs = rs + c2S;
                rc = rc + cV;
                rc = rc + c2V;
            } else {
                rt = rt - c2T;
                rs = rs - c2S;
                rc = rc - cV;
                rc = rc - c2V;
            }
        } else {
            rc = own1 ? rc + cV : rc - cV;
        }
    }
    // QUICK vertical fluxes at the layer's interfaces; when W is exactly zero
    // the reference's carried Tm multiplies a zero, so any finite start value
    // gives the same bits.
    double TmT = 0.0, TmS = 0.0, tvt, tvb, svt, svb;
    hi_vert2<true>(m, s, s.abT, s.abS, n, k + 1, a, TmT, TmS, tvt, svt);
    hi_vert2<true>(m, s, s.abT, s.abS, n, k + 2, a, TmT, TmS, tvb, svb);
    const double tab = To, sab = So;
    const double wdiff = PL(s.W, k, n, KC) - PL(s.W, k + 1, n, KC);
    const double jc = a * m.dsig[k];
    const double aj = m.area[n] * jc;
    if constexpr (sizeof(R) == 4) {  // mixed path: two reciprocals instead of eight FP64 divisions
        const double iaj = 1.0 / aj, ijc = 1.0 / jc;
        rt = rt * iaj - tab * rc * iaj + tab * wdiff * ijc + (tvt - tvb) * ijc;
        rs = rs * iaj - sab * rc * iaj + sab * wdiff * ijc + (svt - svb) * ijc;
    } else {
        rt = rt / aj - tab * rc / aj + tab * wdiff / jc + (tvt - tvb) / jc;
        rs = rs / aj - sab * rc / aj + sab * wdiff / jc + (svt - svb) / jc;
    }
    const double tdm = PL(s.TF, k, n, KC) - Ta, sdm = PL(s.SF, k, n, KC) - Sa;
    if (cl_relax && m.index[n] == 2) {  // climatology relaxation on the current de-meaned tracer
        if (ob_inflow(m, s, p0, p1, Dmin) > 0.0) {
            rt = rt + relax2clim_ac * (PL(s.Tclim, k, n, KC) - (tdm + Ta));
            rs = rs + relax2clim_ac * (PL(s.Sclim, k, n, KC) - (sdm + Sa));
        } else {
            rt = clim_relax * (PL(s.Tclim, k, n, KC) - (tdm + Ta));
            rs = clim_relax * (PL(s.Sclim, k, n, KC) - (sdm + Sa));
        }
    }
    PL(s.T_old, k, n, KC) = tdm + Ta;
    PL(s.S_old, k, n, KC) = sdm + Sa;
    PL(s.TF, k, n, KC) = tdm + rt * dt + Ta;
    PL(s.SF, k, n, KC) = sdm + rs * dt + Sa;
}

// hi_vert2 on a node column held in LDS (X, Y: T and S; Wc: W, slot z-1 of
// 1-based level z; om: 1-sigma of the levels), both signs of W in one
// instruction stream: the downward branch is the upward one on the mirrored
// operands (levels nz, nz-1, nz-2 instead of nz-1, nz, nz+1) with the depth
// differences negated (x - y = -(y - x) exactly), so every value equals
// hi_vert2's bit for bit; TmX/TmY keep their values when W is zero (or NaN).
template <bool kChained, typename R>
__device__ __forceinline__ void hi_vert2_col(const double *om, const R *X, const R *Y, const R *Wc, int K, int nz,
                                             double a, double &TmX, double &TmY, double &fx, double &fy) {
    if (nz == 1) {
        const double w1 = Wc[0];
        fx = -w1 * (X[0] - 0.0);
        fy = -w1 * (Y[0] - 0.0);
        return;
    }
    if (nz == K) {
        fx = 0.0;
        fy = 0.0;
        return;
    }
    const double w = Wc[nz - 1];
    if (w > 0.0 || w < 0.0) {
        const bool up = w > 0.0;
        const int iP = up ? nz - 1 : nz, iQ = up ? nz : nz - 1, iR = up ? nz + 1 : nz - 2;  // 1-based levels
        if (up ? nz == K - 1 : nz == 2) {
            TmX = X[iQ - 1] - 0.0;
            TmY = Y[iQ - 1] - 0.0;
        } else {
            auto Zc = [&](int z) { return 0.5 * (om[z - 1] * a + om[z] * a); };
            const double sg = up ? 1.0 : -1.0;
            const double zb = om[nz - 1] * a;
            const double aa = sg * (zb - Zc(iP)), bb = sg * (Zc(iQ) - zb), cc = sg * (Zc(iR) - zb);
            double dg, db;
            if (kChained) {
                dg = aa * bb / (cc + aa) / (bb - cc);
                db = -aa * cc / (bb + aa) / (bb - cc);
            } else {
                dg = aa * bb / ((cc + aa) * (bb - cc));
                db = -aa * cc / ((bb + aa) * (bb - cc));
            }
            const double da = 1.0 - dg - db;
            TmX = (X[iP - 1] - 0.0) * da + (X[iQ - 1] - 0.0) * db + (X[iR - 1] - 0.0) * dg;
            TmY = (Y[iP - 1] - 0.0) * da + (Y[iQ - 1] - 0.0) * db + (Y[iR - 1] - 0.0) * dg;
        }
    }
    fx = -TmX * w;
    fy = -TmY * w;
}

// solve_tracer_muscl as k_tr_muscl, with each lane owning W consecutive
// layers of its node column (LV<R>; nsigma a multiple of W): the block's
// columns of AB-interpolated T and S, their node gradients and W sit in LDS
// at a stride of K slots, neighbours outside the block and the element
// velocities come through 16-byte loads, so the edge records, orientation
// and addresses of an edge serve W layers.  Each edge is taken from this
// node's side: the face value of the 15 %-upwind blend is formed from the
// own and the other node's reconstruction (x + y = y + x and x - y =
// x + (-y) exactly, so no operand needs reordering by edge_nodes), and the
// QUICK vertical fluxes are formed once per interface of the lane's run.
// Per layer the arithmetic and the order of the edge sum are k_tr_muscl's:
// the results are identical bit for bit (a flux through an interface where
// W is exactly zero can differ in the sign of that zero).
// cap: the most CSR slots of any block, for which the LDS is sized (the
// layout below follows it); a block with more sets a flag and does nothing.
template <typename R>
__global__ __launch_bounds__(256) void k_tr_muscl_v(DevMesh m, DevStateT<R> s, double dt, double Dmin, int cl_relax,
                                                    double clim_relax, double relax2clim_ac, int cap) {
    constexpr int W = LV<R>::W;
    extern __shared__ __align__(16) unsigned char lds_[];
    const int KC = m.K, K = m.K, L = m.L;
    const ColFlat f = col_flat(m.N, K / W);  // f.k: the lane's run of W slots
    if (f.nc <= 0) return;
    const int2 jr = csr_block(m.n2e_ptr, f);
    const int nrec = jr.y - jr.x, T = f.cpb * K;
    if (nrec > cap) {
        if (threadIdx.x == 0) atomicOr(s.err + 3, 1);
        return;
    }
    EdgeGeo<R> *rec = (EdgeGeo<R> *)lds_;
    int *oc = (int *)(rec + cap);  // the other end's column in this block (-1: outside), sign: node is edge_nodes(1)
    R *cT = (R *)(lds_ + (((size_t)cap * (sizeof(EdgeGeo<R>) + sizeof(int)) + 15) & ~(size_t)15));
    R *cS = cT + T, *cTx = cS + T, *cTy = cTx + T, *cSx = cTy + T, *cSy = cSx + T, *cW = cSy + T;
    double *om = (double *)(cW + T);
    {
        const EdgeGeo<R> *eg = (const EdgeGeo<R> *)m.egeo;
        for (int t = threadIdx.x; t < nrec; t += blockDim.x) {
            const int cc = m.n2e[jr.x + t];
            const EdgeGeo<R> g = eg[cc >= 0 ? cc : ~cc];
            rec[t] = g;
            const int on = (cc >= 0 ? g.n2K : g.n1K) / KC - f.c0;
            const int col = (on >= 0 && on < f.nc) ? on : -1;
            oc[t] = cc >= 0 ? col + 2 : -(col + 2);  // |oc| - 2 = column, sign = orientation
        }
        for (int t = threadIdx.x; t < K; t += blockDim.x) om[t] = m.oms[t];
    }
    const int n = f.c0 + f.c, k0 = f.k * W, ck0 = f.c * K + k0;
    const size_t nk0 = (size_t)n * KC + k0;
    if (f.on()) {
        stv(cT + ck0, ldv_glb(s.abT + nk0));
        stv(cS + ck0, ldv_glb(s.abS + nk0));
        stv(cTx + ck0, ldv_glb(s.gTx + nk0));
        stv(cTy + ck0, ldv_glb(s.gTy + nk0));
        stv(cSx + ck0, ldv_glb(s.gSx + nk0));
        stv(cSy + ck0, ldv_glb(s.gSy + nk0));
        stv(cW + ck0, ldv_glb(s.W + nk0));
    }
    __syncthreads();
    if (!f.on()) return;
    const int nl = min(W, L - k0);  // layers of this lane (the column's last lane: W-1)
    const int mb = m.memb[n];
    const double Ta = s.tsa[2 * mb], Sa = s.tsa[2 * mb + 1];
    const double a = s.a[n];
    const int p0 = m.n2e_ptr[n], p1 = m.n2e_ptr[n + 1];
    const LV<R> To = ldv_lds(cT + ck0), So = ldv_lds(cS + ck0);
    const LV<R> gTxo = ldv_lds(cTx + ck0), gTyo = ldv_lds(cTy + ck0), gSxo = ldv_lds(cSx + ck0), gSyo = ldv_lds(cSy + ck0);
    // horizontal fluxes in the storage type (FP32 in the mixed-precision path)
    R rt[W], rs[W], rc[W];
#pragma unroll
    for (int i = 0; i < W; ++i) rt[i] = rs[i] = rc[i] = R(0);
    for (int j = p0 - jr.x; j < p1 - jr.x; ++j) {
        const EdgeGeo<R> &g = rec[j];
        const int o = oc[j];
        const bool own1 = o > 0;  // this node is edge_nodes(1)
        const int col = (own1 ? o : -o) - 2;
        const R sg = own1 ? R(1) : R(-1);
        const R cdx = g.cx, cdy = g.cy, cdz = g.cz, cdw = g.cw, ddy = g.dy, dxa = g.dxa, dxb = g.dxb;
        LV<R> Tx, Sx, gTxx, gTyx, gSxx, gSyx;
        if (col >= 0) {  // the other end is a column of this block
            const int i0 = col * K + k0;
            Tx = ldv_lds(cT + i0);
            Sx = ldv_lds(cS + i0);
            gTxx = ldv_lds(cTx + i0);
            gTyx = ldv_lds(cTy + i0);
            gSxx = ldv_lds(cSx + i0);
            gSyx = ldv_lds(cSy + i0);
        } else {
            const size_t ao = (size_t)(own1 ? g.n2K : g.n1K) + k0;
            Tx = ldv_glb(s.abT + ao);
            Sx = ldv_glb(s.abS + ao);
            gTxx = ldv_glb(s.gTx + ao);
            gTyx = ldv_glb(s.gTy + ao);
            gSxx = ldv_glb(s.gSx + ao);
            gSyx = ldv_glb(s.gSy + ao);
        }
        const size_t i1 = (size_t)g.e1K + k0;
        const LV<R> u1 = ldv_glb(s.U_nf + i1), v1 = ldv_glb(s.V_nf + i1), J1 = ldv_glb(s.Je + i1);
        const bool seg2 = g.e2K >= 0;
        LV<R> u2f, v2f, u2, v2, J2;
        if (seg2) {
            const size_t i2 = (size_t)g.e2K + k0;
            u2f = ldv_glb(s.U_nf + i2);
            v2f = ldv_glb(s.V_nf + i2);
            u2 = ldv_glb(s.U_n + i2);
            v2 = ldv_glb(s.V_n + i2);
            J2 = ldv_glb(s.Je + i2);
        }
#pragma unroll
        for (int i = 0; i < W; ++i) {
            // dT = 2 (T2 - T1) from this node's side
            const R dTt = sg * (R(2.0) * (Tx.x[i] - To.x[i])), dTs = sg * (R(2.0) * (Sx.x[i] - So.x[i]));
            // face value of one segment (length dx): the own and the other
            // node's reconstructions (muscl_mean's Tm1/Tm2 by orientation),
            // 15 % from the upwind one (from2: from edge_nodes(2))
            auto face = [&](R To_, R Tx_, R dT, R dx, R gxo, R gyo, R gxx, R gyx, bool from2) {
                const R Ao = div6(dT + dx * gxo + ddy * R(kREarthD) * gyo);
                const R Ax = div6(dT + dx * gxx + ddy * R(kREarthD) * gyx);
                const R Tmo = To_ + sg * Ao, Tmx = Tx_ - sg * Ax;
                const R up = (from2 == own1) ? Tmx : Tmo;
                return R(0.15) * up + R(0.85) * (Tmo + Tmx) / R(2.0);
            };
            const bool f2 = v1.x[i] * cdx - u1.x[i] * cdy > R(0);
            const R mT = face(To.x[i], Tx.x[i], dTt, dxa, gTxo.x[i], gTyo.x[i], gTxx.x[i], gTyx.x[i], f2);
            const R mS = face(So.x[i], Sx.x[i], dTs, dxa, gSxo.x[i], gSyo.x[i], gSxx.x[i], gSyx.x[i], f2);
            const R J1i = J1.x[i], u1i = u1.x[i], v1i = v1.x[i];
            const R cT = (J1i * v1i * mT) * cdx - (J1i * u1i * mT) * cdy;
            const R cS = (J1i * v1i * mS) * cdx - (J1i * u1i * mS) * cdy;
            const R cV = J1i * v1i * cdx - J1i * u1i * cdy;
            if (own1) {
                rt[i] = rt[i] + cT;
                rs[i] = rs[i] + cS;
            } else {
                rt[i] = rt[i] - cT;
                rs[i] = rs[i] - cS;
            }
            if (seg2) {  // second segment
                const bool g2 = v2f.x[i] * cdz - u2f.x[i] * cdw < R(0);
                const R nT = face(To.x[i], Tx.x[i], dTt, dxb, gTxo.x[i], gTyo.x[i], gTxx.x[i], gTyx.x[i], g2);
                const R nS = face(So.x[i], Sx.x[i], dTs, dxb, gSxo.x[i], gSyo.x[i], gSxx.x[i], gSyx.x[i], g2);
                const R J2i = J2.x[i], u2i = u2.x[i], v2i = v2.x[i];
                const R c2T = -(J2i * v2i * nT) * cdz + (J2i * u2i * nT) * cdw;
                const R c2S = -(J2i * v2i * nS) * cdz + (J2i * u2i * nS) * cdw;
                const R c2V = -J2i * v2f.x[i] * cdz + J2i * u2f.x[i] * cdw;
                if (own1) {
                    rt[i] = rt[i] + c2T;
                    rs[i] = rs[i] + c2S;
                    rc[i] = rc[i] + cV;
                    rc[i] = rc[i] + c2V;
                } else {
                    rt[i] = rt[i] - c2T;
                    rs[i] = rs[i] - c2S;
                    rc[i] = rc[i] - cV;
                    rc[i] = rc[i] - c2V;
                }
            } else {
                rc[i] = own1 ? rc[i] + cV : rc[i] - cV;
            }
        }
    }
    // QUICK vertical fluxes through the interfaces k0+1 .. k0+nl+1; the
    // bottom flux of one layer is the top flux of the next
    const R *Tc = cT + f.c * K, *Sc = cS + f.c * K, *Wc = cW + f.c * K;
    double TmT = 0.0, TmS = 0.0, tvt, svt;
    hi_vert2_col<true>(om, Tc, Sc, Wc, K, k0 + 1, a, TmT, TmS, tvt, svt);
    const LV<R> TF0 = ldv_glb(s.TF + nk0), SF0 = ldv_glb(s.SF + nk0);
    const bool relax = cl_relax && m.index[n] == 2;
    const bool inflow = relax && ob_inflow(m, s, p0, p1, Dmin) > 0.0;
    LV<R> Tcl, Scl;
    if (relax) {
        Tcl = ldv_glb(s.Tclim + nk0);
        Scl = ldv_glb(s.Sclim + nk0);
    }
    const double arn = m.area[n];
    LV<R> oTo, oSo, oT, oS;
#pragma unroll
    for (int i = 0; i < W; ++i) {
        const int k = k0 + i;
        double tvb = 0.0, svb = 0.0;
        if (i < nl) hi_vert2_col<true>(om, Tc, Sc, Wc, K, k + 2, a, TmT, TmS, tvb, svb);
        const double tab = To.x[i], sab = So.x[i];
        const double wdiff = Wc[k] - Wc[k < K - 1 ? k + 1 : k];
        const double jc = a * m.dsig[k < L ? k : L - 1];
        const double aj = arn * jc;
        R rti = rt[i], rsi = rs[i];
        if constexpr (sizeof(R) == 4) {  // as the one-level kernel: two reciprocals in the mixed path
            const double iaj = 1.0 / aj, ijc = 1.0 / jc;
            rti = rti * iaj - tab * rc[i] * iaj + tab * wdiff * ijc + (tvt - tvb) * ijc;
            rsi = rsi * iaj - sab * rc[i] * iaj + sab * wdiff * ijc + (svt - svb) * ijc;
        } else {
            rti = rti / aj - tab * rc[i] / aj + tab * wdiff / jc + (tvt - tvb) / jc;
            rsi = rsi / aj - sab * rc[i] / aj + sab * wdiff / jc + (svt - svb) / jc;
        }
        const double tdm = TF0.x[i] - Ta, sdm = SF0.x[i] - Sa;
        if (relax) {  // climatology relaxation on the current de-meaned tracer
            if (inflow) {
                rti = rti + relax2clim_ac * (Tcl.x[i] - (tdm + Ta));
                rsi = rsi + relax2clim_ac * (Scl.x[i] - (sdm + Sa));
            } else {
                rti = clim_relax * (Tcl.x[i] - (tdm + Ta));
                rsi = clim_relax * (Scl.x[i] - (sdm + Sa));
            }
        }
        oTo.x[i] = tdm + Ta;
        oSo.x[i] = sdm + Sa;
        oT.x[i] = tdm + rti * dt + Ta;
        oS.x[i] = sdm + rsi * dt + Sa;
        tvt = tvb;
        svt = svb;
    }
    stv_n(s.T_old + nk0, oTo, nl);
    stv_n(s.S_old + nk0, oSo, nl);
    stv_n(s.TF + nk0, oT, nl);
    stv_n(s.SF + nk0, oS, nl);
}

// ---------------------------------------------------------------------------
// Level-vector variants of the node-gather and column kernels below: each
// lane owns W = LV<R>::W consecutive layers of its node or element column
// (nsigma a multiple of W) and moves them with 16-byte loads and stores, so
// a CSR slot's LDS record, address and neighbour load serve W layers.  Per
// layer the arithmetic and the order of every sum are those of the one-level
// kernel named in each comment: the results are identical bit for bit.  A
// node-gather block with more CSR slots than its LDS capacity `cap` sets a
// flag and does nothing.
// ---------------------------------------------------------------------------

// compute_el2nodes_3D / pulsation velocity as k_el2nodes
template <typename R, bool P>
__global__ __launch_bounds__(256) void k_el2nodes_v(DevMesh m, const R *__restrict__ Ue, const R *__restrict__ Ve,
                                                    R *__restrict__ Un, R *__restrict__ Vn,
                                                    const double *__restrict__ ap, const double *__restrict__ U2,
                                                    int cap, int *err) {
    constexpr int W = LV<R>::W;
    extern __shared__ __align__(16) unsigned char lds_[];
    const int KC = m.K, L = m.L;
    const ColFlat f = col_flat(m.N, KC / W);
    if (f.nc <= 0) return;
    const int2 jr = csr_block(m.n2el_ptr, f);
    if (jr.y - jr.x > cap) {
        if (threadIdx.x == 0) atomicOr(err, 1);
        return;
    }
    PulsRec *rec = (PulsRec *)lds_;
    for (int t = threadIdx.x; t < jr.y - jr.x; t += blockDim.x) {
        PulsRec r;
        const int e = m.n2el[jr.x + t];
        r.elK = m.n2elK[jr.x + t];
        r.ar = m.elem_area[e];
        if (P) {
            r.ap = ap[e];
            r.u2 = U2[2 * e];
            r.v2 = U2[2 * e + 1];
        }
        rec[t] = r;
    }
    __syncthreads();
    if (!f.on()) return;
    const int n = f.c0 + f.c, k0 = f.k * W, nl = min(W, L - k0);
    const int p0 = m.n2el_ptr[n] - jr.x, p1 = m.n2el_ptr[n + 1] - jr.x;
    double sa = 0.0;
    for (int j = p0; j < p1; ++j) sa = sa + rec[j].ar;
    const double tvi = 1.0 / sa;
    double su[W], sv[W];
#pragma unroll
    for (int i = 0; i < W; ++i) su[i] = sv[i] = 0.0;
    for (int j = p0; j < p1; ++j) {
        const PulsRec r = rec[j];
        const LV<R> ue = ldv_glb(Ue + r.elK + k0), ve = ldv_glb(Ve + r.elK + k0);
#pragma unroll
        for (int i = 0; i < W; ++i) {
            R uei = ue.x[i], vei = ve.x[i];
            if (P) {
                const double un = uei, vn = vei;
                uei = (R)(un - r.ap * r.u2);
                vei = (R)(vn - r.ap * r.v2);
            }
            su[i] = su[i] + uei * r.ar;
            sv[i] = sv[i] + vei * r.ar;
        }
    }
    LV<R> ou, ov;
#pragma unroll
    for (int i = 0; i < W; ++i) {
        ou.x[i] = su[i] * tvi;
        ov.x[i] = sv[i] * tvi;
    }
    stv_n(Un + (size_t)n * KC + k0, ou, nl);
    stv_n(Vn + (size_t)n * KC + k0, ov, nl);
}

// vert_vel_sigma's transport divergence as k_vertvel_div
template <typename R>
__global__ __launch_bounds__(256) void k_vertvel_div_v(DevMesh m, DevStateT<R> s, int cap) {
    constexpr int W = LV<R>::W;
    extern __shared__ __align__(16) unsigned char lds_[];
    const int KC = m.K, L = m.L;
    const ColFlat f = col_flat(m.N, KC / W);
    if (f.nc <= 0) return;
    const int2 jr = csr_block(m.n2e_ptr, f);
    if (jr.y - jr.x > cap) {
        if (threadIdx.x == 0) atomicOr(s.err + 3, 1);
        return;
    }
    DivRec<R> *rec = (DivRec<R> *)lds_;
    {
        const EdgeGeo<R> *eg = (const EdgeGeo<R> *)m.egeo;
        for (int t = threadIdx.x; t < jr.y - jr.x; t += blockDim.x) {
            const int cc = m.n2e[jr.x + t];
            const EdgeGeo<R> g = eg[cc >= 0 ? cc : ~cc];
            DivRec<R> r;
            r.e1K = g.e1K;
            r.e2K = g.e2K;
            r.pos = cc >= 0;
            r.cx = g.cx;
            r.cy = g.cy;
            r.cz = g.cz;
            r.cw = g.cw;
            r.vf1 = (R)s.V_filt[g.e1];
            r.uf1 = (R)s.U_filt[g.e1];
            r.vf2 = g.e2K >= 0 ? (R)s.V_filt[g.e2] : R(0);
            r.uf2 = g.e2K >= 0 ? (R)s.U_filt[g.e2] : R(0);
            rec[t] = r;
        }
    }
    __syncthreads();
    if (!f.on()) return;
    const int n = f.c0 + f.c, k0 = f.k * W, nl = min(W, L - k0);
    const int p0 = m.n2e_ptr[n] - jr.x, p1 = m.n2e_ptr[n + 1] - jr.x;
    LV<R> c;  // storage type: FP32 transport fluxes in the mixed-precision path
#pragma unroll
    for (int i = 0; i < W; ++i) c.x[i] = R(0);
    for (int j = p0; j < p1; ++j) {
        const DivRec<R> g = rec[j];
        const LV<R> J1 = ldv_glb(s.Je + g.e1K + k0), V1 = ldv_glb(s.V_nf + g.e1K + k0), U1 = ldv_glb(s.U_nf + g.e1K + k0);
        const bool two = g.e2K >= 0;
        LV<R> J2, V2, U2;
        if (two) {
            J2 = ldv_glb(s.Je + g.e2K + k0);
            V2 = ldv_glb(s.V_nf + g.e2K + k0);
            U2 = ldv_glb(s.U_nf + g.e2K + k0);
        }
#pragma unroll
        for (int i = 0; i < W; ++i) {
            R fl = J1.x[i] * ((V1.x[i] - g.vf1) * g.cx - (U1.x[i] - g.uf1) * g.cy);
            if (two) fl = fl + J2.x[i] * (-(V2.x[i] - g.vf2) * g.cz + (U2.x[i] - g.uf2) * g.cw);
            c.x[i] = g.pos ? c.x[i] + fl : c.x[i] - fl;
        }
    }
    stv_n(s.W + (size_t)n * KC + k0, c, nl);
}

// node-averaged gradients of the AB-interpolated T and S as k_tr_gradn_muscl
template <typename R>
__global__ __launch_bounds__(256) void k_tr_gradn_muscl_v(DevMesh m, DevStateT<R> s, int cap) {
    constexpr int W = LV<R>::W;
    extern __shared__ __align__(16) unsigned char lds_[];
    const int KC = m.K, L = m.L;
    const ColFlat f = col_flat(m.N, KC / W);
    if (f.nc <= 0) return;
    const int2 jr = csr_block(m.n2el_ptr, f);
    if (jr.y - jr.x > cap) {
        if (threadIdx.x == 0) atomicOr(s.err + 3, 1);
        return;
    }
    GradRec *rec = (GradRec *)lds_;
    for (int t = threadIdx.x; t < jr.y - jr.x; t += blockDim.x) {
        const int el = m.n2el[jr.x + t];
        const int4 q = m.enK[el];
        const double4 gx = m.gsx[el], gy = m.gsy[el];
        GradRec r;
        r.qK[0] = q.x;
        r.qK[1] = q.y;
        r.qK[2] = q.z;
        r.qK[3] = q.w;
        r.ea = m.elem_area[el];
        r.gx[0] = gx.x;
        r.gx[1] = gx.y;
        r.gx[2] = gx.z;
        r.gx[3] = gx.w;
        r.gy[0] = gy.x;
        r.gy[1] = gy.y;
        r.gy[2] = gy.z;
        r.gy[3] = gy.w;
        rec[t] = r;
    }
    __syncthreads();
    if (!f.on()) return;
    const int n = f.c0 + f.c, k0 = f.k * W, nl = min(W, L - k0);
    const int p0 = m.n2el_ptr[n] - jr.x, p1 = m.n2el_ptr[n + 1] - jr.x;
    double tvol = 0.0, tx[W], ty[W], sx[W], sy[W];
#pragma unroll
    for (int i = 0; i < W; ++i) tx[i] = ty[i] = sx[i] = sy[i] = 0.0;
    for (int j = p0; j < p1; ++j) {
        const GradRec &r = rec[j];
        const double ea = r.ea;
        const LV<R> a0 = ldv_glb(s.abT + r.qK[0] + k0), a1 = ldv_glb(s.abT + r.qK[1] + k0);
        const LV<R> a2 = ldv_glb(s.abT + r.qK[2] + k0), a3 = ldv_glb(s.abT + r.qK[3] + k0);
        const LV<R> b0 = ldv_glb(s.abS + r.qK[0] + k0), b1 = ldv_glb(s.abS + r.qK[1] + k0);
        const LV<R> b2 = ldv_glb(s.abS + r.qK[2] + k0), b3 = ldv_glb(s.abS + r.qK[3] + k0);
        tvol = tvol + ea;
#pragma unroll
        for (int i = 0; i < W; ++i) {
            const double t0 = a0.x[i] - 0.0, t1 = a1.x[i] - 0.0, t2 = a2.x[i] - 0.0, t3 = a3.x[i] - 0.0;
            const double s0 = b0.x[i] - 0.0, s1 = b1.x[i] - 0.0, s2 = b2.x[i] - 0.0, s3 = b3.x[i] - 0.0;
            tx[i] = tx[i] + sum4a(r.gx, t0, t1, t2, t3) * ea;
            ty[i] = ty[i] + sum4a(r.gy, t0, t1, t2, t3) * ea;
            sx[i] = sx[i] + sum4a(r.gx, s0, s1, s2, s3) * ea;
            sy[i] = sy[i] + sum4a(r.gy, s0, s1, s2, s3) * ea;
        }
    }
    LV<R> o0, o1, o2, o3;
#pragma unroll
    for (int i = 0; i < W; ++i) {
        o0.x[i] = tx[i] / tvol;
        o1.x[i] = ty[i] / tvol;
        o2.x[i] = sx[i] / tvol;
        o3.x[i] = sy[i] / tvol;
    }
    const size_t nk0 = (size_t)n * KC + k0;
    stv_n(s.gTx + nk0, o0, nl);
    stv_n(s.gTy + nk0, o1, nl);
    stv_n(s.gSx + nk0, o2, nl);
    stv_n(s.gSy + nk0, o3, nl);
}

// vel_gradients (3-D, free slip) as k_velgrad3d
template <typename R>
__global__ __launch_bounds__(256) void k_velgrad3d_v(DevMesh m, DevStateT<R> s) {
    constexpr int W = LV<R>::W;
    extern __shared__ __align__(16) unsigned char lds_[];
    const int KC = m.K, L = m.L;
    const ColFlat f = col_flat(m.E, KC / W);
    VelGradRec *rec = (VelGradRec *)lds_;
    if (f.nc <= 0) return;
    if ((int)threadIdx.x < f.nc) {
        const int e = f.c0 + threadIdx.x;
        const int4 q = m.en[e];
        const int4 nb = m.elnb[e];
        const int4 ed = m.eledges[e];
        const double4 gx = m.gvx[e], gy = m.gvy[e];
        const int nbs[4] = {nb.x, nb.y, nb.z, nb.w};
        const int eds[4] = {ed.x, ed.y, ed.z, ed.w};
        const double gxs[4] = {gx.x, gx.y, gx.z, gx.w}, gys[4] = {gy.x, gy.y, gy.z, gy.w};
        const double ec = m.elem_cos[e];
        VelGradRec r;
        r.jend = (q.x == q.w) ? 3 : 4;
        for (int j = 0; j < 4; ++j) {
            r.gx[j] = gxs[j];
            r.gy[j] = gys[j];
            r.nbK[j] = nbs[j] >= 0 ? nbs[j] * KC : -1;
            r.zc0[j] = 0.0;
            r.zc1[j] = 0.0;
            if (j < r.jend && nbs[j] < 0) {
                const double2 dd = m.edxdy[eds[j]];
                r.zc0[j] = dd.x * ec;
                r.zc1[j] = dd.y;
            }
        }
        rec[threadIdx.x] = r;
    }
    __syncthreads();
    if (!f.on()) return;
    const int e = f.c0 + f.c, k0 = f.k * W, nl = min(W, L - k0);
    const size_t ek0 = (size_t)e * KC + k0;
    const VelGradRec &r = rec[f.c];
    const LV<R> u = ldv_glb(s.U_n + ek0), v = ldv_glb(s.V_n + ek0);
    double ux[W], uy[W], vx[W], vy[W];
#pragma unroll
    for (int i = 0; i < W; ++i) ux[i] = uy[i] = vx[i] = vy[i] = 0.0;
    for (int j = 0; j < r.jend; ++j) {
        if (r.nbK[j] >= 0) {
            const LV<R> un = ldv_glb(s.U_n + r.nbK[j] + k0), vn = ldv_glb(s.V_n + r.nbK[j] + k0);
#pragma unroll
            for (int i = 0; i < W; ++i) {
                const double du = (double)un.x[i] - (double)u.x[i];
                const double dv = (double)vn.x[i] - (double)v.x[i];
                ux[i] = ux[i] + r.gx[j] * du;
                uy[i] = uy[i] + r.gy[j] * du;
                vx[i] = vx[i] + r.gx[j] * dv;
                vy[i] = vy[i] + r.gy[j] * dv;
            }
        } else {
            const double zc0 = r.zc0[j], zc1 = r.zc1[j];
#pragma unroll
            for (int i = 0; i < W; ++i) {
                const double uu = u.x[i], vv = v.x[i];
                double un = -2 * (uu * zc1 - vv * zc0) / (zc0 * zc0 + zc1 * zc1);
                ux[i] = ux[i] + r.gx[j] * un * zc1;
                uy[i] = uy[i] + r.gy[j] * un * zc1;
                vx[i] = vx[i] - r.gx[j] * un * zc0;
                vy[i] = vy[i] - r.gy[j] * un * zc0;
            }
        }
    }
    LV<R> o0, o1, o2, o3;
#pragma unroll
    for (int i = 0; i < W; ++i) {
        o0.x[i] = ux[i];
        o1.x[i] = uy[i];
        o2.x[i] = vx[i];
        o3.x[i] = vy[i];
    }
    stv_n(s.gux + ek0, o0, nl);
    stv_n(s.guy + ek0, o1, nl);
    stv_n(s.gvx + ek0, o2, nl);
    stv_n(s.gvy + ek0, o3, nl);
}

// element Jacobians Je as k_je
template <typename R>
__global__ __launch_bounds__(256) void k_je_v(DevMesh m, DevStateT<R> s) {
    constexpr int W = LV<R>::W;
    const int KC = m.K, L = m.L;
    const ColFlat f = col_flat(m.E, KC / W);
    if (!f.on()) return;
    const int e = f.c0 + f.c, k0 = f.k * W, nl = min(W, L - k0);
    const int4 q = m.en[e];
    const double4 w = m.wcv[e];
    const double c[4] = {w.x, w.y, w.z, w.w};
    const double a0 = s.a[q.x], a1 = s.a[q.y], a2 = s.a[q.z], a3 = s.a[q.w];
    LV<R> o;
#pragma unroll
    for (int i = 0; i < W; ++i) {
        const double ds = m.dsig[k0 + i < L ? k0 + i : L - 1];
        o.x[i] = sum4a(c, a0 * ds, a1 * ds, a2 * ds, a3 * ds);
    }
    stv_n(s.Je + (size_t)e * KC + k0, o, nl);
}

// update_3D_vel as k_update3d: the layers' U_rhs*Je_old products go to LDS
// (stride K), one thread per column sums them in layer order
template <typename R>
__global__ __launch_bounds__(256) void k_update3d_v(DevMesh m, DevStateT<R> s, double Dmin, double Mt_inv) {
    constexpr int W = LV<R>::W;
    extern __shared__ __align__(16) unsigned char lds_[];
    const int L = m.L, KC = m.K;
    const ColFlat f = col_flat(m.E, KC / W);
    double *pu = (double *)lds_, *pv = pu + f.cpb * KC;
    double *sc = pv + f.cpb * KC;  // per column: U2p, V2p, uf, vf
    if (f.nc <= 0) return;
    const int e = f.c0 + f.c, k0 = f.k * W;
    const size_t ek0 = (size_t)e * KC + k0;
    double ak[W], je[W];
    LV<R> ur, vr;
    if (f.on()) {
        const int4 q = m.en[e];
        const double4 w = m.wcv[e];
        const double c[4] = {w.x, w.y, w.z, w.w};
        const double o0 = s.a_old[q.x], o1 = s.a_old[q.y], o2 = s.a_old[q.z], o3 = s.a_old[q.w];
        ur = ldv_glb(s.U_rhs + ek0);
        vr = ldv_glb(s.V_rhs + ek0);
        const LV<R> jv = ldv_glb(s.Je + ek0);
#pragma unroll
        for (int i = 0; i < W; ++i) {
            const double ds = m.dsig[k0 + i < L ? k0 + i : L - 1];
            ak[i] = sum4a(c, o0 * ds, o1 * ds, o2 * ds, o3 * ds);
            je[i] = jv.x[i];
            pu[f.c * KC + k0 + i] = (double)ur.x[i] * ak[i];
            pv[f.c * KC + k0 + i] = (double)vr.x[i] * ak[i];
        }
    }
    __syncthreads();
    if ((int)threadIdx.x < f.nc) {
        const int ci = threadIdx.x, ec = f.c0 + ci;
        const int4 q = m.en[ec];
        const double4 w = m.wcv[ec];
        const double dmi = 1.0 / dmaxr(Dmin, sum4(w, m.depth[q.x] + s.eta_n[q.x], m.depth[q.y] + s.eta_n[q.y],
                                                  m.depth[q.z] + s.eta_n[q.z], m.depth[q.w] + s.eta_n[q.w]));
        double su = 0.0, sv = 0.0;
        for (int j = 0; j < L; ++j) {
            su = su + pu[ci * KC + j];
            sv = sv + pv[ci * KC + j];
        }
        const double uf = (s.U_filt[ec] * Mt_inv) * dmi;
        const double vf = (s.V_filt[ec] * Mt_inv) * dmi;
        s.U_filt[ec] = uf;
        s.V_filt[ec] = vf;
        sc[4 * ci] = su * dmi;
        sc[4 * ci + 1] = sv * dmi;
        sc[4 * ci + 2] = uf;
        sc[4 * ci + 3] = vf;
    }
    __syncthreads();
    if (!f.on()) return;
    const int nl = min(W, L - k0);
    const double U2p = sc[4 * f.c], V2p = sc[4 * f.c + 1], uf = sc[4 * f.c + 2], vf = sc[4 * f.c + 3];
    const double u2 = s.U2[2 * e], v2 = s.U2[2 * e + 1];
    LV<R> onf, vnf, on, vn;
#pragma unroll
    for (int i = 0; i < W; ++i) {
        const double a = ak[i] / je[i];
        const double uri = ur.x[i], vri = vr.x[i];
        onf.x[i] = a * uri + uf - U2p;
        vnf.x[i] = a * vri + vf - V2p;
        on.x[i] = a * uri + u2 - U2p;
        vn.x[i] = a * vri + v2 - V2p;
    }
    stv_n(s.U_nf + ek0, onf, nl);
    stv_n(s.V_nf + ek0, vnf, nl);
    stv_n(s.U_n + ek0, on, nl);
    stv_n(s.V_n + ek0, vn, nl);
}

// ---------------------------------------------------------------------------
// surface boundary conditions from atmospheric fields (key_atm with the
// CORE bulk formulas, src/fv_sbc.F90:368-867)
// ---------------------------------------------------------------------------
__device__ __forceinline__ double sbc_sign(double a, double b) { return copysign(fabs(a), b); }  // Fortran SIGN

// L&Y stability functions (src/fv_sbc.F90:616-660)
__device__ __forceinline__ double psi_h(double zta) {
    double X2 = sqrt(fabs(1.0 - 16.0 * zta));
    X2 = X2 > 1.0 ? X2 : 1.0;
    const double stabit = 0.5 + sbc_sign(0.5, zta);
    return -5.0 * zta * stabit + (1.0 - stabit) * (2.0 * log((1.0 + X2) / 2.0));
}
__device__ __forceinline__ double psi_m(double zta) {
    const double pi = 3.141592653589793;
    double X2 = sqrt(fabs(1.0 - 16.0 * zta));
    X2 = X2 > 1.0 ? X2 : 1.0;
    const double X = sqrt(X2);
    const double stabit = 0.5 + sbc_sign(0.5, zta);
    return -5.0 * zta * stabit
           + (1.0 - stabit) * (2.0 * log((1.0 + X) / 2.0) + log((1.0 + X2) / 2.0) - 2.0 * atan(X) + pi / 2.0);
}

// core_coeff_2z (src/fv_sbc.F90:485-614): transfer coefficients of
// momentum, heat and evaporation for T/q at zt and wind at zu, after five
// Monin-Obukhov iterations; the air temperature and humidity shifted to zu
__device__ void core_coeff_2z(double zt, double zu, double sst, double T_zt, double q_sat, double q_zt, double dU,
                              double &Cd, double &Ch, double &Ce, double &T_zu, double &q_zu) {
    const double grav = 9.8, kappa = 0.4;
    const double dU10 = dU > 0.5 ? dU : 0.5;
    double dT = T_zt - sst, dq;
    double stab = 0.5 + sbc_sign(0.5, dT);
    double Cd_n10 = 1E-3 * (2.7 / dU10 + 0.142 + dU10 / 13.09);
    double sqrt_Cd_n10 = sqrt(Cd_n10);
    double Ce_n10 = 1E-3 * (34.6 * sqrt_Cd_n10);
    double Ch_n10 = 1E-3 * sqrt_Cd_n10 * (18.0 * stab + 32.7 * (1.0 - stab));
    Cd = Cd_n10;
    Ce = Ce_n10;
    Ch = Ch_n10;
    double sqrt_Cd = sqrt(Cd);
    T_zu = T_zt;
    q_zu = q_zt;
    const double lzu = log(zu / 10.0), lzt = log(zt / zu);
    for (int it = 0; it < 5; ++it) {
        dT = T_zu - sst;
        dq = q_zu - q_sat;
        const double T_vpot = T_zu * (1.0 + 0.608 * q_zu);
        const double U_star = sqrt_Cd * dU10;
        const double T_star = Ch / sqrt_Cd * dT;
        const double q_star = Ce / sqrt_Cd * dq;
        const double Lmo = (U_star * U_star) / (kappa * grav / T_vpot * (T_star * (1.0 + 0.608 * q_zu) + 0.608 * T_zu * q_star));
        double zeta_u = zu / Lmo;
        zeta_u = sbc_sign(fabs(zeta_u) < 10.0 ? fabs(zeta_u) : 10.0, zeta_u);
        double zeta_t = zt / Lmo;
        zeta_t = sbc_sign(fabs(zeta_t) < 10.0 ? fabs(zeta_t) : 10.0, zeta_t);
        const double zpsi_hu = psi_h(zeta_u), zpsi_ht = psi_h(zeta_t), zpsi_m = psi_m(zeta_u);
        double U_n10 = dU10 / (1.0 + sqrt_Cd_n10 / kappa * (lzu - zpsi_m));
        U_n10 = 0.25 > U_n10 ? 0.25 : U_n10;
        T_zu = T_zt - T_star / kappa * (lzt + zpsi_hu - zpsi_ht);
        q_zu = q_zt - q_star / kappa * (lzt + zpsi_hu - zpsi_ht);
        stab = 0.5 + sbc_sign(0.5, q_zu);
        q_zu = stab * q_zu;
        Cd_n10 = 1E-3 * (2.7 / U_n10 + 0.142 + U_n10 / 13.09);
        sqrt_Cd_n10 = sqrt(Cd_n10);
        Ce_n10 = 1E-3 * (34.6 * sqrt_Cd_n10);
        stab = 0.5 + sbc_sign(0.5, zeta_u);
        Ch_n10 = 1E-3 * sqrt_Cd_n10 * (18.0 * stab + 32.7 * (1.0 - stab));
        double xct = 1.0 + sqrt_Cd_n10 / kappa * (lzu - zpsi_m);
        Cd = Cd_n10 / (xct * xct);
        sqrt_Cd = sqrt(Cd);
        const double xlogt = lzu - zpsi_hu;
        xct = 1.0 + Ch_n10 * xlogt / kappa / sqrt_Cd_n10;
        Ch = Ch_n10 * sqrt_Cd / sqrt_Cd_n10 / xct;
        xct = 1.0 + Ce_n10 * xlogt / kappa / sqrt_Cd_n10;
        Ce = Ce_n10 * sqrt_Cd / sqrt_Cd_n10 / xct;
    }
}

// sbc_do per node: the fields at the clock from their time coefficients
// (data_timeinterp, src/fv_sbc.F90:843-867), then apply_atm_fluxes (368-470)
// for task type 3 without sea ice: the wind relative to the top layer's node
// velocity, the surface temperature of the top layer, stress at the node,
// evaporation, sensible, latent and longwave heat, net shortwave, pressure.
// Un-suffixed literals of the reference (air density, albedo, emissivity)
// are default REAL (single precision).
template <typename R>
__global__ __launch_bounds__(256) void k_sbc_node(DevMesh m, DevStateT<R> s, double now_date, double prec_coef,
                                                  int net_flux) {
    const int n = TID();
    if (n >= m.N) return;
    const int KC = m.K;
    const double rhoa = (double)1.3f, cpa = 1005.0, Lv = 2.5e6, Stef = 5.67e-8, albo = (double)0.1f;
    const double emiss_wat = (double)0.97f;
    const double4 *c = (const double4 *)(s.sbc_coef + 16 * (size_t)n);
    const double4 a0 = c[0], a1 = c[1], b0 = c[2], b1 = c[3];
    const double xw = now_date * a0.x + b0.x, yw = now_date * a0.y + b0.y, humi = now_date * a0.z + b0.z;
    const double qsr_in = now_date * a0.w + b0.w, qlw = now_date * a1.x + b1.x, tair = now_date * a1.y + b1.y;
    const double prec = now_date * a1.z + b1.z, mslp = now_date * a1.w + b1.w;
    s.windx[n] = xw;
    s.windy[n] = yw;
    const double wdx = xw - (double)PL(s.Unode, 0, n, KC), wdy = yw - (double)PL(s.Vnode, 0, n, KC);
    const double wndm = sqrt(wdx * wdx + wdy * wdy);
    const double tsur = (double)PL(s.TF, 0, n, KC);
    const double zst = tsur + 273.15;
    const double q_sat = 0.98 * 640380.0 / rhoa * exp(-5107.4 / zst);
    double Cd, Ch, Ce, t_zu, q_zu;
    core_coeff_2z(2.0, 10.0, zst, tair, q_sat, humi, wndm, Cd, Ch, Ce, t_zu, q_zu);
    const double rtmp = rhoa * wndm * Cd;
    s.tau_node[2 * (size_t)n] = rtmp * wdx;
    s.tau_node[2 * (size_t)n + 1] = rtmp * wdy;
    double zevap = rhoa * Ce * (q_sat - q_zu) * wndm;
    zevap = 0.0 > zevap ? 0.0 : zevap;
    const double zqsb = rhoa * cpa * Ch * (zst - t_zu) * wndm;
    const double zqlw = net_flux == 0 ? (qlw - emiss_wat * Stef * zst * zst * zst * zst) : qlw;
    const double zqla = Lv * zevap;
    double qns = zqlw - zqsb - zqla;
    if (tsur < -2.0 && qns < 0.0) qns = 0.0;
    s.qns[n] = qns;
    s.emp[n] = zevap - prec / prec_coef;
    s.qsr[n] = (1.0 - albo) * qsr_in;
    s.mslp[n] = mslp;
}

// the node stress on the elements: sum(w_cv(1:4)*taux_node(elem2D_nodes))
// in order (src/fv_sbc.F90:475-479)
template <typename R>
__global__ __launch_bounds__(256) void k_sbc_elem(DevMesh m, DevStateT<R> s) {
    const int e = TID();
    if (e >= m.E) return;
    const int4 q = m.en[e];
    const double4 w = m.wcv[e];
    const double2 *t = (const double2 *)s.tau_node;
    const double2 t0 = t[q.x], t1 = t[q.y], t2 = t[q.z], t3 = t[q.w];
    double sx = w.x * t0.x, sy = w.x * t0.y;
    sx = sx + w.y * t1.x;
    sy = sy + w.y * t1.y;
    sx = sx + w.z * t2.x;
    sy = sy + w.z * t2.y;
    sx = sx + w.w * t3.x;
    sy = sy + w.w * t3.y;
    s.taux[e] = sx;
    s.tauy[e] = sy;
}

// ---------------------------------------------------------------------------
// open-boundary T/S climatology (key_obc): obc_do's data_timeinterp and
// apply_obc / profile_interp (src/fv_obc.F90:376-419, 531-539, 577-623)
// ---------------------------------------------------------------------------
// binarysearch over the climatology depths (fv_obc.F90:669-714): 1-based
// index of the depth within the REAL(4) tolerance 1e-9 of v, else of the last
// one above it, 0 if none
__device__ __forceinline__ int obc_search(int len, const double *arr, double v) {
    const double d = (double)1e-9f;
    int left = 1, right = len;
    while (left <= right) {
        const int middle = (left + right + 1) >> 1;   // nint((left+right)/2.0) for positive sums
        if (fabs(arr[middle - 1] - v) <= d) return middle;
        if (arr[middle - 1] > v) right = middle - 1;
        else left = middle + 1;
    }
    return right;
}

// One thread per (open-boundary node, layer): the T and S profiles at the
// clock, interpolated linearly to the layer's mid-depth Z = ((1-sigma_k) a +
// (1-sigma_k+1) a)/2 (constant below the last depth), into Tclim/Sclim.  A
// layer above the first depth (no bracketing record) sets the flag err and
// keeps its value.
template <typename R>
__global__ __launch_bounds__(256) void k_obc_apply(DevMesh m, DevStateT<R> s, double now_date, int ndep, int *err) {
    const int t = TID(), L = m.L, KC = m.K, nb = m.nobn;
    if (t >= nb * L) return;
    const int i = t / L, k = t - i * L;
    const int n = s.obc_node[i];
    const double a = s.a[n];
    const double Z = 0.5 * (m.oms[k] * a + m.oms[k + 1] * a);
    const int d = obc_search(ndep, s.obc_depth, Z);
    if (d < 1) {
        atomicOr(err, 1);
        return;
    }
    int dp1;
    double delta_d;
    if (d < ndep) {
        dp1 = d + 1;
        delta_d = s.obc_depth[d] - s.obc_depth[d - 1];
    } else {
        dp1 = d;
        delta_d = 1.0;
    }
    const size_t half = 2 * (size_t)nb * ndep;   // coefficients a, then b
    for (int f = 0; f < 2; ++f) {
        const size_t j1 = f + 2 * ((size_t)i + (size_t)nb * (d - 1)), j2 = f + 2 * ((size_t)i + (size_t)nb * (dp1 - 1));
        const double data1 = now_date * s.obc_coef[j1] + s.obc_coef[half + j1];
        const double data2 = now_date * s.obc_coef[j2] + s.obc_coef[half + j2];
        const double cf_a = (data2 - data1) / delta_d;
        const double cf_b = data1 - cf_a * s.obc_depth[d - 1];
        R *clim = f == 0 ? s.Tclim : s.Sclim;
        PL(clim, k, n, KC) = (R)(cf_a * Z + cf_b);
    }
}

// mask_ad (adv_tracer_mask, fv_main.F90:1309-1335) with H_advtr_crit = 0.1,
// H_advtr_min = 0 (fv_var.F90:49-51)
template <typename R>
__device__ __forceinline__ double mask_ad(const DevMesh &m, const DevStateT<R> &s, int n, double Dmin) {
    const double dm = dmaxr(Dmin, m.depth[n] + s.eta_n[n]);
    if (dm > 0.1) return 1.0;
    if (dm < 0.0) return 0.0;
    return (dm - 0.0) / (0.1 - 0.0);
}

// solve_tracer_miura for T (fv_tracer.F90:930-1316), one thread per node and
// layer; S is not advected under this scheme.  The reconstruction keeps the
// reference's unit metric factors a = b = 1 and its first-segment y term
// without *dt.  The implicit vertical diffusion of T and S follows in
// k_tr_vdiff.
template <typename R>
__global__ __launch_bounds__(256) void k_tr_miura(DevMesh m, DevStateT<R> s, double dt, double Dmin, double scale_area, int cl_relax,
                           double clim_relax, double relax2clim_ac) {
    const int KC = m.K;
    int n, k;
    if (!ent_layer(m.N, m.L, n, k)) return;
    const int N = m.N, E = m.E;
    const double Ta = s.tsa[2 * m.memb[n]];
    const double a = s.a[n];
    const int p0 = m.n2e_ptr[n], p1 = m.n2e_ptr[n + 1];
    const double ab = 1.0;
    double rt = 0.0, rc = 0.0;
    for (int j = p0; j < p1; ++j) {
        const int cc = m.n2e[j];
        const int ed = cc >= 0 ? cc : ~cc;
        const int2 nd = m.ednodes[ed];
        const int2 et = m.edtri[ed];
        const double4 cd = m.ecross[ed];
        const double2 dd = m.edxdy[ed];
        const int e1 = et.x, e2 = et.y;
        const double crt1 = 0.5 * (mask_ad(m, s, nd.x, Dmin) + mask_ad(m, s, nd.y, Dmin));
        double Kh = m.elem_area[e1];
        if (e2 >= 0) Kh = 0.5 * (Kh + m.elem_area[e2]);
        Kh = 0.010 * Kh / (m.memb_sa ? m.memb_sa[m.memb[n]] : scale_area);  // the node's member's scale_area
        const double T1 = PL(s.tcur, k, nd.x, KC), T2 = PL(s.tcur, k, nd.y, KC);
        const double tx1 = PL(s.gTx, k, nd.x, KC), tx2 = PL(s.gTx, k, nd.y, KC);
        const double ty1 = PL(s.gTy, k, nd.x, KC), ty2 = PL(s.gTy, k, nd.y, KC);
        const double2 g1 = elem_grad(m, s.tcur, 0.0, e1, k);
        double Tx, Ty;
        double2 g2 = make_double2(0.0, 0.0);
        if (e2 >= 0) {
            g2 = elem_grad(m, s.tcur, 0.0, e2, k);
            Tx = 0.5 * Kh * (g1.x + g2.x);
            Ty = 0.5 * Kh * (g1.y + g2.y);
        } else {
            Tx = Kh * g1.x;
            Ty = Kh * g1.y;
        }
        const double dX1 = cd.x, dY1 = cd.y;
        const double u = PL(s.U_nf, k, e1, KC), v = PL(s.V_nf, k, e1, KC), J1 = PL(s.Je, k, e1, KC);
        double Tm;
        if (v * dX1 - u * dY1 > 0) {
            Tm = T2 + (0.5 * (-dd.x * ab + dX1 - u * dt) * tx2 + 0.5 * (-dd.y * kREarthD + dY1 - v) * ty2) * crt1;
        } else {
            Tm = T1 + (0.5 * (dd.x * ab + dX1 - u * dt) * tx1 + 0.5 * (dd.y * kREarthD + dY1 - v * dt) * ty1) * crt1;
        }
        const double c1 = (J1 * v * Tm - Ty) * dX1 - (J1 * u * Tm - Tx) * dY1;
        const double cV = J1 * v * dX1 - J1 * u * dY1;
        rt = (cc >= 0) ? rt + c1 : rt - c1;
        if (e2 >= 0) {
            const double dX2 = cd.z, dY2 = cd.w;
            const double Tx2 = 0.5 * Kh * (g1.x + g2.x), Ty2 = 0.5 * Kh * (g1.y + g2.y);
            const double u2 = PL(s.U_nf, k, e2, KC), v2 = PL(s.V_nf, k, e2, KC), J2 = PL(s.Je, k, e2, KC);
            double Tn;
            if (v2 * dX2 - u2 * dY2 < 0) {
                Tn = T2 + (0.5 * (-dd.x * ab + dX2 - u2 * dt) * tx2 + 0.5 * (-dd.y * kREarthD + dY2 - v2 * dt) * ty2) *
                              crt1;
            } else {
                Tn = T1 + (0.5 * (dd.x * ab + dX2 - u2 * dt) * tx1 + 0.5 * (dd.y * kREarthD + dY2 - v2 * dt) * ty1) *
                              crt1;
            }
            const double c2 = -(J2 * v2 * Tn - Ty2) * dX2 + (J2 * u2 * Tn - Tx2) * dY2;
            const double c2V = -J2 * v2 * dX2 + J2 * u2 * dY2;
            if (cc >= 0) {
                rt = rt + c2;
                rc = rc + cV;
                rc = rc + c2V;
            } else {
                rt = rt - c2;
                rc = rc - cV;
                rc = rc - c2V;
            }
        } else {
            rc = (cc >= 0) ? rc + cV : rc - cV;
        }
    }
    double Tmv = 0.0;
    const double tvt = hi_vert<false>(m, s, s.tcur, 0.0, n, k + 1, a, Tmv);
    const double tvb = hi_vert<false>(m, s, s.tcur, 0.0, n, k + 2, a, Tmv);
    const double tab = PL(s.abT, k, n, KC);
    const double wdiff = PL(s.W, k, n, KC) - PL(s.W, k + 1, n, KC);
    const double jc = a * m.dsig[k];
    rt = ((rt - tab * rc) / m.area[n] + tab * wdiff + tvt - tvb) / jc;
    const double tdm = PL(s.tcur, k, n, KC);
    if (cl_relax && m.index[n] == 2) {
        if (ob_inflow(m, s, p0, p1, Dmin) > 0.0) rt = rt + relax2clim_ac * (PL(s.Tclim, k, n, KC) - (tdm + Ta));
        else rt = clim_relax * (PL(s.Tclim, k, n, KC) - (tdm + Ta));
    }
    PL(s.T_old, k, n, KC) = tdm + Ta;
    PL(s.TF, k, n, KC) = tdm + rt * dt + Ta;
}

// tracer_impl_vert_diff of T and S (fv_tracer.F90:1319-1494) on tiles of 64
// node columns, after every advection scheme, and alone for tracer_adv = 2
// without MUSCL_CALL (fv_main.F90:1788-1800).
template <typename R, bool ATM>
__device__ __forceinline__ void tr_vdiff_body(DevMesh m, DevStateT<R> s, double dt, AtmPar ap) {
    extern __shared__ __align__(16) unsigned char lds_[];
    const int K = m.K, L = m.L, KP = K + 1;
    R *tT = (R *)lds_, *tS = tT + kColB * KP, *tK = tS + kColB * KP, *tTp = tK + kColB * KP, *tSp = tTp + kColB * KP;
    const int c0 = col_first(), nc = min(kColB, m.N - c0);
    if (nc <= 0) return;
    tile_load(tT, s.TF, c0, nc, K);
    tile_load(tS, s.SF, c0, nc, K);
    tile_load(tK, s.Kv, c0, nc, K);
    __syncthreads();
    if ((int)threadIdx.x < nc) {
        const int c = threadIdx.x, n = c0 + c;
        const int mb = m.memb[n];
        vdiff_column<R, ATM>(m, s, tT + c * KP, tS + c * KP, tK + c * KP, tTp + c * KP, tSp + c * KP, s.a[n],
                             s.tsa[2 * mb], s.tsa[2 * mb + 1], dt, ap, n);
    }
    __syncthreads();
    tile_store(s.TF, tT, c0, nc, K, L);
    tile_store(s.SF, tS, c0, nc, K, L);
}
template <typename R>
__global__ __launch_bounds__(kColT) void k_tr_vdiff(DevMesh m, DevStateT<R> s, double dt, AtmPar ap) {
    tr_vdiff_body<R, false>(m, s, dt, ap);
}
// the same with surface forcing (key_atm)
template <typename R>
__global__ __launch_bounds__(kColT) void k_tr_vdiff_atm(DevMesh m, DevStateT<R> s, double dt, AtmPar ap) {
    tr_vdiff_body<R, true>(m, s, dt, ap);
}

// ---------------------------------------------------------------------------
// learned closure: features, training samples, fused inference
// ---------------------------------------------------------------------------
// Features of one (node, level 1..K) sample, from the quantities d3_end uses
// (src/fv_mixing.F90:111-194): log10 shear^2 and the buoyancy frequency N^2
// (asinh-compressed) at the level's interface (the nearest interior one at
// the surface and bottom), log10 of the column depth, log10 of d3_end's
// bottom boundary value bott_R*|u_b|^2, and the sigma level.
template <typename R>
__device__ void column_features(const DevMesh &m, const DevStateT<R> &s, int n, int nz, double Dmin, double rho0,
                                double z0b_min, float *f) {
    const int KC = m.K;
    const int N = m.N, K = m.K, L = m.L;
    const double a = s.a[n];
    const double dmean = dmaxr(Dmin, m.depth[n] + s.eta_n[n]);
    auto uv = [&](int k, double &u, double &v) {  // node averages of compute_el2nodes_3D
        u = PL(s.Unode, k, n, KC);
        v = PL(s.Vnode, k, n, KC);
    };
    const int iz = nz < 2 ? 2 : (nz > K - 1 ? K - 1 : nz);  // 1-based interface
    double u1, v1, u2, v2, ub, vb;
    uv(iz - 2, u1, v1);
    uv(iz - 1, u2, v2);
    uv(L - 1, ub, vb);
    const double d1 = m.sigma[iz - 2] - m.sigma[iz - 1], d2 = m.sigma[iz - 1] - m.sigma[iz];
    const double d = 0.5 * (d1 + d2);
    const double sh2 = ((u1 - u2) * (u1 - u2) + (v1 - v2) * (v1 - v2)) / (d * d) / (dmean * dmean);
    const double zb = m.oms[iz - 1] * a;
    const double rho_up = densityJM(PL(s.TF, iz - 2, n, KC), PL(s.SF, iz - 2, n, KC), -zb, rho0);
    const double rho_dn = densityJM(PL(s.TF, iz - 1, n, KC), PL(s.SF, iz - 1, n, KC), -zb, rho0);
    const double n2 = -kG * (rho_up - rho_dn) / d / (rho0 * dmean);
    const double delta_b = a * m.dsig[L - 1] + z0b_min;
    const double lg = log(delta_b / z0b_min);
    const double bott = 16.6 * (kCka * kCka) / (lg * lg) * (ub * ub + vb * vb);
    // the transcendental feature maps in FP32: the network sees BF16 inputs
    f[0] = log10f((float)(sh2 + 1e-12));
    f[1] = asinhf((float)(n2 * 1e5));
    f[2] = log10f((float)dmean);
    f[3] = log10f((float)(bott + 1e-12));
    f[4] = (float)m.sigma[nz - 1];
}

// Training samples from the current state, taken right after d3_end: raw
// features X[i][5] and targets T[i][2] = log10(snu), log10(Kv) for every
// level of the columns n = 0, stride, 2*stride, ... (i = (n/stride)*K + nz-1).
template <typename R>
__global__ __launch_bounds__(256) void k_mlp_collect(DevMesh m, DevStateT<R> s, double Dmin, double rho0,
                                                     double z0b_min, int stride, long cap, float *X, float *T) {
    const int KC = m.K;
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int K = m.K;
    const long col = i / K;
    const int nz = (int)(i % K) + 1;
    const long n = col * stride;
    if (n >= m.N || i >= cap) return;
    float f[kMlpF];
    column_features(m, s, (int)n, nz, Dmin, rho0, z0b_min, f);
    for (int j = 0; j < kMlpF; ++j) X[i * kMlpF + j] = f[j];
    T[i * 2] = (float)log10((double)PL(s.snu, nz - 1, n, KC));
    T[i * 2 + 1] = (float)log10((double)PL(s.Kv, nz - 1, n, KC));
}

// The closure step.  Persistent blocks of kMlpT threads, one per compute
// unit: the network's forward fragments and biases go to LDS once per block.
// Each wavefront then walks its own contiguous range of column groups
// (64/K whole node columns, one lane per (column, level 1..K) sample), with
// no block-wide barrier, so one wavefront's matrix products overlap another's
// feature arithmetic.  Per column the depth factors and the bottom-stress
// feature are formed once; per level the density coefficients and velocity;
// per sample the five features (as column_features), normalised and checked
// against the training box; the wavefront then runs its 64 samples through
// the network as four 16-sample tiles on the matrix cores.  snu and Kv =
// 10^(de-normalised outputs), snu held in [snul, 0.2] as d3_end holds it
// (fv_mixing.F90:282-285), go straight to the column slots, for every column
// whose samples all lie in the box; any other column is flagged and left to
// d3_end.
struct MlpWaveLds {
    float xs[64 * kMlpF];
    float ys[64 * 2];
    EosLevel lv[64];
    double2 uv[64];
    double col_a[64 / kMlpMinK], col_dm[64 / kMlpMinK], col_f3[64 / kMlpMinK];
    double col_idm2[64 / kMlpMinK], col_irdm[64 / kMlpMinK];   // 1/dmean^2, 1/(rho0*dmean)
    int col_ood[64 / kMlpMinK];
};
static_assert(sizeof(MlpWaveLds) == kMlpWaveLds, "closure LDS record");

// LDS writes of this wavefront visible to its other lanes, and no code motion
// of its memory operations across this point
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <typename R>
__global__ __launch_bounds__(kMlpT) void k_mlp_closure(DevMesh m, DevStateT<R> s, const fcgm::bf16x8 *__restrict__ frag,
                                                       const float *__restrict__ bias, MlpNorm nm, double Dmin,
                                                       double rho0, double z0b_min, double snul, int *flag) {
    extern __shared__ __align__(16) unsigned char lds_[];
    fcgm::bf16x8 *fl = (fcgm::bf16x8 *)lds_;
    float *bl = (float *)(fl + kFragFwd * 64);
    double *sg = (double *)(bl + kMlpBias), *om = sg + kMlpMaxK;   // sigma and 1-sigma of the K levels
    double *rd = om + kMlpMaxK, *rd2 = rd + kMlpMaxK;   // 1/d, 1/d^2 of interface iz at iz-1
    MlpWaveLds &w = ((MlpWaveLds *)(rd2 + kMlpMaxK))[threadIdx.x >> 6];
    const int K = m.K, L = m.L, KC = K;
    const int tid = threadIdx.x, lane = tid & 63;
    for (int i = tid; i < kFragFwd * 64; i += kMlpT) fl[i] = frag[i];
    for (int i = tid; i < kMlpBias; i += kMlpT) bl[i] = bias[i];
    for (int i = tid; i < K; i += kMlpT) {
        sg[i] = m.sigma[i];
        om[i] = m.oms[i];
        if (i >= 1 && i <= K - 2) {  // interface iz = i + 1 between layers iz-1 and iz
            const double d1 = m.sigma[i - 1] - m.sigma[i], d2 = m.sigma[i] - m.sigma[i + 1];
            const double d = 0.5 * (d1 + d2);
            rd[i] = 1.0 / d;
            rd2[i] = 1.0 / (d * d);
        }
    }
    const double dsb = m.dsig[L - 1];
    // normalisation by multiplication
    float xis[kMlpF];
    for (int j = 0; j < kMlpF; ++j) xis[j] = 1.0f / nm.x_std[j];
    __syncthreads();
    const int cpw = 64 / K;  // whole columns per wavefront
    const int ngroups = (m.N + cpw - 1) / cpw;
    const int nwave = (int)gridDim.x * (kMlpT / 64);
    const int per = (ngroups + nwave - 1) / nwave;
    const int g0 = ((int)xcd_block() * (kMlpT / 64) + (tid >> 6)) * per, g1 = min(ngroups, g0 + per);
    const int c = lane / K, k = lane - (lane / K) * K;
    // 16-sample tiles that hold samples: 64/K columns of K levels
    const int ntile = (cpw * K + 15) / 16;
    // HBM operands of a group, loaded one group ahead so that their latency
    // hides behind the previous group's matrix products
    struct Pre {
        R t, sa, u, v;                   // this lane's level
        double a, dep, eta, ub, vb, ac;  // column lane (lane < columns of the group)
    };
    // (every lane loads, from a clamped in-range slot where it has no sample,
    // so the loads stay unconditional and nothing waits for them before use)
    auto fetch = [&](int g) {
        Pre q;
        const int c0 = g * cpw;
        const int n = min(c0 + min(c, cpw - 1), m.N - 1), kl = min(k, L - 1);
        q.t = PL(s.TF, kl, n, KC);
        q.sa = PL(s.SF, kl, n, KC);
        q.u = PL(s.Unode, kl, n, KC);
        q.v = PL(s.Vnode, kl, n, KC);
        const int nn = min(c0 + min(lane, cpw - 1), m.N - 1);
        q.a = s.a[nn];
        q.dep = m.depth[nn];
        q.eta = s.eta_n[nn];
        q.ub = PL(s.Unode, L - 1, nn, KC);
        q.vb = PL(s.Vnode, L - 1, nn, KC);
        q.ac = m.ac[nn];
        return q;
    };
    Pre nxt = fetch(min(g0, ngroups - 1));
    for (int g = g0; g < g1; ++g) {
        wave_sync();  // the previous group's LDS reads are done
        const Pre cur = nxt;
        const int c0 = g * cpw, nc = min(cpw, m.N - c0);
        const bool on = c < nc;
        const int n = c0 + c, nz = k + 1;
        if (lane < nc) {
            const double a = cur.a;
            const double dmean = dmaxr(Dmin, cur.dep + cur.eta);
            const double ub = cur.ub, vb = cur.vb;
            const double delta_b = a * dsb + z0b_min;
            const double lg = log(delta_b / z0b_min);
            w.col_a[lane] = a;
            w.col_dm[lane] = dmean;
            w.col_idm2[lane] = 1.0 / (dmean * dmean);
            w.col_irdm[lane] = 1.0 / (rho0 * dmean);
            w.col_f3[lane] = log10f((float)(16.6 * (kCka * kCka) / (lg * lg) * (ub * ub + vb * vb) + 1e-12));
            // columns of the open-boundary sponge layer (ac < 1), whose damped
            // dynamics the network is not trained for, keep d3_end
            w.col_ood[lane] = cur.ac < 1.0 ? 1 : 0;
        }
        if (on && k < L) {  // each level's T, S and velocity once
            w.lv[lane] = eos_level(cur.t, cur.sa);
            w.uv[lane] = make_double2(cur.u, cur.v);
        }
        nxt = fetch(min(g + 1, g1 - 1));
        wave_sync();
        // a group whose columns all lie in the sponge layer skips the network
        if (__all(!on || w.col_ood[c] != 0)) {
            if (on && k == 0) flag[n] = 1;
            continue;
        }
        float fx[kMlpF] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
        if (on) {
            // the features of column_features, with its divisions by the
            // layer spacing and the column depth as multiplications by
            // reciprocals formed once per level or column, and the two
            // interface densities' difference over one common denominator
            const double a = w.col_a[c], dmean = w.col_dm[c];
            const int iz = nz < 2 ? 2 : (nz > K - 1 ? K - 1 : nz);  // 1-based interface
            const int up = c * K + iz - 2, dn = up + 1;                 // its two layers
            const double2 w1 = w.uv[up], w2 = w.uv[dn];
            const double sh2 = ((w1.x - w2.x) * (w1.x - w2.x) + (w1.y - w2.y) * (w1.y - w2.y)) * rd2[iz - 1] *
                               w.col_idm2[c];
            const double pz = -(om[iz - 1] * a);
            const EosLevel eu = w.lv[up], ed = w.lv[dn];
            const double bu = eu.c0 - pz * eu.c1 + pz * pz * eu.c2, bd = ed.c0 - pz * ed.c1 + pz * pz * ed.c2;
            const double hu = bu + 0.1 * pz, hd = bd + 0.1 * pz;
            const double drho = (eu.rp * bu * hd - ed.rp * bd * hu) / (hu * hd);   // rho_up - rho_dn
            const double n2 = -kG * drho * rd[iz - 1] * w.col_irdm[c];
            fx[0] = log10f((float)(sh2 + 1e-12));
            fx[1] = asinhf((float)(n2 * 1e5));
            fx[2] = log10f((float)dmean);
            fx[3] = (float)w.col_f3[c];
            fx[4] = (float)sg[nz - 1];
            bool ood = false;
            for (int j = 0; j < kMlpF; ++j) {
                fx[j] = (fx[j] - nm.x_mean[j]) * xis[j];
                ood = ood || !(fx[j] >= nm.x_lo[j] && fx[j] <= nm.x_hi[j]);
            }
            if (ood) atomicOr(&w.col_ood[c], 1);
        }
        for (int j = 0; j < kMlpF; ++j) w.xs[lane * kMlpF + j] = fx[j];
        wave_sync();
        if (ntile == 4) {  // the wavefront's samples as 16-sample tiles
            fcgm::bf16x8 x[4];
            for (int t = 0; t < 4; ++t) x[t] = fcgm::load_bx(&w.xs[16 * t * kMlpF], lane);
            fcgm::f32x4 y[4];
            fcgm::infer_tiles<4>(fl, bl, x, lane, y);
            if (lane < 16)
                for (int t = 0; t < 4; ++t) {
                    w.ys[(16 * t + lane) * 2] = y[t][0];
                    w.ys[(16 * t + lane) * 2 + 1] = y[t][1];
                }
        } else {  // at most 48 samples (24 levels: two columns)
            fcgm::bf16x8 x[3];
            for (int t = 0; t < 3; ++t) x[t] = fcgm::load_bx(&w.xs[16 * t * kMlpF], lane);
            fcgm::f32x4 y[3];
            fcgm::infer_tiles<3>(fl, bl, x, lane, y);
            if (lane < 16)
                for (int t = 0; t < 3; ++t) {
                    w.ys[(16 * t + lane) * 2] = y[t][0];
                    w.ys[(16 * t + lane) * 2 + 1] = y[t][1];
                }
        }
        wave_sync();
        if (!on) continue;
        const bool ood = w.col_ood[c] != 0;
        if (k == 0) flag[n] = ood ? 1 : 0;
        if (ood) continue;
        const float ls = w.ys[lane * 2] * nm.y_std[0] + nm.y_mean[0];
        const float lk = w.ys[lane * 2 + 1] * nm.y_std[1] + nm.y_mean[1];
        double sv = (double)exp10f(ls);
        sv = sv < snul ? snul : (sv > 0.2 ? 0.2 : sv);
        PL(s.snu, nz - 1, n, KC) = (R)sv;
        PL(s.Kv, nz - 1, n, KC) = (R)exp10f(lk);
    }
}

#if FCG_STORAGE_DOUBLE
__global__ __launch_bounds__(256) void k_count_flags(const int *flag, int n, unsigned long long *count) {
    unsigned long long c = 0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) c += flag[i] != 0;
    for (int off = 32; off > 0; off >>= 1) c += __shfl_down(c, off);
    if ((threadIdx.x & 63) == 0 && c) atomicAdd(count, c);
}
#endif

// ---------------------------------------------------------------------------
// diagnostics: energy (fv_average_dynamic.F90:476-554), volume, tracer
// content, extrema -- fixed-order block tree, deterministic run to run
// ---------------------------------------------------------------------------
constexpr int kDiag = 12;
template <typename R>
__global__ __launch_bounds__(256) void k_diag_partial(DevMesh m, DevStateT<R> s, double Dmin, int n0, int n1, int e0,
                                                      int e1, double *out) {
    const int KC = m.K;
    __shared__ double sm[kDiag][256];
    double v[kDiag] = {0, 0, 0, 0, 0, 0, INFINITY, -INFINITY, INFINITY, -INFINITY, INFINITY, -INFINITY};
    // v: 0 PE, 1 area, 2 KE, 3 volume, 4 heat, 5 salt, 6/7 eta, 8/9 T, 10/11 S over nodes [n0, n1),
    // elements [e0, e1) (one ensemble member, or all)
    const int N = m.N, L = m.L;
    const int stride = gridDim.x * blockDim.x;
    for (int n = n0 + blockIdx.x * blockDim.x + threadIdx.x; n < n1; n += stride) {
        double ar = m.area[n], et = s.eta_n[n];
        v[0] += 0.5 * kG * ar * (et * et);
        v[1] += ar;
        v[3] += ar * (m.depth[n] + et);
        v[6] = fmin(v[6], et);
        v[7] = fmax(v[7], et);
        double a = s.a[n];
        for (int k = 0; k < L; ++k) {
            double jc = a * m.dsig[k];
            double t = PL(s.TF, k, n, KC), sv = PL(s.SF, k, n, KC);
            v[4] += ar * jc * t;
            v[5] += ar * jc * sv;
            v[8] = fmin(v[8], t);
            v[9] = fmax(v[9], t);
            v[10] = fmin(v[10], sv);
            v[11] = fmax(v[11], sv);
        }
    }
    for (int e = e0 + blockIdx.x * blockDim.x + threadIdx.x; e < e1; e += stride) {
        const int4 q = m.en[e];
        const double4 w = m.wcv[e];
        double f0 = dmaxr(Dmin, m.depth[q.x] + s.eta_n[q.x]), f1 = dmaxr(Dmin, m.depth[q.y] + s.eta_n[q.y]);
        double f2 = dmaxr(Dmin, m.depth[q.z] + s.eta_n[q.z]), f3 = dmaxr(Dmin, m.depth[q.w] + s.eta_n[q.w]);
        double u = s.U2[2 * e], vv = s.U2[2 * e + 1];
        v[2] += 0.5 * sum4(w, f0, f1, f2, f3) * m.elem_area[e] * (u * u + vv * vv);
    }
    for (int c = 0; c < kDiag; ++c) sm[c][threadIdx.x] = v[c];
    __syncthreads();
    for (int st = blockDim.x / 2; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) {
            for (int c = 0; c < 6; ++c) sm[c][threadIdx.x] += sm[c][threadIdx.x + st];
            for (int c = 6; c < kDiag; c += 2) {
                sm[c][threadIdx.x] = fmin(sm[c][threadIdx.x], sm[c][threadIdx.x + st]);
                sm[c + 1][threadIdx.x] = fmax(sm[c + 1][threadIdx.x], sm[c + 1][threadIdx.x + st]);
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int c = 0; c < kDiag; ++c) out[kDiag * blockIdx.x + c] = sm[c][0];
}

#if FCG_STORAGE_DOUBLE
__global__ __launch_bounds__(256) void k_diag_final(const double *part, int nb, double *res) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double v[kDiag];
    for (int c = 0; c < kDiag; ++c) v[c] = part[c];
    for (int b = 1; b < nb; ++b) {
        for (int c = 0; c < 6; ++c) v[c] += part[kDiag * b + c];
        for (int c = 6; c < kDiag; c += 2) {
            v[c] = fmin(v[c], part[kDiag * b + c]);
            v[c + 1] = fmax(v[c + 1], part[kDiag * b + c + 1]);
        }
    }
    for (int c = 0; c < kDiag; ++c) res[c] = v[c];
}
#endif

#if FCG_STORAGE_DOUBLE
__global__ __launch_bounds__(256) void k_mask_count(const double *mask, int E, int *out) {
    __shared__ int sm[256];
    int c = 0;
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < E; e += gridDim.x * blockDim.x) c += mask[e] != 0.0;
    sm[threadIdx.x] = c;
    __syncthreads();
    for (int st = blockDim.x / 2; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) sm[threadIdx.x] += sm[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(out, sm[0]);
}
#endif


// The state kernels of this build's storage type (except the learned
// closure's, below); the C-ABI unit launches them through their declarations.
template <typename R>
void instantiate_kernels() {
    (void)&k_pressure<R>;
    (void)&k_pressure_atm<R>;
    (void)&k_bpg<R>;
    (void)&k_d3end_prep<R>;
    (void)&k_d3end<R>;
    (void)&k_d3end_fin<R>;
    (void)&k_velgrad3d<R>;
    (void)&k_velrhs<R>;
    (void)&k_velrhs_atm<R>;
    (void)&k_velvisc<R>;
    (void)&k_velvisc_im<R>;
    (void)&k_bt_elem1<R>;
    (void)&k_bt_node<R>;
    (void)&k_bt_edge_ssh<R>;
    (void)&k_bt_edge_mom<R>;
    (void)&k_bt_velgrad<R>;
    (void)&k_bt_puls<R>;
    (void)&k_el2nodes<R, false>;
    (void)&k_el2nodes<R, true>;
    (void)&k_bt_p1edge<R>;
    (void)&k_bt_p1elem<R>;
    (void)&k_bt_elem2<R>;
    (void)&k_filt2x_3d_c<R>;
    (void)&k_filt2x_3d_uv<R>;
    (void)&k_bt_filt2x_c<R>;
    (void)&k_bh3_edge<R>;
    (void)&k_bh3_elem<R>;
    (void)&k_bh3_grad<R>;
    (void)&k_bh_puls<R>;
    (void)&k_bh2_edge<R>;
    (void)&k_bh2_elem<R>;
    (void)&k_jac_node<R>;
    (void)&k_jac_restart<R>;
    (void)&k_je<R>;
    (void)&k_update3d<R>;
    (void)&k_vertvel_div<R>;
    (void)&k_vertvel_scan<R>;
    (void)&k_minmax_partial<R>;
    (void)&k_tr_ab<R>;
    (void)&k_velrhs_filt<R>;
    (void)&k_velrhs_atm_filt<R>;
    (void)&k_velrhs_v<R>;
    (void)&k_velrhs_v_atm<R>;
    (void)&k_velrhs_v_filt<R>;
    (void)&k_velrhs_v_atm_filt<R>;
    (void)&k_tr_rivers<R>;
    (void)&k_tr_upwind<R>;
    (void)&k_tr_gradn_muscl<R>;
    (void)&k_tr_gradn_miura<R>;
    (void)&k_tr_muscl<R>;
    (void)&k_tr_muscl_v<R>;
    (void)&k_el2nodes_v<R, false>;
    (void)&k_el2nodes_v<R, true>;
    (void)&k_vertvel_div_v<R>;
    (void)&k_tr_gradn_muscl_v<R>;
    (void)&k_velgrad3d_v<R>;
    (void)&k_sbc_node<R>;
    (void)&k_sbc_elem<R>;
    (void)&k_obc_apply<R>;
    (void)&k_je_v<R>;
    (void)&k_update3d_v<R>;
    (void)&k_tr_miura<R>;
    (void)&k_tr_vdiff<R>;
    (void)&k_tr_vdiff_atm<R>;
    (void)&k_diag_partial<R>;
}
// The learned closure's kernels of both storage types are built with the FP64
// build's options (no contraction): with fused multiply-adds the inference
// kernel's feature code needs more than the 256 registers a lane has.
template <typename R>
void instantiate_closure_kernels() {
    (void)&k_mlp_collect<R>;
    (void)&k_mlp_closure<R>;
}
#if FCG_STORAGE_DOUBLE
template void instantiate_kernels<double>();
template void instantiate_closure_kernels<double>();
template void instantiate_closure_kernels<float>();
#else
template void instantiate_kernels<float>();
#endif

}  // namespace fcgk
