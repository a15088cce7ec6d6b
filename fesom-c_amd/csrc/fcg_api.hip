// C-ABI of libfesomc_gpu: context set-up, state transfer and the per-call-site
// entry points of oce_timestep (src/fv_main.F90:1877-2106).  See fesomc_gpu.h.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <unordered_map>
#include <vector>

#include "fcg_internal.h"
#include "fcg_mlp.h"

namespace fcgk {
template <typename R> __global__ void k_pressure(DevMesh, DevStateT<R>, double, double);
template <typename R> __global__ void k_pressure_atm(DevMesh, DevStateT<R>, double, double);
template <typename R> __global__ void k_bpg(DevMesh, DevStateT<R>);
template <typename R> __global__ void k_d3end_prep(DevMesh, DevStateT<R>, double, double, double, const int *);
template <typename R> __global__ void k_d3end(DevMesh, DevStateT<R>, double, double, double, const int *);
template <typename R> __global__ void k_d3end_fin(DevMesh, DevStateT<R>, double, double, double, double, const int *);
template <typename R> __global__ void k_velgrad3d(DevMesh, DevStateT<R>);
template <typename R> __global__ void k_velgrad3d_v(DevMesh, DevStateT<R>);
template <typename R> __global__ void k_sbc_node(DevMesh, DevStateT<R>, double, double, int);
template <typename R> __global__ void k_sbc_elem(DevMesh, DevStateT<R>);
template <typename R> __global__ void k_obc_apply(DevMesh, DevStateT<R>, double, int, int *);
template <typename R> __global__ void k_je_v(DevMesh, DevStateT<R>);
template <typename R> __global__ void k_update3d_v(DevMesh, DevStateT<R>, double, double);
template <typename R> __global__ void k_vertvel_div_v(DevMesh, DevStateT<R>, int);
template <typename R> __global__ void k_tr_gradn_muscl_v(DevMesh, DevStateT<R>, int);
template <typename R, bool P>
__global__ void k_el2nodes_v(DevMesh, const R *, const R *, R *, R *, const double *, const double *, int, int *);
template <typename R> __global__ void k_velrhs(DevMesh, DevStateT<R>, double, double, int, int, double);
template <typename R> __global__ void k_velrhs_atm(DevMesh, DevStateT<R>, double, double, int, int, double);
template <typename R> __global__ void k_velrhs_filt(DevMesh, DevStateT<R>, double, double, int, int, double, double);
template <typename R>
__global__ void k_velrhs_atm_filt(DevMesh, DevStateT<R>, double, double, int, int, double, double);
template <typename R> __global__ void k_velrhs_v(DevMesh, DevStateT<R>, double, double, int, int, double);
template <typename R> __global__ void k_velrhs_v_atm(DevMesh, DevStateT<R>, double, double, int, int, double);
template <typename R> __global__ void k_velrhs_v_filt(DevMesh, DevStateT<R>, double, double, int, int, double, double);
template <typename R>
__global__ void k_velrhs_v_atm_filt(DevMesh, DevStateT<R>, double, double, int, int, double, double);
template <typename R> __global__ void k_velvisc(DevMesh, DevStateT<R>, double, double, double, int);
template <typename R> __global__ void k_velvisc_im(DevMesh, DevStateT<R>, double, double, double);
template <typename R> __global__ void k_bt_elem1(DevMesh, DevStateT<R>, double);
template <typename R> __global__ void k_bt_node(DevMesh, DevStateT<R>, double, const double *, int, double,
                                                const double *, double);
template <typename R> __global__ void k_bt_edge_ssh(DevMesh, DevStateT<R>);
template <typename R> __global__ void k_bt_edge_mom(DevMesh, DevStateT<R>, double);
template <typename R> __global__ void k_bt_velgrad(DevMesh, DevStateT<R>);
template <typename R> __global__ void k_bt_puls(DevMesh, DevStateT<R>, double);
template <typename R, bool P>
__global__ void k_el2nodes(DevMesh, const R *, const R *, R *, R *, const double *, const double *);
template <typename R> __global__ void k_bt_p1edge(DevMesh, DevStateT<R>);
template <typename R> __global__ void k_bt_p1elem(DevMesh, DevStateT<R>, double);
template <typename R>
__global__ void k_bt_elem2(DevMesh, DevStateT<R>, double, double, double, double, double, int, int, int, double, R *);
template <typename R> __global__ void k_filt2x_3d_c(DevMesh, DevStateT<R>, double);
template <typename R> __global__ void k_filt2x_3d_uv(DevMesh, DevStateT<R>);
template <typename R> __global__ void k_bt_filt2x_c(DevMesh, DevStateT<R>, double);
template <typename R>
__global__ void k_bh3_edge(DevMesh, DevStateT<R>, const R *, const R *, const R *, const R *, const R *, const R *, int,
                           int);
template <typename R> __global__ void k_bh3_elem(DevMesh, DevStateT<R>, R *, R *, int);
template <typename R> __global__ void k_bh3_grad(DevMesh, DevStateT<R>, const R *, const R *, int);
template <typename R> __global__ void k_bh_puls(DevMesh, DevStateT<R>);
template <typename R> __global__ void k_bh2_edge(DevMesh, DevStateT<R>, const R *, const R *);
template <typename R> __global__ void k_bh2_elem(DevMesh, DevStateT<R>);
template <typename R> __global__ void k_jac_node(DevMesh, DevStateT<R>, double);
template <typename R> __global__ void k_je(DevMesh, DevStateT<R>);
template <typename R> __global__ void k_jac_restart(DevMesh, DevStateT<R>, double);
template <typename R> __global__ void k_update3d(DevMesh, DevStateT<R>, double, double);
template <typename R> __global__ void k_vertvel_div(DevMesh, DevStateT<R>);
template <typename R> __global__ void k_vertvel_scan(DevMesh, DevStateT<R>);
template <typename R> __global__ void k_minmax_partial(DevMesh, const R *, const R *, double *);
__global__ void k_minmax_final(const double *, int, int, double *);
template <typename R> __global__ void k_tr_ab(DevMesh, DevStateT<R>, size_t);
template <typename R> __global__ void k_tr_upwind(DevMesh, DevStateT<R>, double, double, int, double, double);
template <typename R> __global__ void k_tr_gradn_muscl(DevMesh, DevStateT<R>);
template <typename R> __global__ void k_tr_muscl(DevMesh, DevStateT<R>, double, double, int, double, double);
template <typename R> __global__ void k_tr_muscl_v(DevMesh, DevStateT<R>, double, double, int, double, double, int);
template <typename R> __global__ void k_tr_gradn_miura(DevMesh, DevStateT<R>);
template <typename R> __global__ void k_tr_miura(DevMesh, DevStateT<R>, double, double, double, int, double, double);
template <typename R> __global__ void k_tr_rivers(DevMesh, DevStateT<R>, double);
template <typename R> __global__ void k_tr_vdiff(DevMesh, DevStateT<R>, double, AtmPar);
template <typename R> __global__ void k_tr_vdiff_atm(DevMesh, DevStateT<R>, double, AtmPar);
template <typename R> __global__ void k_diag_partial(DevMesh, DevStateT<R>, double, int, int, int, int, double *);
__global__ void k_diag_final(const double *, int, double *);
__global__ void k_mask_count(const double *, int, int *);
template <typename R>
__global__ void k_mlp_collect(DevMesh, DevStateT<R>, double, double, double, int, long, float *, float *);
template <typename R>
__global__ void k_mlp_closure(DevMesh, DevStateT<R>, const fcgm::bf16x8 *, const float *, MlpNorm, double, double,
                              double, double, int *);
__global__ void k_count_flags(const int *, int, unsigned long long *);

}  // namespace fcgk

static const double kPiHost = 3.14159265358979;
static const double kTidePeriod[12] = {12.4200, 12.00, 12.6584, 11.9673, 23.9344, 25.8194,
                                       24.0659, 26.8684, 327.85, 661.31, 4383.05, 6.2103};

int fcg_fail(fcg_ctx *ctx, int code, const std::string &msg) {
    if (ctx) ctx->err = msg;
    return code;
}

int fcg_fail_hip(fcg_ctx *ctx, hipError_t e, const char *what, const char *file, int line) {
    char buf[512];
    std::snprintf(buf, sizeof buf, "HIP error %d (%s) in %s at %s:%d", (int)e, hipGetErrorString(e), what, file, line);
    return fcg_fail(ctx, FCG_E_HIP, buf);
}

namespace {

constexpr int kBlk = 256;
inline unsigned nblk(size_t n) { return (unsigned)((n + kBlk - 1) / kBlk); }
constexpr int kMinmaxChunks = 512;  // blocks per ensemble member in the T/S extrema reduction

// Column-tile kernels (BPG, vertical viscosity, tracer vertical diffusion): 64
// columns per 256-thread block, LDS tiles of K+1 slots per column.
// Column-aligned flat kernels: 256-thread blocks of cpb = 256/S whole columns
// of S slots, one thread per (column, slot).
constexpr int kColB = 64, kColT = 256, kFlatT = 256;
inline unsigned ncolblk(int ents) { return (unsigned)((ents + kColB - 1) / kColB); }
inline size_t tile_bytes(const DevMesh &m, size_t esz) { return (size_t)kColB * (m.K + 1) * esz; }
inline int cpb_of(int S, int T = kFlatT) { return T / S; }
inline unsigned nflatblk(int ents, int S, int T = kFlatT) {
    return (unsigned)((ents + cpb_of(S, T) - 1) / cpb_of(S, T));
}
inline size_t flat_bytes(int S, size_t esz, int T = kFlatT) { return (size_t)cpb_of(S, T) * S * esz; }
// node-gather block sizes (threads) the launch table may choose
inline int gather_slot(int T) { return T == 1024 ? 2 : (T == 512 ? 1 : 0); }
// k_velrhs: per column one element record and four edge records, for T-thread
// blocks (wider blocks hold more columns, so more neighbours come from LDS)
inline size_t velrhs_bytes(const DevMesh &m, size_t esz, int T = kFlatT) {
    return (size_t)cpb_of(m.L, T) * (kVelrhsElemBytes + 4 * velrhs_edge_bytes(esz)) +
           (size_t)cpb_of(m.L, T) * m.L * 9 * esz;
}
// the level-vector k_velrhs: W = 16/esz layers per lane, K/W lanes per
// column; column arrays with a stride of K slots
inline int velrhs_v_cpb(const DevMesh &m, size_t esz, int T) { return T / (m.K / (int)(16 / esz)); }
inline size_t velrhs_v_bytes(const DevMesh &m, size_t esz, int T) {
    const size_t cpb = (size_t)velrhs_v_cpb(m, esz, T);
    return cpb * (kVelrhsElemBytes + 4 * velrhs_edge_bytes(esz)) + cpb * m.K * 5 * esz;
}
// its block: 256 threads, or 128 where three 256-thread blocks' LDS would not
// share one compute unit (160 KB; the kernel's registers allow three)
inline int velrhs_v_threads(const DevMesh &m, size_t esz) {
    return velrhs_v_bytes(m, esz, 256) <= (160u << 10) / 3 ? 256 : 128;
}
inline size_t lds_p1edge(const DevMesh &m) {
    return 4 * flat_bytes(m.L, 8) + (size_t)cpb_of(m.L) * sizeof(P1Edge);
}
// node-gather blocks: LDS records for every CSR slot of the block's nodes
template <typename R>
inline size_t lds_muscl(const fcg_ctx *c, int T = kFlatT) {
    const int cap = c->cap_n2e_t[gather_slot(T)];
    return (size_t)cap * sizeof(EdgeGeo<R>) + (size_t)((cap + 1) & ~1) * sizeof(int) +
           6 * flat_bytes(c->m.L, sizeof(R), T);
}
// level-vector MUSCL: the block's CSR edge records and orientations, then
// (16-byte aligned) seven column arrays of K slots and the K levels' 1-sigma
inline size_t muscl_v_cols(size_t nrec, size_t esz) {
    const size_t geo = esz == 8 ? sizeof(EdgeGeo<double>) : sizeof(EdgeGeo<float>);
    return (nrec * (geo + sizeof(int)) + 15) & ~(size_t)15;
}
inline int vec_threads(const fcg_ctx *c, size_t (*bytes)(const fcg_ctx *, int)) {
    return bytes(c, 256) <= (160u << 10) / 3 ? 256 : 128;   // three 256-thread blocks' LDS per unit, else 128
}
inline size_t lds_muscl_v(const fcg_ctx *c, int T) {
    const int W = 16 / (int)c->esize, cpb = T / (c->m.K / W);
    return muscl_v_cols((size_t)c->cap_n2e_v[T == 256 ? 1 : 0], c->esize) + 7 * (size_t)cpb * c->m.K * c->esize +
           (size_t)c->m.K * sizeof(double);
}
// the other level-vector kernels run 256-thread blocks: cpb = 256/(K/W)
// columns, LDS records for the block's CSR slots (node kernels) or columns
inline int vec_cpb(const fcg_ctx *c) { return 256 / (c->m.K / (16 / (int)c->esize)); }
inline size_t lds_el2n_v(const fcg_ctx *c) { return (size_t)c->cap_n2el_v[1] * sizeof(PulsRec); }
inline size_t lds_gradn_v(const fcg_ctx *c) { return (size_t)c->cap_n2el_v[1] * sizeof(GradRec); }
inline size_t lds_vdiv_v(const fcg_ctx *c) {
    return (size_t)c->cap_n2e_v[1] * (c->esize == 8 ? sizeof(DivRec<double>) : sizeof(DivRec<float>));
}
inline size_t lds_velgrad_v(const fcg_ctx *c) { return (size_t)vec_cpb(c) * sizeof(VelGradRec); }
// a level-vector variant runs when asked for (the default where nsigma is a
// multiple of W) and its LDS fits
static int levels_of(const fcg_ctx *ctx, const char *kern, int W);
inline bool vec_on(const fcg_ctx *c, const char *kern, size_t lds) {
    const int W = 16 / (int)c->esize;
    return levels_of(c, kern, W) == W && lds <= (size_t)c->lds_limit;
}
template <typename R>
inline size_t lds_vdiv(const fcg_ctx *c, int T = kFlatT) { return (size_t)c->cap_n2e_t[gather_slot(T)] * sizeof(DivRec<R>); }
inline size_t lds_gradn(const fcg_ctx *c, int T = kFlatT) { return (size_t)c->cap_n2el_t[gather_slot(T)] * sizeof(GradRec); }
inline size_t lds_el2n(const fcg_ctx *c, int T = kFlatT) { return (size_t)c->cap_n2el_t[gather_slot(T)] * sizeof(PulsRec); }

// largest LDS request of the tile kernels for this mesh and storage type
inline size_t lds_max(const DevMesh &m, size_t esz) {
    size_t b = std::max(2 * tile_bytes(m, 8), tile_bytes(m, 8) + 2 * tile_bytes(m, esz));  // bpg, velvisc
    b = std::max(b, 5 * tile_bytes(m, esz));                // vdiff
    b = std::max(b, 5 * flat_bytes(m.K, esz));              // d3_end prep
    b = std::max(b, 4 * (size_t)m.K * kD3endT * esz);       // d3_end iteration
    return std::max(b, lds_p1edge(m));
}

template <typename T>
int dalloc(fcg_ctx *ctx, T **p, size_t n) {
    void *q = nullptr;
    size_t bytes = std::max<size_t>(n, 1) * sizeof(T);
    FCG_CHECK(hipMalloc(&q, bytes));
    FCG_CHECK(hipMemset(q, 0, bytes));
    ctx->allocs.push_back(q);
    *p = (T *)q;
    return FCG_OK;
}

int dalloc_bytes(fcg_ctx *ctx, void **p, size_t bytes) {
    void *q = nullptr;
    bytes = std::max<size_t>(bytes, 8);
    FCG_CHECK(hipMalloc(&q, bytes));
    FCG_CHECK(hipMemset(q, 0, bytes));
    ctx->allocs.push_back(q);
    *p = q;
    return FCG_OK;
}

// host doubles -> a 3-D device field in the context's storage type
int upload3(fcg_ctx *ctx, void *dev, const std::vector<double> &h) {
    if (ctx->mixed) {
        std::vector<float> f(h.begin(), h.end());
        FCG_CHECK(hipMemcpy(dev, f.data(), f.size() * sizeof(float), hipMemcpyHostToDevice));
    } else {
        FCG_CHECK(hipMemcpy(dev, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
    }
    return FCG_OK;
}

template <typename T>
int upload(fcg_ctx *ctx, T **p, const std::vector<T> &h) {
    int rc = dalloc(ctx, p, h.size());
    if (rc) return rc;
    if (!h.empty()) FCG_CHECK(hipMemcpy(*p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
    return FCG_OK;
}

// EdgeGeo<R> of every edge (see fcg_internal.h), uploaded to m.egeo
template <typename R>
int upload_egeo(fcg_ctx *ctx, const std::vector<int2> &ednodes, const std::vector<int2> &edtri,
                const std::vector<double4> &ecross, const std::vector<double2> &edxdy,
                const std::vector<double> &elem_cos) {
    const int K = ctx->m.K;
    std::vector<EdgeGeo<R>> g(ednodes.size());
    for (size_t d = 0; d < g.size(); ++d) {
        const int2 nd = ednodes[d], et = edtri[d];
        const double4 c = ecross[d];
        const double2 dd = edxdy[d];
        EdgeGeo<R> &r = g[d];
        r.n1K = nd.x * K;
        r.n2K = nd.y * K;
        r.e1 = et.x;
        r.e2 = et.y;
        r.e1K = et.x * K;
        r.e2K = et.y >= 0 ? et.y * K : -1;
        r.cx = (R)c.x;
        r.cy = (R)c.y;
        r.cz = (R)c.z;
        r.cw = (R)c.w;
        r.dy = (R)dd.y;
        r.dxa = (R)(dd.x * (kREarthD * elem_cos[et.x]));
        r.dxb = et.y >= 0 ? (R)(dd.x * (kREarthD * elem_cos[et.y])) : R(0);
    }
    void *p = nullptr;
    int rc = dalloc_bytes(ctx, &p, g.size() * sizeof(EdgeGeo<R>));
    if (rc) return rc;
    FCG_CHECK(hipMemcpy(p, g.data(), g.size() * sizeof(EdgeGeo<R>), hipMemcpyHostToDevice));
    ctx->m.egeo = p;
    return FCG_OK;
}

#define TRY(x)                 \
    do {                       \
        int rc__ = (x);        \
        if (rc__) return rc__; \
    } while (0)

// 3-D field descriptor for Fortran-layout transfer
struct Field {
    double *dev;       // float storage when rtype and the context is mixed precision
    long levels;   // levels per column (L or K), or components for 2-D vectors
    long ents;     // entities
    bool interleaved;  // device stores (comp, ent) interleaved like Fortran (2,E)
    bool nodal;        // node field (else element field)
    bool r3;           // 3-D field (columns of K slots)
    bool rtype;        // held in the context's storage type (3-D fields and the barotropic momentum path)
};

bool find_field(fcg_ctx *c, const char *name, Field &f) {
    const long N = c->m.N, E = c->m.E, L = c->m.L, K = c->m.K;
    DevState &s = c->s;
    struct Ent { const char *n; double *p; long lev; bool nodal; bool il; };
    const Ent tab[] = {
        {"eta_n", s.eta_n, 1, true, false},    {"eta_n_1", s.eta_n_1, 1, true, false}, {"eta_n_2", s.eta_n_2, 1, true, false},
        {"etaAB", s.etaAB, 1, true, false},    {"ssh_rhs", s.ssh_rhs, 1, true, false}, {"windx", s.windx, 1, true, false},
        {"windy", s.windy, 1, true, false},    {"U_n_2D", s.U2, 2, false, true},       {"U_n_1", s.U2_1, 2, false, true},
        {"U_n_2", s.U2_2, 2, false, true},     {"UAB", s.UAB, 2, false, true},         {"U_rhs_2D_3D", s.U_rhs_2D_3D, 2, false, true},
        {"Bar_pr_2D", s.Bar_pr_2D, 2, false, true}, {"vel_grad", s.vel_grad, 4, false, true},
        {"U_filt_2D", s.U_filt, 1, false, false}, {"V_filt_2D", s.V_filt, 1, false, false},
        {"mask_wd", s.mask_wd, 1, false, false}, {"taux", s.taux, 1, false, false},   {"tauy", s.tauy, 1, false, false},
        {"C_d_el", s.C_d_el, 1, false, false}, {"dmean_n", s.dmean_n, 1, false, false},
        {"U_n", s.U_n, L, false, false},       {"V_n", s.V_n, L, false, false},       {"U_rhs", s.U_rhs, L, false, false},
        {"V_rhs", s.V_rhs, L, false, false},   {"U_rhsAB", s.U_rhsAB, L, false, false}, {"V_rhsAB", s.V_rhsAB, L, false, false},
        {"U_n_filt", s.U_nf, L, false, false}, {"V_n_filt", s.V_nf, L, false, false}, {"Bar_pru_3D", s.Bpu, L, false, false},
        {"Bar_prv_3D", s.Bpv, L, false, false}, {"Je", s.Je, L, false, false},
        {"vel_grad_ux", s.gux, L, false, false}, {"vel_grad_uy", s.guy, L, false, false},
        {"vel_grad_vx", s.gvx, L, false, false}, {"vel_grad_vy", s.gvy, L, false, false},
        {"TF", s.TF, L, true, false},          {"SF", s.SF, L, true, false},          {"T_old", s.T_old, L, true, false},
        {"S_old", s.S_old, L, true, false},    {"Tclim", s.Tclim, L, true, false},    {"Sclim", s.Sclim, L, true, false},
        {"Unode_p", s.Unode_p, L, true, false}, {"Vnode_p", s.Vnode_p, L, true, false},
        {"Unode", s.Unode, L, true, false},    {"Vnode", s.Vnode, L, true, false},
        {"hpressure", s.hp, K, true, false},   {"Kv", s.Kv, K, true, false},          {"snu", s.snu, K, true, false},
        {"bt", s.bt, K, true, false},          {"Wvel", s.W, K, true, false},
        {"qsr", s.qsr, 1, true, false},        {"qns", s.qns, 1, true, false},        {"emp", s.emp, 1, true, false},
        {"mslp", s.mslp, 1, true, false},      {"rho_c", s.rho_c, L, true, false},
        {"runoff", s.runoff, 1, true, false},  {"tau_node", s.tau_node, 2, true, true},
    };
    for (const Ent &t : tab)
        if (!std::strcmp(t.n, name)) {
            if (!t.p) return false;  // not held in this configuration (rho_c without key_atm)
            const bool r3 = !t.il && (t.lev == L || t.lev == K) && std::strcmp(t.n, "hpressure") != 0
                            && std::strcmp(t.n, "rho_c") != 0;
            const bool r2 = !std::strcmp(t.n, "UAB") || !std::strcmp(t.n, "U_rhs_2D_3D")
                            || !std::strcmp(t.n, "Bar_pr_2D") || !std::strcmp(t.n, "vel_grad");
            f = Field{t.p, t.lev, t.nodal ? N : E, t.il, t.nodal, r3, r3 || r2};
            return true;
        }
    return false;
}

// Launch block size of a kernel (without its storage-type argument): the
// autotuned value when one was set, else 256.
static int block_of(const fcg_ctx *ctx, const char *kern) {
    const std::string k(kern, std::strcspn(kern, "<"));
    auto it = ctx->launch_block.find(k);
    return it == ctx->launch_block.end() ? 256 : it->second;
}

// Levels per lane of a kernel with a level-vector variant (key "<kernel>:levels"
// of the launch table): the tuned value when one was set, else W =
// 16/sizeof(R) -- one 16-byte access per W levels -- when nsigma is a
// multiple of W, else 1 (the one-level kernel).
static int levels_of(const fcg_ctx *ctx, const char *kern, int W) {
    if (ctx->m.K % W) return 1;
    auto it = ctx->launch_block.find(std::string(kern) + ":levels");
    return it == ctx->launch_block.end() ? W : it->second;
}

// Tidal elevation at the open boundary for each substep of the coming
// barotropic loop: compute_ssh_rhs_elem, src/fv_average_dynamic.F90:190-216,
// evaluated on the host with the same libm cos as a CPU build.
void fill_tides(fcg_ctx *c, double *dst) {
    const int nb = c->m.nobn, Mt = c->p.Mt;
    if (!nb || !c->p.tf_presence) return;
    for (int sub = 1; sub <= Mt; ++sub) {
        double time_2D = c->time + c->p.dt_2D * sub;
        for (int i = 0; i < nb; ++i) {
            double v = 0.0;
            for (int j = 0; j < 12; ++j)
                v = v + c->amp_factor[j] * c->ampt[(size_t)j * nb + i]
                        * std::cos((time_2D - c->p.time_jd0 * 86400.0) * 2.0 * kPiHost / (kTidePeriod[j] * 3600.0)
                                   - (c->fazt[(size_t)j * nb + i] + c->phy_factor[j]));
            if (!c->p.restart)
                if (c->time_jd - c->p.time_jd0 < 10.0) v = v * (c->time_jd - c->p.time_jd0) / 10.0;
            dst[(size_t)(sub - 1) * nb + i] = v;
        }
    }
}

template <typename R>
int enqueue_substeps(fcg_ctx *ctx, const DevStateT<R> &s0) {
    const DevMesh &m = ctx->m;
    const fcg_params &p = ctx->p;
    hipStream_t st = ctx->stream;
    // two AB3 velocity buffers, alternating by substep; the last substep's is
    // the state's own (UAB as the reference holds it after the loop)
    R *const uab[2] = {s0.UAB, s0.UAB + 2 * (size_t)m.E};
    for (int sub = 1; sub <= p.Mt; ++sub) {
        DevStateT<R> s = s0;
        s.UAB = uab[(p.Mt - sub) & 1];
        const double *tide = (p.tf_presence && m.nobn) ? s.tide + (size_t)(sub - 1) * m.nobn : nullptr;
        // the AB3 velocity and element depth of later substeps come from the
        // previous substep's k_bt_elem2
        if (sub == 1) hipLaunchKernelGGL(fcgk::k_bt_elem1<R>, dim3(nblk(m.E)), dim3(kBlk), 0, st, m, s, p.Dmin);
        hipLaunchKernelGGL(fcgk::k_bt_edge_ssh<R>, dim3(nblk(m.D)), dim3(kBlk), 0, st, m, s);
        hipLaunchKernelGGL(fcgk::k_bt_node<R>, dim3(nblk(m.N)), dim3(kBlk), 0, st, m, s, p.dt_2D, tide, p.key_atm,
                           1.0 / ctx->density_0, p.key_rivers ? (const double *)s.runoff : nullptr, p.Dmin);
        hipLaunchKernelGGL(fcgk::k_bt_velgrad<R>, dim3(nblk(m.E)), dim3(kBlk), 0, st, m, s);
        if (p.mom_adv_2D == 2)
            hipLaunchKernelGGL(fcgk::k_bt_edge_mom<R>, dim3(nblk(m.D)), dim3(kBlk), 0, st, m, s, p.Dmin);
        if (sub == 1) {
            hipLaunchKernelGGL(fcgk::k_bt_puls<R>, dim3(nblk(m.E)), dim3(kBlk), 0, st, m, s, p.Dmin);
            const int be = block_of(ctx, "k_el2nodes");
            if (vec_on(ctx, "k_el2nodes", lds_el2n_v(ctx)))
                hipLaunchKernelGGL((fcgk::k_el2nodes_v<R, true>), dim3((m.N + vec_cpb(ctx) - 1) / vec_cpb(ctx)),
                                   dim3(256), lds_el2n_v(ctx), st, m, (const R *)s.U_n, (const R *)s.V_n, s.Unode_p,
                                   s.Vnode_p, (const double *)s.puls_ap, (const double *)s.U2, ctx->cap_n2el_v[1],
                                   s.err + 3);
            else
                hipLaunchKernelGGL((fcgk::k_el2nodes<R, true>), dim3(nflatblk(m.N, m.L, be)), dim3(be), lds_el2n(ctx, be),
                                   st, m, (const R *)s.U_n, (const R *)s.V_n, s.Unode_p, s.Vnode_p,
                                   (const double *)s.puls_ap, (const double *)s.U2);
            hipLaunchKernelGGL(fcgk::k_bt_p1edge<R>, dim3(nflatblk(m.D, m.L)), dim3(kFlatT), lds_p1edge(m), st, m, s);
            if (p.bih_3D) {  // biharmonic_viscosity_3D_to_2D (src/fv_main.F90:1501, fv_viscosity.F90:964-1220);
                             // k_bt_p1elem adds the column sum
                const unsigned gd = (unsigned)(((size_t)m.D * m.L + kFlatT - 1) / kFlatT);
                const unsigned ge = (unsigned)(((size_t)m.E * m.L + kFlatT - 1) / kFlatT);
                hipLaunchKernelGGL(fcgk::k_bh_puls<R>, dim3(ge), dim3(kFlatT), 0, st, m, s);
                hipLaunchKernelGGL(fcgk::k_bh3_grad<R>, dim3(ge), dim3(kFlatT), 0, st, m, s, (const R *)s.bhpu,
                                   (const R *)s.bhpv, 1);
                hipLaunchKernelGGL(fcgk::k_bh3_edge<R>, dim3(gd), dim3(kFlatT), 0, st, m, s, (const R *)s.bhpu,
                                   (const R *)s.bhpv, (const R *)s.bhgux, (const R *)s.bhguy, (const R *)s.bhgvx,
                                   (const R *)s.bhgvy, 1, 0);
                hipLaunchKernelGGL(fcgk::k_bh3_elem<R>, dim3(ge), dim3(kFlatT), 0, st, m, s, s.bhcu, s.bhcv, 2);
                hipLaunchKernelGGL(fcgk::k_bh3_grad<R>, dim3(ge), dim3(kFlatT), 0, st, m, s, (const R *)s.bhcu,
                                   (const R *)s.bhcv, 0);
                hipLaunchKernelGGL(fcgk::k_bh3_edge<R>, dim3(gd), dim3(kFlatT), 0, st, m, s, (const R *)s.bhcu,
                                   (const R *)s.bhcv, (const R *)s.bhgux, (const R *)s.bhguy, (const R *)s.bhgvx,
                                   (const R *)s.bhgvy, 0, 0);
                hipLaunchKernelGGL(fcgk::k_bh3_elem<R>, dim3(ge), dim3(kFlatT), 0, st, m, s, s.bhub, s.bhvb, 0);
            }
            hipLaunchKernelGGL(fcgk::k_bt_p1elem<R>, dim3(nblk(m.E)), dim3(kBlk), 0, st, m, s,
                               p.filt_3D ? p.tau_inv_filt : 0.0);
        }
        if (p.bih_2D) {  // biharmonic_2D (src/fv_average_dynamic.F90:348, fv_viscosity.F90:338-568): first
                         // pass on UAB, Ul*Ah*acc, its gradients; k_bt_elem2 adds the second pass
            hipLaunchKernelGGL(fcgk::k_bh2_edge<R>, dim3(nblk(m.D)), dim3(kBlk), 0, st, m, s, (const R *)s.UAB,
                               (const R *)s.vel_grad);
            hipLaunchKernelGGL(fcgk::k_bh2_elem<R>, dim3(nblk(m.E)), dim3(kBlk), 0, st, m, s);
            DevStateT<R> sg = s;  // vel_gradients_2D(Ul)
            sg.UAB = s.bh_ul;
            sg.vel_grad = s.bh_g2;
            hipLaunchKernelGGL(fcgk::k_bt_velgrad<R>, dim3(nblk(m.E)), dim3(kBlk), 0, st, m, sg);
            hipLaunchKernelGGL(fcgk::k_bh2_edge<R>, dim3(nblk(m.D)), dim3(kBlk), 0, st, m, s, (const R *)s.bh_ul,
                               (const R *)s.bh_g2);
        }
        if (p.filt_bi_2D)  // src/fv_average_dynamic.F90:347 (tau_inv: fv_viscosity.F90:695)
            hipLaunchKernelGGL(fcgk::k_bt_filt2x_c<R>, dim3(nblk(m.E)), dim3(kBlk), 0, st, m, s,
                               p.dt_2D * p.tau_c / 3600.0 / 24.0);
        hipLaunchKernelGGL(fcgk::k_bt_elem2<R>, dim3(nblk(m.E)), dim3(kBlk), 0, st, m, s, p.dt_2D, p.Dmin,
                           ctx->density_0, p.C_d, p.Cd_bz, p.wet_dry_on, p.mom_adv_2D == 2 ? 1 : 0, sub == 1 ? 1 : 0,
                           p.filt_2D ? p.tau_inv_filt : 0.0, sub < p.Mt ? uab[(p.Mt - sub - 1) & 1] : (R *)nullptr);
    }
    return FCG_OK;
}

}  // namespace

// Launch a kernel on the context stream, bracketed by probe events when the
// kernel's name (without its storage-type argument) is the probed one.

static bool probe_is(const std::string &probe, const char *kern) {
    if (probe.empty()) return false;
    size_t n = std::strcspn(kern, "<");
    return probe.size() == n && probe.compare(0, n, kern, n) == 0;
}
#define LAUNCH(ctx, kern, grid, ...) LAUNCH_EX(ctx, kern, grid, kBlk, 0, __VA_ARGS__)
// one thread per item, block size from the context's launch table
#define LAUNCH1(ctx, kern, items, ...)                                                     \
    do {                                                                                   \
        const int bs__ = block_of(ctx, #kern);                                             \
        LAUNCH_EX(ctx, kern, (unsigned)(((size_t)(items) + bs__ - 1) / bs__), bs__, 0, __VA_ARGS__); \
    } while (0)
// one thread per (entity, layer), consecutive threads along a column
#define LAUNCHLV(ctx, kern, ents, levels, ...) LAUNCH1(ctx, kern, (size_t)(ents) * (size_t)(levels), __VA_ARGS__)
#define LAUNCH_EX(ctx, kern, grid, block, lds, ...) LAUNCH_NAMED(ctx, #kern, fcgk::kern, grid, block, lds, __VA_ARGS__)
// the same for a kernel expression whose probe name is given (a template
// with several arguments)
#define LAUNCH_NAMED(ctx, name, kern, grid, block, lds, ...)                               \
    do {                                                                                   \
        bool pr__ = probe_is((ctx)->probe_name, name);                                     \
        hipEvent_t a__ = nullptr, b__ = nullptr;                                           \
        if (pr__) {                                                                        \
            (void)hipEventCreate(&a__);                                                    \
            (void)hipEventCreate(&b__);                                                    \
            (void)hipEventRecord(a__, (ctx)->stream);                                      \
        }                                                                                  \
        hipLaunchKernelGGL(kern, dim3(grid), dim3(block), (lds), (ctx)->stream, __VA_ARGS__);       \
        if (pr__) {                                                                        \
            (void)hipEventRecord(b__, (ctx)->stream);                                      \
            (ctx)->probe_ev.emplace_back(a__, b__);                                        \
        }                                                                                  \
    } while (0)

extern "C" int fcg_probe(fcg_ctx *ctx, const char *kernel) {
    if (!ctx) return FCG_E_ARG;
    ctx->probe_name = kernel ? kernel : "";
    return FCG_OK;
}

extern "C" int fcg_probe_read(fcg_ctx *ctx, double *total_ms, int *launches) {
    if (!ctx || !total_ms || !launches) return FCG_E_ARG;
    FCG_CHECK(hipStreamSynchronize(ctx->stream));
    double t = 0.0;
    for (auto &pr : ctx->probe_ev) {
        float ms = 0.0f;
        FCG_CHECK(hipEventElapsedTime(&ms, pr.first, pr.second));
        t += ms;
        (void)hipEventDestroy(pr.first);
        (void)hipEventDestroy(pr.second);
    }
    *total_ms = t;
    *launches = (int)ctx->probe_ev.size();
    ctx->probe_ev.clear();
    return FCG_OK;
}

extern "C" const char *fcg_version(void) { return "fesomc_gpu 0.1 gfx950 fp64+graph"; }

// node_member: ensemble member of each (union) node, members numbered 0..nmemb-1
// and contiguous in the union numbering; nullptr for a single model.
// member_sa: scale_area of each member (fv_ini.F90:81), nullptr for a single model.
static int create_impl(const fcg_mesh *mesh, const fcg_params *par, const int *node_member, int nmemb,
                       fcg_ctx **out, const double *member_sa = nullptr) {
    if (!mesh || !par || !out || nmemb < 1) return FCG_E_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return FCG_E_NOGPU;
    auto *ctx = new fcg_ctx();
    ctx->p = *par;
    const fcg_params &p = ctx->p;
    if (mesh->nsigma < 4 || mesh->nsigma != p.nsigma || p.Mt < 1 || p.dt <= 0 || p.dt_2D <= 0) {
        fcg_fail(ctx, FCG_E_ARG, "inconsistent nsigma/Mt/dt (nsigma >= 4 required)");
        *out = ctx;
        return FCG_E_ARG;
    }
    if (!(p.tracer_adv == 1 || p.tracer_adv == 2 || p.tracer_adv == 4) ||
        !(p.precision == FCG_FP64 || p.precision == FCG_MIXED) ||
        !(p.closure == FCG_CLOSURE_D3END || p.closure == FCG_CLOSURE_MLP)) {
        fcg_fail(ctx, FCG_E_ARG, "option not available in this build (tracer_adv/precision/closure)");
        *out = ctx;
        return FCG_E_ARG;
    }
    if (p.key_atm && !(p.type_swr_body == 1 || p.type_swr_body == 2)) {  // src/fv_read_run.F90:799-802
        fcg_fail(ctx, FCG_E_ARG, "type_swr_body should be 1 or 2");
        *out = ctx;
        return FCG_E_ARG;
    }
    FCG_CHECK(hipGetDevice(&ctx->device));
    FCG_CHECK(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
    DevMesh &m = ctx->m;
    const int N = mesh->nod2D, E = mesh->elem2D, D = mesh->edge2D, K = mesh->nsigma, L = K - 1;
    m.N = N; m.E = E; m.D = D; m.Din = mesh->edge2D_in; m.K = K; m.L = L; m.nobn = mesh->nobn;
    ctx->density_0 = mesh->density_0;
    ctx->scale_area = mesh->scale_area;
    // ---- locality order.  Device arrays hold nodes, elements and edges in
    // Morton (Z-curve) order of their positions, so the neighbour gathers of a
    // wavefront touch a compact patch of memory.  Every gather list stays sorted
    // by the reference's ids, so summation order -- and FP64 bits -- are unchanged.
    auto z = [](int v) { return v > 0 ? v - 1 : -1; };
    std::vector<int> pn(N), pe(E), pd(D);  // device index -> reference index (0-based)
    for (int i = 0; i < N; ++i) pn[i] = i;
    for (int i = 0; i < E; ++i) pe[i] = i;
    for (int i = 0; i < D; ++i) pd[i] = i;
    if (mesh->coord_nod2D) {
        const double *xy = mesh->coord_nod2D;
        double x0 = xy[0], x1 = xy[0], y0 = xy[1], y1 = xy[1];
        for (int n = 0; n < N; ++n) {
            x0 = std::min(x0, xy[2 * n]); x1 = std::max(x1, xy[2 * n]);
            y0 = std::min(y0, xy[2 * n + 1]); y1 = std::max(y1, xy[2 * n + 1]);
        }
        const double sx = x1 > x0 ? 65535.0 / (x1 - x0) : 0.0, sy = y1 > y0 ? 65535.0 / (y1 - y0) : 0.0;
        auto spread = [](uint64_t v) {
            v &= 0xffffull;
            v = (v | (v << 8)) & 0x00ff00ffull;
            v = (v | (v << 4)) & 0x0f0f0f0full;
            v = (v | (v << 2)) & 0x33333333ull;
            v = (v | (v << 1)) & 0x55555555ull;
            return v;
        };
        auto key = [&](double x, double y) {
            uint64_t qx = (uint64_t)std::lround((x - x0) * sx), qy = (uint64_t)std::lround((y - y0) * sy);
            return spread(qx) | (spread(qy) << 1);
        };
        std::vector<uint64_t> kn(N), ke(E), kd(D);
        // ensemble members stay contiguous: member id above the Morton key
        auto mem = [&](int n0) { return node_member ? (uint64_t)node_member[n0] << 32 : 0ull; };
        for (int n = 0; n < N; ++n) kn[n] = key(xy[2 * n], xy[2 * n + 1]) | mem(n);
        for (int e = 0; e < E; ++e) {
            const int *q = mesh->elem2D_nodes + 4 * e;
            int nq = (q[0] == q[3]) ? 3 : 4;
            double cx = 0, cy = 0;
            for (int j = 0; j < nq; ++j) { cx += xy[2 * (q[j] - 1)]; cy += xy[2 * (q[j] - 1) + 1]; }
            ke[e] = key(cx / nq, cy / nq) | mem(q[0] - 1);
        }
        for (int d = 0; d < D; ++d) {
            int a0 = mesh->edge_nodes[2 * d] - 1, a1 = mesh->edge_nodes[2 * d + 1] - 1;
            kd[d] = key(0.5 * (xy[2 * a0] + xy[2 * a1]), 0.5 * (xy[2 * a0 + 1] + xy[2 * a1 + 1])) | mem(a0);
        }
        std::stable_sort(pn.begin(), pn.end(), [&](int i, int j) { return kn[i] < kn[j]; });
        std::stable_sort(pe.begin(), pe.end(), [&](int i, int j) { return ke[i] < ke[j]; });
        std::stable_sort(pd.begin(), pd.end(), [&](int i, int j) { return kd[i] < kd[j]; });
    }
    std::vector<int> qn(N), qe(E), qd(D);  // reference index -> device index
    for (int i = 0; i < N; ++i) qn[pn[i]] = i;
    for (int i = 0; i < E; ++i) qe[pe[i]] = i;
    for (int i = 0; i < D; ++i) qd[pd[i]] = i;
    ctx->pn = pn;
    ctx->pe = pe;
    // member layout in device order: offsets of each member's nodes / elements
    m.nmemb = nmemb;
    std::vector<int> memb(N, 0);
    ctx->noff.assign(nmemb + 1, 0);
    ctx->eoff.assign(nmemb + 1, 0);
    {
        std::vector<int> ncount(nmemb, 0), ecount(nmemb, 0);
        int prev = 0;
        for (int i = 0; i < N; ++i) {
            const int mb = node_member ? node_member[pn[i]] : 0;
            if (mb < prev || mb >= nmemb) {
                fcg_fail(ctx, FCG_E_MESH, "ensemble members are not contiguous");
                *out = ctx;
                return FCG_E_MESH;
            }
            prev = mb;
            memb[i] = mb;
            ncount[mb]++;
        }
        prev = 0;
        for (int i = 0; i < E; ++i) {
            const int mb = node_member ? node_member[mesh->elem2D_nodes[4 * pe[i]] - 1] : 0;
            if (mb < prev) {
                fcg_fail(ctx, FCG_E_MESH, "ensemble members are not contiguous");
                *out = ctx;
                return FCG_E_MESH;
            }
            prev = mb;
            ecount[mb]++;
        }
        for (int b = 0; b < nmemb; ++b) {
            ctx->noff[b + 1] = ctx->noff[b] + ncount[b];
            ctx->eoff[b + 1] = ctx->eoff[b] + ecount[b];
        }
    }
    auto QN = [&](int v1) { return v1 > 0 ? qn[v1 - 1] : -1; };  // 1-based reference id -> device index
    auto QE = [&](int v1) { return v1 > 0 ? qe[v1 - 1] : -1; };
    auto QD = [&](int v1) { return v1 > 0 ? qd[v1 - 1] : -1; };
    // ---- element constants
    std::vector<int4> en(E), nb(E), eled(E), ied(E);
    std::vector<double4> wcv(E), gsx(E), gsy(E), gvx(E), gvy(E);
    std::vector<double> accB(E), acc_el(E), v_ea(E), v_ec(E), v_co(E), kbi;
    const bool filt_bi = p.filt_bi_2D || p.filt_bi_3D, bih = p.bih_2D || p.bih_3D;
    std::vector<double> sarea(nmemb, mesh->scale_area), bh_ah, bh_ah_acc;
    std::vector<int4> el_bed;
    if (member_sa) sarea.assign(member_sa, member_sa + nmemb);
    if (filt_bi) kbi.resize(E);
    if (bih) {
        bh_ah.resize(E);
        bh_ah_acc.resize(E);
        el_bed.resize(E);
    }
    for (int i = 0; i < E; ++i) {
        const int e = pe[i];
        const int *q = mesh->elem2D_nodes + 4 * e;
        en[i] = make_int4(QN(q[0]), QN(q[1]), QN(q[2]), QN(q[3]));
        const int *b = mesh->elem_neighbors + 4 * e;
        nb[i] = make_int4(QE(b[0]), QE(b[1]), QE(b[2]), QE(b[3]));
        const int *d = mesh->elem_edges + 4 * e;
        eled[i] = make_int4(QD(d[0]), QD(d[1]), QD(d[2]), QD(d[3]));
        const double *w = mesh->w_cv + 4 * e;
        wcv[i] = make_double4(w[0], w[1], w[2], w[3]);
        const double *g = mesh->gradient_sca + 8 * e;
        gsx[i] = make_double4(g[0], g[1], g[2], g[3]);
        gsy[i] = make_double4(g[4], g[5], g[6], g[7]);
        const double *v = mesh->gradient_vec + 8 * e;
        gvx[i] = make_double4(v[0], v[1], v[2], v[3]);
        gvy[i] = make_double4(v[4], v[5], v[6], v[7]);
        double acc = w[0] * mesh->ac[q[0] - 1];
        acc = acc + w[1] * mesh->ac[q[1] - 1];
        acc = acc + w[2] * mesh->ac[q[2] - 1];
        acc = acc + w[3] * mesh->ac[q[3] - 1];
        acc_el[i] = acc;
        accB[i] = std::pow(acc, p.Bp_bz) * 1.0;  // mask_bpr = 1 (Mask_Bar_pr = .false., src/fv_var.F90:57)
        ied[i] = make_int4(kEmpty, kEmpty, kEmpty, kEmpty);
        v_ea[i] = mesh->elem_area[e];
        const double sa = sarea[node_member ? node_member[q[0] - 1] : 0];
        if (filt_bi)  // the factor of K_bi (fv_viscosity.F90:599, 706)
            kbi[i] = std::sqrt(sa / mesh->elem_area[e]);
        if (bih) {  // Ah of biharmonic_viscosity/_2D and of the 3-D-to-2-D form (fv_viscosity.F90:232, 465, 1088-1089)
            bh_ah[i] = -p.Abh0 * std::sqrt(mesh->elem_area[e] / sa) / sa;
            bh_ah_acc[i] = -acc * p.Abh0 * std::sqrt(mesh->elem_area[e] / sa) / sa;
            // boundary edges in ascending reference id: after all internal edges
            // in the serial loops (ed = edge2D_in+1..edge2D)
            int b[4], nbnd = 0;
            const int nq = (d[0] == d[3]) ? 3 : 4;
            for (int j = 0; j < nq; ++j)
                if (mesh->edge_tri[2 * (d[j] - 1) + 1] <= 0) b[nbnd++] = d[j];
            std::sort(b, b + nbnd);
            int bs[4] = {kEmpty, kEmpty, kEmpty, kEmpty};
            for (int j = 0; j < nbnd; ++j) bs[j] = QD(b[j]);
            el_bed[i] = make_int4(bs[0], bs[1], bs[2], bs[3]);
        }
        v_ec[i] = mesh->elem_cos[e];
        v_co[i] = mesh->coriolis[e];
    }
    // ---- edge constants + gathers (lists in ascending reference edge id)
    std::vector<int2> ednodes(D), edtri(D);
    std::vector<double4> ecross(D);
    std::vector<double2> edxdy(D);
    std::vector<BhEdge> bhe(bih ? D : 0);
    for (int i = 0; i < D; ++i) {
        const int d = pd[i];
        ednodes[i] = make_int2(QN(mesh->edge_nodes[2 * d]), QN(mesh->edge_nodes[2 * d + 1]));
        edtri[i] = make_int2(QE(mesh->edge_tri[2 * d]), QE(mesh->edge_tri[2 * d + 1]));
        const double *c = mesh->edge_cross_dxdy + 4 * d;
        ecross[i] = make_double4(c[0], c[1], c[2], c[3]);
        edxdy[i] = make_double2(mesh->edge_dxdy[2 * d], mesh->edge_dxdy[2 * d + 1]);
        if (bih) {  // biharmonic edge geometry (fv_viscosity.F90:148-156, 208-210, 381-392)
            BhEdge &g = bhe[i];
            const int t1 = mesh->edge_tri[2 * d] - 1, t2 = mesh->edge_tri[2 * d + 1] - 1;
            const double dx = mesh->edge_dxdy[2 * d], dy = mesh->edge_dxdy[2 * d + 1];
            g = BhEdge{};
            g.ye = dy * kREarthD;
            g.acc = (mesh->ac[mesh->edge_nodes[2 * d] - 1] + mesh->ac[mesh->edge_nodes[2 * d + 1] - 1]) / 2.0;
            if (t2 >= 0) {
                g.xe = dx * kREarthD * 0.5 * (mesh->elem_cos[t1] + mesh->elem_cos[t2]);
                g.tx = -c[0] + c[2];
                g.ty = -c[1] + c[3];
                g.tl = std::sqrt(g.tx * g.tx + g.ty * g.ty);
                g.tx = g.tx / g.tl;
                g.ty = g.ty / g.tl;
            } else {
                g.xb = dx * kREarthD * mesh->elem_cos[t1];
                g.xe = dx * kREarthD;
                g.xbd = g.xe * mesh->elem_cos[t1];
            }
        }
    }
    std::vector<std::vector<int>> n2e(N);
    std::vector<int> icnt(E, 0);
    for (int d = 0; d < D; ++d) {  // reference order
        const int dv = qd[d];
        const int a = QN(mesh->edge_nodes[2 * d]), b = QN(mesh->edge_nodes[2 * d + 1]);
        n2e[a].push_back(dv);
        n2e[b].push_back(~dv);
        if (d < m.Din) {
            const int t1 = QE(mesh->edge_tri[2 * d]), t2 = QE(mesh->edge_tri[2 * d + 1]);
            if (t1 < 0 || t2 < 0 || icnt[t1] >= 4 || icnt[t2] >= 4) {
                fcg_fail(ctx, FCG_E_MESH, "internal edge without two elements, or element with more than 4 edges");
                *out = ctx;
                return FCG_E_MESH;
            }
            (&ied[t1].x)[icnt[t1]++] = ~dv;  // element is edge_tri(1): receives -flux
            (&ied[t2].x)[icnt[t2]++] = dv;   // element is edge_tri(2): receives +flux
        }
    }
    std::vector<int> n2e_ptr(N + 1, 0), n2e_idx;
    for (int n = 0; n < N; ++n) {
        n2e_ptr[n + 1] = n2e_ptr[n] + (int)n2e[n].size();
        n2e_idx.insert(n2e_idx.end(), n2e[n].begin(), n2e[n].end());
    }
    std::vector<std::vector<int>> n2el(N);
    for (int e = 0; e < E; ++e) {  // reference order: nod_in_elem2D ascending element id
        const int *q = mesh->elem2D_nodes + 4 * e;
        int nq = (q[0] == q[3]) ? 3 : 4;
        for (int j = 0; j < nq; ++j) n2el[QN(q[j])].push_back(qe[e]);
    }
    std::vector<int> n2el_ptr(N + 1, 0), n2el_idx;
    for (int n = 0; n < N; ++n) {
        n2el_ptr[n + 1] = n2el_ptr[n] + (int)n2el[n].size();
        n2el_idx.insert(n2el_idx.end(), n2el[n].begin(), n2el[n].end());
    }
    for (int t = 0; t < 2; ++t) {  // level-vector node blocks: cpb = T/(K/W) nodes, W = 16/storage width
        const int W = p.precision == FCG_MIXED ? 4 : 2;   // the storage width is set further down
        const int cpb = std::max(1, (128 << t) / std::max(1, K / W));
        for (int c0 = 0; c0 < N; c0 += cpb) {
            const int c1 = std::min(N, c0 + cpb);
            ctx->cap_n2e_v[t] = std::max(ctx->cap_n2e_v[t], n2e_ptr[c1] - n2e_ptr[c0]);
            ctx->cap_n2el_v[t] = std::max(ctx->cap_n2el_v[t], n2el_ptr[c1] - n2el_ptr[c0]);
        }
    }
    for (int t = 0; t < 3; ++t) {  // most CSR slots in one node-gather block of cpb = T/L consecutive nodes
        const int cpb = std::max(1, (256 << t) / L);
        for (int c0 = 0; c0 < N; c0 += cpb) {
            const int c1 = std::min(N, c0 + cpb);
            ctx->cap_n2e_t[t] = std::max(ctx->cap_n2e_t[t], n2e_ptr[c1] - n2e_ptr[c0]);
            ctx->cap_n2el_t[t] = std::max(ctx->cap_n2el_t[t], n2el_ptr[c1] - n2el_ptr[c0]);
        }
    }
    // ---- node constants
    std::vector<double> depth(N), area(N), ac(N), atp1(N), atp2(N);
    std::vector<int> index(N), obidx(N, -1);
    for (int i = 0; i < N; ++i) {
        const int n = pn[i];
        depth[i] = mesh->depth[n];
        area[i] = mesh->area[n];
        ac[i] = mesh->ac[n];
        atp1[i] = mesh->a_tp[n];
        atp2[i] = mesh->a_tp[N + n];
        index[i] = mesh->index_nod2D[n];
    }
    for (int i = 0; i < mesh->nobn; ++i) {
        obidx[QN(mesh->in_obn[i])] = i;
        ctx->in_obn0.push_back(mesh->in_obn[i] - 1);
    }
    std::vector<double> sigma(mesh->sigma, mesh->sigma + K), dsig(L), oms(K);
    for (int k = 0; k < L; ++k) dsig[k] = sigma[k] - sigma[k + 1];
    for (int k = 0; k < K; ++k) oms[k] = 1.0 - sigma[k];
    if (mesh->nobn > 0) {
        ctx->ampt.assign(mesh->ampt, mesh->ampt + 12 * (size_t)mesh->nobn);
        ctx->fazt.assign(mesh->fazt, mesh->fazt + 12 * (size_t)mesh->nobn);
    }
    ctx->amp_factor.assign(12, 1.0);
    ctx->phy_factor.assign(12, 0.0);
    if (mesh->amp_factor) ctx->amp_factor.assign(mesh->amp_factor, mesh->amp_factor + 12);
    if (mesh->phy_factor) ctx->phy_factor.assign(mesh->phy_factor, mesh->phy_factor + 12);
    int rc = 0;
#define UP(dst, src) if ((rc = upload(ctx, &(dst), src))) { *out = ctx; return rc; }
    UP(m.en, en) UP(m.elnb, nb) UP(m.eledges, eled) UP(m.el_ied, ied) UP(m.wcv, wcv) UP(m.gsx, gsx) UP(m.gsy, gsy)
    UP(m.gvx, gvx) UP(m.gvy, gvy) UP(m.accB, accB) UP(m.acc_el, acc_el)
    UP(m.elem_area, v_ea) UP(m.elem_cos, v_ec) UP(m.coriolis, v_co)
    if (filt_bi) {
        UP(m.kbi, kbi)
    }
    if (bih) {
        UP(m.bh_ah, bh_ah) UP(m.bh_ah_acc, bh_ah_acc) UP(m.el_bed, el_bed) UP(m.bhe, bhe)
    }
    if (member_sa) {
        UP(m.memb_sa, sarea)
    }
    UP(m.ednodes, ednodes) UP(m.edtri, edtri) UP(m.ecross, ecross) UP(m.edxdy, edxdy)
    UP(m.depth, depth) UP(m.area, area) UP(m.ac, ac) UP(m.atp1, atp1) UP(m.atp2, atp2) UP(m.index, index)
    UP(m.obidx, obidx) UP(m.n2e_ptr, n2e_ptr) UP(m.n2e, n2e_idx) UP(m.n2el_ptr, n2el_ptr) UP(m.n2el, n2el_idx)
    UP(m.sigma, sigma) UP(m.dsig, dsig) UP(m.oms, oms) UP(m.memb, memb) UP(m.memb_noff, ctx->noff)
    // ---- state
    DevState &s = ctx->s;
    // 3-D fields by column, K slots per column (layer fields use the first L)
    const size_t NL = (size_t)N * K, EL = (size_t)E * K, NK = (size_t)N * K, EK = (size_t)E * K;
#define DA(ptr, n) if ((rc = dalloc(ctx, &(ptr), (n)))) { *out = ctx; return rc; }
    // 3-D fields: element size of the precision mode (the pointer is retyped on use)
#define DA3(ptr, n) if ((rc = dalloc_bytes(ctx, (void **)&(ptr), (n) * ctx->esize))) { *out = ctx; return rc; }
    ctx->mixed = p.precision == FCG_MIXED;
    ctx->esize = ctx->mixed ? sizeof(float) : sizeof(double);
    {  // column offsets of element nodes and node elements; per-edge gather operands
        std::vector<int4> enK(E);
        for (int e = 0; e < E; ++e) enK[e] = make_int4(en[e].x * K, en[e].y * K, en[e].z * K, en[e].w * K);
        std::vector<int> n2elK(n2el_idx.size());
        for (size_t j = 0; j < n2el_idx.size(); ++j) n2elK[j] = n2el_idx[j] * K;
        UP(m.enK, enK) UP(m.n2elK, n2elK)
        if (ctx->mixed) {  // the barotropic momentum path's geometry in FP32
            auto f4 = [](const std::vector<double4> &v) {
                std::vector<float4> o(v.size());
                for (size_t i = 0; i < v.size(); ++i)
                    o[i] = make_float4((float)v[i].x, (float)v[i].y, (float)v[i].z, (float)v[i].w);
                return o;
            };
            std::vector<float2> dxdy_f(edxdy.size());
            for (size_t i = 0; i < edxdy.size(); ++i) dxdy_f[i] = make_float2((float)edxdy[i].x, (float)edxdy[i].y);
            UP(m.gsx_f, f4(gsx)) UP(m.gsy_f, f4(gsy)) UP(m.gvx_f, f4(gvx)) UP(m.gvy_f, f4(gvy))
            UP(m.ecross_f, f4(ecross)) UP(m.edxdy_f, dxdy_f)
        }
        rc = ctx->mixed ? upload_egeo<float>(ctx, ednodes, edtri, ecross, edxdy, v_ec)
                        : upload_egeo<double>(ctx, ednodes, edtri, ecross, edxdy, v_ec);
        if (rc) {
            *out = ctx;
            return rc;
        }
    }
    {
        int lds_limit = 0;
        FCG_CHECK(hipDeviceGetAttribute(&lds_limit, hipDeviceAttributeMaxSharedMemoryPerBlock, ctx->device));
        const size_t gather = std::max(std::max(lds_muscl<double>(ctx), lds_vdiv<double>(ctx)),
                                       std::max(lds_gradn(ctx), lds_el2n(ctx)));
        ctx->lds_limit = lds_limit;
        FCG_CHECK(hipDeviceGetAttribute(&ctx->n_cu, hipDeviceAttributeMultiprocessorCount, ctx->device));
        if (m.K > kFlatT || std::max(lds_max(m, ctx->esize), gather) > (size_t)lds_limit) {
            fcg_fail(ctx, FCG_E_ARG, "nsigma too large for the LDS column tiles of this device");
            *out = ctx;
            return FCG_E_ARG;
        }
    }
    DA(s.eta_n, N) DA(s.eta_n_1, N) DA(s.eta_n_2, N) DA(s.etaAB, N) DA(s.eta_g, N) DA(s.ssh_rhs, N)
    DA(s.U2, 2 * E) DA(s.U2_1, 2 * E) DA(s.U2_2, 2 * E) DA(s.UAB, 4 * E) DA(s.aux_el, E) DA(s.vel_grad, 4 * E)
    DA(s.U_rhs_2D_3D, 2 * E) DA(s.Bar_pr_2D, 2 * E) DA(s.adn, N) DA(s.U_filt, E) DA(s.V_filt, E) DA(s.dmean_n, E)
    DA(s.mask_wd, E) DA(s.taux, E) DA(s.tauy, E) DA(s.C_d_el, E) DA(s.windx, N) DA(s.windy, N)
    DA(s.a, N) DA(s.a_old, N) DA(s.ubot, 2 * E) DA(s.fl_ssh, D) DA(s.fl_uv, 2 * (size_t)D) DA(s.fl_p1, 4 * (size_t)D) DA(s.qsr, N) DA(s.qns, N) DA(s.emp, N) DA(s.mslp, N) DA(s.runoff, N) DA(s.node2d, 4 * (size_t)N) DA(s.tide, (size_t)std::max(1, p.Mt * mesh->nobn))
    if (p.key_atm) {
        DA(s.rho_c, (size_t)N * L)
    }
    DA3(s.U_n, EL) DA3(s.V_n, EL) DA3(s.U_rhs, EL) DA3(s.V_rhs, EL) DA3(s.U_rhsAB, EL) DA3(s.V_rhsAB, EL)
    DA3(s.U_nf, EL) DA3(s.V_nf, EL) DA3(s.gux, EL) DA3(s.guy, EL) DA3(s.gvx, EL) DA3(s.gvy, EL)
    DA3(s.Bpu, EL) DA3(s.Bpv, EL) DA3(s.Je, EL) DA(s.puls_ap, E)
    if (p.filt_bi_3D) {
        DA3(s.fbu, EL) DA3(s.fbv, EL) DA3(s.uvu, EL) DA3(s.uvv, EL)
    }
    if (p.filt_bi_2D) {
        DA(s.fb2c, 2 * E)
    }
    if (p.bih_3D) {
        const size_t DK = (size_t)D * K;
        DA3(s.bhfu, DK) DA3(s.bhfv, DK) DA3(s.bhcu, EL) DA3(s.bhcv, EL)
        DA3(s.bhgux, EL) DA3(s.bhguy, EL) DA3(s.bhgvx, EL) DA3(s.bhgvy, EL)
        DA3(s.bhpu, EL) DA3(s.bhpv, EL) DA3(s.bhub, EL) DA3(s.bhvb, EL)
    }
    if (p.bih_2D) {
        DA(s.bh_fl2, 2 * (size_t)D) DA3(s.bh_ul, 2 * (size_t)E) DA3(s.bh_g2, 4 * (size_t)E)
    }
    DA3(s.vxx, EK) DA3(s.vyu, EK) DA3(s.vyv, EK)
    DA3(s.TF, NL) DA3(s.SF, NL) DA3(s.T_old, NL) DA3(s.S_old, NL) DA3(s.Tclim, NL) DA3(s.Sclim, NL)
    DA3(s.abT, NL) DA3(s.abS, NL) DA3(s.Unode_p, NL) DA3(s.Vnode_p, NL) DA3(s.Unode, NL) DA3(s.Vnode, NL) DA3(s.tcp, NL) DA3(s.ttp, NL) DA3(s.tsp, NL)
    if (p.tracer_adv == 4 || (p.tracer_adv == 2 && p.muscl_call)) {
        DA3(s.gTx, NL) DA3(s.gTy, NL)
    }
    if (p.tracer_adv == 4) {
        DA3(s.tcur, NL)
    }
    if (p.tracer_adv == 2 && p.muscl_call) {
        DA3(s.gSx, NL) DA3(s.gSy, NL)
    }
    DA(s.hp, NK) DA3(s.Kv, NK) DA3(s.snu, NK) DA3(s.bt, NK) DA3(s.W, NK)
    DA3(s.mx_p, NK) DA3(s.mx_q, NK) DA3(s.mx_c2, NK) DA3(s.mx_xx, NK) DA3(s.mx_yy, NK) DA3(s.mx_bA, NK) DA3(s.mx_bB, NK) DA3(s.mx_s, NK)
    ctx->red_blocks = 512;
    DA(s.red, std::max<size_t>(16 * (size_t)ctx->red_blocks, 4 * (size_t)kMinmaxChunks * m.nmemb))
    DA(s.scal, 32) DA(s.err, 4) DA(s.tsa, 2 * (size_t)m.nmemb)
    if (p.closure == FCG_CLOSURE_MLP) {
        DA(ctx->mlp_flag, N) DA(ctx->mlp_count, 1)
    }
    FCG_CHECK(hipHostMalloc((void **)&ctx->tide_pinned,
                            sizeof(double) * kTideRing * (size_t)std::max(1, p.Mt * mesh->nobn)));
    for (int b = 0; b < kTideRing; ++b) FCG_CHECK(hipEventCreateWithFlags(&ctx->tide_ev[b], hipEventDisableTiming));
    // ---- cold start: array_setup + jacobian + initial_state (src/fv_main.F90:157-168, 320-362)
    std::vector<double> eta(N, 0.0), ea(N), a1(N), a2(N);
    for (int n = 0; n < N; ++n) {
        a1[n] = std::max(depth[n] + 0.0, p.Dmin);   // first jacobian, eta = 0 (fv_main.F90:162)
        double d = depth[n] + eta[n];
        if (d < 0.0) eta[n] = -depth[n];           // dry start, fv_ini.F90:127-136
    }
    {
        const double de = 0.614, ga = 0.088, ep = 0.013;
        const double fe = 1.0 - de - ga - ep;
        for (int n = 0; n < N; ++n) {
            ea[n] = de * eta[n] + fe * eta[n] + ga * eta[n] + ep * eta[n];
            a2[n] = std::max(depth[n] + eta[n], p.Dmin);  // second jacobian (fv_main.F90:362)
        }
    }
    std::vector<double> cst(NK, p.snul), bt0(NK, 1.0e-8), ones(E, 1.0), cd(E, p.C_d);
    FCG_CHECK(hipMemcpy(s.eta_n, eta.data(), N * 8, hipMemcpyHostToDevice));
    FCG_CHECK(hipMemcpy(s.eta_n_1, eta.data(), N * 8, hipMemcpyHostToDevice));
    FCG_CHECK(hipMemcpy(s.eta_n_2, eta.data(), N * 8, hipMemcpyHostToDevice));
    FCG_CHECK(hipMemcpy(s.etaAB, ea.data(), N * 8, hipMemcpyHostToDevice));
    FCG_CHECK(hipMemcpy(s.a_old, a1.data(), N * 8, hipMemcpyHostToDevice));
    FCG_CHECK(hipMemcpy(s.a, a2.data(), N * 8, hipMemcpyHostToDevice));
    if ((rc = upload3(ctx, s.snu, cst)) || (rc = upload3(ctx, s.Kv, cst)) || (rc = upload3(ctx, s.bt, bt0))) {
        *out = ctx;
        return rc;
    }
    FCG_CHECK(hipMemcpy(s.mask_wd, ones.data(), E * 8, hipMemcpyHostToDevice));
    FCG_CHECK(hipMemcpy(s.C_d_el, cd.data(), E * 8, hipMemcpyHostToDevice));
    if (p.wet_dry_on) {  // wad_mask at set-up (fv_main.F90:320-324)
        std::vector<double> mk(E);
        for (int e = 0; e < E; ++e) {
            const int4 q = en[e];
            double mn = std::min(std::min(std::min(depth[q.x], depth[q.y]), depth[q.z]), depth[q.w]);
            double mx = std::max(std::max(std::max(ea[q.x], ea[q.y]), ea[q.z]), ea[q.w]);
            mk[e] = (mn + mx > p.Dmin) ? 1.0 : 0.0;
        }
        FCG_CHECK(hipMemcpy(s.mask_wd, mk.data(), E * 8, hipMemcpyHostToDevice));
    }
    if (ctx->mixed)
        hipLaunchKernelGGL(fcgk::k_je<float>, dim3(nflatblk(E, L)), dim3(kFlatT), 8 * cpb_of(L) * sizeof(double),
                           ctx->stream, m, retype<float>(s));
    else
        hipLaunchKernelGGL(fcgk::k_je<double>, dim3(nflatblk(E, L)), dim3(kFlatT), 8 * cpb_of(L) * sizeof(double),
                           ctx->stream, m, s);
    FCG_CHECK(hipGetLastError());
    FCG_CHECK(hipStreamSynchronize(ctx->stream));
    ctx->time_jd = p.time_jd0;
    ctx->time = p.time_jd0 * 86400.0;
    ctx->n_dt = 0;
    ctx->lfirst = 1;
    *out = ctx;
    return FCG_OK;
#undef UP
#undef DA
#undef DA3
}

extern "C" int fcg_create(const fcg_mesh *mesh, const fcg_params *par, fcg_ctx **out) {
    return create_impl(mesh, par, nullptr, 1, out);
}

// Ensemble: the members' meshes joined into one disjoint mesh (nodes and
// elements concatenated in member order; internal edges of all members
// first, then their boundary edges, each member keeping its own order), so
// every stage is one launch over all members and each member's gathers run
// in its own reference order.
extern "C" int fcg_create_ensemble(const fcg_mesh *members, int nmemb, const fcg_params *par, fcg_ctx **out) {
    if (!members || nmemb < 1 || !par || !out) return FCG_E_ARG;
    *out = nullptr;
    const fcg_mesh &m0 = members[0];
    long N = 0, E = 0, D = 0, Din = 0, nob = 0;
    for (int b = 0; b < nmemb; ++b) {
        const fcg_mesh &mb = members[b];
        if (mb.nsigma != m0.nsigma || mb.density_0 != m0.density_0) return FCG_E_ARG;
        for (int k = 0; k < m0.nsigma; ++k)
            if (mb.sigma[k] != m0.sigma[k]) return FCG_E_ARG;
        if ((mb.coord_nod2D == nullptr) != (m0.coord_nod2D == nullptr)) return FCG_E_ARG;
        N += mb.nod2D;
        E += mb.elem2D;
        D += mb.edge2D;
        Din += mb.edge2D_in;
        nob += mb.nobn;
    }
    std::vector<int> en(4 * E), edn(2 * D), edt(2 * D), eled(4 * E), nbr(4 * E), idx(N), obn(nob), node_member(N);
    std::vector<double> depth(N), area(N), ac(N), atp(2 * N), coord(2 * N), ea(E), ec(E), cor(E), wcv(4 * E),
        gsca(8 * E), gvec(8 * E), dxdy(2 * D), cross(4 * D), ampt(12 * std::max(nob, 1L)), fazt(12 * std::max(nob, 1L));
    long no = 0, eo = 0, dio = 0, dbo = Din, oo = 0;
    for (int b = 0; b < nmemb; ++b) {
        const fcg_mesh &mb = members[b];
        const int n = mb.nod2D, e = mb.elem2D, d = mb.edge2D, di = mb.edge2D_in, ob = mb.nobn;
        auto uedge = [&](int d0) { return d0 < di ? dio + d0 : dbo + (d0 - di); };  // 0-based member edge -> union
        auto off = [](int v, long o) { return v > 0 ? (int)(v + o) : v; };          // 1-based id, <=0 kept
        for (int i = 0; i < n; ++i) {
            node_member[no + i] = b;
            idx[no + i] = mb.index_nod2D[i];
            depth[no + i] = mb.depth[i];
            area[no + i] = mb.area[i];
            ac[no + i] = mb.ac[i];
            atp[no + i] = mb.a_tp[i];
            atp[N + no + i] = mb.a_tp[n + i];
            if (mb.coord_nod2D) {
                coord[2 * (no + i)] = mb.coord_nod2D[2 * i];
                coord[2 * (no + i) + 1] = mb.coord_nod2D[2 * i + 1];
            }
        }
        for (int i = 0; i < e; ++i) {
            for (int j = 0; j < 4; ++j) {
                en[4 * (eo + i) + j] = off(mb.elem2D_nodes[4 * i + j], no);
                nbr[4 * (eo + i) + j] = off(mb.elem_neighbors[4 * i + j], eo);
                const int ed = mb.elem_edges[4 * i + j];
                eled[4 * (eo + i) + j] = ed > 0 ? (int)(uedge(ed - 1) + 1) : ed;
                wcv[4 * (eo + i) + j] = mb.w_cv[4 * i + j];
            }
            for (int j = 0; j < 8; ++j) {
                gsca[8 * (eo + i) + j] = mb.gradient_sca[8 * i + j];
                gvec[8 * (eo + i) + j] = mb.gradient_vec[8 * i + j];
            }
            ea[eo + i] = mb.elem_area[i];
            ec[eo + i] = mb.elem_cos[i];
            cor[eo + i] = mb.coriolis[i];
        }
        for (int i = 0; i < d; ++i) {
            const long u = uedge(i);
            for (int j = 0; j < 2; ++j) {
                edn[2 * u + j] = off(mb.edge_nodes[2 * i + j], no);
                edt[2 * u + j] = off(mb.edge_tri[2 * i + j], eo);
                dxdy[2 * u + j] = mb.edge_dxdy[2 * i + j];
            }
            for (int j = 0; j < 4; ++j) cross[4 * u + j] = mb.edge_cross_dxdy[4 * i + j];
        }
        for (int i = 0; i < ob; ++i) {
            obn[oo + i] = off(mb.in_obn[i], no);
            for (int j = 0; j < 12; ++j) {
                ampt[(size_t)j * nob + oo + i] = mb.ampt[(size_t)j * ob + i];
                fazt[(size_t)j * nob + oo + i] = mb.fazt[(size_t)j * ob + i];
            }
        }
        no += n;
        eo += e;
        dio += di;
        dbo += d - di;
        oo += ob;
    }
    fcg_mesh u = m0;
    u.nod2D = (int)N;
    u.elem2D = (int)E;
    u.edge2D = (int)D;
    u.edge2D_in = (int)Din;
    u.nobn = (int)nob;
    u.elem2D_nodes = en.data();
    u.edge_nodes = edn.data();
    u.edge_tri = edt.data();
    u.elem_edges = eled.data();
    u.elem_neighbors = nbr.data();
    u.index_nod2D = idx.data();
    u.in_obn = obn.data();
    u.depth = depth.data();
    u.area = area.data();
    u.elem_area = ea.data();
    u.elem_cos = ec.data();
    u.coriolis = cor.data();
    u.w_cv = wcv.data();
    u.gradient_sca = gsca.data();
    u.gradient_vec = gvec.data();
    u.edge_dxdy = dxdy.data();
    u.edge_cross_dxdy = cross.data();
    u.ac = ac.data();
    u.a_tp = atp.data();
    u.ampt = ampt.data();
    u.fazt = fazt.data();
    u.coord_nod2D = m0.coord_nod2D ? coord.data() : nullptr;
    std::vector<double> msa(nmemb);
    for (int b = 0; b < nmemb; ++b) msa[b] = members[b].scale_area;
    return create_impl(&u, par, node_member.data(), nmemb, out, msa.data());
}

extern "C" int fcg_ensemble_size(const fcg_ctx *ctx) { return ctx ? ctx->m.nmemb : 0; }

extern "C" void fcg_destroy(fcg_ctx *ctx) {
    if (!ctx) return;
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    if (ctx->bt_graph) (void)hipGraphExecDestroy(ctx->bt_graph);
    for (void *p : ctx->allocs) (void)hipFree(p);
    if (ctx->tide_pinned) (void)hipHostFree(ctx->tide_pinned);
    for (int b = 0; b < kTideRing; ++b)
        if (ctx->tide_ev[b]) (void)hipEventDestroy(ctx->tide_ev[b]);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

extern "C" int fcg_last_error(const fcg_ctx *ctx, char *buf, int len) {
    if (!buf || len <= 0) return FCG_E_ARG;
    const char *t = ctx ? ctx->err.c_str() : "no context";
    std::snprintf(buf, (size_t)len, "%s", t);
    return FCG_OK;
}

extern "C" int fcg_sync(fcg_ctx *ctx) {
    if (!ctx) return FCG_E_ARG;
    FCG_CHECK(hipStreamSynchronize(ctx->stream));
    return FCG_OK;
}

// Column quantities of the jacobian (src/fv_main.F90:1228-1252) that the device
// forms where they are consumed instead of storing them: Jc, Jc_old, zbar and
// Z, from the stored column depth a = max(depth+eta_n, Dmin) (a_old for Jc_old)
// and sigma, with the reference's expressions.  Read-only.
static int derived_get(fcg_ctx *ctx, const char *name, double *host, long count, bool &handled) {
    const bool jc = !std::strcmp(name, "Jc"), jco = !std::strcmp(name, "Jc_old");
    const bool zb = !std::strcmp(name, "zbar"), zz = !std::strcmp(name, "Z");
    handled = jc || jco || zb || zz;
    if (!handled) return FCG_OK;
    const long N = ctx->m.N, K = ctx->m.K, L = ctx->m.L;
    const long lev = zb ? K : L;
    if (count != lev * N) return fcg_fail(ctx, FCG_E_ARG, std::string("size mismatch for ") + name);
    FCG_CHECK(hipStreamSynchronize(ctx->stream));
    std::vector<double> a(N), sig(K);
    FCG_CHECK(hipMemcpy(a.data(), jco ? ctx->s.a_old : ctx->s.a, sizeof(double) * N, hipMemcpyDeviceToHost));
    FCG_CHECK(hipMemcpy(sig.data(), ctx->m.sigma, sizeof(double) * K, hipMemcpyDeviceToHost));
    std::vector<double> zbar(K);
    for (long i = 0; i < N; ++i) {
        double *h = host + (size_t)ctx->pn[i] * lev;
        const double ai = a[i];
        if (jc || jco) {
            for (long k = 0; k < L; ++k) h[k] = ai * (sig[k] - sig[k + 1]);
        } else {
            for (long k = 0; k < K; ++k) zbar[k] = (1.0 - sig[k]) * ai;
            if (zb)
                for (long k = 0; k < K; ++k) h[k] = zbar[k];
            else
                for (long k = 0; k < L; ++k) h[k] = 0.5 * (zbar[k] + zbar[k + 1]);
        }
    }
    return FCG_OK;
}

static int transfer(fcg_ctx *ctx, const char *name, double *host, long count, bool to_dev) {
    if (!ctx || !name || !host) return FCG_E_ARG;
    {
        bool handled = false;
        if (!to_dev) {
            const int rc = derived_get(ctx, name, host, count, handled);
            if (handled) return rc;
        } else if (!std::strcmp(name, "Jc") || !std::strcmp(name, "Jc_old") || !std::strcmp(name, "zbar")
                   || !std::strcmp(name, "Z")) {
            return fcg_fail(ctx, FCG_E_ARG, std::string(name) + " is derived from eta_n and read-only");
        }
    }
    Field f;
    if (!find_field(ctx, name, f)) return fcg_fail(ctx, FCG_E_ARG, std::string("unknown field ") + name);
    if (count != f.levels * f.ents)
        return fcg_fail(ctx, FCG_E_ARG, std::string("size mismatch for ") + name);
    FCG_CHECK(hipStreamSynchronize(ctx->stream));
    // device entity i holds reference entity perm[i]
    const std::vector<int> &perm = f.nodal ? ctx->pn : ctx->pe;
    // Fortran (levels, ents) column-major  <->  device: columns of K slots for
    // 3-D fields (levels = L or K), interleaved [ent][comp] for the (2,E)/(4,E)
    // 2-D vectors, plain arrays for 2-D scalars
    const bool col3 = f.r3 || f.levels == ctx->m.K;
    const long slots = col3 ? (long)ctx->m.K : f.levels;
    const size_t dcount = (size_t)slots * f.ents;
    auto dev_at = [&](long lev, long i) -> size_t { return (size_t)i * slots + lev; };
    std::vector<double> tmp(dcount, 0.0);
    const bool f32 = f.rtype && ctx->mixed;  // device holds float: round on upload, widen on download
    std::vector<float> tmpf(f32 ? dcount : 0);
    if (to_dev) {
        if (col3 && f.levels < slots) {  // keep the unused slot of layer columns as it is on the device
            if (f32) {
                FCG_CHECK(hipMemcpy(tmpf.data(), f.dev, sizeof(float) * dcount, hipMemcpyDeviceToHost));
            } else {
                FCG_CHECK(hipMemcpy(tmp.data(), f.dev, sizeof(double) * dcount, hipMemcpyDeviceToHost));
            }
            if (f32)
                for (size_t j = 0; j < dcount; ++j) tmp[j] = (double)tmpf[j];
        }
        for (long i = 0; i < f.ents; ++i)
            for (long k = 0; k < f.levels; ++k) tmp[dev_at(k, i)] = host[(size_t)perm[i] * f.levels + k];
        if (f32) {
            for (size_t j = 0; j < dcount; ++j) tmpf[j] = (float)tmp[j];
            FCG_CHECK(hipMemcpy(f.dev, tmpf.data(), sizeof(float) * dcount, hipMemcpyHostToDevice));
        } else {
            FCG_CHECK(hipMemcpy(f.dev, tmp.data(), sizeof(double) * dcount, hipMemcpyHostToDevice));
        }
    } else {
        if (f32) {
            FCG_CHECK(hipMemcpy(tmpf.data(), f.dev, sizeof(float) * dcount, hipMemcpyDeviceToHost));
            for (size_t j = 0; j < dcount; ++j) tmp[j] = (double)tmpf[j];
        } else {
            FCG_CHECK(hipMemcpy(tmp.data(), f.dev, sizeof(double) * dcount, hipMemcpyDeviceToHost));
        }
        for (long i = 0; i < f.ents; ++i)
            for (long k = 0; k < f.levels; ++k) host[(size_t)perm[i] * f.levels + k] = tmp[dev_at(k, i)];
    }
    return FCG_OK;
}

// ---- surface boundary conditions from atmospheric fields (key_atm) -----
extern "C" int fcg_sbc_set_coef(fcg_ctx *ctx, const double *coef_a, const double *coef_b) {
    if (!ctx || !coef_a || !coef_b) return FCG_E_ARG;
    if (!ctx->p.key_atm) return fcg_fail(ctx, FCG_E_ARG, "fcg_sbc_set_coef: the context runs without key_atm");
    const int N = ctx->m.N;
    if (!ctx->s.sbc_coef) {  // allocated with the first coefficients
        FCG_CHECK(hipSetDevice(ctx->device));
        int r = dalloc(ctx, &ctx->s.sbc_coef, 16 * (size_t)N);
        if (r) return r;
        r = dalloc(ctx, &ctx->s.tau_node, 2 * (size_t)N);
        if (r) return r;
    }
    // (8, nod2D) per array in mesh node order -> device order, a then b per node
    std::vector<double> h(16 * (size_t)N);
    for (int i = 0; i < N; ++i) {
        const size_t r = (size_t)ctx->pn[i];
        for (int f = 0; f < 8; ++f) {
            h[16 * (size_t)i + f] = coef_a[8 * r + f];
            h[16 * (size_t)i + 8 + f] = coef_b[8 * r + f];
        }
    }
    FCG_CHECK(hipMemcpyAsync(ctx->s.sbc_coef, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    FCG_CHECK(hipStreamSynchronize(ctx->stream));
    return FCG_OK;
}

template <typename R>
static int sbc_do_impl(fcg_ctx *ctx, const DevStateT<R> &s, double prec_coef, int net_flux) {
    LAUNCH(ctx, k_sbc_node<R>, nblk(ctx->m.N), ctx->m, s, ctx->time_jd, prec_coef, net_flux);
    LAUNCH(ctx, k_sbc_elem<R>, nblk(ctx->m.E), ctx->m, s);
    FCG_CHECK(hipGetLastError());
    return FCG_OK;
}

extern "C" int fcg_sbc_do(fcg_ctx *ctx, double nm_prec_coef, int nm_net_flux) {
    if (!ctx) return FCG_E_ARG;
    if (!ctx->p.key_atm || !ctx->s.sbc_coef)
        return fcg_fail(ctx, FCG_E_ARG, "fcg_sbc_do: no atmospheric coefficients (key_atm and fcg_sbc_set_coef)");
    if (!(nm_prec_coef != 0.0) || (nm_net_flux != 0 && nm_net_flux != 1))
        return fcg_fail(ctx, FCG_E_ARG, "fcg_sbc_do: nm_prec_coef must be nonzero and nm_net_flux 0 or 1");
    return ctx->mixed ? sbc_do_impl(ctx, retype<float>(ctx->s), nm_prec_coef, nm_net_flux)
                      : sbc_do_impl(ctx, ctx->s, nm_prec_coef, nm_net_flux);
}

// ---- open-boundary T/S climatology (key_obc) -----------------------------
extern "C" int fcg_obc_set_coef(fcg_ctx *ctx, int ndepth, const double *depth, const double *coef_a,
                                const double *coef_b) {
    if (!ctx || ndepth < 1 || !depth || !coef_a || !coef_b) return FCG_E_ARG;
    const int nb = ctx->m.nobn, N = ctx->m.N;
    if (nb == 0) return fcg_fail(ctx, FCG_E_ARG, "fcg_obc_set_coef: the mesh has no open-boundary nodes");
    for (int d = 1; d < ndepth; ++d)
        if (!(depth[d] > depth[d - 1]))
            return fcg_fail(ctx, FCG_E_ARG, "fcg_obc_set_coef: the depths must increase");
    FCG_CHECK(hipSetDevice(ctx->device));
    if (ctx->obc_ndep != ndepth) {  // (re)allocated for this profile length
        if (ctx->s.obc_coef) {
            FCG_CHECK(hipStreamSynchronize(ctx->stream));
            FCG_CHECK(hipFree(ctx->s.obc_coef));
            FCG_CHECK(hipFree(ctx->s.obc_depth));
            ctx->allocs.erase(std::remove(ctx->allocs.begin(), ctx->allocs.end(), (void *)ctx->s.obc_coef),
                              ctx->allocs.end());
            ctx->allocs.erase(std::remove(ctx->allocs.begin(), ctx->allocs.end(), (void *)ctx->s.obc_depth),
                              ctx->allocs.end());
            ctx->s.obc_coef = nullptr;
            ctx->s.obc_depth = nullptr;
        }
        int r = dalloc(ctx, &ctx->s.obc_coef, 4 * (size_t)nb * ndepth);
        if (r) return r;
        r = dalloc(ctx, &ctx->s.obc_depth, (size_t)ndepth);
        if (r) return r;
        ctx->obc_ndep = ndepth;
    }
    if (!ctx->s.obc_node) {
        std::vector<int> qn(N), node(nb);
        for (int i = 0; i < N; ++i) qn[ctx->pn[i]] = i;
        for (int i = 0; i < nb; ++i) node[i] = qn[ctx->in_obn0[i]];
        int r = dalloc(ctx, &ctx->s.obc_node, (size_t)nb);
        if (r) return r;
        FCG_CHECK(hipMemcpy(ctx->s.obc_node, node.data(), nb * sizeof(int), hipMemcpyHostToDevice));
    }
    const size_t n = 2 * (size_t)nb * ndepth;
    std::vector<double> h(2 * n);
    std::copy(coef_a, coef_a + n, h.begin());
    std::copy(coef_b, coef_b + n, h.begin() + n);
    FCG_CHECK(hipMemcpyAsync(ctx->s.obc_coef, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    FCG_CHECK(hipMemcpyAsync(ctx->s.obc_depth, depth, ndepth * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    FCG_CHECK(hipStreamSynchronize(ctx->stream));
    return FCG_OK;
}

template <typename R>
static int obc_do_impl(fcg_ctx *ctx, const DevStateT<R> &s, double time_jd) {
    const int items = ctx->m.nobn * ctx->m.L;
    FCG_CHECK(hipMemsetAsync(s.err + 3, 0, sizeof(int), ctx->stream));
    LAUNCH(ctx, k_obc_apply<R>, nblk(items), ctx->m, s, time_jd, ctx->obc_ndep, s.err + 3);
    int bad = 0;
    FCG_CHECK(hipMemcpyAsync(&bad, s.err + 3, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    FCG_CHECK(hipStreamSynchronize(ctx->stream));
    if (bad) return fcg_fail(ctx, FCG_E_ARG, "fcg_obc_do: a layer lies above the first depth of the climatology");
    return FCG_OK;
}

extern "C" int fcg_obc_do(fcg_ctx *ctx, double time_jd) {
    if (!ctx) return FCG_E_ARG;
    if (!ctx->s.obc_coef) return fcg_fail(ctx, FCG_E_ARG, "fcg_obc_do: no climatology coefficients (fcg_obc_set_coef)");
    return ctx->mixed ? obc_do_impl(ctx, retype<float>(ctx->s), time_jd) : obc_do_impl(ctx, ctx->s, time_jd);
}

extern "C" int fcg_set_state(fcg_ctx *ctx, const char *name, const double *host, long count) {
    return transfer(ctx, name, const_cast<double *>(host), count, true);
}
extern "C" int fcg_get_state(fcg_ctx *ctx, const char *name, double *host, long count) {
    return transfer(ctx, name, host, count, false);
}

extern "C" int fcg_set_clock(fcg_ctx *ctx, double time_jd, int lfirst, int n_dt) {
    if (!ctx) return FCG_E_ARG;
    ctx->time_jd = time_jd;
    ctx->time = time_jd * 86400.0;
    ctx->lfirst = lfirst;
    ctx->n_dt = n_dt;
    return FCG_OK;
}

extern "C" int fcg_advance_clock(fcg_ctx *ctx) {
    if (!ctx) return FCG_E_ARG;
    ctx->time_jd = ctx->time_jd + ctx->p.dt / 86400.0;  // src/fv_main.F90:400-401
    ctx->time = ctx->time_jd * 86400.0;
    ctx->n_dt++;
    return FCG_OK;
}

template <typename R>
static int pressure_impl(fcg_ctx *ctx, const DevStateT<R> &s) {
    const double gd0 = kG / ctx->density_0;
    const DevMesh &m = ctx->m;
    if (ctx->p.key_atm)
        LAUNCH_EX(ctx, k_pressure_atm<R>, nflatblk(m.N, m.L), kFlatT, flat_bytes(m.L, 8), m, s, gd0, ctx->density_0);
    else
        LAUNCH_EX(ctx, k_pressure<R>, nflatblk(m.N, m.L), kFlatT, flat_bytes(m.L, 8), m, s, gd0, ctx->density_0);
    LAUNCH_EX(ctx, k_bpg<R>, ncolblk(m.E), kColT, 2 * tile_bytes(m, 8), m, s);
    FCG_CHECK(hipGetLastError());
    return FCG_OK;
}

extern "C" int fcg_pressure(fcg_ctx *ctx) {
    if (!ctx) return FCG_E_ARG;
    return ctx->mixed ? pressure_impl(ctx, retype<float>(ctx->s)) : pressure_impl(ctx, ctx->s);
}

// d3_end (src/fv_mixing.F90:73-347): right-hand side, iterated column
// solve, closure values; `only` restricts it to flagged columns.
template <typename R>
static void d3end(fcg_ctx *ctx, const DevStateT<R> &s, const int *only) {
    const fcg_params &p = ctx->p;
    const DevMesh &m = ctx->m;
    LAUNCH_EX(ctx, k_d3end_prep<R>, nflatblk(m.N, m.K), kFlatT, 5 * flat_bytes(m.K, sizeof(R)), m, s, p.dt, p.Dmin,
              ctx->density_0, only);
    LAUNCH_EX(ctx, k_d3end<R>, (unsigned)((m.N + kD3endT - 1) / kD3endT), kD3endT,
              4 * (size_t)m.K * kD3endT * sizeof(R), m, s, p.dt, p.Dmin, p.z0b_min, only);
    LAUNCH_EX(ctx, k_d3end_fin<R>, nflatblk(m.N, m.K), kFlatT, flat_bytes(m.K, sizeof(R)), m, s, p.Dmin, p.z0b_min,
              p.z0s_min, p.snul, only);
}

template <typename R>
static int vertical_mixing_impl(fcg_ctx *ctx, const DevStateT<R> &s) {
    const fcg_params &p = ctx->p;
    // z0b_min/z0s_min/snul as fv_run.dat gives them
    // compute_el2nodes_3D(U_n, V_n, Unode, Vnode), src/fv_main.F90:2006
    const int be = block_of(ctx, "k_el2nodes");
    if (vec_on(ctx, "k_el2nodes", lds_el2n_v(ctx)))
        LAUNCH_NAMED(ctx, "k_el2nodes", (fcgk::k_el2nodes_v<R, false>), (ctx->m.N + vec_cpb(ctx) - 1) / vec_cpb(ctx),
                     256, lds_el2n_v(ctx), ctx->m, (const R *)s.U_n, (const R *)s.V_n, s.Unode, s.Vnode,
                     (const double *)nullptr, (const double *)nullptr, ctx->cap_n2el_v[1], s.err + 3);
    else
        LAUNCH_NAMED(ctx, "k_el2nodes", (fcgk::k_el2nodes<R, false>), nflatblk(ctx->m.N, ctx->m.L, be), be,
                     lds_el2n(ctx, be), ctx->m, (const R *)s.U_n, (const R *)s.V_n, s.Unode, s.Vnode,
                     (const double *)nullptr, (const double *)nullptr);
    if (p.closure == FCG_CLOSURE_MLP) {
        // learned closure; columns outside its training box go back to d3_end
        if (!ctx->mlp) return fcg_fail(ctx, FCG_E_ARG, "closure = MLP but no network attached (fcg_set_mlp)");
        const fcg_mlp *net = ctx->mlp;
        // persistent blocks, one per compute unit (the network's fragments fill
        // most of a unit's LDS); each wavefront takes whole columns
        if (ctx->m.K < kMlpMinK || ctx->m.K > kMlpMaxK)
            return fcg_fail(ctx, FCG_E_ARG, "the learned closure needs 4 <= nsigma <= 64");
        const int cpw = 64 / ctx->m.K;
        const int groups = (ctx->m.N + cpw - 1) / cpw;
        const int blocks = std::min((groups + kMlpT / 64 - 1) / (kMlpT / 64), ctx->n_cu);
        LAUNCH_EX(ctx, k_mlp_closure<R>, blocks, kMlpT, kMlpClosureLds, ctx->m, s,
                  (const fcgm::bf16x8 *)net->frag, (const float *)net->bias, net->norm, p.Dmin, ctx->density_0,
                  p.z0b_min, p.snul, ctx->mlp_flag);
        d3end(ctx, s, (const int *)ctx->mlp_flag);
        hipLaunchKernelGGL(fcgk::k_count_flags, dim3(64), dim3(kBlk), 0, ctx->stream, (const int *)ctx->mlp_flag,
                           ctx->m.N, ctx->mlp_count);
    } else {
        d3end(ctx, s, nullptr);
    }
    FCG_CHECK(hipGetLastError());
    return FCG_OK;
}

extern "C" int fcg_vertical_mixing(fcg_ctx *ctx) {
    if (!ctx) return FCG_E_ARG;
    return ctx->mixed ? vertical_mixing_impl(ctx, retype<float>(ctx->s)) : vertical_mixing_impl(ctx, ctx->s);
}

template <typename R>
static int compute_vel_rhs_impl(fcg_ctx *ctx, const DevStateT<R> &s) {
    const fcg_params &p = ctx->p;
    double ff = 1.5 + kEps;
    if (ctx->lfirst) {  // src/fv_3D_vel.F90:124-128
        ff = 1.0;
        ctx->lfirst = 0;
    }
    if (vec_on(ctx, "k_velgrad3d", lds_velgrad_v(ctx)))
        LAUNCH_NAMED(ctx, "k_velgrad3d", fcgk::k_velgrad3d_v<R>, (ctx->m.E + vec_cpb(ctx) - 1) / vec_cpb(ctx), 256,
                     lds_velgrad_v(ctx), ctx->m, s);
    else
        LAUNCH_EX(ctx, k_velgrad3d<R>, nflatblk(ctx->m.E, ctx->m.L), kFlatT, cpb_of(ctx->m.L) * sizeof(VelGradRec),
                  ctx->m, s);
    if (p.bih_3D) {  // src/fv_3D_vel.F90:113, the first pass, Ah and the gradients of U_c; k_velrhs
                     // adds the second pass's edge fluxes after the filters (fv_viscosity.F90:104-336)
        const DevMesh &m = ctx->m;
        const unsigned gd = (unsigned)(((size_t)m.D * m.L + kFlatT - 1) / kFlatT);
        const unsigned ge = (unsigned)(((size_t)m.E * m.L + kFlatT - 1) / kFlatT);
        LAUNCH_EX(ctx, k_bh3_edge<R>, gd, kFlatT, 0, m, s, s.U_n, s.V_n, s.gux, s.guy, s.gvx, s.gvy, 1, 0);
        LAUNCH_EX(ctx, k_bh3_elem<R>, ge, kFlatT, 0, m, s, s.bhcu, s.bhcv, 1);
        LAUNCH_EX(ctx, k_bh3_grad<R>, ge, kFlatT, 0, m, s, s.bhcu, s.bhcv, 0);
        LAUNCH_EX(ctx, k_bh3_edge<R>, gd, kFlatT, 0, m, s, s.bhcu, s.bhcv, s.bhgux, s.bhguy, s.bhgvx, s.bhgvy, 1, 1);
    }
    if (p.filt_bi_3D) {  // src/fv_3D_vel.F90:112 (tau_inv: fv_viscosity.F90:586); UV_rhs enters k_velrhs
        const unsigned g = (unsigned)(((size_t)ctx->m.E * ctx->m.L + kFlatT - 1) / kFlatT);  // dense (element, layer)
        LAUNCH_EX(ctx, k_filt2x_3d_c<R>, g, kFlatT, 0, ctx->m, s, p.dt * p.tau_c / 3600.0 / 24);
        LAUNCH_EX(ctx, k_filt2x_3d_uv<R>, g, kFlatT, 0, ctx->m, s);
    }
    const int vt = block_of(ctx, "k_velrhs");
    constexpr int W = 16 / (int)sizeof(R);
    if (levels_of(ctx, "k_velrhs", W) == W &&
        velrhs_v_bytes(ctx->m, sizeof(R), velrhs_v_threads(ctx->m, sizeof(R))) <= (size_t)ctx->lds_limit) {  // W layers per lane
        const int tv = velrhs_v_threads(ctx->m, sizeof(R)), cpb = velrhs_v_cpb(ctx->m, sizeof(R), tv);
        const unsigned g = (unsigned)((ctx->m.E + cpb - 1) / cpb);
        const size_t lds = velrhs_v_bytes(ctx->m, sizeof(R), tv);
        const int hadv = p.mom_adv_3D == 1 ? 1 : 0, vadv = p.vert_adv == 1 ? 1 : 0;
        if (p.filt_3D && p.key_atm)
            LAUNCH_NAMED(ctx, "k_velrhs", fcgk::k_velrhs_v_atm_filt<R>, g, tv, lds, ctx->m, s, p.dt, ff, hadv, vadv,
                         1.0 / ctx->density_0, p.tau_inv_filt);
        else if (p.filt_3D)
            LAUNCH_NAMED(ctx, "k_velrhs", fcgk::k_velrhs_v_filt<R>, g, tv, lds, ctx->m, s, p.dt, ff, hadv, vadv,
                         1.0 / ctx->density_0, p.tau_inv_filt);
        else if (p.key_atm)
            LAUNCH_NAMED(ctx, "k_velrhs", fcgk::k_velrhs_v_atm<R>, g, tv, lds, ctx->m, s, p.dt, ff, hadv, vadv,
                         1.0 / ctx->density_0);
        else
            LAUNCH_NAMED(ctx, "k_velrhs", fcgk::k_velrhs_v<R>, g, tv, lds, ctx->m, s, p.dt, ff, hadv, vadv,
                         1.0 / ctx->density_0);
    } else if (p.filt_3D && p.key_atm)
        LAUNCH_EX(ctx, k_velrhs_atm_filt<R>, nflatblk(ctx->m.E, ctx->m.L, vt), vt, velrhs_bytes(ctx->m, sizeof(R), vt),
                  ctx->m, s, p.dt, ff, p.mom_adv_3D == 1 ? 1 : 0, p.vert_adv == 1 ? 1 : 0, 1.0 / ctx->density_0,
                  p.tau_inv_filt);
    else if (p.filt_3D)
        LAUNCH_EX(ctx, k_velrhs_filt<R>, nflatblk(ctx->m.E, ctx->m.L, vt), vt, velrhs_bytes(ctx->m, sizeof(R), vt),
                  ctx->m, s, p.dt, ff, p.mom_adv_3D == 1 ? 1 : 0, p.vert_adv == 1 ? 1 : 0, 1.0 / ctx->density_0,
                  p.tau_inv_filt);
    else if (p.key_atm)
        LAUNCH_EX(ctx, k_velrhs_atm<R>, nflatblk(ctx->m.E, ctx->m.L, vt), vt, velrhs_bytes(ctx->m, sizeof(R), vt),
                  ctx->m, s, p.dt, ff, p.mom_adv_3D == 1 ? 1 : 0, p.vert_adv == 1 ? 1 : 0, 1.0 / ctx->density_0);
    else
        LAUNCH_EX(ctx, k_velrhs<R>, nflatblk(ctx->m.E, ctx->m.L, vt), vt, velrhs_bytes(ctx->m, sizeof(R), vt), ctx->m,
                  s, p.dt, ff, p.mom_adv_3D == 1 ? 1 : 0, p.vert_adv == 1 ? 1 : 0, 1.0 / ctx->density_0);
    if (p.im_vert_visc)  // src/fv_3D_vel.F90:183
        LAUNCH_EX(ctx, k_velvisc_im<R>, ncolblk(ctx->m.E), kColT,
                  tile_bytes(ctx->m, 8) + 2 * tile_bytes(ctx->m, sizeof(R)), ctx->m, s, p.dt, ctx->density_0,
                  p.Cd_bz);
    if (p.im_vert_visc_fic)  // src/fv_3D_vel.F90:184-186
        LAUNCH_EX(ctx, k_velvisc<R>, ncolblk(ctx->m.E), kColT, tile_bytes(ctx->m, 8) + 2 * tile_bytes(ctx->m, sizeof(R)),
                  ctx->m, s, p.dt, ctx->density_0, p.Cd_bz, p.fic_point);
    FCG_CHECK(hipGetLastError());
    return FCG_OK;
}

extern "C" int fcg_compute_vel_rhs(fcg_ctx *ctx) {
    if (!ctx) return FCG_E_ARG;
    return ctx->mixed ? compute_vel_rhs_impl(ctx, retype<float>(ctx->s)) : compute_vel_rhs_impl(ctx, ctx->s);
}

template <typename R>
static int barotropic_impl(fcg_ctx *ctx, const DevStateT<R> &s) {
    const fcg_params &p = ctx->p;
    if (p.tf_presence && ctx->m.nobn) {
        // the step's open-boundary values go through a ring of pinned staging
        // buffers; a buffer is refilled once the copy issued from it kTideRing
        // steps ago has completed, so the host never drains the stream
        const int b = ctx->tide_slot;
        ctx->tide_slot = (b + 1) % kTideRing;
        double *buf = ctx->tide_pinned + (size_t)b * p.Mt * ctx->m.nobn;
        if (ctx->tide_ev_used[b]) FCG_CHECK(hipEventSynchronize(ctx->tide_ev[b]));
        fill_tides(ctx, buf);
        FCG_CHECK(hipMemcpyAsync(s.tide, buf, sizeof(double) * p.Mt * ctx->m.nobn, hipMemcpyHostToDevice,
                                 ctx->stream));
        FCG_CHECK(hipEventRecord(ctx->tide_ev[b], ctx->stream));
        ctx->tide_ev_used[b] = true;
    }
    if (!ctx->bt_graph_ok) {
        hipGraph_t g = nullptr;
        FCG_CHECK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        enqueue_substeps(ctx, s);
        FCG_CHECK(hipStreamEndCapture(ctx->stream, &g));
        FCG_CHECK(hipGraphInstantiate(&ctx->bt_graph, g, nullptr, nullptr, 0));
        FCG_CHECK(hipGraphDestroy(g));
        ctx->bt_graph_ok = true;
    }
    FCG_CHECK(hipGraphLaunch(ctx->bt_graph, ctx->stream));
    FCG_CHECK(hipGetLastError());
    return FCG_OK;
}

extern "C" int fcg_barotropic_timestep(fcg_ctx *ctx) {
    if (!ctx) return FCG_E_ARG;
    return ctx->mixed ? barotropic_impl(ctx, retype<float>(ctx->s)) : barotropic_impl(ctx, ctx->s);
}

// the element Jacobians (jacobian, src/fv_main.F90:1240-1252), level vectors where possible
template <typename R>
static void je_launch(fcg_ctx *ctx, const DevStateT<R> &s) {
    if (vec_on(ctx, "k_je", 0))
        LAUNCH_NAMED(ctx, "k_je", fcgk::k_je_v<R>, (ctx->m.E + vec_cpb(ctx) - 1) / vec_cpb(ctx), 256, 0, ctx->m, s);
    else
        LAUNCH_EX(ctx, k_je<R>, nflatblk(ctx->m.E, ctx->m.L), kFlatT, 8 * cpb_of(ctx->m.L) * sizeof(double), ctx->m, s);
}

template <typename R>
static int jacobian_impl(fcg_ctx *ctx, const DevStateT<R> &s) {
    LAUNCH1(ctx, k_jac_node<R>, ctx->m.N, ctx->m, s, ctx->p.Dmin);
    je_launch(ctx, s);
    FCG_CHECK(hipGetLastError());
    return FCG_OK;
}

extern "C" int fcg_jacobian(fcg_ctx *ctx) {
    if (!ctx) return FCG_E_ARG;
    return ctx->mixed ? jacobian_impl(ctx, retype<float>(ctx->s)) : jacobian_impl(ctx, ctx->s);
}

template <typename R>
static int hot_start_impl(fcg_ctx *ctx, const DevStateT<R> &s) {
    hipLaunchKernelGGL(fcgk::k_jac_restart<R>, dim3(nblk(ctx->m.N)), dim3(kBlk), 0, ctx->stream, ctx->m, s, ctx->p.Dmin);
    je_launch(ctx, s);
    FCG_CHECK(hipGetLastError());
    FCG_CHECK(hipStreamSynchronize(ctx->stream));
    return FCG_OK;
}

extern "C" int fcg_hot_start(fcg_ctx *ctx, double time_jd) {
    if (!ctx) return FCG_E_ARG;
    const int rc = ctx->mixed ? hot_start_impl(ctx, retype<float>(ctx->s)) : hot_start_impl(ctx, ctx->s);
    if (rc) return rc;
    return fcg_set_clock(ctx, time_jd, 0, 0);  // read_restart_separate: lfirst = .false.
}

template <typename R>
static int update_3d_vel_impl(fcg_ctx *ctx, const DevStateT<R> &s) {
    const size_t lv = (2 * (size_t)ctx->m.K + 4) * vec_cpb(ctx) * sizeof(double);
    if (vec_on(ctx, "k_update3d", lv))
        LAUNCH_NAMED(ctx, "k_update3d", fcgk::k_update3d_v<R>, (ctx->m.E + vec_cpb(ctx) - 1) / vec_cpb(ctx), 256, lv,
                     ctx->m, s, ctx->p.Dmin, 1.0 / (double)ctx->p.Mt);
    else
        LAUNCH_EX(ctx, k_update3d<R>, nflatblk(ctx->m.E, ctx->m.L), kFlatT,
                  2 * flat_bytes(ctx->m.L, 8) + 14 * (size_t)cpb_of(ctx->m.L) * 8, ctx->m, s, ctx->p.Dmin,
                  1.0 / (double)ctx->p.Mt);
    FCG_CHECK(hipGetLastError());
    return FCG_OK;
}

extern "C" int fcg_update_3d_vel(fcg_ctx *ctx) {
    if (!ctx) return FCG_E_ARG;
    return ctx->mixed ? update_3d_vel_impl(ctx, retype<float>(ctx->s)) : update_3d_vel_impl(ctx, ctx->s);
}

template <typename R>
static int vert_vel_sigma_impl(fcg_ctx *ctx, const DevStateT<R> &s) {
    const int bv = block_of(ctx, "k_vertvel_div");
    if (vec_on(ctx, "k_vertvel_div", lds_vdiv_v(ctx)))
        LAUNCH_NAMED(ctx, "k_vertvel_div", fcgk::k_vertvel_div_v<R>, (ctx->m.N + vec_cpb(ctx) - 1) / vec_cpb(ctx), 256,
                     lds_vdiv_v(ctx), ctx->m, s, ctx->cap_n2e_v[1]);
    else
        LAUNCH_EX(ctx, k_vertvel_div<R>, nflatblk(ctx->m.N, ctx->m.L, bv), bv, lds_vdiv<R>(ctx, bv), ctx->m, s);
    LAUNCH_EX(ctx, k_vertvel_scan<R>, nflatblk(ctx->m.N, ctx->m.L), kFlatT, flat_bytes(ctx->m.L, sizeof(R)), ctx->m, s);
    FCG_CHECK(hipGetLastError());
    return FCG_OK;
}

extern "C" int fcg_vert_vel_sigma(fcg_ctx *ctx) {
    if (!ctx) return FCG_E_ARG;
    return ctx->mixed ? vert_vel_sigma_impl(ctx, retype<float>(ctx->s)) : vert_vel_sigma_impl(ctx, ctx->s);
}

template <typename R>
static int solve_tracers_impl(fcg_ctx *ctx, const DevStateT<R> &s) {
    const fcg_params &p = ctx->p;
    const DevMesh &m = ctx->m;
    const size_t NL = (size_t)m.N * m.K;  // every slot of the tracer columns
    // T_aver/S_aver per ensemble member (one member outside an ensemble)
    const int nb = kMinmaxChunks;
    hipLaunchKernelGGL(fcgk::k_minmax_partial<R>, dim3(nb, m.nmemb), dim3(kBlk), 0, ctx->stream, m, (const R *)s.TF,
                       (const R *)s.SF, s.red);
    hipLaunchKernelGGL(fcgk::k_minmax_final, dim3(m.nmemb), dim3(kBlk), 0, ctx->stream, (const double *)s.red,
                       nb, m.nmemb, s.tsa);
    const int relax = p.cl_relax ? 1 : 0;
    // scheme dispatch of src/fv_main.F90:1788-1800; node x layer kernels, then
    // the column-sequential implicit vertical diffusion
    if (p.tracer_adv == 1) {
        LAUNCH1(ctx, k_tr_ab<R>, NL, m, s, NL);
        LAUNCHLV(ctx, k_tr_upwind<R>, m.N, m.L, m, s, p.dt, p.Dmin, relax, p.clim_relax, p.relax2clim_ac);
    } else if (p.tracer_adv == 2 && p.muscl_call) {
        LAUNCH1(ctx, k_tr_ab<R>, NL, m, s, NL);
        const int bg = block_of(ctx, "k_tr_gradn_muscl"), bm = block_of(ctx, "k_tr_muscl");
        if (vec_on(ctx, "k_tr_gradn_muscl", lds_gradn_v(ctx)))
            LAUNCH_NAMED(ctx, "k_tr_gradn_muscl", fcgk::k_tr_gradn_muscl_v<R>, (m.N + vec_cpb(ctx) - 1) / vec_cpb(ctx),
                         256, lds_gradn_v(ctx), m, s, ctx->cap_n2el_v[1]);
        else
            LAUNCH_EX(ctx, k_tr_gradn_muscl<R>, nflatblk(m.N, m.L, bg), bg, lds_gradn(ctx, bg), m, s);
        constexpr int W = 16 / (int)sizeof(R);
        const int tv = vec_threads(ctx, lds_muscl_v);
        if (levels_of(ctx, "k_tr_muscl", W) == W && lds_muscl_v(ctx, tv) <= (size_t)ctx->lds_limit) {
            const int cpb = tv / (m.K / W);
            LAUNCH_NAMED(ctx, "k_tr_muscl", fcgk::k_tr_muscl_v<R>, (unsigned)((m.N + cpb - 1) / cpb), tv,
                         lds_muscl_v(ctx, tv), m, s, p.dt, p.Dmin, relax, p.clim_relax, p.relax2clim_ac,
                         ctx->cap_n2e_v[tv == 256 ? 1 : 0]);
        } else {
            LAUNCH_EX(ctx, k_tr_muscl<R>, nflatblk(m.N, m.L, bm), bm, lds_muscl<R>(ctx, bm), m, s, p.dt, p.Dmin, relax,
                      p.clim_relax, p.relax2clim_ac);
        }
    } else if (p.tracer_adv == 4) {
        LAUNCH1(ctx, k_tr_ab<R>, NL, m, s, NL);
        LAUNCHLV(ctx, k_tr_gradn_miura<R>, m.N, m.L, m, s);
        LAUNCHLV(ctx, k_tr_miura<R>, m.N, m.L, m, s, p.dt, p.Dmin, ctx->scale_area, relax, p.clim_relax, p.relax2clim_ac);
    }
    if (p.key_rivers) LAUNCH1(ctx, k_tr_rivers<R>, m.N, m, s, p.dt);
    AtmPar ap;
    ap.on = p.key_atm;
    ap.type_swr = p.type_swr_body;
    ap.wet_dry = p.wet_dry_on;
    ap.Rjc = p.Rjc;
    ap.ajc = p.ajc;
    ap.bjc = p.bjc;
    ap.rho0 = ctx->density_0;
    ap.Dmin = p.Dmin;
    if (p.key_atm)
        LAUNCH_EX(ctx, k_tr_vdiff_atm<R>, ncolblk(m.N), kColT, 5 * tile_bytes(m, sizeof(R)), m, s, p.dt, ap);
    else
        LAUNCH_EX(ctx, k_tr_vdiff<R>, ncolblk(m.N), kColT, 5 * tile_bytes(m, sizeof(R)), m, s, p.dt, ap);
    FCG_CHECK(hipGetLastError());
    return FCG_OK;
}

extern "C" int fcg_solve_tracers(fcg_ctx *ctx) {
    if (!ctx) return FCG_E_ARG;
    return ctx->mixed ? solve_tracers_impl(ctx, retype<float>(ctx->s)) : solve_tracers_impl(ctx, ctx->s);
}

extern "C" int fcg_oce_timestep(fcg_ctx *ctx) {
    TRY(fcg_advance_clock(ctx));
    TRY(fcg_pressure(ctx));
    TRY(fcg_vertical_mixing(ctx));
    TRY(fcg_compute_vel_rhs(ctx));
    TRY(fcg_barotropic_timestep(ctx));
    TRY(fcg_jacobian(ctx));
    TRY(fcg_update_3d_vel(ctx));
    TRY(fcg_vert_vel_sigma(ctx));
    TRY(fcg_solve_tracers(ctx));
    return FCG_OK;
}

extern "C" int fcg_run(fcg_ctx *ctx, int nsteps) {
    if (!ctx || nsteps < 0) return FCG_E_ARG;
    for (int i = 0; i < nsteps; ++i) TRY(fcg_oce_timestep(ctx));
    return FCG_OK;
}

template <typename R>
static int diagnostics_impl(fcg_ctx *ctx, const DevStateT<R> &s, fcg_diag *out, int member) {
    const DevMesh &m = ctx->m;
    int nb = ctx->red_blocks;
    // nodes/elements of one member, or of all
    const int n0 = member < 0 ? 0 : ctx->noff[member], n1 = member < 0 ? m.N : ctx->noff[member + 1];
    const int e0 = member < 0 ? 0 : ctx->eoff[member], e1 = member < 0 ? m.E : ctx->eoff[member + 1];
    hipLaunchKernelGGL(fcgk::k_diag_partial<R>, dim3(nb), dim3(kBlk), 0, ctx->stream, m, s, ctx->p.Dmin, n0, n1, e0,
                       e1, s.red);
    hipLaunchKernelGGL(fcgk::k_diag_final, dim3(1), dim3(64), 0, ctx->stream, (const double *)s.red, nb, s.scal + 4);
    FCG_CHECK(hipGetLastError());
    double r[12];
    int wet = 0, err = 0, nonfinite = 0, shape = 0;
    FCG_CHECK(hipMemsetAsync(s.err + 1, 0, sizeof(int), ctx->stream));
    hipLaunchKernelGGL(fcgk::k_mask_count, dim3(64), dim3(kBlk), 0, ctx->stream, (const double *)s.mask_wd + e0,
                       e1 - e0, s.err + 1);
    FCG_CHECK(hipMemcpyAsync(r, s.scal + 4, sizeof r, hipMemcpyDeviceToHost, ctx->stream));
    FCG_CHECK(hipMemcpyAsync(&err, s.err, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    FCG_CHECK(hipMemcpyAsync(&wet, s.err + 1, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    FCG_CHECK(hipMemcpyAsync(&nonfinite, s.err + 2, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    FCG_CHECK(hipMemcpyAsync(&shape, s.err + 3, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    FCG_CHECK(hipStreamSynchronize(ctx->stream));
    // a node-gather block that found more CSR slots than its LDS was sized
    // for skipped its work: a set-up fault of this library, never a result
    if (shape) return fcg_fail(ctx, FCG_E_HIP, "a node-gather block exceeded its LDS capacity (launch shape)");
    std::memset(out, 0, sizeof *out);
    out->energy = (r[0] + r[2]) / r[1];
    out->volume = r[3];
    out->heat = r[4];
    out->salt = r[5];
    out->eta_min = r[6];
    out->eta_max = r[7];
    out->T_min = r[8];
    out->T_max = r[9];
    out->S_min = r[10];
    out->S_max = r[11];
    out->time_jd = ctx->time_jd;
    out->n_dt = ctx->n_dt;
    out->n_wet_elem = wet;
    // the reference's STOP conditions first (a NaN passes its tests), then a
    // non-finite elevation or tracer
    out->status = err ? err : (nonfinite ? FCG_E_NAN : FCG_OK);
    if (ctx->mlp_count) {
        unsigned long long fb = 0;
        FCG_CHECK(hipMemcpy(&fb, ctx->mlp_count, sizeof fb, hipMemcpyDeviceToHost));
        out->mlp_fallback = fb;
    }
    return FCG_OK;
}

extern "C" int fcg_diagnostics(fcg_ctx *ctx, fcg_diag *out) {
    if (!ctx || !out) return FCG_E_ARG;
    return ctx->mixed ? diagnostics_impl(ctx, retype<float>(ctx->s), out, -1) : diagnostics_impl(ctx, ctx->s, out, -1);
}

extern "C" int fcg_diagnostics_member(fcg_ctx *ctx, int member, fcg_diag *out) {
    if (!ctx || !out || member < 0 || member >= ctx->m.nmemb) return FCG_E_ARG;
    return ctx->mixed ? diagnostics_impl(ctx, retype<float>(ctx->s), out, member)
                      : diagnostics_impl(ctx, ctx->s, out, member);
}

extern "C" int fcg_set_mlp(fcg_ctx *ctx, fcg_mlp *net) {
    if (!ctx) return FCG_E_ARG;
    if (ctx->p.closure != FCG_CLOSURE_MLP) return fcg_fail(ctx, FCG_E_ARG, "context was created with closure = d3_end");
    if (net && kMlpClosureLds > (size_t)ctx->lds_limit)
        return fcg_fail(ctx, FCG_E_ARG, "the learned closure's LDS exceeds the device limit");
    if (net && (ctx->m.K < kMlpMinK || ctx->m.K > kMlpMaxK))
        return fcg_fail(ctx, FCG_E_ARG, "the learned closure needs 4 <= nsigma <= 64");
    ctx->mlp = net;
    return FCG_OK;
}

template <typename R>
static int mlp_collect_impl(fcg_ctx *ctx, const DevStateT<R> &s, float *X, float *T, long cap, int stride,
                            long *count) {
    const long cols = (ctx->m.N + stride - 1) / stride;
    const long n = std::min<long>(cap, cols * ctx->m.K);
    if (n > 0)
        hipLaunchKernelGGL(fcgk::k_mlp_collect<R>, dim3((unsigned)((n + kBlk - 1) / kBlk)), dim3(kBlk), 0,
                           ctx->stream, ctx->m, s, ctx->p.Dmin, ctx->density_0, ctx->p.z0b_min, stride, n, X, T);
    FCG_CHECK(hipGetLastError());
    FCG_CHECK(hipStreamSynchronize(ctx->stream));
    *count = n;
    return FCG_OK;
}

extern "C" int fcg_mlp_collect(fcg_ctx *ctx, float *X, float *T, long cap, int stride, long *count) {
    if (!ctx || !X || !T || !count || cap < 0 || stride < 1) return FCG_E_ARG;
    return ctx->mixed ? mlp_collect_impl(ctx, retype<float>(ctx->s), X, T, cap, stride, count)
                      : mlp_collect_impl(ctx, ctx->s, X, T, cap, stride, count);
}

// Launch shapes for the autotuner: block size 64, 128 or 256 threads (for
// the layer-parallel kernels: 64 entities x 1, 2 or 4 layers per block); for
// k_velrhs and the node-gather kernels 256, 512 or 1024 threads of whole
// columns, when their LDS fits.
extern "C" int fcg_set_launch(fcg_ctx *ctx, const char *kernel, int block) {
    if (!ctx || !kernel) return FCG_E_ARG;
    static const char *gather[] = {"k_tr_muscl", "k_tr_gradn_muscl", "k_vertvel_div", "k_el2nodes"};
    for (const char *g : gather)
        if (!std::strcmp(g, kernel)) {
            if (!(block == 256 || block == 512 || block == 1024)) return FCG_E_ARG;
            const size_t lds = std::max(std::max(ctx->mixed ? lds_muscl<float>(ctx, block) : lds_muscl<double>(ctx, block),
                                                 ctx->mixed ? lds_vdiv<float>(ctx, block) : lds_vdiv<double>(ctx, block)),
                                        std::max(lds_gradn(ctx, block), lds_el2n(ctx, block)));
            if (lds > (size_t)ctx->lds_limit)
                return fcg_fail(ctx, FCG_E_ARG, std::string(kernel) + ": LDS of this block size exceeds the device limit");
            ctx->launch_block[kernel] = block;
            if (!std::strcmp(kernel, "k_el2nodes") && ctx->bt_graph) {  // captured in the barotropic graph
                FCG_CHECK(hipStreamSynchronize(ctx->stream));
                FCG_CHECK(hipGraphExecDestroy(ctx->bt_graph));
                ctx->bt_graph = nullptr;
                ctx->bt_graph_ok = false;
            }
            return FCG_OK;
        }
    {  // levels per lane: 1, or W = 16/(storage width) when nsigma is a multiple of W
        const char *colon = std::strchr(kernel, ':');
        if (colon && !std::strcmp(colon, ":levels")) {
            const std::string k(kernel, colon - kernel);
            const int W = 16 / (int)ctx->esize;
            static const char *vec[] = {"k_velrhs", "k_tr_muscl", "k_el2nodes", "k_vertvel_div", "k_tr_gradn_muscl",
                                        "k_velgrad3d", "k_je", "k_update3d"};
            bool known = false;
            for (const char *v : vec) known = known || k == v;
            if (!known) return fcg_fail(ctx, FCG_E_ARG, "kernel without a level-vector variant: " + k);
            if (!(block == 1 || block == W)) return FCG_E_ARG;
            if (block == W && ctx->m.K % W)
                return fcg_fail(ctx, FCG_E_ARG, k + ": " + std::to_string(W) + " levels per lane need nsigma to be a multiple of it");
            const size_t lds = k == "k_velrhs"   ? velrhs_v_bytes(ctx->m, ctx->esize, velrhs_v_threads(ctx->m, ctx->esize))
                               : k == "k_tr_muscl" ? lds_muscl_v(ctx, vec_threads(ctx, lds_muscl_v))
                               : k == "k_el2nodes" ? lds_el2n_v(ctx)
                               : k == "k_vertvel_div" ? lds_vdiv_v(ctx)
                               : k == "k_tr_gradn_muscl" ? lds_gradn_v(ctx)
                               : k == "k_velgrad3d" ? lds_velgrad_v(ctx)
                               : k == "k_update3d" ? (2 * (size_t)ctx->m.K + 4) * vec_cpb(ctx) * sizeof(double)
                                                    : 0;
            if (block == W && lds > (size_t)ctx->lds_limit)
                return fcg_fail(ctx, FCG_E_ARG, k + ": LDS of the level-vector variant exceeds the device limit");
            ctx->launch_block[kernel] = block;
            if (k == "k_el2nodes" && ctx->bt_graph) {  // captured in the barotropic graph
                FCG_CHECK(hipStreamSynchronize(ctx->stream));
                FCG_CHECK(hipGraphExecDestroy(ctx->bt_graph));
                ctx->bt_graph = nullptr;
                ctx->bt_graph_ok = false;
            }
            return FCG_OK;
        }
    }
    if (!std::strcmp(kernel, "k_velrhs")) {
        if (!(block == 256 || block == 512 || block == 1024)) return FCG_E_ARG;
        if (velrhs_bytes(ctx->m, ctx->esize, block) > (size_t)ctx->lds_limit)
            return fcg_fail(ctx, FCG_E_ARG, "k_velrhs: LDS of this block size exceeds the device limit");
        ctx->launch_block[kernel] = block;
        return FCG_OK;
    }
    if (!(block == 64 || block == 128 || block == 256)) return FCG_E_ARG;
    static const char *tunable[] = {"k_jac_node", "k_tr_ab", "k_tr_upwind", "k_tr_gradn_miura", "k_tr_miura"};
    for (const char *t : tunable)
        if (!std::strcmp(t, kernel)) {
            ctx->launch_block[kernel] = block;
            return FCG_OK;
        }
    return fcg_fail(ctx, FCG_E_ARG, std::string("kernel without a tunable launch shape: ") + kernel);
}

extern "C" int fcg_get_launch(const fcg_ctx *ctx, const char *kernel) {
    if (!ctx || !kernel) return 0;
    const char *colon = std::strchr(kernel, ':');
    if (colon && !std::strcmp(colon, ":levels"))
        return levels_of(ctx, std::string(kernel, colon - kernel).c_str(), 16 / (int)ctx->esize);
    return block_of(ctx, kernel);
}
