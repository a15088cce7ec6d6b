// Internal declarations shared by the HIP translation units of libfesomc_gpu.
// Device layout: per-entity constants as small AoS vectors (int4/double4 per
// element, int2/double4 per edge) so a wavefront reads them coalesced; 3-D
// fields by column, X[i*K + k] (K = nsigma slots per entity), so that one
// gather of a neighbour reads a contiguous run of its levels.  The scratch of
// column-sequential solves (d3_end, the tracer Thomas sweep) is stored by
// level, X[k*n_entities + i], read across a wavefront of columns.
#pragma once
#include <hip/hip_runtime.h>

#include <cstring>

#include <cstdint>
#include <string>
#include <unordered_map>
#include <vector>

#include "fesomc_gpu.h"

#define FCG_CHECK(expr)                                                            \
    do {                                                                           \
        hipError_t e__ = (expr);                                                   \
        if (e__ != hipSuccess) return fcg_fail_hip(ctx, e__, #expr, __FILE__, __LINE__); \
    } while (0)

constexpr double kG = 9.8;
constexpr double kPref = 101325.0;    // reference atmospheric pressure P_ref (src/fv_var.F90:17)
constexpr double kREarthD = 6400000.0;
constexpr double kBeta = 0.281105;    // AB3 (src/fv_var.F90:108)
constexpr double kEps = 0.01;         // AB2 offset (src/fv_var.F90:110)
constexpr double kCka = 0.4;
constexpr double kPrNum = 0.0001;
constexpr double kBetaScale = 3.9;
constexpr int kEmpty = INT32_MIN;     // unused slot in a fixed-width edge list

// Per-edge operands of the node gathers over edges (tracer fluxes, transport
// divergence) in the storage type R: column offsets (entity id * K) and ids of
// the edge's nodes and elements, the cross distances edge_cross_dxdy, the
// edge's dy and the metric x-length dx*R_earth*cos of each segment, formed
// at set-up with the same operations the flux formulas apply to them.
template <typename R>
struct EdgeGeo {
    int n1K, n2K, e1K, e2K;   // e2K < 0: boundary edge (no second element)
    int e1, e2;
    R cx, cy, cz, cw, dy, dxa, dxb;
};

// LDS records of the node-gather blocks (one per CSR slot of the block's
// nodes): element operands of the node-averaged tracer gradient, element
// column offset and area of compute_el2nodes_3D, edge operands of the
// transport divergence (with the element's filtered 2-D velocity).
struct GradRec {
    int qK[4];            // column offsets of the element's nodes
    double ea;            // elem_area
    double gx[4], gy[4];  // gradient_sca
};
struct AreaRec {
    int elK;
    double ar;
};
// surface forcing of the tracer vertical diffusion (key_atm)
struct AtmPar {
    int on, type_swr, wet_dry;
    double Rjc, ajc, bjc, rho0, Dmin;
};
struct PulsRec {                 // AreaRec + the element's pulsation-velocity operands
    int elK;
    double ar, ap, u2, v2;
};
template <typename R>
struct DivRec {
    int e1K, e2K, pos;    // pos: node is edge_nodes(1)
    R cx, cy, cz, cw, vf1, uf1, vf2, uf2;
};

// per-column LDS records of vel_gradients (3-D) and of the P1 3D-to-2D
// advection (per internal edge slot)
struct VelGradRec {
    int jend, nbK[4];            // neighbour column offsets, < 0: free-slip boundary side
    double gx[4], gy[4], zc0[4], zc1[4];
};
struct P1Edge {
    int n1K, n2K, e1K, e2K;      // column offsets of the edge's nodes and elements
    double xe, ye, an1, an2;
};

// Geometry of one edge for the biharmonic operators (fv_viscosity.F90:148-156,
// 208-210, 381-392): internal edges: the normal (ye, -xe) with the mean cosine
// of the two elements and the unit centre-to-centre direction (tx, ty) over
// its length tl; boundary edges: xb = dx*R*cos of the element (3-D loops) and
// xe = dx*R, xbd = xe*cos (2-D loop); acc = (ac1+ac2)/2.
struct BhEdge {
    double xe, ye, tx, ty, tl, acc, xb, xbd;
};

// Device mirror of the mesh and state of one model instance.
struct DevMesh {
    int N = 0, E = 0, D = 0, Din = 0, L = 0, K = 0, nobn = 0;
    // element constants
    int4 *en = nullptr;        // nodes (0-based); triangle: .w == .x
    int4 *elnb = nullptr;      // neighbour elements, -1 none (free-slip mirror)
    int4 *eledges = nullptr;   // edges in node rotation order (for free-slip)
    int4 *el_ied = nullptr;    // internal edges, ascending id; e>=0: element is edge_tri(2) (+), ~e: edge_tri(1) (-)
    double4 *wcv = nullptr;    // w_cv
    double4 *gsx = nullptr, *gsy = nullptr;   // gradient_sca(1:4), (5:8)
    double4 *gvx = nullptr, *gvy = nullptr;   // gradient_vec(1:4), (5:8)
    double *elem_area = nullptr, *elem_cos = nullptr, *coriolis = nullptr;
    double *accB = nullptr;    // (sum w_cv*ac)**Bp_bz * mask_bpr, host-evaluated with libm pow
    double *kbi = nullptr;     // sqrt(scale_area/elem_area), scale_area of the element's member (filt_bi)
    // biharmonic options (bih_2D, bih_3D): per element the boundary edges in
    // ascending reference id (kEmpty after the last) and the viscosity factors
    // Ah = -Abh0*sqrt(elem_area/scale_area)/scale_area and its 3-D-to-2-D form
    // with the sponge weight in front; per edge the corrected-Laplacian geometry
    int4 *el_bed = nullptr;
    double *bh_ah = nullptr, *bh_ah_acc = nullptr;
    struct BhEdge *bhe = nullptr;
    double *acc_el = nullptr;  // sum w_cv*ac
    // edge constants
    int2 *ednodes = nullptr;   // 0-based
    int2 *edtri = nullptr;     // 0-based, -1 none
    double4 *ecross = nullptr; // edge_cross_dxdy
    double2 *edxdy = nullptr;  // edge_dxdy
    // FP32 copies of the geometry of the barotropic momentum path (gradients of
    // the AB3 velocity and of the elevation, the 2-D momentum edge fluxes),
    // read by the mixed-precision kernels instead of the FP64 arrays above
    float4 *gsx_f = nullptr, *gsy_f = nullptr, *gvx_f = nullptr, *gvy_f = nullptr, *ecross_f = nullptr;
    float2 *edxdy_f = nullptr;
    // node constants
    double *depth = nullptr, *area = nullptr, *ac = nullptr, *atp1 = nullptr, *atp2 = nullptr;
    int *index = nullptr, *obidx = nullptr;   // obidx: position in the OB list or -1
    // CSR gathers
    int *n2e_ptr = nullptr, *n2e = nullptr;   // node -> edges, ascending; e>=0: node is edge_nodes(1) (+), ~e: (2) (-)
    int *n2el_ptr = nullptr, *n2el = nullptr; // node -> elements, ascending (nod_in_elem2D)
    int *n2elK = nullptr;      // n2el * K: column offsets of those elements
    int4 *enK = nullptr;       // en * K: column offsets of the element's nodes
    void *egeo = nullptr;      // EdgeGeo<R>[D] of the context's storage type
    // ensemble members: nodes and elements of one member are contiguous
    int nmemb = 1;
    int *memb = nullptr;       // N: member of each node
    int *memb_noff = nullptr;  // nmemb+1: first node of each member
    double *memb_sa = nullptr; // nmemb: scale_area of each member (null for a single model)
    // vertical grid
    double *sigma = nullptr;   // K
    double *dsig = nullptr;    // L: sigma(k)-sigma(k+1)
    double *oms = nullptr;     // K: 1-sigma(k)
};

// Device state.  The elevation and volume chain of the barotropic loop (eta,
// etaAB, depths, the ssh_rhs sums, the Dmin clamps, masks, the 2-D velocity
// and its history), the hydrostatic pressure and the reductions are FP64 in
// every precision mode; the 3-D fields and the barotropic momentum path (AB3
// velocity, its gradients, the edge momentum fluxes, the 3-D-to-2-D
// tendencies and the bottom velocity cache) use the storage type R (double
// for the FP64 drop-in, float for the mixed-precision path).
template <typename R>
struct DevStateT {
    // 2-D barotropic
    double *eta_n = nullptr, *eta_n_1 = nullptr, *eta_n_2 = nullptr, *etaAB = nullptr, *eta_g = nullptr;
    double *ssh_rhs = nullptr;            // flux sum of the last substep (diagnostic)
    double *U2 = nullptr, *U2_1 = nullptr, *U2_2 = nullptr;   // (2,E) interleaved [2e],[2e+1]
    R *UAB = nullptr;                                          // 2E
    double *aux_el = nullptr;                                  // E
    R *vel_grad = nullptr;                                     // 4E
    R *U_rhs_2D_3D = nullptr, *Bar_pr_2D = nullptr;            // 2E
    double *adn = nullptr;   // N: max(Dmin, AB3 elevation + depth) of the coming substep
    double *U_filt = nullptr, *V_filt = nullptr, *dmean_n = nullptr; // E
    double *mask_wd = nullptr, *taux = nullptr, *tauy = nullptr, *C_d_el = nullptr; // E
    double *windx = nullptr, *windy = nullptr;                 // N
    double *a = nullptr, *a_old = nullptr;                     // N: max(depth+eta_n, Dmin) now / previous jacobian
    R *ubot = nullptr;                                         // 2E: U_n, V_n of the bottom layer (substep cache)
    double *fl_ssh = nullptr;                                  // D: per-edge transport
    R *fl_uv = nullptr;                                        // 2D: per-edge 2-D momentum flux
    double *fl_p1 = nullptr;                                   // 4D: layer sums of the P1 3-D-to-2-D advection
    double *puls_ap = nullptr;                                 // E: a of U_puls = U_n - a*U_n_2D (first substep)
    double *qsr = nullptr, *qns = nullptr, *emp = nullptr, *mslp = nullptr;  // N: surface forcing (key_atm)
    double *sbc_coef = nullptr;    // 16N: per node the 8 fields' time coefficients a, then b (key_atm, bulk)
    double *tau_node = nullptr;    // 2N: bulk wind stress at the nodes (key_atm, bulk)
    double *obc_coef = nullptr;    // 4*nobn*ndep: OB climatology time coefficients (a then b, (2, nobn, ndep) each)
    double *obc_depth = nullptr;   // ndep: depths of the climatology profiles [m]
    int *obc_node = nullptr;       // nobn: device node of each open-boundary node
    double *runoff = nullptr;                                  // N: river discharge [m3/s], 0 off the mouths (key_rivers)
    double *rho_c = nullptr;                                   // L*N (n*L + k): density anomaly of pressure (key_atm)
    double *node2d = nullptr;                                  // 4N: etaAB, eta_g, eta_n_1, eta_n per node (substep)
    double *tide = nullptr;                                    // Mt*nobn prescribed OB elevation per substep
    // 3-D, stored by column: K slots per node/element (layer fields use L)
    R *U_n = nullptr, *V_n = nullptr, *U_rhs = nullptr, *V_rhs = nullptr;   // L*E
    R *U_rhsAB = nullptr, *V_rhsAB = nullptr, *U_nf = nullptr, *V_nf = nullptr; // L*E
    R *gux = nullptr, *guy = nullptr, *gvx = nullptr, *gvy = nullptr;         // L*E
    R *Bpu = nullptr, *Bpv = nullptr, *Je = nullptr;                          // L*E
    // twice-applied filters (filt_bi): first-pass differences U_c, V_c and the
    // 3-D tendency UV_rhs (L*E, null when filt_bi_3D is off); the 2-D first
    // pass (2E, null when filt_bi_2D is off)
    R *fbu = nullptr, *fbv = nullptr, *uvu = nullptr, *uvv = nullptr;
    double *fb2c = nullptr;
    // biharmonic options: per (edge, layer) fluxes; per (element, layer) the
    // first-pass field U_c, its gradients and the pulsation field (then the
    // 3-D-to-2-D result); 2-D: per edge fluxes, the first-pass field Ul and its
    // gradients (null when the option is off)
    R *bhfu = nullptr, *bhfv = nullptr, *bhcu = nullptr, *bhcv = nullptr;
    R *bhgux = nullptr, *bhguy = nullptr, *bhgvx = nullptr, *bhgvy = nullptr;
    R *bhpu = nullptr, *bhpv = nullptr, *bhub = nullptr, *bhvb = nullptr;
    double *bh_fl2 = nullptr;
    R *bh_ul = nullptr, *bh_g2 = nullptr;

    R *vxx = nullptr, *vyu = nullptr, *vyv = nullptr;                         // K*E Thomas scratch
    R *TF = nullptr, *SF = nullptr, *T_old = nullptr, *S_old = nullptr;       // L*N
    R *Tclim = nullptr, *Sclim = nullptr, *abT = nullptr, *abS = nullptr;     // L*N
    R *Unode_p = nullptr, *Vnode_p = nullptr;                                 // L*N
    R *Unode = nullptr, *Vnode = nullptr;                                     // L*N node-averaged U_n, V_n
    R *tcp = nullptr, *ttp = nullptr, *tsp = nullptr;                         // L*N Thomas scratch
    R *gTx = nullptr, *gTy = nullptr, *gSx = nullptr, *gSy = nullptr;         // L*N node-averaged tracer gradients (MUSCL/Miura)
    R *tcur = nullptr;                                                        // L*N de-meaned T before the Miura update
    double *hp = nullptr;                                                          // K*N hydrostatic pressure (FP64)
    R *Kv = nullptr, *snu = nullptr, *bt = nullptr, *W = nullptr; // K*N
    R *mx_p = nullptr, *mx_q = nullptr, *mx_c2 = nullptr, *mx_xx = nullptr, *mx_yy = nullptr; // K*N d3_end scratch
    R *mx_bA = nullptr, *mx_bB = nullptr, *mx_s = nullptr;                    // K*N
    // reductions
    double *red = nullptr;       // partial results
    double *scal = nullptr;      // diagnostics (from [4])
    double *tsa = nullptr;       // per member: [2m] T_aver, [2m+1] S_aver
    int *err = nullptr;          // device STOP flag
};
using DevState = DevStateT<double>;

// The same allocations seen with the mixed-precision storage type (the
// pointer layout of DevStateT does not depend on R).
template <typename R>
inline DevStateT<R> retype(const DevState &s) {
    static_assert(sizeof(DevStateT<R>) == sizeof(DevState), "state layout");
    DevStateT<R> t;
    std::memcpy(static_cast<void *>(&t), static_cast<const void *>(&s), sizeof s);
    return t;
}

// LDS bytes per column of k_velrhs's element and edge records (checked
// against the record types where they are defined)
constexpr size_t kVelrhsElemBytes = 160;
constexpr size_t velrhs_edge_bytes(size_t esz) { return esz == 8 ? 104 : 64; }

constexpr int kTideRing = 4;   // pinned staging slots of the tidal open-boundary values
constexpr int kD3endT = 64;    // node columns per block of the d3_end iteration (its column state in LDS)

struct fcg_ctx {
    fcg_params p{};
    bool mixed = false;    // FCG_MIXED: 3-D fields stored as float
    size_t esize = 8;      // bytes per 3-D field element
    double density_0 = 0, scale_area = 0;
    DevMesh m;
    DevState s;
    hipStream_t stream = nullptr;
    int device = 0;
    // host clock (src/fv_main.F90:400-401, 1663)
    double time_jd = 0, time = 0;
    int n_dt = 0, lfirst = 1;
    // host copies needed for tides and the Fortran-layout transfer
    std::vector<int> in_obn0;            // 0-based OB nodes (reference ids)
    std::vector<int> pn, pe;             // device index -> reference index (nodes, elements)
    std::vector<int> noff, eoff;         // per member: first device node / element (nmemb+1 each)
    std::vector<double> ampt, fazt, amp_factor, phy_factor; // (nobn,12), (12)
    double *tide_pinned = nullptr;       // kTideRing x Mt*nobn staging of the open-boundary values
    hipEvent_t tide_ev[4] = {};          // completion of the copy issued from each ring slot
    bool tide_ev_used[4] = {};
    int tide_slot = 0;
    // barotropic graph
    hipGraphExec_t bt_graph = nullptr;
    bool bt_graph_ok = false;
    int red_blocks = 0;
    // node-gather blocks (cpb = 256/L nodes): most CSR slots of any block
    // most CSR slots (node -> edge, node -> element) in one node-gather block
    // of 256, 512 and 1024 threads (256/L, 512/L, 1024/L node columns)
    int cap_n2e_t[3] = {0, 0, 0}, cap_n2el_t[3] = {0, 0, 0};
    // the same for the level-vector node kernels: blocks of 128 and 256
    // threads, K/W lanes per column (W = 16/storage width)
    int cap_n2e_v[2] = {0, 0}, cap_n2el_v[2] = {0, 0};
    int obc_ndep = 0;            // depths of the open-boundary climatology profiles on the device
    int lds_limit = 0;           // bytes of LDS one block may request on this device
    int n_cu = 1;                // compute units of the device (persistent launches)
    // learned closure
    fcg_mlp *mlp = nullptr;      // attached network (not owned)
    int *mlp_flag = nullptr;     // N: column sent back to d3_end this step
    unsigned long long *mlp_count = nullptr;  // such columns summed over steps
    std::string probe_name;
    std::unordered_map<std::string, int> launch_block;  // autotuned block sizes
    std::vector<std::pair<hipEvent_t, hipEvent_t>> probe_ev;
    std::vector<void *> allocs;
    std::string err;
};

int fcg_fail_hip(fcg_ctx *ctx, hipError_t e, const char *what, const char *file, int line);
int fcg_fail(fcg_ctx *ctx, int code, const std::string &msg);
