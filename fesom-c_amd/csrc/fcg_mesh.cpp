// Host-side mesh preprocessing for hosts that do not run the reference's own
// Fortran set-up.  Restates, in the reference's operation order (so geometry
// agrees bit for bit and integer numbering exactly):
//   read_mesh_ser element ordering     src/fv_main.F90:787-945
//   SET_SIGMA (ind_sigma 1, 2)          src/fv_vert_coord.F90:135-208
//   test_elem                           src/fv_mesh_array.F90:1-82
//   find_edges                          src/fv_mesh_array.F90:395-718
//   elem_center / edge_center / cdiff   src/fv_mesh_array.F90:720-789,1141-1152
//   find_elem_neighbors                 src/fv_mesh_array.F90:791-946
//   mesh_arrays1, mesh_arrays2          src/fv_mesh_array.F90:948-1386
//   initial_state constants             src/fv_ini.F90:81,94-107,187-240
// Built with -ffp-contract=off.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "fesomc_gpu.h"

namespace {

constexpr double kPi = 3.14159265358979;
constexpr double kRad = kPi / 180.0;
constexpr double kREarth = 6400000.0;
constexpr double kOmega = 2.0 * kPi / (3600.0 * 24.0);

inline double f4(double x) { return (double)(float)x; }

// Jackett-McDougall EOS: src/fv_pressure_new.F90:340-414 (the initial_state
// reference density only; the device has its own copy).
double density_jm(double t, double s, double pz, double rho0) {
    double ss3 = std::sqrt(s * s * s);
    double bulk = 19092.56 + t * (209.8925 - t * (3.041638 - t * (-f4(1.852732e-3) - t * (f4(1.361629e-5)))))
                + s * (104.4077 - t * (6.500517 - t * (.1553190 - t * (-f4(2.326469e-4)))))
                + ss3 * (-5.587545 + t * (0.7390729 - t * (f4(1.909078e-2))))
                - pz * (f4(4.721788e-1) + t * (f4(1.028859e-2) + t * (-f4(2.512549e-4) - t * (f4(5.939910e-7)))))
                - pz * s * (-f4(1.571896e-2) - t * (f4(2.598241e-4) + t * (-f4(7.267926e-6))))
                - pz * ss3 * f4(2.042967e-3)
                + pz * pz * (f4(1.045941e-5) - t * (f4(5.782165e-10) - t * (f4(1.296821e-7))))
                + pz * pz * s * (-f4(2.595994e-7) + t * (-f4(1.248266e-9) + t * (-f4(3.508914e-9))));
    double rhopot = 999.842594
                  + t * (f4(6.793952e-2) + t * (-f4(9.095290e-3) + t * (f4(1.001685e-4)
                    + t * (-f4(1.120083e-6) + t * f4(6.536332e-9)))))
                  + s * (0.824493 + t * (-f4(4.08990e-3) + t * (f4(7.64380e-5)
                    + t * (-f4(8.24670e-7) + t * f4(5.38750e-9)))))
                  + ss3 * (-f4(5.72466e-3) + t * (f4(1.02270e-4) + t * (-f4(1.65460e-6))))
                  + f4(4.8314e-4) * s * s;
    return rhopot * bulk / (bulk + 0.1 * pz) - rho0;
}

}  // namespace

struct fcg_mesh_obj {
    int nod2D = 0, elem2D = 0, edge2D = 0, edge2D_in = 0, nsigma = 0, nobn = 0;
    double cyc = 0;  // cyclic length [rad]
    std::vector<double> xy;                // (2,N) radians
    std::vector<int> index, in_obn;        // 1-based ids in in_obn
    std::vector<int> en;                   // (4,E) 0-based during build
    std::vector<int> ed_nodes, ed_tri;     // (2,D) 0-based, -999
    std::vector<int> el_edges, el_nb;      // (4,E)
    std::vector<double> depth, sigma, area, elem_area, elem_cos, coriolis, w_cv, gsca, gvec, edxdy, ecross, ac,
        a_tp, ampt, fazt, amp_factor, phy_factor;
    double density_0 = 0, scale_area = 0;
    // 1-based Fortran copies handed out through fcg_mesh_view
    std::vector<int> f_en, f_ednodes, f_edtri, f_eledges, f_elnb;
};

namespace {

struct Builder {
    fcg_mesh_obj &m;
    explicit Builder(fcg_mesh_obj &mm) : m(mm) {}
    double X(int n) const { return m.xy[2 * n]; }
    double Y(int n) const { return m.xy[2 * n + 1]; }
    bool tri(int e) const { return m.en[4 * e] == m.en[4 * e + 3]; }

    void wrap(double &d) const {
        if (d > m.cyc / 2.0) d -= m.cyc;
        if (d < -m.cyc / 2.0) d += m.cyc;
    }

    void center(int e, double &x, double &y, double &s) const {
        const int *q = &m.en[4 * e];
        double ax[4], ay[4];
        for (int k = 0; k < 4; ++k) { ax[k] = X(q[k]); ay[k] = Y(q[k]); }
        if (q[0] == q[3]) {
            double amin = std::fmin(std::fmin(ax[0], ax[1]), ax[2]);
            for (int k = 0; k < 3; ++k) if (ax[k] - amin > m.cyc / 2.0) ax[k] -= m.cyc;
            x = ((ax[0] + ax[1]) + ax[2]) / 3.0;
            y = ((ay[0] + ay[1]) + ay[2]) / 3.0;
            s = 0.5 * std::fabs((ax[1] - ax[0]) * (ay[2] - ay[0]) - (ay[1] - ay[0]) * (ax[2] - ax[0]));
        } else {
            double amin = std::fmin(std::fmin(ax[0], ax[1]), std::fmin(ax[2], ax[3]));
            for (int k = 0; k < 4; ++k) if (ax[k] - amin > m.cyc / 2.0) ax[k] -= m.cyc;
            double xa = ax[0] + ax[2] + ax[3], ya = ay[0] + ay[2] + ay[3];
            double xb = ax[0] + ax[1] + ax[2], yb = ay[0] + ay[1] + ay[2];
            double s1 = std::fabs((ax[2] - ax[3]) * (ay[0] - ay[3]) - (ay[2] - ay[3]) * (ax[0] - ax[3]));
            double s2 = std::fabs((ax[2] - ax[1]) * (ay[0] - ay[1]) - (ay[2] - ay[1]) * (ax[0] - ax[1]));
            x = (xa * s1 + xb * s2) / 3.0 / (s1 + s2);
            y = (ya * s1 + yb * s2) / 3.0 / (s1 + s2);
            s = 0.5 * (s1 + s2);
        }
    }

    void edge_center(int a, int b, double &x, double &y) const {
        double a0 = X(a), b0 = X(b);
        if (a0 - b0 > m.cyc / 2.0) a0 -= m.cyc;
        if (a0 - b0 < -m.cyc / 2.0) b0 -= m.cyc;
        x = 0.5 * (a0 + b0);
        y = 0.5 * (Y(a) + Y(b));
    }

    std::string test_elem() {
        for (int e = 0; e < m.elem2D; ++e) {
            int *q = &m.en[4 * e];
            double bx = X(q[1]) - X(q[0]), by = Y(q[1]) - Y(q[0]);
            double cx = X(q[2]) - X(q[0]), cy = Y(q[2]) - Y(q[0]);
            wrap(bx);
            wrap(cx);
            double r = bx * cy - by * cx;
            if (q[0] == q[3]) {
                if (r > 0.0) std::swap(q[1], q[2]);
            } else {
                double dx = X(q[3]) - X(q[0]), dy = Y(q[3]) - Y(q[0]);
                wrap(dx);
                double r1 = bx * dy - by * dx;
                if (r1 * r < 0.0) return "test_elem: node list problem in element " + std::to_string(e + 1);
                if (r > 0) std::swap(q[1], q[3]);
            }
        }
        return "";
    }

    std::string find_edges() {
        const int N = m.nod2D, E = m.elem2D;
        // elements around each node, ascending element id
        std::vector<std::vector<int>> ne(N);
        for (int e = 0; e < E; ++e)
            for (int q = 0; q < (tri(e) ? 3 : 4); ++q) ne[m.en[4 * e + q]].push_back(e);
        // neighbouring nodes in the order first met (quads: edge-connected only)
        std::vector<std::vector<int>> nn(N);
        std::vector<char> seen(N, 0);
        for (int n = 0; n < N; ++n) {
            for (int e : ne[n]) {
                const int *q = &m.en[4 * e];
                int cand[3], nc = 0;
                if (tri(e)) {
                    for (int k = 0; k < 3; ++k) if (q[k] != n) cand[nc++] = q[k];
                } else {
                    int pos = q[0] == n ? 0 : q[1] == n ? 1 : q[2] == n ? 2 : 3;
                    // neighbours along the quad's edges, in the reference's pick order
                    static const int pick[4][2] = {{1, 3}, {0, 2}, {1, 3}, {0, 2}};
                    cand[nc++] = q[pick[pos][0]];
                    cand[nc++] = q[pick[pos][1]];
                }
                for (int k = 0; k < nc; ++k)
                    if (!seen[cand[k]]) { seen[cand[k]] = 1; nn[n].push_back(cand[k]); }
            }
            for (int v : nn[n]) seen[v] = 0;
        }
        // edges (n, v>n): internal (two elements) first, then boundary
        auto elems_of = [&](int n, int v, int *two) {
            int f = 0;
            for (int e : ne[n]) {
                const int *q = &m.en[4 * e];
                for (int k = 0; k < (tri(e) ? 3 : 4); ++k)
                    if (q[k] == v) { if (f < 2) two[f] = e; ++f; break; }
            }
            return f;
        };
        m.ed_nodes.clear();
        m.ed_tri.clear();
        for (int pass = 0; pass < 2; ++pass) {
            for (int n = 0; n < N; ++n)
                for (int v : nn[n]) {
                    if (v < n) continue;
                    int two[2] = {-999, -999};
                    int f = elems_of(n, v, two);
                    if ((pass == 0 && f == 2) || (pass == 1 && f == 1)) {
                        m.ed_nodes.push_back(n);
                        m.ed_nodes.push_back(v);
                        m.ed_tri.push_back(two[0]);
                        m.ed_tri.push_back(pass == 0 ? two[1] : -999);
                    }
                }
            if (pass == 0) m.edge2D_in = (int)m.ed_tri.size() / 2;
        }
        m.edge2D = (int)m.ed_tri.size() / 2;
        // left-element orientation
        for (int d = 0; d < m.edge2D; ++d) {
            int a = m.ed_nodes[2 * d], b = m.ed_nodes[2 * d + 1], e = m.ed_tri[2 * d];
            double xc, yc, s;
            center(e, xc, yc, s);
            xc -= X(a);
            yc -= Y(a);
            double xe = X(b) - X(a), ye = Y(b) - Y(a);
            wrap(xe);
            wrap(xc);
            if (xc * ye - yc * xe > 0.0) {
                if (m.ed_tri[2 * d + 1] >= 0) std::swap(m.ed_tri[2 * d], m.ed_tri[2 * d + 1]);
                else std::swap(m.ed_nodes[2 * d], m.ed_nodes[2 * d + 1]);
            }
        }
        // element edges, listed in node rotation order
        m.el_edges.assign(4 * (size_t)E, 0);
        std::vector<int> cnt(E, 0);
        for (int d = 0; d < m.edge2D; ++d)
            for (int k = 0; k < 2; ++k) {
                int e = m.ed_tri[2 * d + k];
                if (e >= 0) m.el_edges[4 * e + cnt[e]++] = d;
            }
        for (int e = 0; e < E; ++e) {
            const int *q = &m.en[4 * e];
            int q1 = tri(e) ? 3 : 4;
            int own[4];
            std::memcpy(own, &m.el_edges[4 * e], sizeof own);
            auto joins = [&](int d, int u, int v) {
                int a = m.ed_nodes[2 * d], b = m.ed_nodes[2 * d + 1];
                return (a == u && b == v) || (a == v && b == u);
            };
            for (int s = 0; s < q1; ++s) {
                int u = q[s], v = q[(s + 1) % q1];
                for (int k = 0; k < q1; ++k)
                    if (joins(own[k], u, v)) { m.el_edges[4 * e + s] = own[k]; break; }
            }
            if (q1 == 3) m.el_edges[4 * e + 3] = m.el_edges[4 * e];
        }
        return "";
    }

    void neighbors() {
        const int E = m.elem2D;
        m.el_nb.assign(4 * (size_t)E, 0);
        for (int e = 0; e < E; ++e) {
            const int *d = &m.el_edges[4 * e];
            int q = (d[0] == d[3]) ? 3 : 4;
            for (int j = 0; j < q; ++j) {
                int o = m.ed_tri[2 * d[j]];
                if (o == e) o = m.ed_tri[2 * d[j] + 1];
                m.el_nb[4 * e + j] = o;
            }
            if (q == 3) m.el_nb[4 * e + 3] = m.el_nb[4 * e];
        }
    }

    void arrays1(const fcg_mesh_cfg &c) {
        const int E = m.elem2D, N = m.nod2D;
        m.elem_area.assign(E, 0.0);
        m.elem_cos.assign(E, 0.0);
        m.coriolis.assign(E, 0.0);
        m.area.assign(N, 0.0);
        m.w_cv.assign(4 * (size_t)E, 0.0);
        for (int e = 0; e < E; ++e) {
            double ax, ay, aa;
            center(e, ax, ay, aa);
            m.coriolis[e] = 2.0 * kOmega * std::sin(ay);
            m.elem_cos[e] = std::cos(ay);
            m.elem_area[e] = aa;
        }
        if (c.cartesian) for (int e = 0; e < E; ++e) m.elem_cos[e] = 1.0;
        if (c.coriolis_tf) {
            if (c.cartesian)
                for (int e = 0; e < E; ++e) m.coriolis[e] = 2.0 * kOmega * std::sin(c.lat_cartesian_deg * kPi / 180.0);
        } else {
            for (int e = 0; e < E; ++e) m.coriolis[e] = 0.0;
        }
        for (int e = 0; e < E; ++e) m.elem_area[e] = m.elem_area[e] * m.elem_cos[e];
        for (int e = 0; e < E; ++e) {
            const int *q = &m.en[4 * e];
            if (tri(e)) {
                for (int j = 0; j < 3; ++j) {
                    m.area[q[j]] = m.area[q[j]] + m.elem_area[e] / 3.0;
                    m.w_cv[4 * e + j] = 1.0 / 3.0;
                }
            } else {
                double rcx, rcy, aa;
                center(e, rcx, rcy, aa);
                for (int j = 0; j < 4; ++j) {
                    int prv = q[(j + 3) % 4], cur = q[j], nxt = q[(j + 1) % 4];
                    double rrx = X(prv) - X(cur), rry = Y(prv) - Y(cur);
                    wrap(rrx);
                    double rlx = X(nxt) - X(cur), rly = Y(nxt) - Y(cur);
                    wrap(rlx);
                    double r0x = rcx - X(cur), r0y = rcy - Y(cur);
                    wrap(r0x);
                    double part = std::fabs(rlx * r0y - rly * r0x) + std::fabs(rrx * r0y - rry * r0x);
                    m.area[cur] = m.area[cur] + 0.25 * m.elem_cos[e] * part;
                    m.w_cv[4 * e + j] = 0.25 * m.elem_cos[e] * part / m.elem_area[e];
                }
            }
        }
        for (int e = 0; e < E; ++e) m.elem_area[e] = m.elem_area[e] * kREarth * kREarth;
        for (int n = 0; n < N; ++n) m.area[n] = m.area[n] * kREarth * kREarth;
    }

    void arrays2() {
        const int E = m.elem2D, D = m.edge2D;
        std::vector<double> cx(E), cy(E);
        for (int e = 0; e < E; ++e) {
            double s;
            center(e, cx[e], cy[e], s);
        }
        m.edxdy.assign(2 * (size_t)D, 0.0);
        m.ecross.assign(4 * (size_t)D, 0.0);
        for (int d = 0; d < D; ++d) {
            int a = m.ed_nodes[2 * d], b = m.ed_nodes[2 * d + 1];
            double dx = X(b) - X(a);
            wrap(dx);
            m.edxdy[2 * d] = dx;
            m.edxdy[2 * d + 1] = Y(b) - Y(a);
        }
        for (int d = 0; d < D; ++d) {
            double ex, ey;
            edge_center(m.ed_nodes[2 * d], m.ed_nodes[2 * d + 1], ex, ey);
            for (int side = 0; side < 2; ++side) {
                int e = m.ed_tri[2 * d + side];
                if (e < 0) continue;
                double bx = cx[e] - ex, by = cy[e] - ey;
                wrap(bx);
                bx = bx * m.elem_cos[e];
                m.ecross[4 * d + 2 * side] = bx * kREarth;
                m.ecross[4 * d + 2 * side + 1] = by * kREarth;
            }
        }
        m.gsca.assign(8 * (size_t)E, 0.0);
        m.gvec.assign(8 * (size_t)E, 0.0);
        for (int e = 0; e < E; ++e) {
            const int *q = &m.en[4 * e];
            double *g = &m.gsca[8 * e];
            if (tri(e)) {
                double dX31 = X(q[2]) - X(q[0]);
                wrap(dX31);
                dX31 = m.elem_cos[e] * dX31;
                double dX21 = X(q[1]) - X(q[0]);
                wrap(dX21);
                dX21 = m.elem_cos[e] * dX21;
                double dY31 = Y(q[2]) - Y(q[0]), dY21 = Y(q[1]) - Y(q[0]);
                double f = -0.5 * kREarth / m.elem_area[e];
                g[0] = (-dY31 + dY21) * f;
                g[1] = dY31 * f;
                g[2] = -dY21 * f;
                g[3] = 0.0;
                g[4] = (dX31 - dX21) * f;
                g[5] = -dX31 * f;
                g[6] = dX21 * f;
                g[7] = 0.0;
            } else {
                double f = 0.5 * kREarth / m.elem_area[e];
                for (int j = 0; j < 4; ++j) {
                    int prv = q[(j + 3) % 4], cur = q[j], nxt = q[(j + 1) % 4];
                    double ax = X(cur) - X(prv), ay = Y(cur) - Y(prv);
                    wrap(ax);
                    double bx = X(nxt) - X(cur), by = Y(nxt) - Y(cur);
                    wrap(bx);
                    double sx = ax + bx, sy = ay + by;
                    sx = sx * m.elem_cos[e];
                    g[j] = -sy * f;
                    g[j + 4] = sx * f;
                }
            }
        }
        for (int e = 0; e < E; ++e) {
            int q = tri(e) ? 3 : 4;
            double x[4] = {0, 0, 0, 0}, y[4] = {0, 0, 0, 0};
            for (int j = 0; j < q; ++j) {
                int o = m.el_nb[4 * e + j];
                if (o >= 0) {
                    x[j] = cx[o] - cx[e];
                    wrap(x[j]);
                    y[j] = cy[o] - cy[e];
                } else {
                    int d = m.el_edges[4 * e + j];
                    double bx, by;
                    edge_center(m.ed_nodes[2 * d], m.ed_nodes[2 * d + 1], bx, by);
                    x[j] = (bx - cx[e]);
                    wrap(x[j]);
                    x[j] = 2.0 * x[j];
                    y[j] = 2.0 * (by - cy[e]);
                }
            }
            double cxx = 0, cxy = 0, cyy = 0;
            for (int j = 0; j < 4; ++j) {
                x[j] = x[j] * m.elem_cos[e] * kREarth;
                y[j] = y[j] * kREarth;
            }
            for (int j = 0; j < 4; ++j) {
                cxx += x[j] * x[j];
                cxy += x[j] * y[j];
                cyy += y[j] * y[j];
            }
            double det = cxy * cxy - cxx * cyy;
            for (int j = 0; j < 4; ++j) {
                m.gvec[8 * e + j] = (cxy * y[j] - cyy * x[j]) / det;
                m.gvec[8 * e + 4 + j] = (cxy * x[j] - cxx * y[j]) / det;
            }
        }
    }
};

int set_sigma(fcg_mesh_obj &m, const fcg_mesh_cfg &c) {
    const int K = c.nsigma;
    m.sigma.assign(K, 0.0);
    if (c.ind_sigma == 1) {
        if (c.pow_sig > 1 && K % 2 == 0) return FCG_E_ARG;
        if (c.pow_sig == 1) {
            for (int k = 1; k <= K; ++k) m.sigma[k - 1] = -std::pow((k - 1) / (double)(K - 1), c.pow_sig) + 1.0;
        } else {
            int h = (K + 1) / 2;
            for (int k = 1; k <= h; ++k) m.sigma[k - 1] = -std::pow((k - 1) / (double)(h - 1), c.pow_sig) * 0.5 + 1.0;
            for (int k = h + 1; k <= K; ++k) m.sigma[k - 1] = std::pow((K - k) / (double)(h - 1), c.pow_sig) * 0.5;
        }
    } else if (c.ind_sigma == 2) {
        for (int k = 1; k <= K; ++k) {
            double x1 = std::tanh((c.lev_bot + c.lev_sur) * (K - k) / (K - 1) - c.lev_bot);
            double x2 = std::tanh(c.lev_bot);
            double x3 = x2 + std::tanh(c.lev_sur);
            m.sigma[k - 1] = (x1 + x2) / x3;
        }
    } else {
        return FCG_E_ARG;
    }
    return FCG_OK;
}

}  // namespace

extern "C" int fcg_mesh_build(const fcg_mesh_cfg *cfg, int N, const double *xy, const int *ind, int E,
                              const int *elem, const double *depth, int nobn_in, const double *ob,
                              fcg_mesh_obj **out) {
    if (!cfg || !xy || !ind || !elem || !out || N <= 0 || E <= 0 || cfg->nsigma < 4) return FCG_E_ARG;
    if (!cfg->is_LE && !depth) return FCG_E_ARG;
    auto *m = new fcg_mesh_obj();
    m->nod2D = N;
    m->elem2D = E;
    m->nsigma = cfg->nsigma;
    m->cyc = cfg->cyclic_length_deg * kRad;
    m->xy.resize(2 * (size_t)N);
    m->index.assign(ind, ind + N);
    for (int n = 0; n < N; ++n) {
        if (cfg->cartesian) {
            m->xy[2 * n] = xy[2 * n] / kREarth;
            m->xy[2 * n + 1] = xy[2 * n + 1] / kREarth;
        } else {
            m->xy[2 * n] = xy[2 * n] * kRad;
            m->xy[2 * n + 1] = xy[2 * n + 1] * kRad;
        }
        if (ind[n] == 2) m->in_obn.push_back(n + 1);
    }
    m->nobn = (int)m->in_obn.size();
    // triangles first, each group in file order
    m->en.reserve(4 * (size_t)E);
    for (int pass = 0; pass < 2; ++pass)
        for (int e = 0; e < E; ++e) {
            bool t = elem[4 * e] == elem[4 * e + 3];
            if ((pass == 0) != t) continue;
            for (int k = 0; k < 4; ++k) {
                int v = elem[4 * e + k];
                if (v < 1 || v > N) { delete m; return FCG_E_MESH; }
                m->en.push_back(v - 1);
            }
        }
    m->depth.resize(N);
    for (int n = 0; n < N; ++n) m->depth[n] = cfg->is_LE ? 20.0 : depth[n];
    if (set_sigma(*m, *cfg) != FCG_OK) { delete m; return FCG_E_ARG; }
    Builder b(*m);
    if (!b.test_elem().empty()) { delete m; return FCG_E_MESH; }
    b.find_edges();
    b.neighbors();
    b.arrays1(*cfg);
    b.arrays2();
    // initial_state constants (src/fv_ini.F90:81, 94-107, 187-240)
    double s = 0.0;
    for (int e = 0; e < E; ++e) s = s + m->elem_area[e];
    m->scale_area = s / E;
    m->a_tp.assign(2 * (size_t)N, 0.0);
    for (int n = 0; n < N; ++n) {
        m->a_tp[n] = cfg->T_potential ? 0.918 : 1.0;
        m->a_tp[N + n] = cfg->T_potential ? 0.69 : 0.0;
        if (m->index[n] == 2) { m->a_tp[n] = 1.0; m->a_tp[N + n] = 0.0; }
    }
    m->ac.assign(N, 1.0);
    // Sponge layer (ac_create_Xobn + ac_create, src/fv_main.F90:297-300,
    // src/fv_utilit.F90:190-212): ac = d/SL_radius where the distance d to the
    // nearest open-boundary node (Dist2in1, src/fv_utilit.F90:53-88) is below
    // SL_radius.  Cartesian coordinates are held divided by the earth radius.
    if (cfg->SL_obn && m->nobn > 0) {
        if (!(cfg->SL_radius > 0.0)) { delete m; return FCG_E_ARG; }
        std::vector<double> ox(m->nobn), oy(m->nobn), coy(m->nobn);
        for (int j = 0; j < m->nobn; ++j) {
            ox[j] = m->xy[2 * (size_t)(m->in_obn[j] - 1)];
            oy[j] = m->xy[2 * (size_t)(m->in_obn[j] - 1) + 1];
            coy[j] = std::cos(oy[j]);
        }
        for (int n = 0; n < N; ++n) {
            const double x1 = m->xy[2 * (size_t)n], x2 = m->xy[2 * (size_t)n + 1];
            double d2obn;
            if (cfg->cartesian) {
                double d = (ox[0] - x1) * (ox[0] - x1) + (oy[0] - x2) * (oy[0] - x2);
                for (int j = 1; j < m->nobn; ++j) {
                    const double v = (ox[j] - x1) * (ox[j] - x1) + (oy[j] - x2) * (oy[j] - x2);
                    if (v < d) d = v;
                }
                d2obn = std::sqrt(d) * kREarth;
            } else {
                const double cx2 = std::cos(x2);
                double a = 0.0;
                for (int j = 0; j < m->nobn; ++j) {
                    const double s2 = std::sin((oy[j] - x2) * 0.5), s1 = std::sin((ox[j] - x1) * 0.5);
                    const double v = s2 * s2 + cx2 * coy[j] * s1 * s1;
                    if (j == 0 || v < a) a = v;
                }
                d2obn = 2.0 * kREarth * std::asin(std::min(1.0, std::sqrt(a)));
            }
            if (d2obn < cfg->SL_radius) m->ac[n] = d2obn / cfg->SL_radius;
        }
    }
    m->density_0 = 1000.0;
    m->density_0 = m->density_0 + density_jm(cfg->T_const, cfg->S_const, 0.0, m->density_0);
    m->amp_factor.assign(12, 1.0);
    m->phy_factor.assign(12, 0.0);
    const int nb = m->nobn;
    m->ampt.assign(12 * (size_t)(nb > 0 ? nb : 1), 0.0);
    m->fazt.assign(12 * (size_t)(nb > 0 ? nb : 1), 0.0);
    if (cfg->tf_presence && ob) {
        int nh = 0;
        for (int j = 0; j < 12; ++j) nh += cfg->harmonics[j] ? 1 : 0;
        for (int i = 0; i < nb && i < nobn_in; ++i) {
            int col = 0;
            for (int j = 0; j < 12; ++j) {
                if (!cfg->harmonics[j]) continue;
                double amp = ob[(size_t)i * 2 * nh + col], ph = ob[(size_t)i * 2 * nh + col + 1];
                if (ph < 0.0) ph = 360.0 + ph;
                m->ampt[(size_t)j * nb + i] = amp;      // Fortran (nobn,12)
                m->fazt[(size_t)j * nb + i] = ph * kRad;
                col += 2;
            }
        }
    }
    // Fortran-facing 1-based copies
    auto to1 = [](const std::vector<int> &v) {
        std::vector<int> r(v);
        for (int &x : r) if (x >= 0) x += 1;
        return r;
    };
    m->f_en = to1(m->en);
    m->f_ednodes = to1(m->ed_nodes);
    m->f_edtri = to1(m->ed_tri);
    m->f_eledges = to1(m->el_edges);
    m->f_elnb = to1(m->el_nb);
    *out = m;
    return FCG_OK;
}

extern "C" int fcg_mesh_view(const fcg_mesh_obj *m, fcg_mesh *v) {
    if (!m || !v) return FCG_E_ARG;
    std::memset(v, 0, sizeof *v);
    v->nod2D = m->nod2D;
    v->elem2D = m->elem2D;
    v->edge2D = m->edge2D;
    v->edge2D_in = m->edge2D_in;
    v->nsigma = m->nsigma;
    v->nobn = m->nobn;
    v->elem2D_nodes = m->f_en.data();
    v->edge_nodes = m->f_ednodes.data();
    v->edge_tri = m->f_edtri.data();
    v->elem_edges = m->f_eledges.data();
    v->elem_neighbors = m->f_elnb.data();
    v->index_nod2D = m->index.data();
    v->in_obn = m->in_obn.data();
    v->depth = m->depth.data();
    v->sigma = m->sigma.data();
    v->area = m->area.data();
    v->elem_area = m->elem_area.data();
    v->elem_cos = m->elem_cos.data();
    v->coriolis = m->coriolis.data();
    v->w_cv = m->w_cv.data();
    v->gradient_sca = m->gsca.data();
    v->gradient_vec = m->gvec.data();
    v->edge_dxdy = m->edxdy.data();
    v->edge_cross_dxdy = m->ecross.data();
    v->ac = m->ac.data();
    v->a_tp = m->a_tp.data();
    v->ampt = m->ampt.data();
    v->fazt = m->fazt.data();
    v->amp_factor = m->amp_factor.data();
    v->phy_factor = m->phy_factor.data();
    v->coord_nod2D = m->xy.data();
    v->density_0 = m->density_0;
    v->scale_area = m->scale_area;
    return FCG_OK;
}

extern "C" void fcg_mesh_free(fcg_mesh_obj *m) { delete m; }
