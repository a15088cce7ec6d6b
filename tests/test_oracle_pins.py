"""CPU checks that pin the FP64 oracle (oracle/fesomc_oracle.c) to the
reference: its own golden files and closed-form / invariant properties.

Reference-held pins used here:
* test_cases/LE_medium/postproc/energy_tocheck.dat (first rows, 7 significant
  digits) -- the whole internal step: pressure/BPG, d3_end, momentum,
  barotropic subcycle, upwind tracers, implicit vertical diffusion.
* test_cases/LE_medium/MESH/edges.out, edge_tri.out -- edge numbering and
  orientation of find_edges.
"""
import gzip
import math
import os

import numpy as np
import pytest

from oracle.pyoracle import Oracle, densityJM

from conftest import GOLDEN

LE = os.path.join(GOLDEN, "le_medium")


def digits7_ok(value, ref, units=0.6):
    """|value-ref| within `units` of the 7th significant digit (format 4e16.7)."""
    ulp7 = 10.0 ** (math.floor(math.log10(abs(ref))) - 6)
    return abs(value - ref) <= units * ulp7


@pytest.fixture(scope="module")
def le_oracle(le_case):
    txt, mesh = le_case
    return Oracle(txt, mesh)


def test_le_energy_series_first_rows(le_case):
    """energy(eout) every IREP=100 steps (src/fv_main.F90:614-619) equals the
    reference's energy_tocheck.dat to its 7 printed digits."""
    txt, mesh = le_case
    o = Oracle(txt, mesh)
    ref = np.loadtxt(os.path.join(LE, "energy_tocheck_head200.dat"))
    for row in range(2):
        o.step(100)
        days = o.scalar("time") / 86400.0 - o.scalar("time_jd0")
        e = o.energy()
        assert digits7_ok(e, ref[row, 1]), (row, e, ref[row, 1])
        # time column: same accumulation (src/fv_main.F90:400-401); the producing
        # build differs in the last printed digits of the first rows
        assert abs(days - ref[row, 0]) <= 1e-6 * ref[row, 0]


def test_le_energy_series_30_rows(le_case):
    """The first 30 rows of energy_tocheck.dat (3000 steps, 1.7 days of the
    lock exchange) at the file's 7 printed digits.  Later rows diverge:
    the flow is chaotic and the producing build is unknown."""
    txt, mesh = le_case
    o = Oracle(txt, mesh)
    ref = np.loadtxt(os.path.join(LE, "energy_tocheck_head200.dat"))
    for row in range(30):
        o.step(100)
        e = o.energy()
        assert digits7_ok(e, ref[row, 1]), (row, e, ref[row, 1])


def test_le_edges_match_reference_edge_files(le_oracle):
    """find_edges numbering/orientation equals MESH/edges.out and edge_tri.out."""
    en = le_oracle.get_int("edge_nodes").reshape(-1, 2) + 1
    et = le_oracle.get_int("edge_tri").reshape(-1, 2)
    et = np.where(et >= 0, et + 1, et)
    ref_e = np.loadtxt(gzip.open(os.path.join(LE, "edges.out.gz"), "rt"), dtype=int)
    ref_t = np.loadtxt(gzip.open(os.path.join(LE, "edge_tri.out.gz"), "rt"), dtype=int)
    nums = np.loadtxt(gzip.open(os.path.join(LE, "edgenum.out.gz"), "rt"), dtype=int)
    assert le_oracle.size["D"] == nums[0] and le_oracle.size["D_in"] == nums[1]
    assert np.array_equal(en, ref_e)
    assert np.array_equal(et, ref_t)


@pytest.mark.parametrize("S,T,rho", [(0.0, 5.0, 999.96675), (35.0, 5.0, 1027.67547), (35.0, 25.0, 1023.34306)])
def test_density_surface_eos80_check_values(S, T, rho):
    """At zero pressure densityJM reduces to the UNESCO one-atmosphere EOS-80
    polynomial, whose published check values it must reproduce."""
    assert abs(densityJM(T, S, 0.0, 0.0) - rho) < 6e-6


def test_density_pressure_increases_density():
    """Compressibility: density grows with depth (pz=-Z) at fixed T,S."""
    r = [densityJM(10.0, 30.0, -z, 0.0) for z in (0.0, 10.0, 100.0, 1000.0)]
    assert all(b > a for a, b in zip(r, r[1:]))


def test_jacobian_columns_sum_to_depth(le_oracle):
    """sum_k Jc(k,n) = max(depth+eta, Dmin) (closed form of src/fv_main.F90:1233-1237)."""
    L = le_oracle.size["L"]
    Jc = le_oracle.get("Jc").reshape(-1, L)
    d = np.maximum(le_oracle.get("depth") + le_oracle.get("eta_n"), 0.25)
    assert np.allclose(Jc.sum(1), d, rtol=1e-13, atol=0)


def _linear_field_grad(o, coeff_x, coeff_y):
    xy = o.get("coord").reshape(-1, 2)
    en = o.get_int("elem_nodes").reshape(-1, 4)
    gs = o.get("gradient_sca").reshape(-1, 8)
    R = 6400000.0
    f = coeff_x * xy[:, 0] * R + coeff_y * xy[:, 1] * R
    fx = (gs[:, :4] * f[en]).sum(1)
    fy = (gs[:, 4:] * f[en]).sum(1)
    return fx, fy


def test_gradient_sca_exact_for_linear_fields():
    """gradient_sca differentiates linear fields exactly on triangles and
    parallelograms (cartesian channel with split quads)."""
    from fesomc_amd import meshgen
    txt, mesh, _ = meshgen.config("c1")
    o = Oracle(txt, mesh)
    fx, fy = _linear_field_grad(o, 2.0, -3.0)
    assert np.allclose(fx, 2.0, rtol=1e-9) and np.allclose(fy, -3.0, rtol=1e-9)


def test_lake_at_rest_stays_at_rest(le_case):
    """Uniform T: no pressure gradient, no motion (sigma-PGF is exact for flat
    bottom and uniform density)."""
    txt, mesh = le_case
    o = Oracle(txt, mesh)
    o.set_tracers(np.full(o.size["N"] * o.size["L"], 20.0), None)
    o.step(10)
    assert abs(o.get("eta_n")).max() < 1e-15
    assert abs(o.get("U_n")).max() < 1e-15
    assert abs(o.energy()) < 1e-25


def test_closed_basin_conserves_volume(le_case):
    """No open boundary: sum(area*eta) stays at round-off (antisymmetric edge
    fluxes of compute_ssh_rhs_elem)."""
    txt, mesh = le_case
    o = Oracle(txt, mesh)
    area = o.get("area")
    o.step(20)
    vol = float((area * o.get("eta_n")).sum())
    scale = float((area * np.abs(o.get("eta_n"))).sum())
    assert abs(vol) <= 1e-11 * scale


def test_uniform_tracer_stays_uniform_c1():
    """Upwind advection in advective form keeps a uniform tracer uniform, with
    tides flowing through the open boundary (c1 channel)."""
    from fesomc_amd import meshgen
    txt, mesh, _ = meshgen.config("c1")
    o = Oracle(txt, mesh)
    o.step(20)
    T = o.get("TF")
    S = o.get("SF")
    assert np.abs(T - 10.0).max() < 1e-10 and np.abs(S - 30.0).max() < 1e-10
    assert np.abs(o.get("eta_n")).max() > 1e-6   # the tide is moving water


def test_min_max_principle_upwind():
    """Stratified + noisy T/S, wind, wetting/drying: extrema do not grow beyond
    the initial range by more than the vertical-diffusion round-off."""
    from fesomc_amd import meshgen
    txt, mesh, ex = meshgen.config("c2_small")
    txt = txt.replace("tracer_adv = 2", "tracer_adv = 1")
    o = Oracle(txt, mesh)
    o.set_tracers(ex["T"].reshape(-1), ex["S"].reshape(-1))
    T0 = ex["T"]
    o.step(10)
    T = o.get("TF")
    assert T.max() <= T0.max() + 1e-6 and T.min() >= T0.min() - 1e-6


# ---- surface forcing (key_atm): src/fv_tracer.F90:1376-1413, 1635-1910,
# src/fv_average_dynamic.F90:278-285, src/fv_3D_vel.F90:40-53

def test_sw_cp_unesco_check_values():
    """UNESCO 1983 (Fofonoff and Millard) check value of the heat capacity:
    3849.500 J/(kg degC) at S = 40, T68 = 40 degC (IPTS-68; the routine takes
    ITS-90 and converts T68 = 1.00024 T), P = 10000 dbar (1000 bar); at
    S = 0, T = 0, P = 0 the polynomial reduces to its constant 4217.4."""
    from oracle.pyoracle import sw_cp
    assert sw_cp(40.0, 40.0 / 1.00024, 1000.0) == pytest.approx(3849.500, abs=5e-3)
    assert sw_cp(0.0, 0.0, 0.0) == 4217.4


def _le_atm(le_case, extra):
    import re
    txt, mesh = le_case
    txt, n = re.subn(r"^(\s*key_atm\s*=\s*)\S+", r"\g<1>T", txt, count=1, flags=re.M)
    assert n == 1
    return Oracle(txt + extra, mesh), mesh


def _content(o, field):
    """sum over nodes and layers of area * Jc * field (the tracer content)."""
    L = o.size["L"]
    return float((o.get("area")[:, None] * o.get("Jc").reshape(-1, L) * o.get(field).reshape(-1, L)).sum())


def test_surface_heat_flux_closes_the_heat_budget(le_case):
    """In the closed LE basin advection and diffusion conserve heat content,
    so one step with a uniform non-solar heat flux Q (qsr = 0) adds
    sum(area*dt*Q/(c_p*rho)) of heat content against the same step without
    it (c_p*rho ~ 3.9e6 J/m3/K: the check catches a lost term, a wrong sign,
    a missing dt or Jc factor)."""
    from fesomc_amd.runconfig import RunConfig
    Q = 500.0
    dH = []
    for q in (0.0, Q):
        o, mesh = _le_atm(le_case, "")
        N = o.size["N"]
        o.set("mslp", np.full(N, 101325.0))
        o.set("qns", np.full(N, q))
        h0 = _content(o, "TF")
        o.step(1)
        dH.append(_content(o, "TF") - h0)
    dt = RunConfig(le_case[0]).f("dt")
    want = float((o.get("area") * dt * Q).sum()) / 3.9e6
    assert dH[1] - dH[0] == pytest.approx(want, rel=0.05)


def test_freshwater_flux_closes_the_salt_budget():
    """One step of the stratified channel with a uniform evaporation minus
    precipitation E adds sum(area*dt*S_surface*E/rho_fresh) of salt content
    over the wet nodes against the same step without it (rho_fresh ~ 1000
    kg/m3; the open boundary passes the same fluxes in both runs)."""
    import re
    from fesomc_amd import meshgen
    from fesomc_amd.runconfig import RunConfig
    txt, mesh, ex = meshgen.config("c2_small")
    txt = re.sub(r"^(\s*key_atm\s*=\s*)\S+", r"\g<1>T", txt, count=1, flags=re.M)
    E = 1e-4
    dS = []
    for e in (0.0, E):
        o = Oracle(txt, mesh)
        o.set_tracers(ex["T"].reshape(-1), ex["S"].reshape(-1))
        N = o.size["N"]
        o.set("mslp", np.full(N, 101325.0))
        o.set("emp", np.full(N, e))
        s0 = _content(o, "SF")
        o.step(1)
        dS.append(_content(o, "SF") - s0)
    dt = RunConfig(txt).f("dt")
    wet = o.get("mask_wd_node") != 0
    S_surf = o.get("SF").reshape(-1, o.size["L"])[:, 0]
    want = float((o.get("area") * dt * S_surf * E)[wet].sum()) / 1.0e3
    assert want > 0
    assert dS[1] - dS[0] == pytest.approx(want, rel=0.05)


def test_uniform_air_pressure_exerts_no_force(le_case):
    """A uniform sea-level pressure and freshwater flux add the same value to
    every node's pressure term, whose gradient vanishes: velocities after a
    few steps equal those of the run without surface forcing to round-off."""
    txt, mesh = le_case
    o0 = Oracle(txt, mesh)
    o1, _ = _le_atm(le_case, "")
    N = o1.size["N"]
    o1.set("mslp", np.full(N, 101325.0 + 2000.0))
    o0.step(3)
    o1.step(3)
    u0, u1 = o0.get("U_n"), o1.get("U_n")
    assert np.abs(u1 - u0).max() <= 1e-9 * np.abs(u0).max()


def test_river_runoff_closes_the_volume_budget(le_case):
    """A river of Q m3/s at one node of the closed LE basin raises the total
    volume sum(area*(depth+eta)) by Q per barotropic second: Mt*dt_2D*Q over
    one internal step (compute_ssh_rhs_elem, src/fv_average_dynamic.F90:
    177-183; the edge fluxes cancel in the sum)."""
    import re
    from fesomc_amd.runconfig import RunConfig
    txt, mesh = le_case
    txt = re.sub(r"^(\s*key_rivers\s*=\s*)\S+", r"\g<1>T", txt, count=1, flags=re.M)
    cfg = RunConfig(txt)
    Q = 5.0
    o = Oracle(txt, mesh)
    N = o.size["N"]
    r = np.zeros(N)
    r[N // 3] = Q
    o.set("runoff", r)
    area = o.get("area")
    e0 = o.get("eta_n").copy()
    o.step(1)
    dv = float((area * (o.get("eta_n") - e0)).sum())
    assert dv == pytest.approx(o.size["Mt"] * cfg.f("dt_2D") * Q, rel=1e-9)
    # Without the river the basin keeps its volume: the term is the river's.
    o2 = Oracle(txt, mesh)
    e0 = o2.get("eta_n").copy()
    o2.step(1)
    assert abs(float((area * (o2.get("eta_n") - e0)).sum())) < 1e-9 * o.size["Mt"] * cfg.f("dt_2D") * Q


def test_river_water_dilutes_its_node():
    """A river of Q m3/s at one wet interior node of the stratified channel
    brings Q*dt m3 of water of salinity 0.2 per step: the salt content
    sum(area*Jc*S) gains Q*dt*0.2 against the same step without it.  The
    advective-form transport keeps the node's concentration while its volume
    grows by Q*dt (+Q*dt*S_col, S_col the layer-thickness-weighted column
    mean), and the dilution term of src/fv_main.F90:1805-1822 takes
    Q*dt*(S_col - 0.2) back.  A dropped dilution term leaves +Q*dt*S_col
    (~150 times more), a wrong sign or a lost dt, area or Jc factor misses by
    far more than the tolerance."""
    import re
    from fesomc_amd import meshgen
    from fesomc_amd.runconfig import RunConfig
    txt, mesh, ex = meshgen.config("c2_small")
    txt = re.sub(r"^(\s*key_rivers\s*=\s*)\S+", r"\g<1>T", txt, count=1, flags=re.M)
    Q = 50.0
    dS, node = [], None
    for q in (0.0, Q):
        o = Oracle(txt, mesh)
        o.set_tracers(ex["T"].reshape(-1), ex["S"].reshape(-1))
        N, L = o.size["N"], o.size["L"]
        if node is None:
            depth = o.get("depth")
            node = int(np.argmax(np.where(o.get("ac") >= 1.0, depth, -np.inf)))
        r = np.zeros(N)
        r[node] = q
        o.set("runoff", r)
        S0 = o.get("SF").reshape(N, L)[node].copy()
        s0 = _content(o, "SF")
        o.step(1)
        dS.append(_content(o, "SF") - s0)
    dsig = -np.diff(o.get("sigma"))
    S_col = float((S0 * dsig).sum() / dsig.sum())
    Qdt = Q * RunConfig(txt).f("dt")
    assert S_col > 10.0
    assert dS[1] - dS[0] == pytest.approx(Qdt * 0.2, rel=0.05)


# ---- neighbour-difference filters (A13: src/fv_viscosity.F90:769-939)

def _filter_case():
    import re
    from fesomc_amd import meshgen
    txt, mesh, ex = meshgen.config("c2_small")
    for k in ("filt_2D", "filt_3D"):
        txt, n = re.subn(rf"^(\s*{k}\s*=\s*)\S+", r"\g<1>T", txt, count=1, flags=re.M)
        assert n == 1, k
    return txt, mesh, ex


def _neighbours(o):
    E = o.size["E"]
    en = o.get_int("elem_nodes").reshape(E, 4)
    nb = o.get_int("elem_neighbors").reshape(E, 4)
    jend = np.where(en[:, 0] == en[:, 3], 3, 4)
    return [[int(nb[e, j]) for j in range(jend[e]) if nb[e, j] >= 0] for e in range(E)]


def test_filter_2d_is_a_discrete_laplacian():
    """filt_2D on UAB: zero for a uniform field; for a unit spike at one
    interior element, -n tau area D at the element (n neighbours, D the
    w_cv-weighted depth) and +tau area D at each neighbour, nothing
    elsewhere; for any field, sum over elements of the term / (tau area D)
    vanishes (each interior edge enters once with each sign)."""
    txt, mesh, ex = _filter_case()
    o = Oracle(txt, mesh)
    E = o.size["E"]
    from fesomc_amd.runconfig import RunConfig
    tau = RunConfig(txt).f("tau_inv_filt")
    assert tau > 0
    en = o.get_int("elem_nodes").reshape(E, 4)
    D = (o.get("w_cv").reshape(E, 4) * o.get("depth")[en]).sum(1)
    area = o.get("elem_area")
    nbs = _neighbours(o)
    o.set("UAB", np.tile([0.3, -0.2], E))
    o.filter_only(2)
    assert np.abs(o.get("U_rhs_2D")).max() < 1e-15 * (tau * area * D).max()
    e0 = next(e for e in range(E) if len(nbs[e]) == len(set(nbs[e])) >= 3)
    u = np.zeros(2 * E)
    u[2 * e0] = 1.0
    o.set("UAB", u)
    o.filter_only(2)
    r = o.get("U_rhs_2D").reshape(E, 2)
    want = np.zeros(E)
    want[e0] = -len(nbs[e0]) * tau * area[e0] * D[e0]
    for b in nbs[e0]:
        want[b] = tau * area[b] * D[b]
    np.testing.assert_allclose(r[:, 0], want, rtol=1e-13, atol=0)
    assert not r[:, 1].any()
    rng = np.random.default_rng(5)
    o.set("UAB", rng.standard_normal(2 * E))
    o.filter_only(2)
    lap = o.get("U_rhs_2D").reshape(E, 2) / (tau * area * D)[:, None]
    assert np.abs(lap.sum(0)).max() < 1e-10 * np.abs(lap).sum()


def test_filter_3d_and_3d_to_2d_are_layerwise_laplacians():
    """filt_3D on U_n at every layer: the spike stencil times Je; filt_3D_to_2D
    on the pulsation transports U_puls*Je: its column sum equals the sum over
    layers of the filt_3D stencil applied to U_puls*Je, times Je."""
    txt, mesh, ex = _filter_case()
    o = Oracle(txt, mesh)
    E, L = o.size["E"], o.size["L"]
    from fesomc_amd.runconfig import RunConfig
    tau = RunConfig(txt).f("tau_inv_filt")
    area = o.get("elem_area")
    Je = o.get("Je").reshape(E, L)
    nbs = _neighbours(o)
    e0 = next(e for e in range(E) if len(nbs[e]) == len(set(nbs[e])) >= 3)
    u = np.zeros((E, L))
    u[e0, 2] = 1.0
    o.set("U_n", u.reshape(-1))
    o.set("V_n", np.zeros(E * L))
    o.filter_only(3)
    r = o.get("U_rhs").reshape(E, L)
    want = np.zeros((E, L))
    want[e0, 2] = -len(nbs[e0]) * tau * area[e0] * Je[e0, 2]
    for b in nbs[e0]:
        want[b, 2] = tau * area[b] * Je[b, 2]
    np.testing.assert_allclose(r, want, rtol=1e-13, atol=0)
    assert not o.get("V_rhs").any()
    # 3D_to_2D: spike in the pulsation velocity at layer 1 of e0
    p = np.zeros((E, L))
    p[e0, 1] = 2.0
    o.set("U_puls", p.reshape(-1))
    o.set("V_puls", np.zeros(E * L))
    o.filter_only(4)
    r2 = o.get("U_rhs_2D_3D").reshape(E, 2)
    want2 = np.zeros(E)
    want2[e0] = -len(nbs[e0]) * tau * 2.0 * Je[e0, 1] * area[e0] * Je[e0, 1]
    for b in nbs[e0]:
        want2[b] = tau * 2.0 * Je[e0, 1] * area[b] * Je[b, 1]
    np.testing.assert_allclose(r2[:, 0], want2, rtol=1e-13, atol=0)
