"""GPU parity: the HIP FP64 drop-in against the FP64 oracle on the same
inputs.  The FP64 kernels keep the reference's operation order and use
deterministic gathers in ascending edge order, and both sides evaluate the
step's one logarithm (d3_end's bottom boundary value) correctly rounded, so
the agreement is bit for bit: the comparisons below are exact (RTOL = 0)."""
import os

import numpy as np
import pytest

from oracle.pyoracle import Oracle

pytestmark = pytest.mark.gpu

FIELDS = ["eta_n", "eta_n_1", "eta_n_2", "U_n_2D", "U_n", "V_n", "U_n_filt", "V_n_filt", "Wvel", "TF", "SF",
          "T_old", "S_old", "Kv", "snu", "bt", "U_rhsAB", "V_rhsAB", "Bar_pr_2D", "mask_wd", "etaAB", "C_d_el",
          "Jc", "Jc_old", "Z", "zbar"]
RTOL = 0.0


def _model(txt, mesh, ex=None):
    from fesomc_amd import gpu, meshgen
    from fesomc_amd.runconfig import RunConfig
    g = gpu.Model(RunConfig(txt), mesh)
    o = Oracle(txt, mesh)
    ex = ex or {}
    if ex.get("T") is not None:
        g.set_tracers(ex["T"], ex["S"])
        o.set_tracers(ex["T"].reshape(-1), ex["S"].reshape(-1))
    if ex.get("wind") is not None:
        wx, wy = ex["wind"]
        tx, ty = meshgen.wind_stress(wx, wy)
        for m in (g, o):
            m.set("windx", np.full(mesh.nod2D, wx))
            m.set("windy", np.full(mesh.nod2D, wy))
            m.set("taux", np.full(mesh.elem2D, tx))
            m.set("tauy", np.full(mesh.elem2D, ty))
    return g, o


def _compare(g, o, fields=FIELDS):
    for f in fields:
        a, b = g.get(f), o.get(f)
        scale = float(np.abs(b).max())
        err = float(np.abs(a - b).max())
        assert err <= RTOL * scale, f"{f}: max|gpu-oracle| = {err:.3e} (scale {scale:.3e})"


def test_le_medium_steps_match_oracle(le_case):
    txt, mesh = le_case
    g, o = _model(txt, mesh)
    g.step(5)
    o.step(5)
    _compare(g, o)
    g.close()


def test_le_medium_energy_matches_reference_series(le_case):
    """100 GPU steps reproduce row 1 of the reference's energy_tocheck.dat."""
    from test_oracle_pins import digits7_ok
    txt, mesh = le_case
    from fesomc_amd import gpu
    from fesomc_amd.runconfig import RunConfig
    g = gpu.Model(RunConfig(txt), mesh)
    g.step(100)
    d = g.diagnostics()
    assert d["status"] == 0
    assert digits7_ok(d["energy"], 8.853516e-05)
    g.close()


def test_channel_tides_triangles_match_oracle():
    from fesomc_amd import meshgen
    txt, mesh, ex = meshgen.config("c1")
    g, o = _model(txt, mesh, ex)
    g.step(10)
    o.step(10)
    _compare(g, o)
    assert abs(g.diagnostics()["energy"] - o.energy()) <= 1e-12 * o.energy()
    g.close()


def test_wetdry_stratified_wind_match_oracle():
    """Wetting/drying, stratified T/S with noise, wind stress, upwind tracers."""
    from fesomc_amd import meshgen
    txt, mesh, ex = meshgen.config("c2_small")
    txt = txt.replace("tracer_adv = 2", "tracer_adv = 1").replace("MUSCL_CALL = T", "MUSCL_CALL = F")
    g, o = _model(txt, mesh, ex)
    g.step(10)
    o.step(10)
    _compare(g, o)
    assert np.array_equal(g.get("mask_wd"), o.get("mask_wd")), "wet/dry mask must match bit-exactly"
    g.close()


SCHEMES = {
    "muscl": ("tracer_adv = 2", "MUSCL_CALL = T"),
    "miura": ("tracer_adv = 4", "MUSCL_CALL = F"),
    "none": ("tracer_adv = 2", "MUSCL_CALL = F"),
}


def _with_scheme(txt, scheme):
    adv, muscl = SCHEMES[scheme]
    for k in ("tracer_adv = 1", "tracer_adv = 2", "tracer_adv = 4"):
        txt = txt.replace(k, adv)
    return txt.replace("MUSCL_CALL = T", muscl).replace("MUSCL_CALL = F", muscl)


@pytest.mark.parametrize("scheme", sorted(SCHEMES))
@pytest.mark.parametrize("case", ["c1", "c2_small"])
def test_tracer_schemes_match_oracle(case, scheme):
    """MUSCL (solve_tracer_muscl), Miura (solve_tracer_miura, T only) and the
    no-advection tracer step against the oracle: tides + open-boundary
    climatology relaxation on the channel with triangles (c1), wet/dry +
    stratified T/S + wind on c2_small."""
    from fesomc_amd import meshgen
    txt, mesh, ex = meshgen.config(case)
    txt = _with_scheme(txt, scheme)
    if case == "c1":  # stratified tracers, climatology offset so the OB relaxation acts
        T, S = meshgen.stratified_ts(mesh, 10, seed=3)
        ex = dict(ex, T=T, S=S)
    g, o = _model(txt, mesh, ex)
    if case == "c1":
        for m in (g, o):
            m.set("Tclim", ex["T"] + 0.5)
            m.set("Sclim", ex["S"] - 0.25)
    g.step(8)
    o.step(8)
    _compare(g, o)
    if scheme != "muscl":  # S is not advected under Miura or without advection
        assert np.array_equal(g.get("S_old"), o.get("S_old"))
    g.close()


def test_runs_are_bitwise_deterministic():
    from fesomc_amd import meshgen
    txt, mesh, ex = meshgen.config("c1")
    outs = []
    for _ in range(2):
        g, _o = _model(txt, mesh, ex)
        g.step(5)
        outs.append([g.get(f) for f in ("eta_n", "U_n", "TF", "Wvel")])
        g.close()
    for a, b in zip(*outs):
        assert np.array_equal(a, b)


def test_negative_temperature_stop_is_reported():
    """The reference STOPs when T<-3 after the tracer step
    (src/fv_tracer.F90:1478-1482); the device raises the same condition as a
    status code.  Forced from the host with one cold cell; the oracle stops too."""
    from fesomc_amd import meshgen
    txt, mesh, ex = meshgen.config("c1")
    g, o = _model(txt, mesh, ex)
    T = np.full((g.N, g.L), 10.0)
    T[100, 3] = -3.5
    for m in (g, o):
        m.set("TF", T)
        m.set("T_old", T)
    g.step(1)
    d = g.diagnostics()
    assert d["status"] == -6, d
    with pytest.raises(RuntimeError, match="Negative temperature"):
        o.step(1)
    g.close()


@pytest.mark.parametrize("swr", [1, 2])
def test_surface_forcing_matches_oracle(swr):
    """key_atm on the stratified wet/dry channel: spatially varying shortwave
    and non-solar heat flux, evaporation minus precipitation and sea-level
    pressure (fv_average_dynamic.F90:278-285, fv_3D_vel.F90:40-53,
    fv_tracer.F90:1376-1413), with the Paulson-Simpson (1) or Soloviev (2)
    shortwave profile: bit for bit against the oracle."""
    import re
    from fesomc_amd import meshgen
    txt, mesh, ex = meshgen.config("c2_small")
    txt = re.sub(r"^(\s*key_atm\s*=\s*)\S+", r"\g<1>T", txt, count=1, flags=re.M)
    txt += f"\n type_swr_body = {swr}\n jer_const = T\n R = 0.58\n a = 0.35\n b = 23.0\n"
    g, o = _model(txt, mesh, ex)
    x, y = mesh.xy[:, 0], mesh.xy[:, 1]
    forcing = {"qsr": 300.0 * (0.5 + 0.5 * np.sin(2 * np.pi * x / 20000.0)),
               "qns": -80.0 + 40.0 * np.cos(2 * np.pi * y / 10000.0),
               "emp": 2e-5 * np.sin(2 * np.pi * (x + y) / 15000.0),
               "mslp": 101325.0 + 500.0 * np.sin(2 * np.pi * y / 10000.0)}
    for k, v in forcing.items():
        g.set(k, v)
        o.set(k, v)
    g.step(5)
    o.step(5)
    _compare(g, o, FIELDS + ["rho_c"])
    assert g.diagnostics()["status"] == 0
    g.close()


def test_rivers_match_oracle():
    """key_rivers on the stratified wet/dry channel: discharge at three
    interior nodes adds to their elevation's right-hand side
    (fv_average_dynamic.F90:177-183) and dilutes their salinity towards 0.2
    (fv_main.F90:1805-1822): bit for bit against the oracle."""
    import re
    from fesomc_amd import meshgen
    txt, mesh, ex = meshgen.config("c2_small")
    txt = re.sub(r"^(\s*key_rivers\s*=\s*)\S+", r"\g<1>T", txt, count=1, flags=re.M)
    g, o = _model(txt, mesh, ex)
    depth = o.get("depth")
    mouths = np.argsort(np.where(o.get("ac") >= 1.0, depth, -np.inf))[-3:]
    r = np.zeros(mesh.nod2D)
    r[mouths] = (50.0, 20.0, 5.0)
    g.set("runoff", r)
    o.set("runoff", r)
    S0 = o.get("SF").copy()
    g.step(5)
    o.step(5)
    _compare(g, o, FIELDS)
    assert g.diagnostics()["status"] == 0
    L = o.size["L"]
    assert (o.get("SF").reshape(-1, L)[mouths] < S0.reshape(-1, L)[mouths]).all()
    g.close()


def test_filters_match_oracle():
    """filt_2D and filt_3D (fv_viscosity.F90:769-939) on the stratified wet/dry
    channel, tau_inv_filt = 1e-3/s: the 2-D filter in the barotropic element
    pass, the 3-D one in the momentum right-hand side and, on the pulsation
    transports, in the barotropic forcing of the first substep: bit for bit
    against the oracle, and visibly different from the unfiltered run."""
    import re
    from fesomc_amd import meshgen
    txt, mesh, ex = meshgen.config("c2_small")
    for k, v in (("filt_2D", "T"), ("filt_3D", "T"), ("tau_inv_filt", "1.0e-3")):
        txt, n = re.subn(rf"^(\s*{k}\s*=\s*)\S+", rf"\g<1>{v}", txt, count=1, flags=re.M)
        assert n == 1, k
    g, o = _model(txt, mesh, ex)
    g.step(5)
    o.step(5)
    _compare(g, o, FIELDS)
    assert g.diagnostics()["status"] == 0
    g.close()
    plain, _, _ = meshgen.config("c2_small")
    o0 = Oracle(plain, mesh)
    o0.set_tracers(ex["T"].reshape(-1), ex["S"].reshape(-1))
    wx, wy = ex["wind"]
    tx, ty = meshgen.wind_stress(wx, wy)
    for k, v in (("windx", np.full(mesh.nod2D, wx)), ("windy", np.full(mesh.nod2D, wy)),
                 ("taux", np.full(mesh.elem2D, tx)), ("tauy", np.full(mesh.elem2D, ty))):
        o0.set(k, v)
    o0.step(5)
    assert np.abs(o0.get("U_n") - o.get("U_n")).max() > 1e-6 * np.abs(o.get("U_n")).max()


def test_twice_applied_filters_match_oracle():
    """filt_bi_2D and filt_bi_3D (viscosity_filt2x_2D/_3D, fv_viscosity.F90:
    570-767) on the stratified wet/dry channel at tau_c = 20/day: the 2-D
    filter's first pass before the barotropic element update of every
    substep and its second pass inside it (UV2_rhs), the 3-D one as two
    element-layer passes whose UV_rhs enters the momentum update: bit for bit
    against the oracle, and visibly different from the unfiltered run."""
    import re
    from fesomc_amd import meshgen
    txt, mesh, ex = meshgen.config("c2_small")
    for k, v in (("filt_bi_2D", "T"), ("filt_bi_3D", "T"), ("tau_c", "20.0")):
        txt, n = re.subn(rf"^(\s*{k}\s*=\s*)\S+", rf"\g<1>{v}", txt, count=1, flags=re.M)
        assert n == 1, k
    g, o = _model(txt, mesh, ex)
    print("filt_bi: c2_small, 5 steps", flush=True)
    g.step(5)
    o.step(5)
    _compare(g, o, FIELDS)
    assert g.diagnostics()["status"] == 0
    g.close()
    assert np.abs(o.get("UV_rhs")).max() > 0 and np.abs(o.get("UV2_rhs")).max() > 0
    plain, _, _ = meshgen.config("c2_small")
    o0 = Oracle(plain, mesh)
    o0.set_tracers(ex["T"].reshape(-1), ex["S"].reshape(-1))
    wx, wy = ex["wind"]
    tx, ty = meshgen.wind_stress(wx, wy)
    for k, v in (("windx", np.full(mesh.nod2D, wx)), ("windy", np.full(mesh.nod2D, wy)),
                 ("taux", np.full(mesh.elem2D, tx)), ("tauy", np.full(mesh.elem2D, ty))):
        o0.set(k, v)
    o0.step(5)
    assert np.abs(o0.get("U_n") - o.get("U_n")).max() > 1e-6 * np.abs(o.get("U_n")).max()
    assert np.abs(o0.get("U_n_2D") - o.get("U_n_2D")).max() > 1e-6 * np.abs(o.get("U_n_2D")).max()


def test_biharmonic_viscosity_matches_oracle():
    """bih_2D and bih_3D (biharmonic_2D, biharmonic_viscosity,
    biharmonic_viscosity_3D_to_2D; fv_viscosity.F90:104-568, 964-1220) on
    the stratified wet/dry channel at Abh0 = 1e5: each corrected Laplacian
    as an edge pass and an element gather in ascending reference edge id, the
    gradients of the first result in between, the second pass's sums inside
    the momentum update, the barotropic element update and the first
    substep's 3-D-to-2-D forcing: bit for bit against the oracle after 5
    steps, and visibly different from the run without it."""
    import re
    from fesomc_amd import meshgen
    txt, mesh, ex = meshgen.config("c2_small")
    for k, v in (("bih_2D", "T"), ("bih_3D", "T"), ("Abh0", "1.0e5")):
        txt, n = re.subn(rf"^(\s*{k}\s*=\s*)\S+", rf"\g<1>{v}", txt, count=1, flags=re.M)
        assert n == 1, k
    g, o = _model(txt, mesh, ex)
    print("bih: c2_small, 5 steps", flush=True)
    g.step(5)
    o.step(5)
    _compare(g, o, [f for f in FIELDS if not f.startswith("vel_grad")])
    assert g.diagnostics()["status"] == 0
    g.close()
    plain, _, _ = meshgen.config("c2_small")
    o0 = Oracle(plain, mesh)
    o0.set_tracers(ex["T"].reshape(-1), ex["S"].reshape(-1))
    wx, wy = ex["wind"]
    tx, ty = meshgen.wind_stress(wx, wy)
    for k, v in (("windx", np.full(mesh.nod2D, wx)), ("windy", np.full(mesh.nod2D, wy)),
                 ("taux", np.full(mesh.elem2D, tx)), ("tauy", np.full(mesh.elem2D, ty))):
        o0.set(k, v)
    o0.step(5)
    assert np.abs(o0.get("U_n") - o.get("U_n")).max() > 1e-3 * np.abs(o.get("U_n")).max()
    assert np.abs(o0.get("U_n_2D") - o.get("U_n_2D")).max() > 1e-3 * np.abs(o.get("U_n_2D")).max()


def test_mixed_precision_with_rivers_filters_and_forcing():
    """The FP32-storage variants of the river, filter and forcing kernels
    (k_tr_rivers, k_velrhs_atm_filt, the filter terms of k_bt_elem2 and
    k_bt_p1elem, the twice-applied filters' and the biharmonic passes)
    against the FP64 device
    run on the same case: after 20
    steps the elevation agrees to 1e-4 m RMS and T, S to 1e-3 RMS."""
    import re
    from fesomc_amd import gpu, meshgen
    from fesomc_amd.runconfig import RunConfig
    txt, mesh, ex = meshgen.config("c2_small")
    for k, v in (("filt_2D", "T"), ("filt_3D", "T"), ("tau_inv_filt", "1.0e-3"), ("key_rivers", "T"),
                 ("key_atm", "T"), ("filt_bi_2D", "T"), ("filt_bi_3D", "T"), ("tau_c", "20.0"),
                 ("bih_2D", "T"), ("bih_3D", "T"), ("Abh0", "1.0e5")):
        txt, n = re.subn(rf"^(\s*{k}\s*=\s*)\S+", rf"\g<1>{v}", txt, count=1, flags=re.M)
        assert n == 1, k
    x, y = mesh.xy[:, 0], mesh.xy[:, 1]
    o = Oracle(txt, mesh)
    mouths = np.argsort(np.where(o.get("ac") >= 1.0, o.get("depth"), -np.inf))[-3:]
    r = np.zeros(mesh.nod2D)
    r[mouths] = (50.0, 20.0, 5.0)
    fields = {"qsr": 300.0 * (0.5 + 0.5 * np.sin(2 * np.pi * x / 20000.0)),
              "qns": np.full(mesh.nod2D, -60.0), "emp": np.full(mesh.nod2D, 1e-6),
              "mslp": 101325.0 + 500.0 * np.sin(2 * np.pi * y / 10000.0), "runoff": r}
    wx, wy = ex["wind"]
    tx, ty = meshgen.wind_stress(wx, wy)
    runs = []
    for prec in (gpu.FCG_FP64, gpu.FCG_MIXED):
        m = gpu.Model(RunConfig(txt), mesh, precision=prec)
        m.set_tracers(ex["T"], ex["S"])
        for k, v in fields.items():
            m.set(k, v)
        m.set("windx", np.full(mesh.nod2D, wx))
        m.set("windy", np.full(mesh.nod2D, wy))
        m.set("taux", np.full(mesh.elem2D, tx))
        m.set("tauy", np.full(mesh.elem2D, ty))
        m.step(20)
        assert m.diagnostics()["status"] == 0
        runs.append(m)
    a, b = runs
    assert np.sqrt(np.mean((a.get("eta_n") - b.get("eta_n")) ** 2)) < 1e-4
    assert np.sqrt(np.mean((a.get("TF") - b.get("TF")) ** 2)) < 1e-3
    assert np.sqrt(np.mean((a.get("SF") - b.get("SF")) ** 2)) < 1e-3
    for m in runs:
        m.close()


@pytest.mark.parametrize("fic", ["F", "T"])
def test_implicit_vertical_viscosity_matches_oracle(fic):
    """im_vert_visc = T (momentum_vert_impl_visc, fv_viscosity.F90:1266-1371:
    Thomas solve with the wind stress on the surface row and the explicit
    bottom drag on the bottom row), alone and followed by the _tmp solve
    (im_vert_visc_fic = T, fv_3D_vel.F90:183-186), on the stratified wet/dry
    channel: bit for bit against the oracle."""
    import re
    from fesomc_amd import meshgen
    txt, mesh, ex = meshgen.config("c2_small")
    for k, v in (("im_vert_visc", "T"), ("im_vert_visc_fic", fic)):
        txt, n = re.subn(rf"^(\s*{k}\s*=\s*)\S+", rf"\g<1>{v}", txt, count=1, flags=re.M)
        assert n == 1, k
    g, o = _model(txt, mesh, ex)
    g.step(8)
    o.step(8)
    _compare(g, o)
    assert g.diagnostics()["status"] == 0
    g.close()


def test_negative_salinity_stop_is_reported():
    """The reference STOPs when S<0 after the tracer step
    (src/fv_tracer.F90:1466-1473).  A fresh/salt front (S = 0 west of 5 km,
    30 east) under MUSCL undershoots below zero in the first step, in the
    oracle and on the device alike; the device reports FCG_E_SALINITY and
    holds the same salinity field, negative cells included.  (A negative
    salinity set from the host would not do: densityJM's sqrt(S**3) turns the
    state into NaN, which passes the reference's test; the device reports that
    as FCG_E_NAN, see test_non_finite_state_is_reported.)"""
    from fesomc_amd import meshgen
    txt, mesh, ex = meshgen.config("c1")
    txt = _with_scheme(txt, "muscl")
    g, o = _model(txt, mesh, ex)
    S = np.where(mesh.xy[:, 0:1] < 5000.0, 0.0, 30.0) * np.ones((1, g.L))
    g.set_tracers(None, S)
    o.set_tracers(None, S.reshape(-1))
    g.step(1)
    d = g.diagnostics()
    assert d["status"] == -5, d
    with pytest.raises(RuntimeError, match="Negative salinity"):
        o.step(1)
    sg, so = g.get("SF"), o.get("SF")
    assert sg.min() < 0.0
    assert np.abs(sg - so).max() <= RTOL * np.abs(so).max()
    g.close()


@pytest.mark.parametrize("field", ["TF", "eta_n"])
def test_non_finite_state_is_reported(field):
    """A NaN in the temperature or the elevation passes the reference's
    T < -3 and S < 0 tests; the device reports FCG_E_NAN instead of status 0."""
    from fesomc_amd import meshgen
    txt, mesh, ex = meshgen.config("c1")
    g, _ = _model(txt, mesh, ex)
    a = g.get(field)
    a[len(a) // 2] = np.nan
    g.set(field, a)
    g.step(1)
    d = g.diagnostics()
    assert d["status"] == -4, d
    g.close()


def test_python_host_applies_the_river_table(tmp_path):
    """gpu.Model with key_rivers and a two-record rivers.dat interpolates the
    table after each clock advance (rivers_do, src/fv_main.F90:400-412) and
    hands the discharge to the step: bit for bit against the oracle
    driven by the same table."""
    import re
    from fesomc_amd import gpu, meshgen
    from fesomc_amd.rivers import Rivers
    from fesomc_amd.runconfig import RunConfig
    txt, mesh, ex = meshgen.config("c2_small")
    txt = re.sub(r"^(\s*key_rivers\s*=\s*)\S+", r"\g<1>T", txt, count=1, flags=re.M)
    cfg = RunConfig(txt)
    jd0 = cfg.f("JD0")
    o = Oracle(txt, mesh)
    mouths = np.argsort(np.where(o.get("ac") >= 1.0, o.get("depth"), -np.inf))[-2:] + 1
    table = tmp_path / "rivers.dat"
    table.write_text(f"2\n2\n{jd0 - 1.0} {jd0 + 1.0}\n{mouths[0]} {mouths[1]}\n4.0e4 1.0e4\n8.0e4 2.0e4\n")
    g = gpu.Model(cfg, mesh, rivers=str(table))
    g.set_tracers(ex["T"], ex["S"])
    o.set_tracers(ex["T"].reshape(-1), ex["S"].reshape(-1))
    riv = Rivers(str(table), mesh.nod2D, jd0, jd0)
    t = jd0
    g.step(6)
    for _ in range(6):
        t = t + cfg.f("dt") / 86400.0
        o.set("runoff", riv.do(t))
        o.step(1)
    assert o.get("runoff").max() > 0.0
    assert np.array_equal(g.get("runoff"), o.get("runoff"))
    _compare(g, o, FIELDS)
    g.close()


def test_flooding_and_drying_shoreline_matches_oracle_at_24_levels():
    """Stelling-Duinmeijer wetting and drying with a moving shoreline, at the
    benchmark's column shape (24 sigma levels, MUSCL tracers, stratified T/S,
    wind): c2_small hot-started with a 1 m M2 tide of phase 90 degrees (no
    ten-day ramp), so the intertidal east end floods and then dries again.
    At checkpoints through both phases every field, and the element and node
    masks bit for bit, equal the oracle's."""
    from test_oracle_closed_forms import flooding_case
    txt, mesh, ex = flooding_case(nsigma=24)
    g, o = _model(txt, mesh, ex)
    m0 = o.get("mask_wd_node").copy()
    e0 = o.get("mask_wd").copy()
    done, flips = 0, []
    for stop in (120, 450, 650, 950):
        g.step(stop - done)
        o.step(stop - done)
        done = stop
        print(f"checkpoint step {stop}", flush=True)
        _compare(g, o)
        assert np.array_equal(g.get("mask_wd"), o.get("mask_wd")), stop
        # the device forms the node mask where it is used, from etaAB (wad_mask,
        # src/fv_main.F90:1301-1304); from the device's etaAB it is the oracle's
        node_mask = (o.get("depth") + g.get("etaAB") > 0.25).astype(float)
        assert np.array_equal(node_mask, o.get("mask_wd_node")), stop
        flips.append(int((o.get("mask_wd_node") != m0).sum()))
    assert g.diagnostics()["status"] == 0
    assert max(flips) > 60, flips                # flooded
    assert flips[-1] < max(flips), flips         # and dried again
    assert int((o.get("mask_wd") != e0).sum()) > 0
    g.close()


def test_mixed_precision_with_learned_closure_over_one_m2_cycle():
    """The headline path (FP32 3-D fields and fluxes, learned closure with its
    d3_end fallback) against the FP64 drop-in on the benchmark's physics
    (wet/dry, 24 levels, MUSCL, stratified T/S, wind, 3 km sponge) on a
    121 x 61 channel, both advanced one M2 period (44 712 s, 2236 steps of
    20 s): SSH within 1 cm RMS, T and S within 0.01 degC / psu RMS, as the
    precision target states; both runs end without a STOP."""
    import math
    from fesomc_amd import closure, gpu
    from fesomc_amd.workload import build_model
    steps = int(math.ceil(44712.0 / 20.0))
    runs = []
    for prec, clos in ((gpu.FCG_FP64, gpu.FCG_CLOSURE_D3END), (gpu.FCG_MIXED, gpu.FCG_CLOSURE_MLP)):
        txt, mesh, ex, m = build_model("c2", scale=(121, 61), precision=prec, closure=clos)
        if clos == gpu.FCG_CLOSURE_MLP:
            m.attach_mlp(closure.load(os.path.join(os.path.dirname(gpu.__file__), "closure_c2.npz")))
        assert abs(m.par.dt - 20.0) < 1e-12
        for done in range(0, steps, 500):
            m.step(min(500, steps - done))
            print(f"precision {prec} step {min(done + 500, steps)}/{steps}", flush=True)
        d = m.diagnostics()
        assert d["status"] == 0, d
        runs.append(m)
    a, b = runs
    rms = lambda f: float(np.sqrt(np.mean((a.get(f) - b.get(f)) ** 2)))
    assert rms("eta_n") <= 0.01
    assert rms("TF") <= 0.01
    assert rms("SF") <= 0.01
    for m in runs:
        m.close()
