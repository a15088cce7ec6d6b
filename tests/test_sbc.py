"""Surface boundary conditions from atmospheric fields (key_atm with the CORE
bulk formulas; src/fv_sbc.F90): oracle pins, the host reader of the forcing
files, and the device against the oracle.

The pins state what fixes the restated formulas independently of them: the
stability functions are the integrals of the Businger-Dyer gradient
functions; under exactly neutral conditions the transfer coefficients are
Large & Yeager's (2004) neutral ones; an unstable surface layer transfers
more than a neutral one and a stable one less; with a neutral layer the net
heat is the longwave balance of a grey body of emissivity 0.97 and the
shortwave reaches the ocean less a 0.1 albedo; evaporation cools.  Planted
slips in the oracle (pyoracle.set_mutation, ids 14-16) must fail them.
"""
import math
import os
import re

import numpy as np
import pytest

from oracle import pyoracle
from oracle.pyoracle import Oracle, set_mutation

RHOA = float(np.float32(1.3))         # air density, a REAL(4) literal (src/fv_sbc.F90:381)
STEF = 5.67e-8                        # Stefan-Boltzmann constant
EMISS = float(np.float32(0.97))       # sea-water emissivity (src/fv_sbc.F90:387)


@pytest.fixture(autouse=True)
def _no_mutation():
    set_mutation(0)
    yield
    set_mutation(0)


def _mutated(mid, check):
    set_mutation(mid)
    try:
        check()
    except AssertionError:
        return True
    finally:
        set_mutation(0)
    return False


def q_sat(zst):
    """The bulk formula's sea-surface saturation humidity (src/fv_sbc.F90:424)."""
    return 0.98 * 640380.0 / RHOA * math.exp(-5107.4 / zst)


def ly_neutral_cd(u):
    """Large & Yeager (2004) eq. 6a, 10 m neutral drag coefficient."""
    return 1e-3 * (2.7 / u + 0.142 + u / 13.09)


# ---------------------------------------------------------------------------
# stability functions and transfer coefficients
# ---------------------------------------------------------------------------
def _check_psi():
    from scipy.integrate import quad
    for zeta in (-8.0, -1.0, -0.1, -0.003):
        phi_m = lambda z: (1.0 - 16.0 * z) ** -0.25        # Businger-Dyer (Paulson 1970)
        phi_h = lambda z: (1.0 - 16.0 * z) ** -0.5
        im = quad(lambda z: (1.0 - phi_m(z)) / z, zeta, 0.0)[0]
        ih = quad(lambda z: (1.0 - phi_h(z)) / z, zeta, 0.0)[0]
        assert abs(pyoracle.psi_m(zeta) + im) <= 1e-8 * max(1.0, abs(im)), (zeta, pyoracle.psi_m(zeta), -im)
        assert abs(pyoracle.psi_h(zeta) + ih) <= 1e-8 * max(1.0, abs(ih)), (zeta, pyoracle.psi_h(zeta), -ih)
    for zeta in (0.0, 0.01, 0.5, 3.0, 10.0):                 # phi = 1 + 5 zeta
        assert pyoracle.psi_m(zeta) == pytest.approx(-5.0 * zeta, abs=1e-15)
        assert pyoracle.psi_h(zeta) == pytest.approx(-5.0 * zeta, abs=1e-15)


def test_stability_functions_are_the_businger_dyer_integrals():
    """psi(zeta) = integral_0^zeta (1 - phi(z))/z dz with phi_m = (1-16z)^-1/4,
    phi_h = (1-16z)^-1/2 for unstable layers and phi = 1 + 5z for stable ones."""
    _check_psi()
    assert _mutated(14, _check_psi)


@pytest.mark.parametrize("u", [0.2, 3.0, 10.0, 22.0])
def test_neutral_transfer_coefficients_are_large_yeager(u):
    """T_air = SST and q_air = q_sat: the iteration stays at Large & Yeager's
    neutral coefficients (eqs. 6a-6c, with the 0.5 m/s floor of the wind)."""
    sst = 285.0
    qs = q_sat(sst)
    Cd, Ch, Ce, T_zu, q_zu = pyoracle.core_coeff_2z(2.0, 10.0, sst, sst, qs, qs, u)
    cd = ly_neutral_cd(max(u, 0.5))
    assert Cd == pytest.approx(cd, rel=1e-14)
    assert Ce == pytest.approx(34.6e-3 * math.sqrt(cd), rel=1e-14)
    assert Ch == pytest.approx(18.0e-3 * math.sqrt(cd), rel=1e-14)
    assert T_zu == sst and q_zu == qs


def test_stability_orders_the_coefficients():
    """A sea warmer than the air (unstable) transfers more than a neutral
    layer, a colder one (stable) less; the 10 m air is colder than the 2 m
    air over a warm sea and warmer over a cold one."""
    sst, u = 288.0, 8.0
    qs = q_sat(sst)
    neutral = pyoracle.core_coeff_2z(2.0, 10.0, sst, sst, qs, qs, u)
    unstable = pyoracle.core_coeff_2z(2.0, 10.0, sst, sst - 4.0, qs, 0.8 * qs, u)
    stable = pyoracle.core_coeff_2z(2.0, 10.0, sst, sst + 4.0, qs, 0.8 * qs, u)
    for i in (0, 1, 2):
        assert unstable[i] > neutral[i] > stable[i], i
    assert unstable[3] < sst - 4.0 and stable[3] > sst + 4.0


def test_saturation_humidity_is_physical():
    """Within 10 % of the saturation specific humidity over sea water from
    Bolton's (1980) vapour pressure at 1013 hPa, x 0.98 (salinity)."""
    for tc in (0.0, 10.0, 20.0, 28.0):
        e = 6.112 * math.exp(17.67 * tc / (tc + 243.5))
        q = 0.98 * 0.622 * e / (1013.25 - 0.378 * e)
        assert q_sat(tc + 273.15) == pytest.approx(q, rel=0.1), tc


# ---------------------------------------------------------------------------
# apply_atm_fluxes on an oracle model
# ---------------------------------------------------------------------------
def _atm_oracle(steps=2):
    from fesomc_amd import meshgen
    txt, mesh, ex = meshgen.config("c2_small")
    txt = re.sub(r"^(\s*key_atm\s*=\s*)\S+", r"\g<1>T", txt, count=1, flags=re.M)
    txt += "\n type_swr_body = 1\n jer_const = T\n R = 0.58\n a = 0.35\n b = 23.0\n"
    o = Oracle(txt, mesh)
    o.set_tracers(ex["T"], ex["S"])
    o.step(steps)           # nonzero surface currents
    return o, mesh


def _neutral_coef(o, qlw=320.0, qsr=400.0, prec=3e-5, mslp=101000.0, humi_factor=1.0, at=None):
    """Time coefficients of fields that are neutral at every node (T_air = the
    surface temperature, q_air = q_sat times humi_factor); wind varying."""
    N, L = o.size["N"], o.size["L"]
    TF = o.get("TF").reshape(N, L)
    zst = TF[:, 0] + 273.15
    qs = np.array([q_sat(z) for z in zst])
    rng = np.random.default_rng(3)
    a = np.zeros((N, 8))
    b = np.zeros((N, 8))
    b[:, 0] = 6.0 + rng.uniform(-2, 2, N)
    b[:, 1] = -3.0 + rng.uniform(-2, 2, N)
    b[:, 2] = qs * humi_factor
    b[:, 3] = qsr
    b[:, 4] = qlw
    b[:, 5] = zst
    b[:, 6] = prec
    b[:, 7] = mslp
    if at is not None:     # a time trend on the pressure only (keeps the layer neutral)
        a[:, 7] = 50.0
        b[:, 7] = mslp - 50.0 * at
    return a, b, zst


def _check_neutral_fluxes(o, a, b, zst, prec_coef=1.0):
    N, L = o.size["N"], o.size["L"]
    node = o.apply_atm_fluxes(a, b, nm_prec_coef=prec_coef, nm_net_flux=0, node_out=True)
    U = o.get("Unode").reshape(N, L)[:, 0]
    V = o.get("Vnode").reshape(N, L)[:, 0]
    wdx, wdy = b[:, 0] - U, b[:, 1] - V
    wndm = np.hypot(wdx, wdy)
    cd = np.array([ly_neutral_cd(max(w, 0.5)) for w in wndm])
    assert np.allclose(node[:, 0], cd, rtol=1e-13, atol=0)
    assert np.allclose(node[:, 5], RHOA * wndm * cd * wdx, rtol=1e-12, atol=1e-15)
    assert np.allclose(node[:, 6], RHOA * wndm * cd * wdy, rtol=1e-12, atol=1e-15)
    assert np.array_equal(o.get("windx"), b[:, 0]) and np.array_equal(o.get("windy"), b[:, 1])
    # no sensible or latent heat: the net non-solar heat is the longwave balance
    TF0 = o.get("TF").reshape(N, L)[:, 0]
    qns = np.where((TF0 < -2.0) & (b[:, 4] - EMISS * STEF * zst ** 4 < 0), 0.0, b[:, 4] - EMISS * STEF * zst ** 4)
    assert np.allclose(o.get("qns"), qns, rtol=1e-12, atol=1e-9)
    assert np.allclose(o.get("emp"), -b[:, 6] / prec_coef, rtol=1e-14, atol=0)
    assert np.allclose(o.get("qsr"), 0.9 * b[:, 3], rtol=1e-7)      # ocean albedo 0.1
    assert np.allclose(o.get("mslp"), o.scalar("time_jd") * a[:, 7] + b[:, 7], rtol=1e-15)
    # the element stress: w_cv-weighted node stress
    w = o.get("w_cv").reshape(-1, 4)
    en = o.get_int("elem_nodes").reshape(-1, 4)
    assert np.allclose(o.get("taux"), (w * node[:, 5][en]).sum(axis=1), rtol=1e-12, atol=1e-16)
    assert np.allclose(o.get("tauy"), (w * node[:, 6][en]).sum(axis=1), rtol=1e-12, atol=1e-16)


def test_apply_atm_fluxes_neutral_layer_closed_forms():
    """A neutral surface layer at every node: stress rho_a |du| Cd_n10 du with
    the wind relative to the top layer's current, no turbulent heat, net
    non-solar heat = longwave - 0.97 sigma SST^4, E - P = -P/prec_coef, net
    shortwave = 0.9 x shortwave, pressure at the clock from its coefficients,
    element stress = w_cv average."""
    o, _ = _atm_oracle()
    a, b, zst = _neutral_coef(o, at=o.scalar("time_jd"))
    _check_neutral_fluxes(o, a, b, zst, prec_coef=3600.0)
    assert _mutated(15, lambda: _check_neutral_fluxes(o, a, b, zst))


def _check_evaporation_cools(o, zst, b_dry, a):
    o.apply_atm_fluxes(a, b_dry)
    lw = b_dry[:, 4] - EMISS * STEF * zst ** 4
    assert (o.get("emp") > -b_dry[:, 6]).all()       # evaporation
    assert (o.get("qns") < lw).all()                 # latent heat taken from the ocean


def test_evaporation_cools_the_ocean():
    """Air at the sea's temperature but at half its saturation humidity:
    evaporation is positive and the latent heat makes the net non-solar heat
    smaller than the longwave balance."""
    o, _ = _atm_oracle()
    a, b, zst = _neutral_coef(o, humi_factor=0.5)
    _check_evaporation_cools(o, zst, b, a)
    assert _mutated(16, lambda: _check_evaporation_cools(o, zst, b, a))


# ---------------------------------------------------------------------------
# the host reader (sbc.AtmForcing)
# ---------------------------------------------------------------------------
def _bilinear(c, lon, lat):
    return c[0] + c[1] * lon + c[2] * lat + c[3] * lon * lat


def test_julday_and_namelist(tmp_path):
    from fesomc_amd import sbc
    assert sbc.julday(1948, 1, 1) == 2432552          # JD 2000-01-01 12h = 2451545, 18993 days later
    assert sbc.julday(2000, 1, 1) == 2451545
    p = tmp_path / "namelist_bc.nml"
    p.write_text("&nam_sbc\n nm_sbc = 2 ! from files\n nm_xwind_file = 'a b.nc', nm_prec_coef = 3.6d3\n"
                 " nm_net_flux = 1\n/\n&other\n x = 1\n/\n")
    nm = sbc.read_namelist(str(p))
    assert nm == {"nm_sbc": 2, "nm_xwind_file": "a b.nc", "nm_prec_coef": 3600.0, "nm_net_flux": 1}


@pytest.mark.parametrize("descending", [False, True])
def test_forcing_reader_interpolates_bilinearly_and_in_time(tmp_path, descending):
    """Fields bilinear in (lon, lat) per record: at nodes inside the grid the
    coefficients reproduce each record exactly at its time and the linear
    blend in between; a file with descending latitudes gives the same
    (flip_lat); past the last record the field stays constant (one_field)."""
    from fesomc_amd import sbc
    lon = np.linspace(-1.0, 1.0, 9)
    lat = np.linspace(50.0, 52.0, 7)
    time = np.array([0.0, 6.0, 12.0]) * 3600.0          # seconds since 2000-01-01
    rng = np.random.default_rng(5)
    cs = rng.uniform(-2, 2, (len(sbc.FIELDS), 3, 4))
    LON, LAT = np.meshgrid(lon, lat)
    fields = {f: np.stack([_bilinear(cs[i, t], LON, LAT) for t in range(3)]) for i, f in enumerate(sbc.FIELDS)}
    lat_f = lat[::-1] if descending else lat
    fields_f = {f: v[:, ::-1, :] if descending else v for f, v in fields.items()}
    sbc.write_forcing(str(tmp_path), lon, lat_f, time, fields_f, secstep=86400.0, iyear=2000)
    N = 40
    xs, ys = rng.uniform(-0.95, 0.95, N), rng.uniform(50.05, 51.95, N)
    coord = np.vstack([xs, ys]) * sbc.RAD
    f = sbc.AtmForcing(str(tmp_path), coord)
    t0 = sbc.julday(2000, 1, 1) - 0.5
    assert f.time == pytest.approx(t0 + np.array([0.0, 0.25, 0.5]), abs=1e-12)
    coef = f.update(t0 + 0.1)
    assert coef is not None and (f.t_indx, f.t_indx_p1) == (1, 2)
    for tt, w in ((t0, 0.0), (t0 + 0.1, 0.4), (t0 + 0.25, 1.0)):
        got = f.fields_at(coef, tt)
        for i in range(len(sbc.FIELDS)):
            want = (1 - w) * _bilinear(cs[i, 0], xs, ys) + w * _bilinear(cs[i, 1], xs, ys)
            # t*coef_a + coef_b at t ~ 2.45e6 days cancels ~1e8 down to the
            # field: the reference's formulation leaves ~1e-8 absolute
            assert np.allclose(got[:, i], want, rtol=1e-9, atol=1e-6), (sbc.FIELDS[i], tt)
    assert f.update(t0 + 0.2) is None                  # still inside the pair
    coef = f.update(t0 + 0.3)                          # past the second record
    assert coef is not None and (f.t_indx, f.t_indx_p1) == (2, 3)
    coef = f.update(t0 + 0.6)                          # past the last record
    assert coef is not None and f.one_field and (f.t_indx, f.t_indx_p1) == (3, 3)
    assert np.allclose(coef[0], 0.0)
    assert f.update(t0 + 0.9) is None


def test_forcing_reader_outside_the_grid(tmp_path):
    """Nodes outside the forcing grid take the reference's nearest-edge rules
    (extrp 1: linear in latitude on the edge column; 2: in longitude on the
    edge row; 3: the corner value); a clock before the first record uses it
    as a constant field."""
    from fesomc_amd import sbc
    lon = np.array([0.0, 1.0, 2.0])
    lat = np.array([10.0, 11.0])
    time = np.array([0.0, 1.0])
    LON, LAT = np.meshgrid(lon, lat)
    val = 3.0 * LON + 7.0 * LAT
    fields = {f: np.stack([val, val + 1.0]) for f in sbc.FIELDS}
    sbc.write_forcing(str(tmp_path), lon, lat, time, fields, secstep=1.0, iyear=2000)
    pts = np.array([[-1.0, 10.5],    # west of the grid: column i = 1, linear in lat
                    [2.5, 10.25],    # east: i = N, linear in lat
                    [0.5, 12.0],     # north: row j = N, linear in lon
                    [3.0, 9.0]])     # south-east corner
    f = sbc.AtmForcing(str(tmp_path), pts.T * sbc.RAD)
    t0 = sbc.julday(2000, 1, 1) - 0.5
    coef = f.update(t0 - 1.0)        # before the first record
    got = f.fields_at(coef, t0 - 1.0)[:, 0]
    want = [3.0 * 0.0 + 7.0 * 10.5, 3.0 * 2.0 + 7.0 * 10.25, 3.0 * 0.5 + 7.0 * 11.0, 3.0 * 2.0 + 7.0 * 10.0]
    assert np.allclose(got, want, rtol=1e-12)
    assert np.allclose(coef[0], 0.0)


# ---------------------------------------------------------------------------
# device against the oracle
# ---------------------------------------------------------------------------
@pytest.mark.gpu
def test_bulk_formulas_match_oracle():
    """fcg_sbc_do against the oracle's apply_atm_fluxes on the stratified
    wet/dry channel with a time-varying, non-neutral atmosphere, then three
    forced steps of both.  The fields agree to 1e-12 relative (the formulas'
    exp, log and atan come from the device library and the host libm); the
    state after the forced steps to 1e-9."""
    from fesomc_amd import gpu, meshgen
    txt, mesh, ex = meshgen.config("c2_small")
    txt = re.sub(r"^(\s*key_atm\s*=\s*)\S+", r"\g<1>T", txt, count=1, flags=re.M)
    txt += "\n type_swr_body = 1\n jer_const = T\n R = 0.58\n a = 0.35\n b = 23.0\n"
    from fesomc_amd.runconfig import RunConfig
    g = gpu.Model(RunConfig(txt), mesh)
    g.set_tracers(ex["T"], ex["S"])
    o = Oracle(txt, mesh)
    o.set_tracers(ex["T"], ex["S"])
    g.step(2)
    o.step(2)
    N = mesh.nod2D
    x, y = mesh.xy[:, 0], mesh.xy[:, 1]
    rng = np.random.default_rng(11)
    a = np.zeros((N, 8))
    b = np.zeros((N, 8))
    t = o.scalar("time_jd")
    b[:, 0] = 5.0 + 3.0 * np.sin(2 * np.pi * x / 20000.0)
    b[:, 1] = -2.0 + 2.0 * np.cos(2 * np.pi * y / 10000.0)
    b[:, 2] = 0.007 + 0.002 * rng.uniform(size=N)
    b[:, 3] = 350.0
    b[:, 4] = 300.0 + 20.0 * rng.uniform(size=N)
    b[:, 5] = 283.0 + 3.0 * np.sin(2 * np.pi * (x + y) / 15000.0)
    b[:, 6] = 2e-5
    b[:, 7] = 101000.0
    a[:, 0] = 40.0                    # the wind and the air temperature change in time
    a[:, 5] = -10.0
    b[:, 0] -= 40.0 * t
    b[:, 5] += 10.0 * t
    g.sbc_set_coef(a, b)
    g.sbc_do(3600.0, 0)
    o.apply_atm_fluxes(a, b, nm_prec_coef=3600.0, nm_net_flux=0)
    for f in ("windx", "windy", "taux", "tauy", "qns", "emp", "qsr", "mslp"):
        gv, ov = g.get(f), o.get(f)
        assert np.allclose(gv, ov, rtol=1e-12, atol=1e-12 * np.abs(ov).max()), f
    for _ in range(3):        # forced steps: clock, sbc_do, the step
        g.phase("fcg_advance_clock")
        o.phase(0)
        g.sbc_do(3600.0, 0)
        o.apply_atm_fluxes(a, b, nm_prec_coef=3600.0, nm_net_flux=0)
        for ph in gpu.Model.PHASES[1:]:
            g.phase(ph)
        for ph in range(1, 9):
            o.phase(ph)
    for f in ("eta_n", "TF", "SF", "U_n", "V_n"):
        gv, ov = g.get(f), o.get(f)
        assert np.allclose(gv, ov, rtol=1e-9, atol=1e-9 * np.abs(ov).max()), f
    assert g.diagnostics()["status"] == 0
    g.close()


@pytest.mark.gpu
def test_model_reads_forcing_files_and_steps_like_the_oracle(tmp_path):
    """The host chain of sbc_ini/sbc_do: forcing files on a lon/lat grid over
    the channel (two records six hours apart), read by sbc.AtmForcing and
    attached to the model, which forms the fields at start-up and after every
    clock advance; the oracle takes the same coefficients at the same
    points.  Fields to 1e-12, the state after four forced steps to 1e-9."""
    from fesomc_amd import gpu, meshgen, sbc
    from fesomc_amd.runconfig import RunConfig
    txt, mesh, ex = meshgen.config("c2_small")
    txt = re.sub(r"^(\s*key_atm\s*=\s*)\S+", r"\g<1>T", txt, count=1, flags=re.M)
    txt += "\n type_swr_body = 2\n"
    cfg = RunConfig(txt)
    g = gpu.Model(cfg, mesh)
    g.set_tracers(ex["T"], ex["S"])
    o = Oracle(txt, mesh)
    o.set_tracers(ex["T"], ex["S"])
    N = mesh.nod2D
    coord = np.ctypeslib.as_array(g.view.coord_nod2D, shape=(2 * N,)).reshape(N, 2).T.copy()
    xd, yd = coord[0] / sbc.RAD, coord[1] / sbc.RAD
    lon = np.linspace(xd.min() - 0.02, xd.max() + 0.02, 6)
    lat = np.linspace(yd.min() - 0.02, yd.max() + 0.02, 5)
    LON, LAT = np.meshgrid(lon, lat)
    t_start = o.scalar("time_jd")
    iyear = 2000
    t0 = sbc.julday(iyear, 1, 1) - 0.5
    times = np.array([t_start - t0 - 0.01, t_start - t0 + 0.24])      # days since the epoch
    base = {"xwind": 6.0 + 20.0 * LON, "ywind": -2.0 + 10.0 * LAT, "humi": 0.006 + 0.01 * LAT,
            "qsr": 300.0 + 0 * LON, "qlw": 310.0 + 50.0 * LAT, "tair": 284.0 + 10.0 * LON,
            "prec": 1e-5 + 0 * LON, "mslp": 101300.0 + 1000.0 * LAT}
    fields = {k: np.stack([v, v * (1.0 + 0.05 * (k != "tair"))]) for k, v in base.items()}
    sbc.write_forcing(str(tmp_path), lon, lat, times, fields, secstep=1.0, iyear=iyear, prec_coef=1.0)
    f = sbc.AtmForcing(str(tmp_path), coord)
    assert f.out_of_bounds == 0
    g.attach_sbc(f)
    coef = f._coefficients(t_start)          # the oracle's copy of the same coefficients
    node = o.apply_atm_fluxes(*coef, nm_prec_coef=f.prec_coef, nm_net_flux=f.net_flux, node_out=True)
    for name in ("windx", "taux", "qns", "emp", "qsr", "mslp"):
        assert np.allclose(g.get(name), o.get(name), rtol=1e-12, atol=1e-12 * np.abs(o.get(name)).max()), name
    # the nodal stress (taux_node, tauy_node), which the NetCDF output serves as ids 14 and 15
    tn = g.get("tau_node").reshape(N, 2)
    assert np.allclose(tn, node[:, 5:7], rtol=1e-12, atol=1e-12 * np.abs(node[:, 5:7]).max())
    for _ in range(4):
        g.step(1)
        o.phase(0)
        o.apply_atm_fluxes(*coef, nm_prec_coef=f.prec_coef, nm_net_flux=f.net_flux)
        for ph in range(1, 9):
            o.phase(ph)
    for name in ("eta_n", "TF", "SF", "U_n", "taux", "qns"):
        gv, ov = g.get(name), o.get(name)
        assert np.allclose(gv, ov, rtol=1e-9, atol=1e-9 * np.abs(ov).max()), name
    assert g.diagnostics()["status"] == 0
    # NetCDF output of the nodal stress (ids 14, 15): the device's field times the node mask
    from scipy.io import netcdf_file
    from fesomc_amd import ncout
    (tmp_path / "outputfilelist.dat").write_text(
        "!total number of files:\n1\n!=========== Field File Name ===\nstress.nc\n! snapshot\n1\n! type\n0\n"
        "! position\n0.0, 0.0, 0.0, 0.0\n! frequency\n1\n! special\n0\n! number of variables\n2\n! list\n"
        "14,15" + ",0" * 28 + "\n")
    out = ncout.Output(g, cfg, mesh, workdir=str(tmp_path))
    g.step(1)
    out.step()
    out.close()
    tn = g.get("tau_node").reshape(N, 2)
    mn = np.where(out.depth + g.get("etaAB") > out.Dmin, 1.0, 0.0)
    with netcdf_file(str(tmp_path / "nc_output" / "stress.nc"), "r", mmap=False) as nc:
        np.testing.assert_array_equal(nc.variables["taux_node"][0], tn[:, 0] * mn)
        np.testing.assert_array_equal(nc.variables["tauy_node"][0], tn[:, 1] * mn)
    g.close()
