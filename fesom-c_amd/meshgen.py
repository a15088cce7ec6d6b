"""Seeded synthetic channel meshes and run configurations.

A rectangular channel (x east, y north, cartesian metres) of nx*ny nodes
with quads split into two triangles with probability ``split`` along a
random diagonal.  Elements are listed clockwise (the orientation
``test_elem``, src/fv_mesh_array.F90:1-82, expects).  The west side is an
open boundary (index 2) forced by an M2 tide; the other sides are solid
(index 1).  Depth falls from ``depth_west`` to ``depth_east``, optionally
with an intertidal strip, plus seeded Gaussian bumps.

Configurations (sizes as the benchmark defines them):

* ``c1``: 101x51 nodes (~5.1k), 30 % split, nsigma=10, dt_2D=2 s, dt=20 s,
  upwind tracers, uniform T=10 / S=30, no wetting/drying.
* ``c2``: ~1.0M nodes, 90 % split (~1.9M elements), nsigma=24, MUSCL tracers,
  wetting/drying (depth -2..30 m, 15 % intertidal), stratified T/S with
  seeded smooth noise, seeded uniform wind.
* ``c2_small``: the c2 physics on a small mesh (parity-test size).
"""
from __future__ import annotations

import math

import numpy as np

from .meshio import RawMesh

RHO_AIR = (1.22)


def channel(nx, ny, Lx=20000.0, Ly=10000.0, split=0.3, seed=0, depth_west=30.0, depth_east=5.0,
            intertidal=0.0, depth_dry=-2.0, bumps=3, bump_amp=3.0, bump_sigma=1000.0, tide_amp=1.0):
    rng = np.random.default_rng(seed)
    x = np.linspace(0.0, Lx, nx)
    y = np.linspace(0.0, Ly, ny)
    X, Y = np.meshgrid(x, y)                      # (ny, nx)
    xy = np.stack([X.ravel(), Y.ravel()], axis=1)  # node id = j*nx + i
    N = nx * ny
    ii = np.tile(np.arange(nx), ny)
    jj = np.repeat(np.arange(ny), nx)
    index = np.zeros(N, dtype=np.int32)
    index[(jj == 0) | (jj == ny - 1) | (ii == nx - 1)] = 1
    index[ii == 0] = 2
    # depth
    xs = xy[:, 0] / Lx
    if intertidal > 0.0:
        x_it = 1.0 - intertidal
        d = np.where(xs <= x_it, depth_west + (2.0 - depth_west) * xs / x_it,
                     2.0 + (depth_dry - 2.0) * (xs - x_it) / intertidal)
    else:
        d = depth_west + (depth_east - depth_west) * xs
    cx = rng.uniform(0.2 * Lx, 0.8 * Lx, bumps)
    cy = rng.uniform(0.2 * Ly, 0.8 * Ly, bumps)
    for k in range(bumps):
        d = d + bump_amp * np.exp(-((xy[:, 0] - cx[k]) ** 2 + (xy[:, 1] - cy[k]) ** 2) / (2.0 * bump_sigma ** 2))
    # elements (1-based), clockwise
    ci, cj = np.meshgrid(np.arange(nx - 1), np.arange(ny - 1))
    ci, cj = ci.ravel(), cj.ravel()
    n00 = cj * nx + ci + 1
    n10 = n00 + 1
    n01 = n00 + nx
    n11 = n01 + 1
    sp = rng.random(ci.size) < split
    diag = rng.random(ci.size) < 0.5
    quads = np.stack([n00, n01, n11, n10], axis=1)[~sp]
    sA = sp & diag
    sB = sp & ~diag
    tA1 = np.stack([n00[sA], n01[sA], n11[sA], n00[sA]], axis=1)
    tA2 = np.stack([n00[sA], n11[sA], n10[sA], n00[sA]], axis=1)
    tB1 = np.stack([n00[sB], n01[sB], n10[sB], n00[sB]], axis=1)
    tB2 = np.stack([n01[sB], n11[sB], n10[sB], n01[sB]], axis=1)
    elem = np.concatenate([quads, tA1, tA2, tB1, tB2], axis=0).astype(np.int32)
    # interleave so the file is not grouped by type (the reader sorts triangles first)
    perm = rng.permutation(elem.shape[0])
    elem = elem[perm]
    nob = int((index == 2).sum())
    ob = np.zeros((nob, 2))
    ob[:, 0] = tide_amp
    return RawMesh(xy=xy, index=index, elem=elem, depth=d, ob=ob,
                   meta=dict(nx=nx, ny=ny, Lx=Lx, Ly=Ly, split=split, seed=seed))


FV_RUN_TEMPLATE = """! generated by fesomc_amd.meshgen ({name})
 TITLE = {title}
 dt_2D = {dt_2D}
 dt = {dt}
 nsteps = {nsteps}
 JD0 = 2457266.5
 INDIR = ./
 OUTDIR = ./Output/
 MESHPATH = ./MESH/
 IREP = {irep}
 IREC = 1000
 IRESTART = 100000000
 RESTART = cold_start
 VERBOSITY = 1
 cartesian = T
 Coriolis_TF = {coriolis}
 lat_cartesian = 54.0
 mom_adv_2D = 2
 mom_adv_3D = 1
 vert_adv = 1
 ex_vert_visc = F
 im_vert_visc = F
 im_vert_visc_fic = T
 fic_point = 1
 tracer_adv = {tracer_adv}
 MUSCL_CALL = {muscl}
 mu_w = 0.85
 up_w = 0.15
 Cd = {Cd}
 z0b_min = .01
 z0s_min = .02
 za = .00
 BFC = 1
 tau_inv_filt = 0.000046296
 Abh0 = 1000.0
 C_Smag = .3
 A_hor = 10000.0
 tau_c = .4
 snul = 7.5e-4
 filt_2D = F
 filt_bi_2D = F
 bih_2D = F
 filt_3D = F
 filt_bi_3D = F
 bih_3D = F
 BAROTROPIC = F
 BAROTROPIC_3D = F
 SALINITY_ON = T
 TEMP_ON = T
 S_const = 30.0
 T_const = 10.0
 Vert_turb = 3
 type_vd = sigma
 nsigma = {nsigma}
 ind_sigma = 1
 pow_sig = 1.
 lev_bot = 1.
 lev_sur = 1.
 KSw = 0.5 0.5 1.5 1.5
 KBw = 0.5 0.5 1.5 1.5
 D_ref_max = 40.0
 key_rivers = F
 key_atm = F
 key_obc = F
 key_ic = F
 TF_presence = T
 Harmonics_tf = M2
 T_potential = F
 cyclic_length = 360.0
 Dcr = {Dmin}
 Dmin = {Dmin}
 T_out = F
 WET_DRY_ON = {wetdry}
 key_nc_output = F
 SL_obn = {sl_obn}
 SL_obn_user_def = F
 SL_radius = {sl_radius}
 mixed_BP_grad_OB = F
 Cd_bz = 10.0
 Bp_bz = 5.0
 cl_relax = T
 clim_relax = {clim_relax}
 relax2clim_ac = {relax2clim_ac}
 cl_solid_boder_relax = F
 OB_nudg = F
 comp_sediment = F
"""


def fv_run(name, dt_2D, dt, nsigma, tracer_adv=1, muscl=False, wetdry=False, Dmin=0.25, Cd=0.0025, nsteps=1000,
           irep=100, coriolis=True, sponge=0.0, ob_relax_days=(15.0, 60.0), extra=""):
    """fv_run.dat text; ``sponge`` > 0 turns on the open-boundary sponge layer
    (SL_obn) of that width in metres; ``ob_relax_days`` are the time scales
    (clim_relax, relax2clim_ac) of the T/S relaxation at open-boundary nodes."""
    txt = FV_RUN_TEMPLATE.format(name=name, title=name.upper(), dt_2D=dt_2D, dt=dt, nsteps=nsteps, irep=irep,
                                 tracer_adv=tracer_adv, muscl="T" if muscl else "F", nsigma=nsigma,
                                 wetdry="T" if wetdry else "F", Dmin=Dmin, Cd=Cd,
                                 coriolis="T" if coriolis else "F", sl_obn="T" if sponge > 0 else "F",
                                 sl_radius=sponge if sponge > 0 else 100000.0, clim_relax=ob_relax_days[0],
                                 relax2clim_ac=ob_relax_days[1])
    return txt + extra


def wind_stress(u10x, u10y):
    """Large & Pond (1981) as wind2stress_scalar, src/fv_sbc.F90:1071-1078 (1.22, 0.49, 0.065 are REAL(4))."""
    f4 = lambda v: float(np.float32(v))
    u10 = math.sqrt(u10x * u10x + u10y * u10y)
    if u10 <= 11:
        return 1.2e-3 * f4(1.22) * u10 * u10x, 1.2e-3 * f4(1.22) * u10 * u10y
    c = 1.0e-3 * (f4(0.49) + f4(0.065) * u10)
    return c * f4(1.22) * u10 * u10x, c * f4(1.22) * u10 * u10y


def stratified_ts(mesh: RawMesh, nsigma, seed=0, T_surf=12.0, T_bot=8.0, S_surf=29.0, S_bot=31.0, T_std=0.5,
                  S_std=0.3, H=30.0, modes=6):
    """Linear stratification in depth plus seeded smooth (few-mode) horizontal noise; (N, nsigma-1)."""
    rng = np.random.default_rng(seed + 1000)
    N = mesh.nod2D
    L = nsigma - 1
    sig = 1.0 - np.arange(nsigma) / (nsigma - 1)
    a = np.maximum(mesh.depth, 0.25)
    zb = (1.0 - sig)[None, :] * a[:, None]
    Z = 0.5 * (zb[:, :-1] + zb[:, 1:])
    frac = np.clip(Z / H, 0.0, 1.0)
    T = T_surf + (T_bot - T_surf) * frac
    S = S_surf + (S_bot - S_surf) * frac
    Lx = mesh.meta.get("Lx", np.ptp(mesh.xy[:, 0]) or 1.0)
    Ly = mesh.meta.get("Ly", np.ptp(mesh.xy[:, 1]) or 1.0)

    def smooth():
        f = np.zeros(N)
        for _ in range(modes):
            kx, ky = rng.integers(1, 4, 2)
            ph = rng.uniform(0, 2 * np.pi, 2)
            f += np.cos(2 * np.pi * kx * mesh.xy[:, 0] / Lx + ph[0]) * np.cos(2 * np.pi * ky * mesh.xy[:, 1] / Ly + ph[1])
        return f / (f.std() + 1e-30)

    T = T + T_std * smooth()[:, None]
    S = S + S_std * smooth()[:, None]
    return T.reshape(N, L), S.reshape(N, L)


# Open-boundary treatment of the stratified channels.  The reference applies
# no open-boundary condition to the 3-D momentum and relaxes T/S at
# open-boundary nodes only; with its defaults (no sponge, 15/60-day T/S
# relaxation) the 3-D velocity of the elements at the open boundary grows
# without bound, and T at the nodes next to the open-boundary corners drifts
# below the -3 degC stop, within a few hundred to ~1500 internal steps, in the
# oracle and on the GPU alike (1M-node channel: ~350 steps).  A 3 km sponge
# layer (SL_obn: ac < 1 raises the bottom drag by Cd_bz and damps momentum
# advection and the baroclinic pressure gradient by ac**Bp_bz) and T/S
# nudging at open-boundary nodes on a 0.001-day (86 s) time scale carry the
# channels from 7k to 1M nodes through one M2 cycle.
SPONGE_M = 3000.0
OB_RELAX_DAYS = (0.001, 0.001)


def config(name, seed=0, scale=None):
    """Return (fv_run text, RawMesh, extras) for a named configuration.

    extras: dict with optional 'T', 'S' (N, nsigma-1) initial tracers and
    'wind' (u10x, u10y) uniform wind, applied after initial_state.
    """
    if name == "c1":
        m = channel(101, 51, split=0.3, seed=seed)
        return fv_run("c1", 2.0, 20.0, 10, tracer_adv=1, wetdry=False), m, {}
    if name in ("c2", "c2_small"):
        if name == "c2":
            nx, ny = scale or (1415, 708)
        else:
            nx, ny = (scale or (61, 31))
        m = channel(nx, ny, split=0.9, seed=seed, intertidal=0.15, depth_west=30.0, depth_dry=-2.0)
        dx = 20000.0 / (nx - 1)
        dt_2D = min(2.0, round(0.45 * dx / math.sqrt(9.8 * 34.0), 2))
        dt = 10 * dt_2D
        nsig = 24 if name == "c2" else 8
        txt = fv_run(name, dt_2D, dt, nsig, tracer_adv=2, muscl=True, wetdry=True, Dmin=0.25, sponge=SPONGE_M,
                     ob_relax_days=OB_RELAX_DAYS)
        T, S = stratified_ts(m, nsig, seed=seed)
        rng = np.random.default_rng(seed + 7)
        spd, ang = rng.uniform(0.0, 10.0), rng.uniform(0.0, 2 * np.pi)
        return txt, m, dict(T=T, S=S, wind=(spd * math.cos(ang), spd * math.sin(ang)))
    raise KeyError(name)


# Ensemble at the benchmark's configuration: 64 channels of 50k..400k nodes,
# each with its own number of sigma levels (one of 16, 20, 24, 28, 32),
# tidal amplitude (+-10 %) and wind (+-3 m/s per component about a common
# seeded wind).  "c4_mini" has the same structure with 2k..16k-node members.
ENSEMBLE_SPECS = {"c4": (64, 50_000, 400_000), "c4_mini": (64, 2_000, 16_000)}
NSIGMA_CLASSES = (16, 20, 24, 28, 32)


def ensemble_classes(name="c4", members=None, seed=0):
    """Members of an ensemble with per-member level counts, grouped by nsigma:
    a list of (nsigma, fv_run text, [RawMesh], [extras], [member ids]) in
    ascending nsigma.  Every member shares dt (it fits the finest member) and
    the physics switches; members with the same nsigma can share one
    fcg_create_ensemble context."""
    nm, lo, hi = ENSEMBLE_SPECS[name]
    nm = members or nm
    rng = np.random.default_rng(seed + 4343)
    nodes = rng.uniform(lo, hi, nm)
    nxs = np.maximum(8, np.round(np.sqrt(2.0 * nodes))).astype(int)
    nsig = rng.choice(NSIGMA_CLASSES, nm)
    amps = 1.0 + rng.uniform(-0.1, 0.1, nm)
    base = rng.uniform(0.0, 10.0, 2) - 5.0
    winds = base[None, :] + rng.uniform(-3.0, 3.0, (nm, 2))
    dx_min = 20000.0 / (nxs.max() - 1)
    dt_2D = min(2.0, round(0.45 * dx_min / math.sqrt(9.8 * 34.0), 2))
    def member(b):   # every member draws from its own seeded generators
        nx = int(nxs[b])
        m = channel(nx, max(8, nx // 2), split=0.9, seed=seed + 100 + b, intertidal=0.15, depth_west=30.0,
                    depth_dry=-2.0, tide_amp=float(amps[b]))
        T, S = stratified_ts(m, int(nsig[b]), seed=seed + 200 + b)
        return m, dict(T=T, S=S, wind=(float(winds[b, 0]), float(winds[b, 1])), tide_amp=float(amps[b]))

    import os
    from concurrent.futures import ThreadPoolExecutor
    with ThreadPoolExecutor(min(16, os.cpu_count() or 1)) as ex:
        built = list(ex.map(member, range(nm)))
    out = []
    for ns in sorted(set(int(v) for v in nsig)):
        ids = [b for b in range(nm) if nsig[b] == ns]
        txt = fv_run(f"{name}_s{ns}", dt_2D, 10 * dt_2D, ns, tracer_adv=2, muscl=True, wetdry=True, Dmin=0.25,
                     sponge=SPONGE_M, ob_relax_days=OB_RELAX_DAYS)
        out.append((ns, txt, [built[b][0] for b in ids], [built[b][1] for b in ids], ids))
    return out


def ensemble(name="c4", members=None, seed=0):
    """Ensemble of channels with uneven mesh sizes and perturbed forcing, for
    fcg_create_ensemble.  Returns (fv_run text, [RawMesh], [extras]) where
    every member shares nsigma, dt and the physics switches (dt fits the
    finest member), and differs in resolution, bathymetry seed, tidal
    amplitude (+-10 %), wind (+-3 m/s per component) and T/S noise.

    ``c4_16``: 16 members, 150..400 nodes along the channel (~11k..80k nodes
    each), nsigma 24.  ``c4_small``: 3 members of 30..60 nodes along, nsigma 8.
    (The configuration's ensemble, with its own level count per member, is
    ``ensemble_classes``.)
    """
    rng = np.random.default_rng(seed + 4242)
    if name == "c4_16":
        nm, lo, hi, nsig = members or 16, 150, 400, 24
    elif name == "c4_small":
        nm, lo, hi, nsig = members or 3, 30, 60, 8
    else:
        raise KeyError(name)
    nxs = rng.integers(lo, hi + 1, nm)
    dx_min = 20000.0 / (nxs.max() - 1)
    dt_2D = min(2.0, round(0.45 * dx_min / math.sqrt(9.8 * 34.0), 2))
    txt = fv_run(name, dt_2D, 10 * dt_2D, nsig, tracer_adv=2, muscl=True, wetdry=True, Dmin=0.25, sponge=SPONGE_M,
                     ob_relax_days=OB_RELAX_DAYS)
    base = rng.uniform(0.0, 10.0, 2) - 5.0
    meshes, extras = [], []
    for b, nx in enumerate(nxs):
        ny = max(8, int(nx) // 2)
        amp = 1.0 + rng.uniform(-0.1, 0.1)
        m = channel(int(nx), ny, split=0.9, seed=seed + 100 + b, intertidal=0.15, depth_west=30.0, depth_dry=-2.0,
                    tide_amp=amp)
        T, S = stratified_ts(m, nsig, seed=seed + 200 + b)
        wx, wy = base + rng.uniform(-3.0, 3.0, 2)
        meshes.append(m)
        extras.append(dict(T=T, S=S, wind=(float(wx), float(wy)), tide_amp=amp))
    return txt, meshes, extras
