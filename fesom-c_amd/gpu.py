"""ctypes binding of libfesomc_gpu.so -- the C-ABI declared in include/fesomc_gpu.h.

This is the Python host mirror of the reference call sequence: it reads the
fv_run.dat values (runconfig), builds the mesh through ``fcg_mesh_build``,
initialises tracers as ``initial_state`` does (src/fv_ini.F90:476-546) and
drives the per-call-site entry points of ``oce_timestep``.  There is no CPU
fallback: without the library or without a HIP device every call fails.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import meshio
from .runconfig import RunConfig

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libfesomc_gpu.so")

FCG_FP64, FCG_MIXED = 0, 1
FCG_CLOSURE_D3END, FCG_CLOSURE_MLP = 0, 1
ERRORS = {0: "ok", -1: "bad argument", -2: "mesh check failed", -3: "HIP error", -4: "non-finite state",
          -5: "Negative salinity, time resolution is insufficient, stop",
          -6: "Negative temperature, time resolution is insufficient, stop", -7: "no HIP device"}

EXPORTS = ["fcg_mesh_build", "fcg_mesh_view", "fcg_mesh_free", "fcg_create", "fcg_destroy", "fcg_set_state",
           "fcg_get_state", "fcg_set_clock", "fcg_advance_clock", "fcg_pressure", "fcg_vertical_mixing",
           "fcg_compute_vel_rhs", "fcg_barotropic_timestep", "fcg_jacobian", "fcg_update_3d_vel",
           "fcg_vert_vel_sigma", "fcg_solve_tracers", "fcg_oce_timestep", "fcg_run", "fcg_diagnostics",
           "fcg_sync", "fcg_last_error", "fcg_version", "fcg_probe", "fcg_probe_read", "fcg_mlp_create",
           "fcg_mlp_destroy", "fcg_mlp_get_params", "fcg_mlp_set_params", "fcg_mlp_set_norm", "fcg_mlp_forward",
           "fcg_mlp_gradient", "fcg_mlp_train_step", "fcg_set_mlp", "fcg_mlp_collect",
           "fcg_create_ensemble", "fcg_ensemble_size", "fcg_diagnostics_member", "fcg_set_launch",
           "fcg_get_launch", "fcg_hot_start", "fcg_mlp_gradient_static", "fcg_sbc_set_coef", "fcg_sbc_do",
           "fcg_obc_set_coef", "fcg_obc_do"]

MLP_FEATURES, MLP_PARAMS = 5, 39312


class Params(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("dt", "dt_2D", "time_jd0", "Dmin", "C_d", "z0b_min", "z0s_min", "snul",
                                          "Cd_bz", "Bp_bz", "clim_relax", "relax2clim_ac")] + \
               [(n, C.c_int) for n in ("nsigma", "Mt", "tracer_adv", "muscl_call", "mom_adv_2D", "mom_adv_3D",
                                       "vert_adv", "im_vert_visc", "im_vert_visc_fic", "fic_point", "wet_dry_on", "tf_presence",
                                       "restart", "cl_relax", "precision", "closure", "key_atm", "type_swr_body")] + \
               [(n, C.c_double) for n in ("Rjc", "ajc", "bjc")] + \
               [(n, C.c_int) for n in ("key_rivers", "filt_2D", "filt_3D")] + [("tau_inv_filt", C.c_double)] + \
               [(n, C.c_int) for n in ("filt_bi_2D", "filt_bi_3D")] + [("tau_c", C.c_double)] + \
               [(n, C.c_int) for n in ("bih_2D", "bih_3D")] + [("Abh0", C.c_double)]


_IP = C.POINTER(C.c_int)
_DP = C.POINTER(C.c_double)


class Mesh(C.Structure):
    _fields_ = [(n, C.c_int) for n in ("nod2D", "elem2D", "edge2D", "edge2D_in", "nsigma", "nobn")] + \
               [(n, _IP) for n in ("elem2D_nodes", "edge_nodes", "edge_tri", "elem_edges", "elem_neighbors",
                                   "index_nod2D", "in_obn")] + \
               [(n, _DP) for n in ("depth", "sigma", "area", "elem_area", "elem_cos", "coriolis", "w_cv",
                                   "gradient_sca", "gradient_vec", "edge_dxdy", "edge_cross_dxdy", "ac", "a_tp",
                                   "ampt", "fazt", "amp_factor", "phy_factor", "coord_nod2D")] + \
               [("density_0", C.c_double), ("scale_area", C.c_double)]


class MeshCfg(C.Structure):
    _fields_ = [(n, C.c_int) for n in ("cartesian", "coriolis_tf", "nsigma", "ind_sigma", "is_LE", "T_potential",
                                       "tf_presence")] + \
               [(n, C.c_double) for n in ("lat_cartesian_deg", "cyclic_length_deg", "pow_sig", "lev_bot",
                                          "lev_sur", "T_const", "S_const")] + \
               [("harmonics", C.c_int * 12), ("SL_obn", C.c_int), ("SL_radius", C.c_double)]


class Diag(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("energy", "volume", "heat", "salt", "eta_min", "eta_max", "T_min",
                                          "T_max", "S_min", "S_max", "time_jd")] + \
               [("n_dt", C.c_int), ("n_wet_elem", C.c_int), ("mlp_fallback", C.c_longlong), ("status", C.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib = None


def lib():
    """Load libfesomc_gpu.so; raises if it was not built (no silent fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build()")
        # PyTorch (used here for device buffers) bundles its own HIP runtime;
        # load it first so that the process holds one runtime, shared by both.
        try:
            import torch
            torch.cuda.is_available()
        except ImportError:
            pass
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        L.fcg_mesh_build.restype = C.c_int
        L.fcg_mesh_build.argtypes = [C.POINTER(MeshCfg), C.c_int, _DP, _IP, C.c_int, _IP, _DP, C.c_int, _DP,
                                     C.POINTER(P)]
        L.fcg_mesh_view.restype = C.c_int
        L.fcg_mesh_view.argtypes = [P, C.POINTER(Mesh)]
        L.fcg_mesh_free.argtypes = [P]
        L.fcg_create.restype = C.c_int
        L.fcg_create.argtypes = [C.POINTER(Mesh), C.POINTER(Params), C.POINTER(P)]
        L.fcg_destroy.argtypes = [P]
        for n in ("fcg_set_state", "fcg_get_state"):
            getattr(L, n).restype = C.c_int
            getattr(L, n).argtypes = [P, C.c_char_p, _DP, C.c_long]
        L.fcg_set_clock.restype = C.c_int
        L.fcg_set_clock.argtypes = [P, C.c_double, C.c_int, C.c_int]
        for n in ("fcg_advance_clock", "fcg_pressure", "fcg_vertical_mixing", "fcg_compute_vel_rhs",
                  "fcg_barotropic_timestep", "fcg_jacobian", "fcg_update_3d_vel", "fcg_vert_vel_sigma",
                  "fcg_solve_tracers", "fcg_oce_timestep", "fcg_sync"):
            getattr(L, n).restype = C.c_int
            getattr(L, n).argtypes = [P]
        L.fcg_run.restype = C.c_int
        L.fcg_run.argtypes = [P, C.c_int]
        L.fcg_sbc_set_coef.restype = C.c_int
        L.fcg_sbc_set_coef.argtypes = [P, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.fcg_sbc_do.restype = C.c_int
        L.fcg_sbc_do.argtypes = [P, C.c_double, C.c_int]
        L.fcg_obc_set_coef.restype = C.c_int
        L.fcg_obc_set_coef.argtypes = [P, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                       C.POINTER(C.c_double)]
        L.fcg_obc_do.restype = C.c_int
        L.fcg_obc_do.argtypes = [P, C.c_double]
        L.fcg_diagnostics.restype = C.c_int
        L.fcg_diagnostics.argtypes = [P, C.POINTER(Diag)]
        L.fcg_last_error.restype = C.c_int
        L.fcg_last_error.argtypes = [P, C.c_char_p, C.c_int]
        L.fcg_probe.restype = C.c_int
        L.fcg_probe.argtypes = [P, C.c_char_p]
        L.fcg_probe_read.restype = C.c_int
        L.fcg_probe_read.argtypes = [P, _DP, _IP]
        L.fcg_version.restype = C.c_char_p
        L.fcg_version.argtypes = []
        _FP = C.POINTER(C.c_float)
        L.fcg_mlp_create.restype = C.c_int
        L.fcg_mlp_create.argtypes = [C.c_uint, C.POINTER(P)]
        L.fcg_mlp_destroy.argtypes = [P]
        for n in ("fcg_mlp_get_params", "fcg_mlp_set_params"):
            getattr(L, n).restype = C.c_int
            getattr(L, n).argtypes = [P, _FP, C.c_long]
        L.fcg_mlp_set_norm.restype = C.c_int
        L.fcg_mlp_set_norm.argtypes = [P] + [_FP] * 6
        L.fcg_mlp_forward.restype = C.c_int
        L.fcg_mlp_forward.argtypes = [P, P, P, C.c_long]
        L.fcg_mlp_gradient.restype = C.c_int
        L.fcg_mlp_gradient.argtypes = [P, P, P, C.c_long, _FP]
        L.fcg_mlp_gradient_static.restype = C.c_int
        L.fcg_mlp_gradient_static.argtypes = [P, P, P, C.c_long, P, P]
        L.fcg_mlp_train_step.restype = C.c_int
        L.fcg_mlp_train_step.argtypes = [P, P, P, C.c_long, C.c_float]
        L.fcg_set_mlp.restype = C.c_int
        L.fcg_set_mlp.argtypes = [P, P]
        L.fcg_hot_start.restype = C.c_int
        L.fcg_hot_start.argtypes = [P, C.c_double]
        L.fcg_set_launch.restype = C.c_int
        L.fcg_set_launch.argtypes = [P, C.c_char_p, C.c_int]
        L.fcg_get_launch.restype = C.c_int
        L.fcg_get_launch.argtypes = [P, C.c_char_p]
        L.fcg_create_ensemble.restype = C.c_int
        L.fcg_create_ensemble.argtypes = [C.POINTER(Mesh), C.c_int, C.POINTER(Params), C.POINTER(P)]
        L.fcg_ensemble_size.restype = C.c_int
        L.fcg_ensemble_size.argtypes = [P]
        L.fcg_diagnostics_member.restype = C.c_int
        L.fcg_diagnostics_member.argtypes = [P, C.c_int, C.POINTER(Diag)]
        L.fcg_mlp_collect.restype = C.c_int
        L.fcg_mlp_collect.argtypes = [P, P, P, C.c_long, C.c_int, C.POINTER(C.c_long)]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(_DP)


def _ip(a):
    return a.ctypes.data_as(_IP)


def mesh_cfg(cfg: RunConfig) -> MeshCfg:
    c = MeshCfg()
    c.cartesian = int(cfg.l("cartesian"))
    c.coriolis_tf = int(cfg.l("Coriolis_TF"))
    c.nsigma = cfg.nsigma
    c.ind_sigma = cfg.i("ind_sigma", 1)
    c.is_LE = int(cfg.is_LE)
    c.T_potential = int(cfg.l("T_potential"))
    c.tf_presence = int(cfg.l("TF_presence"))
    c.lat_cartesian_deg = cfg.f("lat_cartesian", 0.0)
    c.cyclic_length_deg = cfg.f("cyclic_length", 360.0)
    c.pow_sig = cfg.f("pow_sig", 1.0)
    c.lev_bot = cfg.f("lev_bot", 1.0)
    c.lev_sur = cfg.f("lev_sur", 1.0)
    c.T_const = cfg.f("T_const", 10.0)
    c.S_const = cfg.f("S_const", 35.0)
    for j, v in enumerate(cfg.harmonics):
        c.harmonics[j] = v
    # src/fv_main.F90:297-300: ac_create runs for SL_obn unless SL_obn_user_def
    c.SL_obn = int(cfg.l("SL_obn") and not cfg.l("SL_obn_user_def"))
    c.SL_radius = cfg.f("SL_radius", 0.0)
    return c


def params(cfg: RunConfig, precision=FCG_FP64, closure=FCG_CLOSURE_D3END) -> Params:
    p = Params()
    p.dt, p.dt_2D, p.time_jd0 = cfg.f("dt"), cfg.f("dt_2D"), cfg.f("JD0", 0.0)
    p.Dmin, p.C_d = cfg.f("Dmin", 0.25), cfg.f("Cd", 0.0)
    p.z0b_min, p.z0s_min, p.snul = cfg.f("z0b_min", 0.01), cfg.f("z0s_min", 0.02), cfg.f("snul", 7.5e-4)
    p.Cd_bz, p.Bp_bz = cfg.f("Cd_bz", 0.0), cfg.f("Bp_bz", 0.0)
    p.clim_relax, p.relax2clim_ac = cfg.clim_relax, cfg.relax2clim_ac
    p.nsigma, p.Mt = cfg.nsigma, cfg.Mt
    p.tracer_adv, p.muscl_call = cfg.i("tracer_adv", 1), int(cfg.l("MUSCL_CALL"))
    p.mom_adv_2D, p.mom_adv_3D, p.vert_adv = cfg.i("mom_adv_2D", 2), cfg.i("mom_adv_3D", 1), cfg.i("vert_adv", 1)
    p.im_vert_visc = int(cfg.l("im_vert_visc"))
    p.im_vert_visc_fic, p.fic_point = int(cfg.l("im_vert_visc_fic", True)), cfg.i("fic_point", 1)
    p.wet_dry_on, p.tf_presence = int(cfg.l("WET_DRY_ON")), int(cfg.l("TF_presence"))
    p.restart, p.cl_relax = int(cfg.restart), int(cfg.l("cl_relax"))
    p.precision, p.closure = precision, closure
    p.key_atm, p.type_swr_body = int(cfg.l("key_atm")), cfg.i("type_swr_body", 2)
    p.Rjc, p.ajc, p.bjc = cfg.f("R", 0.58), cfg.f("a", 0.35), cfg.f("b", 23.0)
    p.key_rivers = int(cfg.l("key_rivers"))
    p.filt_2D, p.filt_3D = int(cfg.l("filt_2D")), int(cfg.l("filt_3D"))
    p.tau_inv_filt = cfg.f("tau_inv_filt", 0.0)
    p.filt_bi_2D, p.filt_bi_3D = int(cfg.l("filt_bi_2D")), int(cfg.l("filt_bi_3D"))
    p.tau_c = cfg.f("tau_c", 0.4)
    p.bih_2D, p.bih_3D = int(cfg.l("bih_2D")), int(cfg.l("bih_3D"))
    p.Abh0 = cfg.f("Abh0", 0.0)
    return p


class GpuError(RuntimeError):
    pass


class Model:
    """One FESOM-C model instance resident on the GPU."""

    rivers = None   # river table of the host (key_rivers), see __init__
    sbc = None      # atmospheric forcing of the host (key_atm bulk formulas), see attach_sbc
    obc = None      # open-boundary T/S climatology of the host (key_obc), see attach_obc

    def __init__(self, cfg: RunConfig, mesh: meshio.RawMesh, precision=FCG_FP64, closure=FCG_CLOSURE_D3END,
                 rivers=None):
        """``rivers``: with key_rivers, the path of the run's MESHPATH/rivers.dat
        (or a ``rivers.Rivers`` table); ``step`` then interpolates it after each
        clock advance and hands the discharge to the step as the field
        "runoff", as the reference's main loop calls rivers_do
        (src/fv_main.F90:400-412).  Without a table the field "runoff" keeps
        what the caller last set (zero at start)."""
        L = lib()
        bad = cfg.unsupported()
        if bad:
            raise GpuError(f"options outside this build's hot path: {bad}")
        self.cfg = cfg
        self.rivers = None
        self._mc = mesh_cfg(cfg)
        xy = np.ascontiguousarray(mesh.xy, dtype=np.float64).reshape(-1)
        ind = np.ascontiguousarray(mesh.index, dtype=np.int32)
        el = np.ascontiguousarray(mesh.elem, dtype=np.int32).reshape(-1)
        dep = np.ascontiguousarray(mesh.depth, dtype=np.float64)
        ob = None if mesh.ob is None else np.ascontiguousarray(mesh.ob, dtype=np.float64).reshape(-1)
        nobn = 0 if mesh.ob is None else int(mesh.ob.shape[0])
        self._mesh = C.c_void_p()
        rc = L.fcg_mesh_build(C.byref(self._mc), mesh.nod2D, _dp(xy), _ip(ind), mesh.elem2D, _ip(el), _dp(dep), nobn,
                              None if ob is None else _dp(ob), C.byref(self._mesh))
        if rc:
            raise GpuError(f"fcg_mesh_build failed: {ERRORS.get(rc, rc)}")
        self.view = Mesh()
        L.fcg_mesh_view(self._mesh, C.byref(self.view))
        v = self.view
        self.N, self.E, self.D, self.D_in = v.nod2D, v.elem2D, v.edge2D, v.edge2D_in
        self.K = v.nsigma
        self.L = self.K - 1
        self.par = params(cfg, precision, closure)
        self.ctx = C.c_void_p()
        rc = L.fcg_create(C.byref(self.view), C.byref(self.par), C.byref(self.ctx))
        if rc:
            msg = self.last_error() if self.ctx else ""
            raise GpuError(f"fcg_create failed: {ERRORS.get(rc, rc)} {msg}")
        self.init_tracers(mesh)
        if rivers is not None and cfg.l("key_rivers"):
            from .rivers import Rivers
            t0 = self.par.time_jd0
            # rivers_ini (src/fv_main.F90:352) reads the table at the start-up clock
            self.rivers = rivers if isinstance(rivers, Rivers) else Rivers(rivers, self.N, self.clock()[1], t0)

    # -- lifetime ---------------------------------------------------------
    def close(self):
        L = lib()
        if getattr(self, "ctx", None):
            L.fcg_destroy(self.ctx)
            self.ctx = None
        if getattr(self, "_mesh", None):
            L.fcg_mesh_free(self._mesh)
            self._mesh = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def last_error(self):
        buf = C.create_string_buffer(512)
        lib().fcg_last_error(self.ctx, buf, 512)
        return buf.value.decode()

    def _ck(self, rc, what):
        if rc:
            raise GpuError(f"{what}: {ERRORS.get(rc, rc)} {self.last_error()}")

    # -- initial state (initial_state, src/fv_ini.F90:476-546) ------------
    def init_tracers(self, mesh):
        cfg = self.cfg
        T = np.zeros((self.N, self.L))
        S = np.zeros((self.N, self.L))
        if cfg.l("TEMP_ON"):
            T[:] = cfg.f("T_const", 10.0)
            if cfg.is_LE:
                T[mesh.xy[:, 0] * (np.pi / 180.0) < 0.0] = 5.0 / 0.28
                T[~(mesh.xy[:, 0] * (np.pi / 180.0) < 0.0)] = 10.0 / 0.28
        if cfg.l("SALINITY_ON"):
            S[:] = cfg.f("S_const", 35.0)
        self.set_tracers(T, S)

    def set_tracers(self, T=None, S=None):
        """Set TF (and T_old, Tclim) / SF (and S_old, Sclim); arrays (nod2D, nsigma-1)."""
        if T is not None:
            for n in ("TF", "T_old", "Tclim"):
                self.set(n, T)
        if S is not None:
            for n in ("SF", "S_old", "Sclim"):
                self.set(n, S)

    # -- state transfer (Fortran layout) ----------------------------------
    def set(self, name, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64).reshape(-1)
        self._ck(lib().fcg_set_state(self.ctx, name.encode(), _dp(a), a.size), f"set {name}")

    def get(self, name, count=None):
        sizes = {"N": self.N, "E": self.E}
        if count is None:
            count = self._count(name)
        out = np.empty(count)
        self._ck(lib().fcg_get_state(self.ctx, name.encode(), _dp(out), count), f"get {name}")
        return out

    def _count(self, name):
        N, E, L, K = self.N, self.E, self.L, self.K
        node1 = {"eta_n", "eta_n_1", "eta_n_2", "etaAB", "ssh_rhs", "windx", "windy", "qsr", "qns", "emp", "mslp",
                 "runoff"}
        el1 = {"U_filt_2D", "V_filt_2D", "mask_wd", "taux", "tauy", "C_d_el", "dmean_n"}
        el2 = {"U_n_2D", "U_n_1", "U_n_2", "UAB", "U_rhs_2D_3D", "Bar_pr_2D"}
        nodeL = {"TF", "SF", "T_old", "S_old", "Tclim", "Sclim", "Unode_p", "Vnode_p", "Unode", "Vnode", "Jc", "Jc_old",
                 "Z", "rho_c"}
        nodeK = {"hpressure", "Kv", "snu", "bt", "Wvel", "zbar"}
        if name in node1:
            return N
        if name in el1:
            return E
        if name in el2:
            return 2 * E
        if name == "vel_grad":
            return 4 * E
        if name == "tau_node":
            return 2 * N
        if name in nodeL:
            return N * L
        if name in nodeK:
            return N * K
        return E * L

    # -- stepping ---------------------------------------------------------
    PHASES = ("fcg_advance_clock", "fcg_pressure", "fcg_vertical_mixing", "fcg_compute_vel_rhs",
              "fcg_barotropic_timestep", "fcg_jacobian", "fcg_update_3d_vel", "fcg_vert_vel_sigma",
              "fcg_solve_tracers")

    def phase(self, name):
        self._ck(getattr(lib(), name)(self.ctx), name)

    # -- surface boundary conditions from atmospheric fields (key_atm) ----
    def sbc_set_coef(self, coef_a, coef_b):
        """getcoeffld's coefficients (nod2D, 8) to the device (fcg_sbc_set_coef)."""
        a = np.ascontiguousarray(coef_a, dtype=np.float64).reshape(-1)
        b = np.ascontiguousarray(coef_b, dtype=np.float64).reshape(-1)
        if a.size != 8 * self.N or b.size != 8 * self.N:
            raise GpuError("sbc coefficients must be (nod2D, 8)")
        self._ck(lib().fcg_sbc_set_coef(self.ctx, _dp(a), _dp(b)), "fcg_sbc_set_coef")

    def sbc_do(self, nm_prec_coef=1.0, nm_net_flux=0):
        """data_timeinterp + apply_atm_fluxes at the device clock (fcg_sbc_do)."""
        self._ck(lib().fcg_sbc_do(self.ctx, float(nm_prec_coef), int(nm_net_flux)), "fcg_sbc_do")

    def attach_sbc(self, forcing):
        """Drive the bulk formulas from ``forcing`` (sbc.AtmForcing, or any
        object with update(time_jd) -> (coef_a, coef_b) | None, prec_coef and
        net_flux): as sbc_ini does at start-up (src/fv_main.F90:343), the
        fields are formed now, and ``step`` calls sbc_do after each clock
        advance (src/fv_main.F90:419)."""
        if not self.cfg.l("key_atm"):
            raise GpuError("attach_sbc: the run has key_atm = F")
        self.sbc = forcing
        t = self.clock()[1]
        coef = forcing.update(t)
        if coef is not None:
            self.sbc_set_coef(*coef)
        self.sbc_do(forcing.prec_coef, forcing.net_flux)

    # -- open-boundary T/S climatology (key_obc) ---------------------------
    def obc_set_coef(self, depth, coef_a, coef_b):
        """getcoeffld's profile coefficients (nobn, ndepth, 2) to the device."""
        d = np.ascontiguousarray(depth, dtype=np.float64)
        a = np.ascontiguousarray(np.transpose(coef_a, (1, 0, 2)), dtype=np.float64).reshape(-1)   # (2, nobn, ndep) F
        b = np.ascontiguousarray(np.transpose(coef_b, (1, 0, 2)), dtype=np.float64).reshape(-1)
        if a.size != 2 * self.view.nobn * d.size or b.size != a.size:
            raise GpuError("obc coefficients must be (nobn, ndepth, 2)")
        self._ck(lib().fcg_obc_set_coef(self.ctx, d.size, _dp(d), _dp(a), _dp(b)), "fcg_obc_set_coef")

    def obc_do(self, time_jd):
        """Tclim/Sclim of the open-boundary nodes at time_jd (fcg_obc_do)."""
        self._ck(lib().fcg_obc_do(self.ctx, float(time_jd)), "fcg_obc_do")

    def attach_obc(self, clim):
        """Drive Tclim/Sclim at the open boundary from ``clim`` (obc.ObcClimatology,
        or any object with update(time_jd) -> (depth, coef_a, coef_b) | None and
        eval_time(time_jd)): as obc_ini does at start-up (src/fv_main.F90:348)
        the coefficients are taken now, and ``step`` calls obc_do after each
        clock advance (src/fv_main.F90:431)."""
        self.obc = clim
        t = self.clock()[1]
        c = clim.update(t)
        if c is not None:
            self.obc_set_coef(*c)

    def step(self, n=1):
        if self.rivers is None and self.sbc is None and self.obc is None:
            self._ck(lib().fcg_run(self.ctx, int(n)), "fcg_run")
            return
        # the river table and the atmospheric records are interpolated on the
        # host after each clock advance, in the reference's order
        # (src/fv_main.F90:400-439: rivers_do, sbc_do, oce_timestep)
        t = self.clock()[1]
        for _ in range(int(n)):
            self.phase("fcg_advance_clock")
            t = t + self.par.dt / 86400.0          # the device clock, src/fv_main.F90:400
            if self.rivers is not None:
                self.set("runoff", self.rivers.do(t))
            if self.sbc is not None:
                coef = self.sbc.update(t)
                if coef is not None:
                    self.sbc_set_coef(*coef)
                self.sbc_do(self.sbc.prec_coef, self.sbc.net_flux)
            if self.obc is not None:
                c = self.obc.update(t)
                if c is not None:
                    self.obc_set_coef(*c)
                self.obc_do(self.obc.eval_time(t))
            for ph in self.PHASES[1:]:
                self.phase(ph)

    def sync(self):
        self._ck(lib().fcg_sync(self.ctx), "fcg_sync")

    def diagnostics(self):
        d = Diag()
        self._ck(lib().fcg_diagnostics(self.ctx, C.byref(d)), "fcg_diagnostics")
        return d.as_dict()

    def energy(self):
        return self.diagnostics()["energy"]

    def clock(self):
        """(n_dt, time_jd) after the last internal step."""
        d = self.diagnostics()
        return d["n_dt"], d["time_jd"]

    def probe(self, kernel):
        self._ck(lib().fcg_probe(self.ctx, (kernel or "").encode()), "fcg_probe")

    def probe_read(self):
        t = C.c_double()
        n = C.c_int()
        self._ck(lib().fcg_probe_read(self.ctx, C.byref(t), C.byref(n)), "fcg_probe_read")
        return t.value, n.value


def mesh_arrays(cfg: RunConfig, mesh: meshio.RawMesh) -> dict:
    """Run only the host mesh preprocessing (no GPU needed) and return its arrays."""
    L = lib()
    mc = mesh_cfg(cfg)
    xy = np.ascontiguousarray(mesh.xy, dtype=np.float64).reshape(-1)
    ind = np.ascontiguousarray(mesh.index, dtype=np.int32)
    el = np.ascontiguousarray(mesh.elem, dtype=np.int32).reshape(-1)
    dep = np.ascontiguousarray(mesh.depth, dtype=np.float64)
    ob = None if mesh.ob is None else np.ascontiguousarray(mesh.ob, dtype=np.float64).reshape(-1)
    nobn = 0 if mesh.ob is None else int(mesh.ob.shape[0])
    h = C.c_void_p()
    rc = L.fcg_mesh_build(C.byref(mc), mesh.nod2D, _dp(xy), _ip(ind), mesh.elem2D, _ip(el), _dp(dep), nobn,
                          None if ob is None else _dp(ob), C.byref(h))
    if rc:
        raise GpuError(f"fcg_mesh_build failed: {ERRORS.get(rc, rc)}")
    v = Mesh()
    L.fcg_mesh_view(h, C.byref(v))
    N, E, D, K = v.nod2D, v.elem2D, v.edge2D, v.nsigma

    def I(p, n):
        return np.ctypeslib.as_array(p, shape=(n,)).copy()

    out = dict(nod2D=N, elem2D=E, edge2D=D, edge2D_in=v.edge2D_in, nobn=v.nobn,
               elem_nodes=I(v.elem2D_nodes, 4 * E), edge_nodes=I(v.edge_nodes, 2 * D), edge_tri=I(v.edge_tri, 2 * D),
               elem_edges=I(v.elem_edges, 4 * E), elem_neighbors=I(v.elem_neighbors, 4 * E),
               index=I(v.index_nod2D, N), depth=I(v.depth, N), sigma=I(v.sigma, K), area=I(v.area, N),
               elem_area=I(v.elem_area, E), elem_cos=I(v.elem_cos, E), coriolis=I(v.coriolis, E),
               w_cv=I(v.w_cv, 4 * E), gradient_sca=I(v.gradient_sca, 8 * E), gradient_vec=I(v.gradient_vec, 8 * E),
               edge_dxdy=I(v.edge_dxdy, 2 * D), edge_cross_dxdy=I(v.edge_cross_dxdy, 4 * D), a_tp=I(v.a_tp, 2 * N), ac=I(v.ac, N),
               density_0=v.density_0, scale_area=v.scale_area)
    if v.nobn:
        out["in_obn"] = I(v.in_obn, v.nobn)
        out["ampt"] = I(v.ampt, 12 * v.nobn)
        out["fazt"] = I(v.fazt, 12 * v.nobn)
    L.fcg_mesh_free(h)
    return out


def _fp(a):
    return np.ascontiguousarray(a, dtype=np.float32).ctypes.data_as(C.POINTER(C.c_float))


class Mlp:
    """Device network of the learned closure (fcg_mlp_*).  Device buffers are
    passed as raw pointers (e.g. ``tensor.data_ptr()`` of a float32 CUDA
    tensor); there is no host fallback."""

    def __init__(self, seed=0):
        self.h = C.c_void_p()
        rc = lib().fcg_mlp_create(seed, C.byref(self.h))
        if rc:
            raise GpuError(f"fcg_mlp_create failed: {ERRORS.get(rc, rc)}")
        self.norm = None

    def close(self):
        if getattr(self, "h", None):
            lib().fcg_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _ck(self, rc, what):
        if rc:
            raise GpuError(f"{what}: {ERRORS.get(rc, rc)}")

    def params(self):
        a = np.zeros(MLP_PARAMS, dtype=np.float32)
        self._ck(lib().fcg_mlp_get_params(self.h, a.ctypes.data_as(C.POINTER(C.c_float)), a.size), "get_params")
        return a

    def set_params(self, a):
        a = np.ascontiguousarray(a, dtype=np.float32)
        self._ck(lib().fcg_mlp_set_params(self.h, a.ctypes.data_as(C.POINTER(C.c_float)), a.size), "set_params")

    def set_norm(self, x_mean, x_std, y_mean, y_std, x_lo, x_hi):
        arrs = [np.ascontiguousarray(v, dtype=np.float32) for v in (x_mean, x_std, y_mean, y_std, x_lo, x_hi)]
        self.norm = arrs
        self._ck(lib().fcg_mlp_set_norm(self.h, *[a.ctypes.data_as(C.POINTER(C.c_float)) for a in arrs]),
                 "set_norm")

    def forward(self, X_ptr, Y_ptr, n):
        self._ck(lib().fcg_mlp_forward(self.h, C.c_void_p(X_ptr), C.c_void_p(Y_ptr), n), "forward")

    def gradient(self, X_ptr, T_ptr, n):
        g = np.zeros(MLP_PARAMS, dtype=np.float32)
        self._ck(lib().fcg_mlp_gradient(self.h, C.c_void_p(X_ptr), C.c_void_p(T_ptr), n,
                                        g.ctypes.data_as(C.POINTER(C.c_float))), "gradient")
        return g

    def gradient_static(self, X_ptr, T_ptr, n, G_ptr, stream=None):
        """fcg_mlp_gradient_static: the gradient into the device buffer at
        G_ptr (39312 floats) without any device allocation, launched on the
        HIP stream handle ``stream`` (None: the network's own stream) and not
        synchronised."""
        self._ck(lib().fcg_mlp_gradient_static(self.h, C.c_void_p(X_ptr), C.c_void_p(T_ptr), n, C.c_void_p(G_ptr),
                                                C.c_void_p(stream) if stream else None), "gradient_static")

    def train_step(self, X_ptr, T_ptr, n, lr):
        self._ck(lib().fcg_mlp_train_step(self.h, C.c_void_p(X_ptr), C.c_void_p(T_ptr), n, lr), "train_step")


def model_attach_mlp(model, net: Mlp):
    model._ck(lib().fcg_set_mlp(model.ctx, net.h), "fcg_set_mlp")
    model._mlp = net  # keep alive


def model_collect(model, X_ptr, T_ptr, cap, stride):
    cnt = C.c_long()
    model._ck(lib().fcg_mlp_collect(model.ctx, C.c_void_p(X_ptr), C.c_void_p(T_ptr), cap, stride, C.byref(cnt)),
              "fcg_mlp_collect")
    return cnt.value


Model.attach_mlp = model_attach_mlp
Model.collect = model_collect


def build_meshes(cfg: RunConfig, meshes, workers=None):
    """The reference preprocessing of each mesh (fcg_mesh_build, which keeps no
    shared state, so the meshes are processed on a thread pool): a list of
    (mesh object, its fcg_mesh view); free with free_meshes."""
    from concurrent.futures import ThreadPoolExecutor
    workers = workers or min(16, os.cpu_count() or 1, max(1, len(meshes)))
    if workers > 1 and len(meshes) > 1:
        with ThreadPoolExecutor(workers) as ex:
            parts = list(ex.map(lambda m: build_meshes(cfg, [m], workers=1), meshes))
        return [p[0] for p in parts]
    L = lib()
    mc = mesh_cfg(cfg)
    out = []
    for mesh in meshes:
        xy = np.ascontiguousarray(mesh.xy, dtype=np.float64).reshape(-1)
        ind = np.ascontiguousarray(mesh.index, dtype=np.int32)
        el = np.ascontiguousarray(mesh.elem, dtype=np.int32).reshape(-1)
        dep = np.ascontiguousarray(mesh.depth, dtype=np.float64)
        ob = None if mesh.ob is None else np.ascontiguousarray(mesh.ob, dtype=np.float64).reshape(-1)
        nobn = 0 if mesh.ob is None else int(mesh.ob.shape[0])
        h = C.c_void_p()
        rc = L.fcg_mesh_build(C.byref(mc), mesh.nod2D, _dp(xy), _ip(ind), mesh.elem2D, _ip(el), _dp(dep), nobn,
                              None if ob is None else _dp(ob), C.byref(h))
        if rc:
            raise GpuError(f"fcg_mesh_build failed: {ERRORS.get(rc, rc)}")
        v = Mesh()
        L.fcg_mesh_view(h, C.byref(v))
        out.append((h, v))
    return out


def free_meshes(bank):
    for h, _ in bank:
        lib().fcg_mesh_free(h)


class Ensemble(Model):
    """Several models with their own meshes (uneven sizes) and forcing,
    advanced together by one context (fcg_create_ensemble).  ``get``/``set``
    address the joined arrays; ``split`` cuts them per member.  ``prebuilt``
    (from build_meshes) reuses mesh objects the caller owns."""

    def __init__(self, cfg: RunConfig, meshes, precision=FCG_FP64, closure=FCG_CLOSURE_D3END, prebuilt=None):
        L = lib()
        bad = cfg.unsupported()
        if bad:
            raise GpuError(f"options outside this build's hot path: {bad}")
        self.cfg = cfg
        self._mc = mesh_cfg(cfg)
        bank = prebuilt if prebuilt is not None else build_meshes(cfg, meshes)
        self._meshes = [] if prebuilt is not None else [h for h, _ in bank]
        self.views = [v for _, v in bank]
        arr = (Mesh * len(self.views))(*self.views)
        self.nodes = [v.nod2D for v in self.views]
        self.elems = [v.elem2D for v in self.views]
        self.N, self.E = sum(self.nodes), sum(self.elems)
        self.D = sum(v.edge2D for v in self.views)
        self.K = self.views[0].nsigma
        self.L = self.K - 1
        self.par = params(cfg, precision, closure)
        self.ctx = C.c_void_p()
        rc = L.fcg_create_ensemble(arr, len(self.views), C.byref(self.par), C.byref(self.ctx))
        if rc:
            msg = self.last_error() if self.ctx else ""
            raise GpuError(f"fcg_create_ensemble failed: {ERRORS.get(rc, rc)} {msg}")
        self.members = len(self.views)
        T = np.concatenate([np.full((v.nod2D, self.L), cfg.f("T_const", 10.0)) for v in self.views])
        S = np.concatenate([np.full((v.nod2D, self.L), cfg.f("S_const", 35.0)) for v in self.views])
        self.set_tracers(T, S)

    def close(self):
        L = lib()
        if getattr(self, "ctx", None):
            L.fcg_destroy(self.ctx)
            self.ctx = None
        for h in getattr(self, "_meshes", []):
            L.fcg_mesh_free(h)
        self._meshes = []

    def split(self, arr, nodal=True):
        """Cut a joined (entities, ...) array into the members' arrays."""
        sizes = self.nodes if nodal else self.elems
        a = np.asarray(arr)
        a = a.reshape(sum(sizes), -1) if a.ndim == 1 else a
        out, o = [], 0
        for n in sizes:
            out.append(a[o:o + n])
            o += n
        return out

    def diagnostics_member(self, b):
        d = Diag()
        self._ck(lib().fcg_diagnostics_member(self.ctx, b, C.byref(d)), "fcg_diagnostics_member")
        return d.as_dict()


def model_set_launch(model, kernel, block):
    model._ck(lib().fcg_set_launch(model.ctx, kernel.encode(), block), f"fcg_set_launch {kernel}")


def model_get_launch(model, kernel):
    return lib().fcg_get_launch(model.ctx, kernel.encode())


Model.set_launch = model_set_launch
Model.get_launch = model_get_launch
