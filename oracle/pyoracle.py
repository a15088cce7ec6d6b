"""ctypes front end of the FP64 CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's CPU-baseline leg, as the checker -- never by the product.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None

SIZES = dict(N=0, E=1, D=2, D_in=3, E_tri=4, nobn=5, L=6, K=7, maxdeg=8, Mt=9, n_dt=10)
SCALARS = dict(time_jd=0, time=1, density_0=2, T_aver=3, S_aver=4, scale_area=5, time_jd0=6)
INTS = dict(elem_nodes=0, edge_nodes=1, edge_tri=2, elem_edges=3, elem_neighbors=4,
            nod_in_elem2D_num=5, nod_in_elem2D=6, index=7, in_obn=8)


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = C.CDLL(_LIB)
        P = C.c_void_p
        dp = C.POINTER(C.c_double)
        ip = C.POINTER(C.c_int)
        L.orc_create.restype = P
        L.orc_create.argtypes = [C.c_char_p, C.c_int, dp, ip, C.c_int, ip, dp, C.c_int, dp]
        L.orc_step.restype = C.c_int
        L.orc_step.argtypes = [P]
        L.orc_phase.restype = C.c_int
        L.orc_phase.argtypes = [P, C.c_int]
        L.orc_energy.restype = C.c_double
        L.orc_energy.argtypes = [P]
        L.orc_status.restype = C.c_int
        L.orc_status.argtypes = [P]
        L.orc_errmsg.restype = C.c_char_p
        L.orc_errmsg.argtypes = [P]
        L.orc_size.restype = C.c_int
        L.orc_size.argtypes = [P, C.c_int]
        L.orc_scalar.restype = C.c_double
        L.orc_scalar.argtypes = [P, C.c_int]
        L.orc_get.restype = C.c_long
        L.orc_get.argtypes = [P, C.c_char_p, dp]
        L.orc_set.restype = C.c_long
        L.orc_set.argtypes = [P, C.c_char_p, dp]
        L.orc_get_int.restype = C.c_long
        L.orc_get_int.argtypes = [P, C.c_int, ip]
        L.orc_set_tracers.argtypes = [P, dp, dp]
        L.orc_destroy.argtypes = [P]
        L.orc_densityJM.restype = C.c_double
        L.orc_densityJM.argtypes = [C.c_double, C.c_double, C.c_double, C.c_double]
        L.orc_sw_cp.restype = C.c_double
        L.orc_filter_only.restype = C.c_int
        L.orc_filter_only.argtypes = [C.c_void_p, C.c_int]
        L.orc_sw_cp.argtypes = [C.c_double, C.c_double, C.c_double]
        L.orc_psi_h.restype = C.c_double
        L.orc_psi_h.argtypes = [C.c_double]
        L.orc_psi_m.restype = C.c_double
        L.orc_psi_m.argtypes = [C.c_double]
        L.orc_core_coeff_2z.restype = None
        L.orc_core_coeff_2z.argtypes = [C.c_double] * 7 + [dp]
        L.orc_apply_atm_fluxes.restype = C.c_int
        L.orc_apply_atm_fluxes.argtypes = [P, dp, dp, C.c_double, C.c_int, dp]
        L.orc_obc_apply.restype = C.c_int
        L.orc_obc_apply.argtypes = [P, C.c_int, dp, dp, dp, C.c_double]
        L.orc_tracer_faces.restype = C.c_int
        L.orc_tracer_faces.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                       C.POINTER(C.c_double)]
        L.orc_wad_mask_only.restype = C.c_int
        L.orc_wad_mask_only.argtypes = [C.c_void_p]
        L.orc_log_cr.restype = C.c_double
        L.orc_log_cr.argtypes = [C.c_double]
        L.orc_set_mutation.restype = C.c_int
        L.orc_set_mutation.argtypes = [C.c_int]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def densityJM(t, s, pz, density_0):
    return lib().orc_densityJM(float(t), float(s), float(pz), float(density_0))


def sw_cp(s, t, p_bar):
    return lib().orc_sw_cp(float(s), float(t), float(p_bar))


class Oracle:
    """One FESOM-C model instance on the CPU (reference layout: level-fastest)."""

    def __init__(self, cfg_text: str, mesh):
        L = lib()
        self._keep = []
        xy = np.ascontiguousarray(mesh.xy, dtype=np.float64).reshape(-1)
        ind = np.ascontiguousarray(mesh.index, dtype=np.int32)
        el = np.ascontiguousarray(mesh.elem, dtype=np.int32).reshape(-1)
        dep = np.ascontiguousarray(mesh.depth, dtype=np.float64)
        ob = None if mesh.ob is None else np.ascontiguousarray(mesh.ob, dtype=np.float64).reshape(-1)
        nobn = 0 if mesh.ob is None else int(mesh.ob.shape[0])
        self.h = L.orc_create(cfg_text.encode(), mesh.nod2D, _dp(xy), _ip(ind), mesh.elem2D, _ip(el), _dp(dep),
                              nobn, None if ob is None else _dp(ob))
        st = L.orc_status(self.h)
        if st:
            raise RuntimeError(f"oracle create failed ({st}): {L.orc_errmsg(self.h).decode()}")
        self.size = {k: L.orc_size(self.h, v) for k, v in SIZES.items()}

    def __del__(self):
        try:
            if getattr(self, "h", None):
                lib().orc_destroy(self.h)
                self.h = None
        except Exception:
            pass

    def step(self, n=1):
        for _ in range(n):
            st = lib().orc_step(self.h)
            if st:
                raise RuntimeError(f"oracle step failed ({st}): {lib().orc_errmsg(self.h).decode()}")

    def phase(self, i):
        st = lib().orc_phase(self.h, int(i))
        if st:
            raise RuntimeError(f"oracle phase {i} failed ({st}): {lib().orc_errmsg(self.h).decode()}")

    def energy(self):
        return lib().orc_energy(self.h)

    def filter_only(self, which):
        """One neighbour-difference filter alone on the current state into its
        zeroed target (2: filt_2D, 3: filt_3D, 4: filt_3D_to_2D; 5: filt_bi_2D
        into UV2_rhs, 6: filt_bi_3D into UV_rhs)."""
        if lib().orc_filter_only(self.h, int(which)):
            raise ValueError(f"no filter {which}")

    def tracer_faces(self, scheme, T, with_rhs=False):
        """Face values of the MUSCL (scheme 2) or Miura (scheme 4) horizontal
        flux for the tracer field T with the current velocities, as an
        (edge, layer, segment) array, and with ``with_rhs`` the advective
        tendency (node, layer); the state is unchanged."""
        D, L, N = self.size["D"], self.size["L"], self.size["N"]
        t = np.ascontiguousarray(T, dtype=np.float64).reshape(-1)
        out = np.zeros(2 * D * L)
        rhs = np.zeros(N * L)
        if lib().orc_tracer_faces(self.h, int(scheme), _dp(t), _dp(out), _dp(rhs) if with_rhs else None):
            raise ValueError(f"no tracer scheme {scheme}")
        return (out.reshape(D, L, 2), rhs.reshape(N, L)) if with_rhs else out.reshape(D, L, 2)

    def wad_mask_only(self):
        """wad_mask alone on the current etaAB."""
        lib().orc_wad_mask_only(self.h)

    def scalar(self, name):
        return lib().orc_scalar(self.h, SCALARS[name])

    def get(self, name):
        c = lib().orc_get(self.h, name.encode(), None)
        if c < 0:
            raise KeyError(name)
        out = np.empty(c)
        lib().orc_get(self.h, name.encode(), _dp(out))
        return out

    def set(self, name, values):
        c = lib().orc_get(self.h, name.encode(), None)
        v = np.ascontiguousarray(values, dtype=np.float64).reshape(-1)
        if c < 0 or v.size != c:
            raise ValueError(f"bad field {name}: {v.size} vs {c}")
        lib().orc_set(self.h, name.encode(), _dp(v))

    def get_int(self, name):
        c = lib().orc_get_int(self.h, INTS[name], None)
        out = np.empty(c, dtype=np.int32)
        lib().orc_get_int(self.h, INTS[name], _ip(out))
        return out

    def apply_atm_fluxes(self, coef_a, coef_b, nm_prec_coef=1.0, nm_net_flux=0, node_out=False):
        """sbc_do's data_timeinterp + apply_atm_fluxes at the oracle's clock
        (src/fv_sbc.F90:1042-1044); coef_a/coef_b (nod2D, 8).  With
        node_out: per node (Cd, Ch, Ce, T_zu, q_zu, taux_node, tauy_node)."""
        N = self.size["N"]
        a = np.ascontiguousarray(coef_a, dtype=np.float64).reshape(N, 8)
        b = np.ascontiguousarray(coef_b, dtype=np.float64).reshape(N, 8)
        out = np.empty((N, 7)) if node_out else None
        r = lib().orc_apply_atm_fluxes(self.h, _dp(a), _dp(b), float(nm_prec_coef), int(nm_net_flux),
                                       None if out is None else _dp(out))
        if r:
            raise RuntimeError("orc_apply_atm_fluxes failed")
        return out

    def obc_apply(self, depth, coef_a, coef_b, time_jd):
        """obc_do's data_timeinterp at time_jd and apply_obc on the OB nodes'
        Tclim/Sclim (src/fv_obc.F90:376-419, 577-623); coef (nobn, ndep, 2)."""
        d = np.ascontiguousarray(depth, dtype=np.float64)
        a = np.ascontiguousarray(np.transpose(coef_a, (1, 0, 2)), dtype=np.float64)   # -> (ndep, nobn, 2)
        b = np.ascontiguousarray(np.transpose(coef_b, (1, 0, 2)), dtype=np.float64)
        if lib().orc_obc_apply(self.h, d.size, _dp(d), _dp(a), _dp(b), float(time_jd)):
            raise ValueError("obc_apply: a layer lies above the first depth of the climatology")

    def set_tracers(self, T=None, S=None):
        t = None if T is None else np.ascontiguousarray(T, dtype=np.float64).reshape(-1)
        s = None if S is None else np.ascontiguousarray(S, dtype=np.float64).reshape(-1)
        lib().orc_set_tracers(self.h, None if t is None else _dp(t), None if s is None else _dp(s))


def psi_h(zeta):
    """L&Y stability function for heat (src/fv_sbc.F90:616-631)."""
    return lib().orc_psi_h(float(zeta))


def psi_m(zeta):
    """L&Y stability function for momentum (src/fv_sbc.F90:633-660)."""
    return lib().orc_psi_m(float(zeta))


def core_coeff_2z(zt, zu, sst, T_zt, q_sat, q_zt, dU):
    """(Cd, Ch, Ce, T_zu, q_zu) of core_coeff_2z (src/fv_sbc.F90:485-614)."""
    out = np.empty(5)
    lib().orc_core_coeff_2z(float(zt), float(zu), float(sst), float(T_zt), float(q_sat), float(q_zt), float(dU),
                            _dp(out))
    return tuple(float(v) for v in out)


def set_mutation(mid):
    """Plant mistake ``mid`` in the oracle (0 removes it); for the pin tests."""
    lib().orc_set_mutation(int(mid))
