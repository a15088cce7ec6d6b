"""Open-boundary T/S climatology (key_obc): the host part of the reference's
fv_obc module (src/fv_obc.F90), serial.

obc_ini (531-576) reads the namelist ``nam_obc`` from ``namelist_bc.nml``
and one NetCDF file (nm_tsobc_file) holding T and S profiles at the
open-boundary nodes: dimensions ``node`` (= nobn, in the mesh's in_obn
order), ``depth`` [m, positive down, ascending] and ``time``; variables
``node``, ``depth``, ``time`` and the two fields (nm_tobc_var, nm_sobc_var,
stored (time, depth, node) in NetCDF order).  getcoeffld (418-520) keeps per
node and depth the coefficients of the linear interpolation in time between
the records that bracket the clock; obc_do (577-623) refreshes them when the
clock passes the second record and forms the profiles at the clock, until
the last record is reached (one_field), after which the profiles stay at
their last time.  The device interpolates them to the layers' depths every
step (fcg_obc_do).

Model times before the first record leave the reference reading record 0
(out of bounds); this module raises ValueError.  Files are read with scipy's
reader of classic NetCDF (CDF-1/CDF-2).
"""
from __future__ import annotations

import os

import numpy as np

from .rivers import binarysearch
from .sbc import julday, read_namelist


class ObcClimatology:
    """obc_ini / getcoeffld / obc_do's record logic for one serial host.
    ``update(time_jd)`` returns (depth, coef_a, coef_b) with coef (nobn,
    ndepth, 2) when new coefficients are due, else None; ``eval_time(t)`` is
    the time at which the profiles are formed (t until one_field, then the
    time of the last interpolation)."""

    def __init__(self, run_dir, nobn, in_obn=None, namelist="namelist_bc.nml"):
        from scipy.io import netcdf_file
        nm = read_namelist(os.path.join(run_dir, namelist), group="nam_obc")
        self.active = int(nm.get("nm_obc", 1)) != -1 and nobn > 0
        self.path = os.path.join(run_dir, str(nm.get("nm_tsobc_file", "tsobc.")).strip())
        self.vars = (str(nm.get("nm_tobc_var", "temp")).strip(), str(nm.get("nm_sobc_var", "salt")).strip())
        secstep = float(np.float32(nm.get("nm_nc_secstep", 86400.0)))
        iyear, imm, idd = int(nm.get("nm_nc_iyear", 1948)), int(nm.get("nm_nc_imm", 1)), int(nm.get("nm_nc_idd", 1))
        with netcdf_file(self.path, "r", mmap=False, maskandscale=False) as f:
            if f.dimensions["node"] != nobn:
                raise ValueError("nc_readTimeGrid: size of OB nodes in file .neq. to mesh")
            self.time = np.array(f.variables["time"][:], dtype=np.float64) / secstep + julday(iyear, imm, idd) - 0.5
            self.depth = np.array(f.variables["depth"][:], dtype=np.float64)
            nodes = np.array(f.variables["node"][:], dtype=np.float64)
        # the reference only warns when the file's node ids differ from in_obn (145-152)
        self.node_mismatch = in_obn is not None and not np.array_equal(nodes, np.asarray(in_obn, dtype=np.float64))
        self.one_field = False
        self.t_indx = self.t_indx_p1 = 0
        self.t_eval = None
        self._first = True

    def _coefficients(self, rdate):
        """getcoeffld (418-520)."""
        from scipy.io import netcdf_file
        nt = self.time.size
        t = binarysearch(self.time, rdate)
        if t < nt:
            tp1 = t + 1
            delta_t = self.time[tp1 - 1] - self.time[t - 1] if t >= 1 else 1.0
        else:                                   # no extrapolation to the future
            tp1, delta_t = t, 1.0
            self.one_field = True
        if rdate < self.time[0]:
            raise ValueError("obc: the clock lies before the first climatology record")
        self.t_indx, self.t_indx_p1 = t, tp1
        out = []
        with netcdf_file(self.path, "r", mmap=False, maskandscale=False) as f:
            for var in self.vars:
                v = f.variables[var]
                d1 = np.array(v[t - 1], dtype=np.float64).T      # (depth, node) -> (node, depth)
                d2 = np.array(v[tp1 - 1], dtype=np.float64).T
                a = (d2 - d1) / delta_t
                out.append((a, d1 - a * self.time[t - 1]))
        coef_a = np.stack([out[0][0], out[1][0]], axis=-1)      # (nobn, ndepth, 2)
        coef_b = np.stack([out[0][1], out[1][1]], axis=-1)
        return self.depth, coef_a, coef_b

    def update(self, time_jd):
        """nc_obc_ini at the first call, then obc_do's record logic (610-616)."""
        if not self.active:
            return None
        if self._first:
            self._first = False
            self.t_eval = time_jd
            return self._coefficients(time_jd)
        if self.one_field:
            return None
        new = None
        if time_jd > self.time[self.t_indx_p1 - 1]:
            new = self._coefficients(time_jd)
        self.t_eval = time_jd               # data_timeinterp(rdate)
        return new

    def eval_time(self, time_jd):
        return self.t_eval if self.t_eval is not None else time_jd


def write_climatology(path_dir, in_obn, depth, time, T, S, name="tsobc.nc", secstep=86400.0, iyear=1948, imm=1,
                      idd=1):
    """Write namelist_bc.nml (group nam_obc) and a climatology file the reader
    above takes: T, S (time, depth, nobn)."""
    from scipy.io import netcdf_file
    os.makedirs(path_dir, exist_ok=True)
    with netcdf_file(os.path.join(path_dir, name), "w", version=2) as nc:
        nc.createDimension("time", len(time))
        nc.createDimension("depth", len(depth))
        nc.createDimension("node", len(in_obn))
        for dim, vals in (("time", time), ("depth", depth), ("node", in_obn)):
            v = nc.createVariable(dim, "d", (dim,))
            v[:] = np.asarray(vals, dtype=np.float64)
        for var, vals in (("temp", T), ("salt", S)):
            v = nc.createVariable(var, "d", ("time", "depth", "node"))
            v[:] = np.asarray(vals, dtype=np.float64)
    text = (f"&nam_obc\n nm_obc = 1\n nm_tsobc_file = '{name}'\n nm_tobc_var = 'temp'\n nm_sobc_var = 'salt'\n"
            f" nm_nc_iyear = {iyear}\n nm_nc_imm = {imm}\n nm_nc_idd = {idd}\n nm_nc_secstep = {secstep}\n/\n")
    nml = os.path.join(path_dir, "namelist_bc.nml")
    with open(nml, "a") as fh:
        fh.write(text)
