"""Atmospheric forcing for the bulk formulas (key_atm, nm_sbc = 2): the host
part of the reference's sbc module (src/fv_sbc.F90), serial.

The reference reads the namelist ``nam_sbc`` from ``namelist_bc.nml`` in the
run directory (sbc_ini, 870-1010), one NetCDF file per field on a regular
longitude/latitude grid with a time axis (nc_readTimeGrid, 162-249; every
file shares the grid and time axis of the x-wind file), and keeps per node
the coefficients of a linear interpolation in time between the two records
that bracket the clock, bilinear in space (getcoeffld, 665-841).  This
module does the same and hands the coefficients to the device
(fcg_sbc_set_coef); the device forms the fields at the clock and applies the
CORE bulk formulas (fcg_sbc_do, data_timeinterp + apply_atm_fluxes).

Fields, in the reference's order (i_xwind .. i_mslp, 125-132): 10 m wind x,
y [m/s], specific humidity [kg/kg], shortwave [W/m2], longwave [W/m2], 2 m
air temperature [K], precipitation [kg/m2/s], sea-level pressure [Pa].

The files are read with scipy's reader of classic NetCDF (CDF-1/CDF-2; the
image holds no NetCDF-4/HDF5 library).  Values are taken as stored, as
nf_get_vara_double returns them (no scale_factor/add_offset).
"""
from __future__ import annotations

import os
import re

import numpy as np

from .rivers import binarysearch

FIELDS = ("xwind", "ywind", "humi", "qsr", "qlw", "tair", "prec", "mslp")
RAD = np.pi / 180.0


def julday(yyyy, mm, dd):
    """Julian day number that begins at noon of the date (src/fv_obc.F90:641-668)."""
    igreg = 15 + 31 * (10 + 12 * 1582)
    jy = yyyy
    if jy == 0:
        raise ValueError("julday: there is no year zero")
    if jy < 0:
        jy += 1
    if mm > 2:
        jm = mm + 1
    else:
        jy -= 1
        jm = mm + 13
    jd = int(365.25 * jy) + int(30.6001 * jm) + dd + 1720995
    if dd + 31 * (mm + 12 * yyyy) >= igreg:
        ja = int(np.float32(0.01) * np.float32(jy))   # 0.01 is a REAL(4) literal
        jd = jd + 2 - ja + int(0.25 * ja)
    return jd


def read_namelist(path, group="nam_sbc"):
    """Values of one Fortran namelist group (key = value, strings quoted,
    logicals .true./.false., d or e exponents, ! comments), keys lower-case."""
    text = open(path).read()
    m = re.search(rf"&{group}\b(.*?)(?:^\s*/|/\s*$|&end)", text, flags=re.S | re.I | re.M)
    if not m:
        raise ValueError(f"{path}: no namelist group {group}")
    body = re.sub(r"!.*", "", m.group(1))
    out = {}
    for key, val in re.findall(r"(\w+)\s*=\s*('[^']*'|\"[^\"]*\"|[^,\s]+)", body):
        v = val.strip()
        if v[0] in "'\"":
            out[key.lower()] = v[1:-1]
        elif v.lower() in (".true.", ".t.", "t"):
            out[key.lower()] = True
        elif v.lower() in (".false.", ".f.", "f"):
            out[key.lower()] = False
        else:
            try:
                out[key.lower()] = int(v)
            except ValueError:
                out[key.lower()] = float(v.lower().replace("d", "e"))
    return out


class AtmForcing:
    """sbc_ini / getcoeffld / sbc_do's record logic for one serial host.

    ``coord``: node coordinates (2, nod2D) in radians, as coord_nod2D;
    ``run_dir``: the directory holding namelist_bc.nml and, relative to it,
    the files it names.  ``update(time_jd)`` returns new (coef_a, coef_b)
    arrays (nod2D, 8) when the clock has passed the second record (and at the
    first call), else None; the device keeps the last coefficients."""

    def __init__(self, run_dir, coord, namelist="namelist_bc.nml"):
        nm = read_namelist(os.path.join(run_dir, namelist))
        self.nm = nm
        if int(nm.get("nm_sbc", 2)) != 2:
            raise ValueError("namelist_bc.nml: nm_sbc must be 2 (fields from NetCDF files); "
                             "the reference STOPs on the constant-field mode (src/fv_sbc.F90:978)")
        self.prec_coef = float(nm.get("nm_prec_coef", 1.0))
        self.net_flux = int(nm.get("nm_net_flux", 0))
        secstep = float(np.float32(nm.get("nm_nc_secstep", 86400.0)))   # a default REAL in the reference
        iyear, imm, idd = int(nm.get("nm_nc_iyear", 1948)), int(nm.get("nm_nc_imm", 1)), int(nm.get("nm_nc_idd", 1))
        self.files, self.vars = [], []
        for f in FIELDS:
            # nc_sbc_ini_fillnames (251-299): the year suffix it builds is blanked, the file is the namelist's name
            self.files.append(os.path.join(run_dir, str(nm[f"nm_{f}_file"]).strip()))
            self.vars.append(str(nm[f"nm_{f}_var"]).strip())
        lon, lat, time = self._read_grid(self.files[0])
        self.time = time / secstep + julday(iyear, imm, idd) - 0.5
        self.flip_lat = lat.size > 1 and lat[0] > lat[-1]
        self.lon = lon
        self.lat = lat[::-1].copy() if self.flip_lat else lat
        x = np.asarray(coord[0], dtype=np.float64) / RAD
        y = np.asarray(coord[1], dtype=np.float64) / RAD
        self.x, self.y = x, y
        self.ii = np.array([self._index(self.lon, v) for v in x], dtype=np.int64)
        self.jj = np.array([self._index(self.lat, v) for v in y], dtype=np.int64)
        self.out_of_bounds = int(((self.ii < 1) | (self.jj < 1)).sum())
        self.one_field = False
        self.t_indx = self.t_indx_p1 = 0
        self._first = True

    @staticmethod
    def _read_grid(path):
        from scipy.io import netcdf_file
        with netcdf_file(path, "r", mmap=False, maskandscale=False) as f:
            for d in ("latitude", "longitude", "time"):
                if d not in f.dimensions:
                    raise ValueError(f"{path}: no dimension {d}")
            return (np.array(f.variables["longitude"][:], dtype=np.float64),
                    np.array(f.variables["latitude"][:], dtype=np.float64),
                    np.array(f.variables["time"][:], dtype=np.float64))

    @staticmethod
    def _index(axis, v):
        """nc_sbc_ini's bracketing index (331-356): 1-based index of the grid
        point at or before v, -1 before the grid, 0 at or past its end."""
        if axis[-1] > v >= axis[0]:
            return binarysearch(axis, v)
        return -1 if v < axis[0] else 0

    def _coefficients(self, rdate):
        """getcoeffld (665-841): (coef_a, coef_b) of every node and field."""
        from scipy.io import netcdf_file
        nt = self.time.size
        t = binarysearch(self.time, rdate)
        if rdate < self.time[0]:          # no extrapolation back in time
            t, tp1, delta_t = 1, 1, 1.0
        elif t < nt:
            tp1 = t + 1
            delta_t = self.time[tp1 - 1] - self.time[t - 1]
        else:                             # no extrapolation to the future
            tp1, delta_t = t, 1.0
            self.one_field = True
        self.t_indx, self.t_indx_p1 = t, tp1
        N = self.x.size
        a = np.empty((N, 8))
        b = np.empty((N, 8))
        i0, j0 = self.ii.copy(), self.jj.copy()
        ip1, jp1 = i0 + 1, j0 + 1
        nlon, nlat = self.lon.size, self.lat.size
        extrp = np.zeros(N, dtype=np.int64)
        m = i0 == 0
        i0[m], ip1[m], extrp[m] = nlon, nlon, extrp[m] + 1
        m = i0 == -1
        i0[m], ip1[m], extrp[m] = 1, 1, extrp[m] + 1
        m = j0 == 0
        j0[m], jp1[m], extrp[m] = nlat, nlat, extrp[m] + 2
        m = j0 == -1
        j0[m], jp1[m], extrp[m] = 1, 1, extrp[m] + 2
        x1, x2 = self.lon[i0 - 1], self.lon[ip1 - 1]
        y1, y2 = self.lat[j0 - 1], self.lat[jp1 - 1]
        x, y = self.x, self.y
        for fld, (path, var) in enumerate(zip(self.files, self.vars)):
            with netcdf_file(path, "r", mmap=False, maskandscale=False) as f:
                v = f.variables[var]
                d1 = np.array(v[t - 1], dtype=np.float64)      # (lat, lon): the Fortran (lon, lat) record
                d2 = np.array(v[tp1 - 1], dtype=np.float64)
            if self.flip_lat:
                d1, d2 = d1[::-1], d2[::-1]
            vals = []
            for d in (d1, d2):
                # every branch is evaluated on all nodes and masked; the
                # branches that divide by a zero grid spacing are not taken
                def s(i, j):
                    return d[j - 1, i - 1]
                out = np.empty(N)
                err = np.seterr(divide="ignore", invalid="ignore")
                e = extrp == 0
                denom = (x2 - x1) * (y2 - y1)
                out[e] = ((s(i0, j0) * (x2 - x) * (y2 - y) + s(ip1, j0) * (x - x1) * (y2 - y)
                           + s(i0, jp1) * (x2 - x) * (y - y1) + s(ip1, jp1) * (x - x1) * (y - y1)) / denom)[e]
                e = extrp == 1
                out[e] = ((s(i0, j0) * (y2 - y) + s(ip1, jp1) * (y - y1)) / (y2 - y1))[e]
                e = extrp == 2
                out[e] = ((s(i0, j0) * (x2 - x) + s(ip1, jp1) * (x - x1)) / (x2 - x1))[e]
                e = extrp == 3
                out[e] = s(i0, j0)[e]
                np.seterr(**err)
                vals.append(out)
            data1, data2 = vals
            a[:, fld] = (data2 - data1) / delta_t
            b[:, fld] = data1 - a[:, fld] * self.time[t - 1]
        return a, b

    def update(self, time_jd):
        """sbc_do's record logic (1034-1040): new coefficients at the first
        call (nc_sbc_ini) and whenever the clock has passed the second record
        (while more than one record is in use), else None."""
        if self._first:
            self._first = False
            return self._coefficients(time_jd)
        if not self.one_field and time_jd > self.time[self.t_indx_p1 - 1]:
            return self._coefficients(time_jd)
        return None

    def fields_at(self, coef, time_jd):
        """data_timeinterp (843-867): the eight fields (nod2D, 8) at time_jd."""
        a, b = coef
        return time_jd * a + b


def write_forcing(path_dir, lon, lat, time, fields, names=None, secstep=86400.0, iyear=1948, imm=1, idd=1,
                  prec_coef=1.0, net_flux=0):
    """Write a forcing set the reader above (and the reference) takes:
    namelist_bc.nml and one classic NetCDF file per field, variables named
    after FIELDS.  ``fields``: dict name -> array (time, lat, lon); ``time``
    in units of secstep since iyear-imm-idd."""
    from scipy.io import netcdf_file
    os.makedirs(path_dir, exist_ok=True)
    names = names or {f: f"{f}.nc" for f in FIELDS}
    for f in FIELDS:
        with netcdf_file(os.path.join(path_dir, names[f]), "w", version=2) as nc:
            nc.createDimension("time", len(time))
            nc.createDimension("latitude", len(lat))
            nc.createDimension("longitude", len(lon))
            for dim, vals in (("time", time), ("latitude", lat), ("longitude", lon)):
                v = nc.createVariable(dim, "d", (dim,))
                v[:] = np.asarray(vals, dtype=np.float64)
            v = nc.createVariable(f, "d", ("time", "latitude", "longitude"))
            v[:] = np.asarray(fields[f], dtype=np.float64)
    lines = ["&nam_sbc", " nm_sbc = 2", f" nm_nc_secstep = {secstep}", f" nm_nc_iyear = {iyear}",
             f" nm_nc_imm = {imm}", f" nm_nc_idd = {idd}", f" nm_prec_coef = {prec_coef}", f" nm_net_flux = {net_flux}"]
    for f in FIELDS:
        lines.append(f" nm_{f}_file = '{names[f]}'")
        lines.append(f" nm_{f}_var = '{f}'")
    lines.append("/")
    with open(os.path.join(path_dir, "namelist_bc.nml"), "w") as fh:
        fh.write("\n".join(lines) + "\n")
