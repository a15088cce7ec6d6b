"""NetCDF output in the layout of the reference's output module
(src/fv_ncoutput.F90): the files, dimensions, mesh variables, attributes and
snapshot/average records that ``outputfilelist.dat`` asks for, so the
reference's post-processing reads them unchanged.

The reference links the NetCDF library; this image has none, so the files are
written by ``ClassicFile`` below in the NetCDF classic format with 64-bit
offsets (CDF-2), which every NetCDF reader opens.  Fields come from any source with ``get(name)`` in
the reference's Fortran layout (``gpu.Model`` on the device; the oracle in the
CPU tests).  Call ``Output.step()`` after every internal step, as the
reference calls ``output_do`` (src/fv_ncoutput.F90:868-889).

Averaged files need the fields of every step on the host (one device-to-host
copy per listed field per step); snapshot files copy only at their output
steps.
"""
from __future__ import annotations

import os

import numpy as np
import struct

R_EARTH = 6400000.0
PI = 3.14159265358979          # the reference's pi (o_PARAM)
RAD = PI / 180.0

# define_varlist (src/fv_ncoutput.F90:195-708):
# id: (name, modelname, standard_name, units, onnode, timedim, vertdim, type_task)
#   onnode 1 nodes / 0 elements; vertdim 0 none / 1 sigma (nsigma) / 2 sigmam1
VARLIST = {
    1: ("eta", "eta_n+depth(depth<0)", "sea surface elevation", "meter", 1, 1, 0, 1),
    2: ("u", "U_n", "velocity in x direction", "m/s", 0, 1, 2, 2),
    3: ("v", "V_n", "velocity in y direction", "m/s", 0, 1, 2, 2),
    4: ("ssu", "U_n(1,:)", "surface velocity in x direction", "m/s", 0, 1, 0, 2),
    5: ("ssv", "V_n(1,:)", "surface velocity in y direction", "m/s", 0, 1, 0, 2),
    6: ("u2d", "U_n_2D(1,:)", "2D velocity in x direction", "m/s", 0, 1, 0, 1),
    7: ("v2d", "U_n_2D(2,:)", "2D velocity in y direction", "m/s", 0, 1, 0, 1),
    8: ("temperature", "TF", "temperature", "grad C", 1, 1, 2, 3),
    9: ("salinity", "SF", "salinity", "psu", 1, 1, 2, 3),
    10: ("sst", "TF(1,:)", "temperature", "grad C", 1, 1, 0, 3),
    11: ("sss", "SF(1,:)", "salinity", "psu", 1, 1, 0, 3),
    12: ("taux", "taux", "wind stress along x direction", "N/m2", 0, 1, 0, 1),
    13: ("tauy", "tauy", "wind stress along y direction", "N/m2", 0, 1, 0, 1),
    14: ("taux_node", "taux_node", "wind stress along x direction on node", "N/m2", 1, 1, 0, 1),
    15: ("tauy_node", "tauy_node", "wind stress along y direction on node", "N/m2", 1, 1, 0, 1),
    16: ("mslp", "mslp", "mean sea level pressure", "Pascals", 1, 1, 0, 1),
    17: ("emp", "emp", "evaporation minus precipitation", "kg/m2/s", 1, 1, 0, 1),
    18: ("qns", "qns", "downward non solar heat over the ocean", "W/m2", 1, 1, 0, 1),
    19: ("qsr", "qsr", "downward solar heat over the ocean", "W/m2", 1, 1, 0, 1),
    20: ("zbar", "zbar", "node depth", "m", 1, 1, 1, 2),
    21: ("Z", "Z", "cell depth", "m", 1, 1, 2, 2),
    22: ("mask_wd_node", "mask_wd_node", "mask (wetting_drying) nodes", "", 1, 1, 0, 1),
    23: ("mask_wd", "mask_wd", "mask (wetting_drying) elements", "", 0, 1, 0, 1),
    24: ("Cd", "C_d_el", "bottom friction elements", "", 0, 1, 0, 1),
    25: ("windx", "winds", "wind along x direction on node", "m/s", 1, 1, 0, 1),
    26: ("windy", "windy", "wind along y direction on node", "m/s", 1, 1, 0, 1),
    27: ("w", "Wvel", "vertical velocity", "m/s", 1, 1, 1, 2),
    28: ("snu", "snu", "snu", "", 1, 1, 1, 3),
    29: ("bt", "bt", "turbulent kinetic energy", "", 1, 1, 1, 3),
    30: ("ri", "Ri", "Richardson number", "", 1, 1, 1, 3),
    31: ("m_ice", "m_ice", "effective ice thickness", "m", 1, 1, 0, 3),
    32: ("m_snow", "m_snow", "effective snow thickness", "m", 1, 1, 0, 3),
    33: ("a_ice", "a_ice", "ice concentation", "percent", 1, 1, 0, 3),
    34: ("t_skin", "t_skin", "temperature of snow/ice top surface", "grad C", 1, 1, 0, 3),
    35: ("runoff_rivers", "runoff_rivers", "runoff rivers", "m3/s", 1, 1, 0, 1),
    36: ("uice", "U_n_ice", "ice velocity in x direction", "m/s", 0, 1, 0, 3),
    37: ("vice", "U_n_ice", "ice velocity in y direction", "m/s", 0, 1, 0, 3),
}
# fields this build does not hold: the Richardson number of d3_end.  The
# surface fields mslp, emp, qns, qsr are the ones the host sets for the step
# or the bulk formulas form (key_atm); the nodal wind stress (taux_node,
# tauy_node) is held where the bulk formulas run.  Ice is off, so its fields
# are zero as in getvar1d (src/fv_ncoutput.F90:773-815); runoff_rivers is
# zero without key_rivers.
NOT_HELD = {30}
ZERO = {31, 32, 33, 34, 36, 37}


class ClassicFile:
    """Minimal NetCDF classic writer (CDF-2: 64-bit offsets), from the
    published format specification: header (dimensions, variables with their
    text attributes, byte offsets), then the fixed-size variables, then the
    records, each record holding one slab of every record variable in
    definition order, padded to 4 bytes.  Records are appended and the record
    count in the header updated in place, so the file is complete after every
    record.  Unwritten fixed-size variables hold the default fill value."""

    _TYPES = {"i": (4, 4, ">i4", -2147483647), "d": (6, 8, ">f8", 9.9692099683868690e+36)}

    def __init__(self, path):
        self.path, self.dims, self.vars = path, [], []

    def dim(self, name, size):
        self.dims.append((name, size))

    def var(self, name, typ, dims, attrs=None, data=None):
        self.vars.append({"name": name, "typ": typ, "dims": dims, "attrs": attrs or {}, "data": data})

    @staticmethod
    def _name(s):
        b = s.encode()
        return struct.pack(">i", len(b)) + b + b"\0" * (-len(b) % 4)

    def enddef(self):
        dsize = dict(self.dims)
        did = {n: i for i, (n, _) in enumerate(self.dims)}
        self.rec_vars = [v for v in self.vars if v["dims"] and dsize[v["dims"][0]] is None]
        for v in self.vars:
            shape = [dsize[d] for d in v["dims"] if dsize[d] is not None]
            v["shape"] = shape
            n = int(np.prod(shape)) if shape else 1
            v["vsize"] = n * self._TYPES[v["typ"]][1]
            v["vsize"] += -v["vsize"] % 4

        def header(begins):
            h = b"CDF\x02" + struct.pack(">i", 0)
            h += struct.pack(">ii", 10, len(self.dims)) if self.dims else b"\0" * 8
            for n, sz in self.dims:
                h += self._name(n) + struct.pack(">i", 0 if sz is None else sz)
            h += b"\0" * 8                                   # no global attributes
            h += struct.pack(">ii", 11, len(self.vars))
            for v, b in zip(self.vars, begins):
                h += self._name(v["name"]) + struct.pack(">i", len(v["dims"]))
                h += b"".join(struct.pack(">i", did[d]) for d in v["dims"])
                if v["attrs"]:
                    h += struct.pack(">ii", 12, len(v["attrs"]))
                    for k, t in v["attrs"].items():
                        tb = t.encode()
                        h += self._name(k) + struct.pack(">ii", 2, len(tb)) + tb + b"\0" * (-len(tb) % 4)
                else:
                    h += b"\0" * 8
                h += struct.pack(">iiq", self._TYPES[v["typ"]][0], v["vsize"], b)
            return h

        hlen = len(header([0] * len(self.vars)))
        off, begins = hlen, []
        fixed = [v for v in self.vars if v not in self.rec_vars]
        for v in self.vars:
            if v in fixed:
                begins.append(off)
                off += v["vsize"]
            else:
                begins.append(0)
        rb = off
        for v, i in zip(self.rec_vars, range(len(self.rec_vars))):
            begins[self.vars.index(v)] = rb
            rb += v["vsize"]
        self.rec_begin = off
        # a file with one record variable holds its records unpadded
        self.rec_size = sum(v["vsize"] for v in self.rec_vars) if len(self.rec_vars) != 1 else \
            int(np.prod(self.rec_vars[0]["shape"] or [1])) * self._TYPES[self.rec_vars[0]["typ"]][1]
        self.nrec = 0
        with open(self.path, "wb") as fh:
            fh.write(header(begins))
            for v in fixed:
                fh.write(self._payload(v, v["data"]))

    def _payload(self, v, data):
        _, size, code, fill = self._TYPES[v["typ"]]
        n = int(np.prod(v["shape"])) if v["shape"] else 1
        a = np.full(n, fill) if data is None else np.asarray(data).reshape(-1)
        if a.size != n:
            raise ValueError(f"{v['name']}: {a.size} values for shape {v['shape']}")
        b = a.astype(code).tobytes()
        return b + b"\0" * (-len(b) % 4) if len(self.rec_vars) != 1 or v not in self.rec_vars else b

    def append(self, values):
        """Write one record: values[name] for every record variable."""
        rec = b"".join(self._payload(v, values[v["name"]]) for v in self.rec_vars)
        with open(self.path, "r+b") as fh:
            fh.seek(self.rec_begin + self.nrec * self.rec_size)
            fh.write(rec)
            self.nrec += 1
            fh.seek(4)
            fh.write(struct.pack(">i", self.nrec))


def read_filelist(path):
    """outputfilelist.dat as output_ini reads it (src/fv_ncoutput.F90:985-1040):
    one list-directed record per line, a comment line before every value."""
    with open(path) as f:
        lines = f.read().splitlines()
    it = iter(lines)

    def nxt():
        return next(it).replace(",", " ").split()

    next(it)
    nfiles = int(nxt()[0])
    files = []
    for _ in range(nfiles):
        next(it)
        d = {"name": next(it).strip().split()[0]}
        next(it)
        d["snapshot"] = int(nxt()[0])
        next(it)
        d["ftype"] = int(nxt()[0])
        next(it)
        d["pos"] = [float(v) for v in nxt()[:4]]
        next(it)
        d["freq"] = int(nxt()[0])
        next(it)
        d["specfreq"] = int(nxt()[0])
        next(it)
        d["numvars"] = int(nxt()[0])
        next(it)
        d["vars"] = [int(v) for v in nxt()[:d["numvars"]]]
        if any(v < 1 for v in d["vars"]):
            raise ValueError(f"inconsist number of variables with provided index, {d['name']}")
        files.append(d)
    return files


def _nearest(px, py, cx, cy, cartesian):
    """ne_node / ne_element (src/fv_ncoutput.F90:1969-2055): index of the
    nearest point, first minimum wins.  output_ini passes the station
    position already multiplied by rad and the search scales it again (by
    1/r_earth on a Cartesian mesh, by rad otherwise); kept as is."""
    x, y = px * RAD, py * RAD
    if cartesian:
        d = np.sqrt((y / R_EARTH - cy) ** 2 + (x / R_EARTH - cx) ** 2)
    else:
        d = np.sin((y * RAD - cy) * 0.5) ** 2 + np.cos(cy) * np.cos(y * RAD) * np.sin((x * RAD - cx) * 0.5) ** 2
        d = np.arcsin(np.minimum(1.0, np.sqrt(d)))
    return int(np.argmin(d))


class Output:
    """The reference's output_ini + output_do over one model.

    source: object with ``get(name)`` (Fortran-layout doubles) for eta_n,
    U_n, V_n, U_n_2D, TF, SF, taux, tauy, mask_wd, etaAB, windx, windy, Z,
    zbar, Wvel, snu, bt; ``clock()`` -> (n_dt, time_jd) after the last step.
    cfg: RunConfig of the run; mesh: RawMesh; workdir: run directory holding
    outputfilelist.dat (files go to workdir/nc_output/, as in the reference).
    """

    def __init__(self, source, cfg, mesh, workdir=".", type_task=3):
        from . import gpu
        self.src, self.cfg = source, cfg
        m = gpu.mesh_arrays(cfg, mesh)
        self.N, self.E, self.K = m["nod2D"], m["elem2D"], len(m["sigma"])
        self.sigma = m["sigma"]
        self.L = self.K - 1
        self.cartesian = bool(cfg.l("cartesian"))
        self.Dmin, self.C_d, self.dt = cfg.f("Dmin", 0.25), cfg.f("Cd", 0.0), cfg.f("dt")
        self.wet_dry = bool(cfg.l("WET_DRY_ON"))
        self.rivers_on = bool(cfg.l("key_rivers"))
        self.depth = m["depth"]
        en = m["elem_nodes"].reshape(self.E, 4)            # 1-based, reference element order
        wcv = m["w_cv"].reshape(self.E, 4)
        xy = np.asarray(mesh.xy, dtype=np.float64)
        coord = xy / R_EARTH if self.cartesian else xy * RAD  # coord_nod2D (read_mesh_ser)
        self.files = [f for f in read_filelist(os.path.join(workdir, "outputfilelist.dat"))]
        for f in self.files:
            f["vars"] = [v for v in f["vars"] if VARLIST[v][7] <= type_task]
            bad = [v for v in f["vars"] if v in NOT_HELD]
            if bad:
                raise ValueError(f"{f['name']}: variables {bad} ({', '.join(VARLIST[v][0] for v in bad)}) "
                                 "belong to code this build does not hold (Richardson number)")
            if f["ftype"] not in (0, 3):
                raise ValueError("output_ini: wrong/not supported output file type")
        # element centres and depth (output_ini, src/fv_ncoutput.F90:965-981)
        q = en - 1
        ex = ((wcv[:, 0] * coord[q[:, 0], 0] + wcv[:, 1] * coord[q[:, 1], 0]) + wcv[:, 2] * coord[q[:, 2], 0]) \
            + wcv[:, 3] * coord[q[:, 3], 0]
        ey = ((wcv[:, 0] * coord[q[:, 0], 1] + wcv[:, 1] * coord[q[:, 1], 1]) + wcv[:, 2] * coord[q[:, 2], 1]) \
            + wcv[:, 3] * coord[q[:, 3], 1]
        dep = self.depth
        de = ((wcv[:, 0] * dep[q[:, 0]] + wcv[:, 1] * dep[q[:, 1]]) + wcv[:, 2] * dep[q[:, 2]]) + wcv[:, 3] * dep[q[:, 3]]
        # nod_in_elem2D (src/fv_mesh_array.F90:874-884): elements of each node in ascending id
        cnt = np.zeros(self.N, dtype=np.int32)
        lists = [[] for _ in range(self.N)]
        for e in range(self.E):
            for j in range(3 if en[e, 0] == en[e, 3] else 4):
                lists[q[e, j]].append(e + 1)
                cnt[q[e, j]] += 1
        maxne = int(cnt.max())
        nie = np.zeros((self.N, maxne), dtype=np.int32)
        for n in range(self.N):
            nie[n, :cnt[n]] = lists[n]
        os.makedirs(os.path.join(workdir, "nc_output"), exist_ok=True)
        for f in self.files:
            if f["ftype"] == 0:
                f["nods"], f["elems"] = np.arange(self.N), np.arange(self.E)
            else:
                f["nods"] = np.array([_nearest(f["pos"][0], f["pos"][1], coord[:, 0], coord[:, 1], self.cartesian)])
                f["elems"] = np.array([_nearest(f["pos"][0], f["pos"][1], ex, ey, self.cartesian)])
            f["path"] = os.path.join(workdir, "nc_output", f["name"])
            f["nc_rec"] = 0
            f["avg"] = {v: None for v in f["vars"]}
            nn, ne = len(f["nods"]), len(f["elems"])
            scale = (lambda a: a * R_EARTH) if self.cartesian else (lambda a: a / RAD)
            nc = ClassicFile(f["path"])
            nc.dim("node", nn)
            nc.dim("nele", ne)
            nc.dim("maxnod", 4)
            nc.dim("maxnodelem", maxne)
            nc.dim("time", None)
            if type_task > 1:
                nc.dim("sigma", self.K)
                nc.dim("sigmam1", self.L)
                nc.var("sigma_lev", "d", ("sigma",), {}, m["sigma"])
            nc.var("nv", "i", ("nele", "maxnod"), {"standard_name": "face_node_connectivity"}, en[f["elems"]])
            nc.var("fvcoastal_mesh", "i", (), {"cf_role": "mesh_topology", "node_coordinates": "lon lat",
                                               "face_node_connectivity": "nv"}, None)
            # lat units: the first len_trim('degrees_east') characters of
            # 'degrees_north', as the reference writes them (src/fv_ncoutput.F90:1183)
            if self.cartesian:
                a_lon, a_lat = {"standard_name": "X", "units": "meters"}, {"standard_name": "Y", "units": "meters"}
                a_lone = {"standard_name": "X_elements", "units": "meters"}
                a_late = {"standard_name": "Y_elements", "units": "meters"}
            else:
                a_lon = {"standard_name": "longitude", "units": "degrees_east"}
                a_lat = {"standard_name": "latitude", "units": "degrees_nort"}
                a_lone = {"standard_name": "longitude_elements", "units": "degrees_east"}
                a_late = {"standard_name": "latitude_elements", "units": "degrees_nort"}
            nc.var("lon", "d", ("node",), a_lon, scale(coord[f["nods"], 0]))
            nc.var("lat", "d", ("node",), a_lat, scale(coord[f["nods"], 1]))
            nc.var("depth", "d", ("node",), {"standard_name": "depth", "units": "m"}, dep[f["nods"]])
            nc.var("depth_elem", "d", ("nele",), {"standard_name": "depth on elem", "units": "m"}, de[f["elems"]])
            nc.var("lon_elem", "d", ("nele",), a_lone, scale(ex[f["elems"]]))
            nc.var("lat_elem", "d", ("nele",), a_late, scale(ey[f["elems"]]))
            nc.var("time", "d", ("time",), {"standard_name": "time", "units": "days since 1900-01-01 00:00:00",
                                             "axis": "T", "title": "Time"})
            nc.var("area", "d", ("node",), {"standard_name": "cell area", "units": "m**2"}, m["area"][f["nods"]])
            nc.var("elem_area", "d", ("nele",), {"standard_name": "elemens area", "units": "m**2"},
                   m["elem_area"][f["elems"]])
            nc.var("w_cv", "d", ("nele", "maxnod"), {"standard_name": " Scv/Sc"}, wcv[f["elems"]])
            nc.var("nod_in_elem2d_num", "i", ("node",), {"standard_name": "node is in N elemets"}, cnt[f["nods"]])
            nc.var("nod_in_elem2d", "i", ("node", "maxnodelem"), {"standard_name": "node is in elemets"},
                   nie[f["nods"]])
            for vid in f["vars"]:
                name, model, std, units, onnode, timedim, vertdim, _ = VARLIST[vid]
                dims = (("time",) if timedim else ()) + (("node",) if onnode == 1 else ("nele",)) + \
                    {0: (), 1: ("sigma",), 2: ("sigmam1",)}[vertdim]
                nc.var(name, "d", dims, {"units": units, "standard_name": std, "modelname": model})
            nc.enddef()
            f["nc"] = nc
        self.n_dt = None

    # -- getvar1d / getvar2d (src/fv_ncoutput.F90:720-866) -----------------
    def _masks(self):
        mask_e = self.src.get("mask_wd")
        if self.wet_dry:    # wad_mask, src/fv_main.F90:1301-1304
            mask_n = np.where(self.depth + self.src.get("etaAB") > self.Dmin, 1.0, 0.0)
        else:
            mask_n = np.ones(self.N)
        return mask_n, mask_e

    def _field(self, vid, cache):
        N, E, L, K = self.N, self.E, self.L, self.K
        g = self.src.get
        if vid == 1:
            eta = g("eta_n")
            a = np.where(self.depth < 0.0, eta + self.depth, eta)
        elif vid in (2, 3):
            a = g("U_n" if vid == 2 else "V_n").reshape(E, L)
        elif vid in (4, 5):
            a = g("U_n" if vid == 4 else "V_n").reshape(E, L)[:, 0].copy()
        elif vid in (6, 7):
            a = g("U_n_2D").reshape(E, 2)[:, vid - 6].copy()
        elif vid in (8, 9):
            a = g("TF" if vid == 8 else "SF").reshape(N, L)
        elif vid in (10, 11):
            a = g("TF" if vid == 10 else "SF").reshape(N, L)[:, 0].copy()
        elif vid in (12, 13):
            a = g("taux" if vid == 12 else "tauy")
        elif vid == 20:
            a = g("zbar").reshape(N, K)
        elif vid == 21:
            a = g("Z").reshape(N, L)
        elif vid == 22:
            a = cache["mask"][0].copy()
        elif vid == 23:
            a = cache["mask"][1].copy()
        elif vid == 24:       # update_2D_vel sets C_d_el to C_d (src/fv_average_dynamic.F90:260-262)
            a = np.full(E, self.C_d)
        elif vid in (25, 26):
            a = g("windx" if vid == 25 else "windy")
        elif vid in (14, 15):  # the bulk formulas' nodal stress (src/fv_sbc.F90:436-437)
            try:
                a = g("tau_node").reshape(N, 2)[:, vid - 14].copy()
            except Exception as err:
                raise ValueError(f"{VARLIST[vid][0]} is formed by the bulk formulas (sbc), which this run "
                                 "does not apply") from err
        elif vid in (16, 17, 18, 19):
            a = g({16: "mslp", 17: "emp", 18: "qns", 19: "qsr"}[vid])
        elif vid == 35:
            # the discharge of every mouth on its node (summed where mouths
            # share one).  The reference's serial writer assigns its per-river
            # array to this nodal slot (src/fv_ncoutput.F90:797-798) and its
            # MPI writer writes zeros (fv_ncoutputmpi.F90:992).
            a = g("runoff") if self.rivers_on else np.zeros(N)
        elif vid in (27, 28, 29):
            a = g({27: "Wvel", 28: "snu", 29: "bt"}[vid]).reshape(N, K)
        elif vid in ZERO:
            a = np.zeros(N if VARLIST[vid][4] == 1 else E)
        else:
            raise ValueError(f"variable {vid} not held")
        mk = cache["mask"][0] if VARLIST[vid][4] == 1 else cache["mask"][1]
        return a * (mk[:, None] if a.ndim == 2 else mk)

    # -- output_do (src/fv_ncoutput.F90:868-889) -----------------------------
    def step(self):
        """Call after every internal step."""
        if self.n_dt is None:
            self.n_dt = self.src.clock()[0]
        else:
            self.n_dt += 1
        n_dt = self.n_dt
        cache = {}
        for f in self.files:
            write = n_dt % f["freq"] == 0 or n_dt == 1
            if not (write or f["snapshot"] == 2):
                continue
            if "mask" not in cache:
                cache["mask"] = self._masks()
            sel = {}
            for vid in f["vars"]:
                if vid not in cache:
                    cache[vid] = self._field(vid, cache)
                idx = f["nods"] if VARLIST[vid][4] == 1 else f["elems"]
                sel[vid] = cache[vid][idx]
                if f["snapshot"] == 2:
                    f["avg"][vid] = sel[vid].copy() if f["avg"][vid] is None else f["avg"][vid] + sel[vid]
            if write:
                self._write(f, sel)

    def _write(self, f, sel):
        """write_data / write_var (src/fv_ncoutput.F90:1653-1757)."""
        f["nc_rec"] += 1
        tshift = f["freq"] * self.dt * 0.5 / 86400.0 if f["snapshot"] == 2 else 0.0
        rec = {"time": np.array([self.src.clock()[1] - 2415020.5 - tshift])}
        for vid in f["vars"]:
            if f["snapshot"] == 2:    # calculateavg, then zeroavg
                val = f["avg"][vid] / float(f["freq"])
                f["avg"][vid] = np.zeros_like(f["avg"][vid])
            else:
                val = sel[vid]
            rec[VARLIST[vid][0]] = val
        f["nc"].append(rec)

    def close(self):
        """Files are complete after every record; nothing is held open."""
        for f in self.files:
            f["nc"] = None
