"""NetCDF output in the reference's layout (fesom-c_amd/ncout.py,
src/fv_ncoutput.F90).  The writer runs over the FP64 oracle on the CPU (the
oracle's own mask_wd_node checks the node mask the writer derives from etaAB)
and over the device model on the GPU; the files are read back with scipy."""
import numpy as np
import pytest
from scipy.io import netcdf_file

from oracle.pyoracle import Oracle, lib

FILELIST = """!total number of files:
3
!=========== Field File Name ===
snap.nc
! If snapshot = 1, averaged needed = 2
1
! Type of file 0 - full domain, 3 - station
0
! Position of station
0.0, 0.0, 0.0, 0.0
! frequency of output (in time speps)
5
! special frequency (not in use)
0
! total number of variables in file
19
!G list of variables
1,2,3,4,5,6,7,8,9,10,11,20,21,22,23,24,27,28,29,0,0,0,0,0,0,0,0,0,0,0
!=========== Field File Name ===
avg.nc
! snapshot/average
2
! type
0
! position
0.0, 0.0, 0.0, 0.0
! frequency
4
! special
0
! number of variables
3
! list
1,8,3,0,0,0,0,0,0,0,0,0,0,0,0,0,0,0,0,0,0,0,0,0,0,0,0,0,0,0
!=========== Field File Name ===
station.nc
! snapshot
1
! station
3
! position (x, y; the reference scales it by rad before the search)
300000.0, 200000.0, 0.0, 0.0
! frequency
3
! special
0
! number
2
! list
1,8,0,0,0,0,0,0,0,0,0,0,0,0,0,0,0,0,0,0,0,0,0,0,0,0,0,0,0,0
"""

SNAP_VARS = {"eta": 1, "u": 2, "v": 3, "ssu": 4, "ssv": 5, "u2d": 6, "v2d": 7, "temperature": 8, "salinity": 9,
             "sst": 10, "sss": 11, "zbar": 20, "Z": 21, "mask_wd_node": 22, "mask_wd": 23, "Cd": 24, "w": 27,
             "snu": 28, "bt": 29}


class OracleSource:
    def __init__(self, o):
        self.o = o

    def get(self, name):
        return self.o.get(name)

    def clock(self):
        return lib().orc_size(self.o.h, 10), self.o.scalar("time_jd")


def _case():
    from fesomc_amd import meshgen
    txt, mesh, ex = meshgen.config("c2_small")
    return txt, mesh, ex


def _oracle(txt, mesh, ex):
    from fesomc_amd import meshgen
    o = Oracle(txt, mesh)
    o.set_tracers(ex["T"].reshape(-1), ex["S"].reshape(-1))
    wx, wy = ex["wind"]
    tx, ty = meshgen.wind_stress(wx, wy)
    o.set("windx", np.full(mesh.nod2D, wx))
    o.set("windy", np.full(mesh.nod2D, wy))
    o.set("taux", np.full(mesh.elem2D, tx))
    o.set("tauy", np.full(mesh.elem2D, ty))
    return o


def _expected(o, N, E, L, K, depth):
    """getvar1d/getvar2d from the oracle's own arrays and masks."""
    mn, me = o.get("mask_wd_node"), o.get("mask_wd")
    eta = o.get("eta_n")
    U, V = o.get("U_n").reshape(E, L), o.get("V_n").reshape(E, L)
    T, S = o.get("TF").reshape(N, L), o.get("SF").reshape(N, L)
    u2 = o.get("U_n_2D").reshape(E, 2)
    return {"eta": np.where(depth < 0, eta + depth, eta) * mn, "u": U * me[:, None], "v": V * me[:, None],
            "ssu": U[:, 0] * me, "ssv": V[:, 0] * me, "u2d": u2[:, 0] * me, "v2d": u2[:, 1] * me,
            "temperature": T * mn[:, None], "salinity": S * mn[:, None], "sst": T[:, 0] * mn, "sss": S[:, 0] * mn,
            "zbar": o.get("zbar").reshape(N, K) * mn[:, None], "Z": o.get("Z").reshape(N, L) * mn[:, None],
            "mask_wd_node": mn * mn, "mask_wd": me * me, "Cd": o.get("C_d_el") * me,
            "w": o.get("Wvel").reshape(N, K) * mn[:, None], "snu": o.get("snu").reshape(N, K) * mn[:, None],
            "bt": o.get("bt").reshape(N, K) * mn[:, None]}


def _run_and_check(tmp_path, make_source, steps=12):
    from fesomc_amd import ncout
    from fesomc_amd.runconfig import RunConfig
    txt, mesh, ex = _case()
    (tmp_path / "outputfilelist.dat").write_text(FILELIST)
    o = _oracle(txt, mesh, ex)
    src, stepper = make_source(txt, mesh, ex)
    out = ncout.Output(src, RunConfig(txt), mesh, workdir=str(tmp_path))
    N, E, L, K = out.N, out.E, out.L, out.K
    depth = out.depth
    want_snap, want_avg, want_sta, acc = [], [], [], None
    for n in range(1, steps + 1):
        stepper()
        out.step()
        o.step(1)
        ex_ = _expected(o, N, E, L, K, depth)
        t = o.scalar("time_jd") - 2415020.5
        if n % 5 == 0 or n == 1:
            want_snap.append((t, ex_))
        acc = {k: ex_[k].copy() for k in ("eta", "temperature", "v")} if acc is None else \
            {k: acc[k] + ex_[k] for k in acc}
        if n % 4 == 0 or n == 1:
            want_avg.append((t - 4 * float(RunConfig(txt).f("dt")) * 0.5 / 86400.0,
                             {k: acc[k] / 4.0 for k in acc}))
            acc = {k: np.zeros_like(v) for k, v in acc.items()}
        if n % 3 == 0 or n == 1:
            want_sta.append((t, ex_))
    out.close()
    return tmp_path, want_snap, want_avg, want_sta, mesh, out


def _check_files(d, want_snap, want_avg, want_sta, mesh, out, rtol):
    with netcdf_file(str(d / "nc_output" / "snap.nc"), "r", mmap=False) as f:
        assert f.dimensions["node"] == out.N and f.dimensions["nele"] == out.E and f.dimensions["sigma"] == out.K
        assert f.variables["time"].units == b"days since 1900-01-01 00:00:00"
        assert f.variables["lon"].standard_name == b"X" and f.variables["nv"].shape == (out.E, 4)
        assert f.variables["temperature"].dimensions == ("time", "node", "sigmam1")
        assert f.variables["u"].dimensions == ("time", "nele", "sigmam1")
        assert f.variables["w"].dimensions == ("time", "node", "sigma")
        mesh_var = f.variables["fvcoastal_mesh"]
        assert mesh_var.cf_role == b"mesh_topology" and mesh_var.face_node_connectivity == b"nv"
        assert f.variables["temperature"].modelname == b"TF" and f.variables["eta"].units == b"meter"
        assert f.variables["w_cv"].standard_name == b" Scv/Sc"
        np.testing.assert_array_equal(f.variables["sigma_lev"][:], out.sigma)
        np.testing.assert_array_equal(f.variables["lon"][:], (mesh.xy[:, 0] / 6400000.0) * 6400000.0)
        assert f.variables["time"].shape[0] == len(want_snap) == 3
        for r, (t, ex_) in enumerate(want_snap):
            assert f.variables["time"][r] == pytest.approx(t, abs=1e-9)
            for name in SNAP_VARS:
                np.testing.assert_allclose(f.variables[name][r], ex_[name], rtol=rtol, atol=rtol * 1e-3,
                                           err_msg=name)
    with netcdf_file(str(d / "nc_output" / "avg.nc"), "r", mmap=False) as f:
        assert f.variables["time"].shape[0] == len(want_avg) == 4
        for r, (t, ex_) in enumerate(want_avg):
            assert f.variables["time"][r] == pytest.approx(t, abs=1e-9)
            for name in ("eta", "temperature", "v"):
                np.testing.assert_allclose(f.variables[name][r], ex_[name], rtol=rtol, atol=rtol * 1e-3,
                                           err_msg=name)
    with netcdf_file(str(d / "nc_output" / "station.nc"), "r", mmap=False) as f:
        assert f.dimensions["node"] == 1 and f.dimensions["nele"] == 1
        rad = 3.14159265358979 / 180.0
        px, py = 300000.0 * rad, 200000.0 * rad
        dn = np.hypot(mesh.xy[:, 0] / 6400000.0 - px / 6400000.0, mesh.xy[:, 1] / 6400000.0 - py / 6400000.0)
        node = int(np.argmin(dn))
        assert f.variables["lon"][0] == (mesh.xy[node, 0] / 6400000.0) * 6400000.0
        for r, (t, ex_) in enumerate(want_sta):
            np.testing.assert_allclose(f.variables["temperature"][r][0], ex_["temperature"][node], rtol=rtol)


def test_filelist_reader_reads_reference_file():
    import os
    from fesomc_amd import ncout
    from conftest import ROOT
    p = os.path.join(ROOT, "tests", "golden", "le_medium", "outputfilelist.dat")
    files = ncout.read_filelist(p)
    assert len(files) == 1
    f = files[0]
    assert f["name"] == "LE.nc" and f["snapshot"] == 1 and f["ftype"] == 0 and f["freq"] == 300
    assert f["vars"] == [1, 2, 3, 8, 9]


def test_writer_over_oracle_matches_oracle_fields(tmp_path):
    """The writer fed by the oracle writes, at the reference's output steps,
    the oracle's own fields times the oracle's own wet/dry masks (snapshots),
    their means over each interval (averages, time shifted by half the
    interval) and the nearest node's values (station)."""
    def make(txt, mesh, ex):
        s = _oracle(txt, mesh, ex)
        return OracleSource(s), lambda: s.step(1)
    _check_files(*_run_and_check(tmp_path, make), rtol=0.0)


@pytest.mark.gpu
def test_writer_over_device_model(tmp_path):
    """The same files from the device model (FP64): equal to the oracle's."""
    from fesomc_amd import gpu, meshgen
    from fesomc_amd.runconfig import RunConfig

    def make(txt, mesh, ex):
        m = gpu.Model(RunConfig(txt), mesh)
        m.set_tracers(ex["T"], ex["S"])
        wx, wy = ex["wind"]
        tx, ty = meshgen.wind_stress(wx, wy)
        m.set("windx", np.full(mesh.nod2D, wx))
        m.set("windy", np.full(mesh.nod2D, wy))
        m.set("taux", np.full(mesh.elem2D, tx))
        m.set("tauy", np.full(mesh.elem2D, ty))
        return m, lambda: m.step(1)
    _check_files(*_run_and_check(tmp_path, make), rtol=1e-12)


FORCED_FILELIST = """!total number of files:
1
!=========== Field File Name ===
forced.nc
! snapshot
1
! type
0
! position
0.0, 0.0, 0.0, 0.0
! frequency
2
! special
0
! number of variables
5
! list
16,17,18,19,35,0,0,0,0,0,0,0,0,0,0,0,0,0,0,0,0,0,0,0,0,0,0,0,0,0
"""


def test_writer_serves_surface_forcing_and_river_fields(tmp_path):
    """With key_atm and key_rivers the writer's mslp, emp, qns, qsr and
    runoff_rivers (ids 16-19, 35) are the fields the step used -- the host's
    surface fields and the per-node discharge -- times the node mask."""
    import re
    from fesomc_amd import meshgen, ncout
    from fesomc_amd.runconfig import RunConfig
    txt, mesh, ex = _case()
    for k in ("key_atm", "key_rivers"):
        txt = re.sub(rf"^(\s*{k}\s*=\s*)\S+", r"\g<1>T", txt, count=1, flags=re.M)
    (tmp_path / "outputfilelist.dat").write_text(FORCED_FILELIST)
    o = _oracle(txt, mesh, ex)
    x, y = mesh.xy[:, 0], mesh.xy[:, 1]
    fields = {"qsr": 200.0 + 50.0 * np.sin(x / 3000.0), "qns": -40.0 + 10.0 * np.cos(y / 2000.0),
              "emp": np.full(mesh.nod2D, 1e-6), "mslp": 101325.0 + 300.0 * np.sin(y / 5000.0)}
    r = np.zeros(mesh.nod2D)
    r[np.argsort(np.where(o.get("ac") >= 1.0, o.get("depth"), -np.inf))[-2:]] = (30.0, 7.0)
    fields["runoff"] = r
    for k, v in fields.items():
        o.set(k, v)
    out = ncout.Output(OracleSource(o), RunConfig(txt), mesh, workdir=str(tmp_path))
    want = []
    for n in range(1, 5):
        o.step(1)
        out.step()
        mn = o.get("mask_wd_node")
        if n % 2 == 0 or n == 1:
            want.append({name: o.get(f) * mn for name, f in (("mslp", "mslp"), ("emp", "emp"), ("qns", "qns"),
                                                            ("qsr", "qsr"), ("runoff_rivers", "runoff"))})
    out.close()
    with netcdf_file(str(tmp_path / "nc_output" / "forced.nc"), "r", mmap=False) as f:
        assert f.variables["time"].shape[0] == len(want) == 3
        for rec, w in enumerate(want):
            for name, a in w.items():
                np.testing.assert_array_equal(f.variables[name][rec], a, err_msg=name)
        assert f.variables["runoff_rivers"][0].max() == 30.0
