"""Fortran host over the C-ABI: the ISO_C_BINDING module and the driver that
replaces fv_main's call sites (fesom-c_amd/fortran).  The driver reads the
reference's own input files (fv_run.dat, nod2d.out, elem2d.out, depth.out,
m2.out) and writes energy.dat in the reference's record format."""
import gzip
import os
import shutil
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, ROOT, have_gpu

EXE = os.path.join(ROOT, "fesom-c_amd", "fortran", "fesomc_run")
LE = os.path.join(GOLDEN, "le_medium")


def _le_rundir(tmp_path):
    d = tmp_path / "le"
    (d / "MESH").mkdir(parents=True)
    for f in ("nod2d", "elem2d", "depth"):
        with gzip.open(os.path.join(LE, f + ".out.gz"), "rb") as src, open(d / "MESH" / (f + ".out"), "wb") as dst:
            shutil.copyfileobj(src, dst)
    # the reference's energy series was produced with Mt = 4 (see test_oracle_pins)
    (d / "fv_run.dat").write_text(open(os.path.join(LE, "fv_run.dat")).read() + "\n Mt = 4\n")
    return d


def _run(args, cwd, timeout=120):
    assert os.path.exists(EXE), "Fortran driver not built (run __graft_entry__.build())"
    return subprocess.run([EXE] + [str(a) for a in args], cwd=cwd, capture_output=True, text=True, timeout=timeout)


def _mesh_line(out):
    line = [ln for ln in out.splitlines() if "mesh:" in ln][0]
    return [int(v) for v in line.split()[-5:]]


def test_driver_le_mesh_matches_reference_edge_files(tmp_path):
    """nod2D/elem2D and the edge counts equal the reference's MESH/edgenum.out."""
    d = _le_rundir(tmp_path)
    r = _run([d, 1, "--mesh-only"], d)
    assert r.returncode == 0, r.stdout + r.stderr
    with gzip.open(os.path.join(LE, "edgenum.out.gz"), "rt") as f:
        edgenum = [int(v) for v in f.read().split()]
    nod, elem, edges, edges_in, nobn = _mesh_line(r.stdout)
    assert (nod, elem, nobn) == (4515, 4352, 0)
    assert [edges, edges_in] == edgenum[:2]


def test_driver_channel_mesh_with_tides(tmp_path):
    """Triangles + quads + m2.out open boundary: the driver's preprocessing
    sizes equal the Python host mirror's on the same files."""
    from fesomc_amd import gpu, meshgen, meshio
    from fesomc_amd.runconfig import RunConfig
    txt, mesh, _ = meshgen.config("c1")
    d = tmp_path / "c1"
    meshio.write_mesh(str(d / "MESH"), mesh)
    (d / "fv_run.dat").write_text(txt)
    r = _run([d, 1, "--mesh-only"], d)
    assert r.returncode == 0, r.stdout + r.stderr
    back = meshio.read_mesh(str(d / "MESH"), n_harm=1)
    ref = gpu.mesh_arrays(RunConfig(txt), back)
    assert _mesh_line(r.stdout) == [ref["nod2D"], ref["elem2D"], ref["edge2D"], ref["edge2D_in"], ref["nobn"]]
    assert ref["nobn"] > 0


@pytest.mark.skipif(have_gpu(), reason="checks the behaviour without a GPU")
def test_driver_without_gpu_stops_with_status(tmp_path):
    d = _le_rundir(tmp_path)
    r = _run([d, 5], d)
    assert r.returncode != 0
    assert "fcg_create failed with status" in r.stdout and "-7" in r.stdout


@pytest.mark.gpu
def test_driver_reproduces_reference_energy_rows(tmp_path):
    """300 internal steps of LE_medium through the Fortran call sites: the
    three energy.dat rows match energy_tocheck.dat at its printed digits."""
    from test_oracle_pins import digits7_ok
    d = _le_rundir(tmp_path)
    r = _run([d, 300], d, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    rows = np.loadtxt(d / "energy.dat", ndmin=2)
    ref = np.loadtxt(os.path.join(LE, "energy_tocheck_head200.dat"), ndmin=2)
    assert rows.shape == (3, 2)
    for k in range(3):
        assert abs(rows[k, 0] - ref[k, 0]) <= 1e-6 * ref[k, 0]
        assert digits7_ok(rows[k, 1], ref[k, 1]), (k, rows[k, 1], ref[k, 1])
    _check_control_block(r.stdout, rows)


def _block_values(out, label):
    vals = []
    for ln in out.splitlines():
        if label in ln:
            vals.append([float(v) for v in ln.split(label, 1)[1].split()])
    return vals


def _check_control_block(out, rows):
    """The stdout block of every IREP-th step (src/fv_main.F90:614-653, serial
    branch, VERBOSITY = 1): one block per energy.dat row, the energy line equal
    to the record, and the extrema of the last block equal to those of the
    FP64 oracle after the same 300 steps (eta_n, U_n_2D, TF, SF)."""
    from oracle.pyoracle import Oracle
    from fesomc_amd import meshio
    from test_oracle_pins import digits7_ok
    energy = _block_values(out, "Energy ENERGY ENERGY :")
    assert len(energy) == 3 and len(_block_values(out, "dt_2D")) == 3
    for k in range(3):
        assert digits7_ok(energy[k][0], rows[k, 1])
    txt = open(os.path.join(LE, "fv_run.dat")).read() + "\n Mt = 4\n"
    o = Oracle(txt, meshio.read_mesh(LE))
    o.step(300)
    want = {"max_min_ssh:": o.get("eta_n"), "max_min_vel 2D:": o.get("U_n_2D"), "T_maxmin=": o.get("TF"),
            "S_maxmin=": o.get("SF")}
    for label, a in want.items():
        got = _block_values(out, label)[-1]
        assert got == pytest.approx([a.max(), a.min()], rel=1e-12, abs=1e-15), label


@pytest.mark.gpu
def test_driver_reads_and_applies_the_river_table(tmp_path):
    """key_rivers with a two-record MESH/rivers.dat: the driver reads the
    table (src/fv_rivers.F90:75-214), interpolates it after each clock
    advance (rivers_do, ramped over the first ten days) and hands the mouths'
    discharge to the step.  After 100 steps its stdout extrema of eta_n,
    U_n_2D, TF and SF equal those of the FP64 oracle driven through the same
    file by the Python host's table (fesom-c_amd/rivers.py).  Salinity is off
    in this case, so SF > 0 is river water alone."""
    import re
    from oracle.pyoracle import Oracle
    from fesomc_amd import meshio
    from fesomc_amd.rivers import Rivers
    d = _le_rundir(tmp_path)
    txt, n = re.subn(r"^(\s*key_rivers\s*=\s*)\S+", r"\g<1>T", (d / "fv_run.dat").read_text(), count=1, flags=re.M)
    assert n == 1
    (d / "fv_run.dat").write_text(txt)
    jd0 = 2457266.5
    table = d / "MESH" / "rivers.dat"
    table.write_text(f"2\n2\n{jd0 - 1.0} {jd0 + 1.0}\n100 2000\n1.0e5 5.0e4\n3.0e5 1.0e5\n")
    r = _run([d, 100], d, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "rivers:" in r.stdout
    o = Oracle(txt, meshio.read_mesh(LE))
    riv = Rivers(str(table), o.size["N"], jd0, jd0)
    t = jd0
    for _ in range(100):
        t = t + 0.5 / 86400.0           # the clock of src/fv_main.F90:400
        o.set("runoff", riv.do(t))
        o.step(1)
    assert o.get("SF").max() > 0.0
    want = {"max_min_ssh:": o.get("eta_n"), "max_min_vel 2D:": o.get("U_n_2D"), "T_maxmin=": o.get("TF"),
            "S_maxmin=": o.get("SF")}
    for label, a in want.items():
        got = _block_values(r.stdout, label)[-1]
        assert got == pytest.approx([a.max(), a.min()], rel=1e-12, abs=1e-15), label


def test_driver_refuses_options_it_does_not_hold(tmp_path):
    """The tidal potential (T_potential, off the hot path) in fv_run.dat stops
    the driver before any set-up, naming the option (no GPU: the check
    precedes the mesh)."""
    d = _le_rundir(tmp_path)
    import re
    txt, n = re.subn(r"^(\s*T_potential\s*=\s*)\S+", r"\g<1>T", (d / "fv_run.dat").read_text(), count=1,
                     flags=re.M)
    assert n == 1
    (d / "fv_run.dat").write_text(txt)
    r = _run([d, 1, "--mesh-only"], d)
    assert r.returncode == 2 and "T_potential" in r.stdout
