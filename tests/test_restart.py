"""Restart files in the reference's format (src/fv_restart.F90:474-613): the
record framing against the Fortran runtime of the toolchain, the Python
reader/writer, and -- on the GPU -- the Fortran driver's IRESTART writes and
hot start: a run restarted from the file written after step 10 reproduces,
bit for bit, the restart a continuous run writes after step 20."""
import os
import shutil
import struct
import subprocess

import numpy as np
import pytest

from conftest import ROOT

FLANG = "/opt/rocm/lib/llvm/bin/flang"


def test_python_restart_round_trip(tmp_path):
    from fesomc_amd import restart
    N, E, K = 7, 9, 5
    rng = np.random.default_rng(0)
    arrays = {k: rng.normal(size=v) for k, v in restart.sizes(N, E, K).items()}
    p = tmp_path / "r.out"
    restart.write_arrays(p, 12, (2.5e6, 2.4e6, 3.5, 0.35), arrays)
    n, clock, back = restart.read(p, N, E, K)
    assert n == 12 and clock == (2.5e6, 2.4e6, 3.5, 0.35)
    for k, v in arrays.items():
        assert np.array_equal(back[k], v), k
    raw = open(p, "rb").read()
    assert struct.unpack_from("<iii", raw, 0) == (4, 12, 4)          # record 1: one 4-byte integer
    assert struct.unpack_from("<i", raw, 12)[0] == 32                 # record 2: four reals


@pytest.mark.skipif(not os.path.exists(FLANG), reason="no Fortran compiler")
def test_framing_matches_the_fortran_runtime(tmp_path):
    """A file written by a Fortran program with the reference's write
    statements parses with the reader (marker width, order, endianness)."""
    from fesomc_amd import restart
    src = tmp_path / "w.f90"
    src.write_text("""program w
  integer :: n = 7
  real(8) :: a(3) = [1d0, 2d0, 3d0], b(2, 2) = reshape([4d0, 5d0, 6d0, 7d0], [2, 2])
  open(25, file='f.out', form='unformatted')
  write(25) n
  write(25) 1.5d0, 2.5d0, 3.5d0, 4.5d0
  write(25) a, b
  close(25)
end program w
""")
    exe = tmp_path / "w"
    subprocess.run([FLANG, "-o", str(exe), str(src)], check=True, cwd=tmp_path, capture_output=True)
    subprocess.run([str(exe)], check=True, cwd=tmp_path)
    recs = restart.read_records(tmp_path / "f.out")
    assert struct.unpack("<i", recs[0]) == (7,)
    assert struct.unpack("<4d", recs[1]) == (1.5, 2.5, 3.5, 4.5)
    assert np.array_equal(np.frombuffer(recs[2], "<f8"), [1, 2, 3, 4, 5, 6, 7])


def _set_keys(path, **kv):
    """fv_run.dat with KEY = value replaced (the scanner takes the first occurrence)."""
    import re
    txt = open(path).read()
    for k, v in kv.items():
        txt, n = re.subn(rf"^(\s*{k}\s*=\s*)\S+", rf"\g<1>{v}", txt, count=1, flags=re.M)
        if n == 0:
            txt += f"\n {k} = {v}\n"
    open(path, "w").write(txt)


@pytest.mark.gpu
def test_driver_hot_start_continues_bit_exactly(tmp_path):
    from fesomc_amd import restart
    from test_fortran_host import _le_rundir, _mesh_line, _run
    a = _le_rundir(tmp_path / "a")
    _set_keys(a / "fv_run.dat", IRESTART=10)
    (a / "restarts").mkdir()
    r = _run([a, 20], a, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    N, E = _mesh_line(r.stdout)[:2]
    import re
    K = int(re.search(r"^\s*nsigma\s*=\s*(\d+)", open(a / "fv_run.dat").read(), re.M).group(1))
    n10, c10, s10 = restart.read(a / "restarts" / "0000000010_restart_0.out", N, E, K)
    n20, c20, s20 = restart.read(a / "restarts" / "0000000020_restart_0.out", N, E, K)
    assert (n10, n20) == (10, 20) and c20[0] > c10[0]
    assert not np.array_equal(s10["eta_n"], s20["eta_n"])
    b = _le_rundir(tmp_path / "b")
    _set_keys(b / "fv_run.dat", IRESTART=10, RESTART="hot_start")
    (b / "restarts").mkdir()
    shutil.copy(a / "restarts" / "0000000010_restart_0.out", b / "restarts" / "restart_0.out")
    r = _run([b, 20], b, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "Reading restart" in r.stdout
    n, c, s = restart.read(b / "restarts" / "0000000010_restart_0.out", N, E, K)
    assert n == 10 and c == c20
    for k in s20:
        assert np.array_equal(s[k], s20[k]), k
