"""CPU checks of the product's host side: the C-ABI library loads and exports
every entry point include/fesomc_gpu.h declares, and its C++ mesh
preprocessing reproduces the reference numbering and geometry bit for bit
(checked against the oracle, which is pinned to the reference edge files)."""
import os
import re

import numpy as np
import pytest

from conftest import ROOT
from oracle.pyoracle import Oracle


def header_functions():
    txt = open(os.path.join(ROOT, "include", "fesomc_gpu.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|void|const char \*)\s*\*?\s*(fcg_\w+)\s*\(", txt, re.M)))


def test_library_exports_every_declared_symbol():
    from fesomc_amd import gpu
    lib = gpu.lib()
    names = header_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(names) == set(gpu.EXPORTS)


def test_create_without_gpu_fails_loudly():
    """No CPU fallback: on a host without a HIP device fcg_create reports FCG_E_NOGPU."""
    from conftest import have_gpu
    if have_gpu():
        pytest.skip("a GPU is present")
    from fesomc_amd import gpu, meshgen
    from fesomc_amd.runconfig import RunConfig
    txt, mesh, _ = meshgen.config("c1")
    with pytest.raises(gpu.GpuError, match="no HIP device"):
        gpu.Model(RunConfig(txt), mesh)


def _compare_mesh(txt, mesh):
    from fesomc_amd import gpu
    from fesomc_amd.runconfig import RunConfig
    o = Oracle(txt, mesh)
    p = gpu.mesh_arrays(RunConfig(txt), mesh)

    def oi(n):
        a = o.get_int(n)
        return np.where(a >= 0, a + 1, a)

    for k in ("elem_nodes", "edge_nodes", "edge_tri", "elem_edges", "elem_neighbors"):
        assert np.array_equal(p[k], oi(k)), k
    for k in ("depth", "sigma", "area", "elem_area", "elem_cos", "coriolis", "w_cv", "gradient_sca", "gradient_vec",
              "edge_dxdy", "edge_cross_dxdy"):
        assert np.array_equal(p[k], o.get(k)), k
    atp = o.get("a_tp").reshape(-1, 2)
    assert np.array_equal(p["a_tp"], np.concatenate([atp[:, 0], atp[:, 1]]))
    assert p["density_0"] == o.scalar("density_0") and p["scale_area"] == o.scalar("scale_area")
    if p["nobn"]:
        for k in ("ampt", "fazt"):
            assert np.array_equal(p[k], o.get(k).reshape(-1, 12).T.reshape(-1)), k


def test_mesh_build_bitexact_le(le_case):
    _compare_mesh(*le_case)


def test_mesh_build_bitexact_channel_tri_quad_ob():
    from fesomc_amd import meshgen
    txt, mesh, _ = meshgen.config("c1")
    _compare_mesh(txt, mesh)


def test_mesh_build_rejects_bad_quad():
    """A quad whose vertices do not form a circle is test_elem's 'Node list problem'."""
    from fesomc_amd import gpu, meshgen
    from fesomc_amd.runconfig import RunConfig
    txt, mesh, _ = meshgen.config("c1")
    el = mesh.elem.copy()
    q = np.nonzero(el[:, 0] != el[:, 3])[0][0]
    el[q] = el[q][[0, 2, 1, 3]]
    mesh.elem = el
    with pytest.raises(gpu.GpuError, match="mesh check failed"):
        gpu.mesh_arrays(RunConfig(txt), mesh)


def test_runconfig_reads_reference_file(le_case):
    from fesomc_amd.runconfig import RunConfig
    txt, _ = le_case
    c = RunConfig(txt)
    assert c.title == "LE" and c.nsigma == 40 and c.Mt == 4 and c.type_task == 3
    assert c.f("dt_2D") == 0.1 and c.f("dt") == 0.5 and not c.l("cartesian")
    base = RunConfig(txt.replace("Mt = 4", ""))
    assert base.Mt == 5   # int(0.5/0.1) in IEEE double
