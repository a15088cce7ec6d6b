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
              "edge_dxdy", "edge_cross_dxdy", "ac"):
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


def test_mesh_build_sponge_layer():
    """SL_obn: ac_create (src/fv_utilit.F90:190-212) gives every node
    d/SL_radius, d its distance to the nearest open-boundary node, where d is
    below SL_radius, and 1 elsewhere.  Pinned by brute force in metres, by
    ac = 0 on the open boundary, and bit for bit between the mesh builder and
    the oracle."""
    from fesomc_amd import meshgen
    txt, mesh, _ = meshgen.config("c2_small")
    assert "SL_obn = T" in txt
    _compare_mesh(txt, mesh)
    from fesomc_amd import gpu
    from fesomc_amd.runconfig import RunConfig
    ac = gpu.mesh_arrays(RunConfig(txt), mesh)["ac"]
    ob = mesh.xy[mesh.index == 2]
    d = np.sqrt(((mesh.xy[:, None, :] - ob[None, :, :]) ** 2).sum(-1)).min(1)
    want = np.where(d < meshgen.SPONGE_M, d / meshgen.SPONGE_M, 1.0)
    np.testing.assert_allclose(ac, want, rtol=1e-12, atol=1e-12)
    assert np.all(ac[mesh.index == 2] == 0.0)
    assert 0 < (ac < 1.0).sum() < mesh.nod2D // 2
    # switched off: ac stays 1
    off = gpu.mesh_arrays(RunConfig(txt.replace("SL_obn = T", "SL_obn = F")), mesh)["ac"]
    assert np.all(off == 1.0)


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
