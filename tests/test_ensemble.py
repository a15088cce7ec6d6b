"""Ensemble of uneven meshes with perturbed forcing in one context: every
member evolves exactly as it does alone (bit for bit in FP64), with its own
T/S means and diagnostics."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FIELDS_N = ("eta_n", "TF", "SF", "Wvel", "Kv", "snu")
FIELDS_E = ("U_n", "V_n", "U_n_filt", "mask_wd")


def _apply_forcing(model, mesh_nodes, mesh_elems, extras):
    from fesomc_amd import meshgen
    wx = np.concatenate([np.full(n, e["wind"][0]) for n, e in zip(mesh_nodes, extras)])
    wy = np.concatenate([np.full(n, e["wind"][1]) for n, e in zip(mesh_nodes, extras)])
    tau = [meshgen.wind_stress(*e["wind"]) for e in extras]
    tx = np.concatenate([np.full(n, t[0]) for n, t in zip(mesh_elems, tau)])
    ty = np.concatenate([np.full(n, t[1]) for n, t in zip(mesh_elems, tau)])
    model.set("windx", wx)
    model.set("windy", wy)
    model.set("taux", tx)
    model.set("tauy", ty)
    model.set_tracers(np.concatenate([e["T"] for e in extras]), np.concatenate([e["S"] for e in extras]))


def test_ensemble_members_match_standalone_runs():
    from fesomc_amd import gpu, meshgen
    from fesomc_amd.runconfig import RunConfig
    txt, meshes, extras = meshgen.ensemble("c4_small")
    cfg = RunConfig(txt)
    ens = gpu.Ensemble(cfg, meshes)
    _apply_forcing(ens, [m.nod2D for m in meshes], [m.elem2D for m in meshes], extras)
    ens.step(8)
    assert ens.diagnostics()["status"] == 0
    for b, (mesh, ex) in enumerate(zip(meshes, extras)):
        one = gpu.Model(cfg, mesh)
        _apply_forcing(one, [mesh.nod2D], [mesh.elem2D], [ex])
        one.step(8)
        for f in FIELDS_N:
            a = ens.split(ens.get(f), nodal=True)[b].reshape(-1)
            assert np.array_equal(a, one.get(f)), (b, f)
        for f in FIELDS_E:
            a = ens.split(ens.get(f), nodal=False)[b].reshape(-1)
            assert np.array_equal(a, one.get(f)), (b, f)
        d1, d2 = ens.diagnostics_member(b), one.diagnostics()
        assert abs(d1["energy"] - d2["energy"]) <= 1e-12 * abs(d2["energy"])
        assert (d1["T_min"], d1["T_max"], d1["n_wet_elem"]) == (d2["T_min"], d2["T_max"], d2["n_wet_elem"])
        one.close()
    ens.close()


def test_ensemble_mixed_precision_close_to_fp64():
    from fesomc_amd import gpu, meshgen
    from fesomc_amd.runconfig import RunConfig
    txt, meshes, extras = meshgen.ensemble("c4_small")
    cfg = RunConfig(txt)
    runs = []
    for prec in (gpu.FCG_FP64, gpu.FCG_MIXED):
        e = gpu.Ensemble(cfg, meshes, precision=prec)
        _apply_forcing(e, [m.nod2D for m in meshes], [m.elem2D for m in meshes], extras)
        e.step(20)
        runs.append(e)
    a, b = runs
    assert np.sqrt(np.mean((a.get("eta_n") - b.get("eta_n")) ** 2)) < 1e-4
    assert np.sqrt(np.mean((a.get("TF") - b.get("TF")) ** 2)) < 1e-3
    for e in runs:
        e.close()
