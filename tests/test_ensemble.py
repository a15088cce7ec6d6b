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


@pytest.mark.parametrize("opts", [(), (("tracer_adv", "4"), ("filt_bi_2D", "T"), ("filt_bi_3D", "T"),
                                         ("tau_c", "20.0"), ("bih_2D", "T"), ("bih_3D", "T"), ("Abh0", "1.0e5"))],
                         ids=["muscl", "miura-filt_bi-bih"])
def test_ensemble_members_match_their_oracle_runs(opts):
    """Every member of the joined ensemble equals the FP64 oracle run of that
    member alone (its own mesh, tide amplitude, wind and T/S), bit for bit;
    also with the options that read each member's own scale_area (the Miura
    horizontal diffusion, the twice-applied filters, the biharmonic factors)."""
    import re
    from oracle.pyoracle import Oracle
    from fesomc_amd import gpu, meshgen
    from fesomc_amd.runconfig import RunConfig
    txt, meshes, extras = meshgen.ensemble("c4_small")
    for k, v in opts:
        txt, n = re.subn(rf"^(\s*{k}\s*=\s*)\S+", rf"\g<1>{v}", txt, count=1, flags=re.M)
        assert n == 1, k
    print("ensemble vs oracle:", opts, flush=True)
    ens = gpu.Ensemble(RunConfig(txt), meshes)
    _apply_forcing(ens, [m.nod2D for m in meshes], [m.elem2D for m in meshes], extras)
    ens.step(6)
    for b, (mesh, ex) in enumerate(zip(meshes, extras)):
        o = Oracle(txt, mesh)
        o.set_tracers(ex["T"].reshape(-1), ex["S"].reshape(-1))
        tx, ty = meshgen.wind_stress(*ex["wind"])
        o.set("windx", np.full(mesh.nod2D, ex["wind"][0]))
        o.set("windy", np.full(mesh.nod2D, ex["wind"][1]))
        o.set("taux", np.full(mesh.elem2D, tx))
        o.set("tauy", np.full(mesh.elem2D, ty))
        o.step(6)
        for f in FIELDS_N:
            assert np.array_equal(ens.split(ens.get(f), nodal=True)[b].reshape(-1), o.get(f)), (b, f)
        for f in FIELDS_E:
            assert np.array_equal(ens.split(ens.get(f), nodal=False)[b].reshape(-1), o.get(f)), (b, f)
    ens.close()


def test_level_classes_run_together_and_match_standalone_members():
    """The configuration's ensemble structure at small member sizes: 64
    channels of 2k..16k nodes with 16..32 sigma levels, one context per level
    count, stepped together (FP64).  Every class ends without a STOP; the
    smallest and the largest member, in their classes, equal standalone runs
    of the same member bit for bit."""
    from fesomc_amd import gpu, meshgen
    from fesomc_amd.runconfig import RunConfig
    from fesomc_amd.workload import EnsembleSet, apply_member_forcing
    classes = meshgen.ensemble_classes("c4_mini")
    assert sum(len(c[2]) for c in classes) == 64 and len(classes) >= 3
    assert {c[0] for c in classes} <= set(meshgen.NSIGMA_CLASSES)
    es = EnsembleSet(classes)
    es.step(4)
    assert es.diagnostics()["status"] == 0
    sizes = [(m.nod2D, ci, j) for ci, c in enumerate(classes) for j, m in enumerate(c[2])]
    for _, ci, j in (min(sizes), max(sizes)):
        ns, txt, meshes, extras, ids = classes[ci]
        one = gpu.Model(RunConfig(txt), meshes[j])
        apply_member_forcing(one, [meshes[j]], [extras[j]])
        one.step(4)
        part = es.parts[ci]
        for f in FIELDS_N:
            assert np.array_equal(part.split(part.get(f), nodal=True)[j].reshape(-1), one.get(f)), (ns, j, f)
        for f in FIELDS_E:
            assert np.array_equal(part.split(part.get(f), nodal=False)[j].reshape(-1), one.get(f)), (ns, j, f)
        one.close()
    es.close()


def test_level_classes_mixed_precision_with_learned_closure():
    """The headline path on the same structure: mixed precision and the
    learned closure in every class (the network trained on the 1M-node
    channel), against the FP64 run of the same ensemble after 20 steps."""
    import os
    from fesomc_amd import closure, gpu, meshgen
    from fesomc_amd.workload import EnsembleSet
    classes = meshgen.ensemble_classes("c4_mini")
    banks = [gpu.build_meshes(__import__("fesomc_amd.runconfig", fromlist=["RunConfig"]).RunConfig(c[1]), c[2])
             for c in classes]
    net = closure.load(os.path.join(os.path.dirname(gpu.__file__), "closure_c2.npz"))
    ref = EnsembleSet(classes, banks=banks)
    head = EnsembleSet(classes, precision=gpu.FCG_MIXED, closure=gpu.FCG_CLOSURE_MLP, banks=banks, net=net)
    ref.step(20)
    head.step(20)
    assert ref.diagnostics()["status"] == 0 and head.diagnostics()["status"] == 0
    for f, tol in (("eta_n", 1e-3), ("TF", 1e-2), ("SF", 1e-2)):
        d = head.get(f) - ref.get(f)
        assert np.sqrt(np.mean(d ** 2)) < tol, f
    assert head.diagnostics()["mlp_fallback"] < 20 * head.N
    ref.close()
    head.close()
    for b in banks:
        gpu.free_meshes(b)
    net.close()
