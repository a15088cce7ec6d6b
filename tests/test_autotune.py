"""Launch shapes: every tunable kernel gives the same results, bit for bit,
at every block size (each thread computes one (entity, layer) or one column
whatever the block), and the tuner keeps a candidate for each kernel."""
import numpy as np
import pytest


def test_set_launch_checks_its_arguments_without_a_gpu():
    from fesomc_amd import gpu
    L = gpu.lib()
    assert L.fcg_set_launch(None, b"k_velrhs", 64) != 0
    assert L.fcg_get_launch(None, b"k_velrhs") == 0


def _model(precision, nsigma=None):
    """The c2 physics on the small channel (8 levels, or nsigma)."""
    from fesomc_amd import gpu, meshgen
    from fesomc_amd.runconfig import RunConfig
    txt, mesh, ex = meshgen.config("c2_small")
    if nsigma is not None:
        txt = txt.replace("nsigma = 8\n", f"nsigma = {nsigma}\n")
        assert f"nsigma = {nsigma}\n" in txt
        ex["T"], ex["S"] = meshgen.stratified_ts(mesh, nsigma)
    m = gpu.Model(RunConfig(txt), mesh, precision=precision)
    m.set_tracers(ex["T"], ex["S"])
    wx, wy = ex["wind"]
    tx, ty = meshgen.wind_stress(wx, wy)
    m.set("windx", np.full(mesh.nod2D, wx))
    m.set("windy", np.full(mesh.nod2D, wy))
    m.set("taux", np.full(mesh.elem2D, tx))
    m.set("tauy", np.full(mesh.elem2D, ty))
    return m


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("block", [64, 128])
def test_block_size_does_not_change_results(precision, block):
    from fesomc_amd import autotune
    a, b = _model(precision), _model(precision)
    for k in autotune.LAYER_KERNELS + autotune.COLUMN_KERNELS:
        b.set_launch(k, block)
        assert b.get_launch(k) == block
    a.step(4)
    b.step(4)
    for f in ("eta_n", "U_n", "V_n", "TF", "SF", "Wvel", "Kv", "snu"):
        assert np.array_equal(a.get(f), b.get(f)), f
    a.close()
    b.close()


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("block", [512, 1024])
def test_wide_blocks_do_not_change_results(precision, block):
    """k_velrhs and the node-gather kernels (MUSCL tracer advection and its
    node gradients, the vertical-velocity divergence, the node averaging,
    also inside the captured barotropic loop) with 512- or 1024-thread blocks
    (more columns per block, so more neighbours read from LDS) give the
    256-thread results bit for bit."""
    from fesomc_amd import autotune, gpu
    a, b = _model(precision), _model(precision)
    a.step(1)
    b.step(1)   # the barotropic graph is captured with the default shapes, then re-captured
    for k in autotune.WIDE_KERNELS:
        try:
            b.set_launch(k, block)
        except gpu.GpuError as err:          # LDS of the block above the device limit
            assert "LDS" in str(err)
            continue
        assert b.get_launch(k) == block
    a.step(4)
    b.step(4)
    for f in ("eta_n", "U_n", "V_n", "U_rhsAB", "TF", "SF", "Wvel", "Unode", "Kv"):
        assert np.array_equal(a.get(f), b.get(f)), f
    a.close()
    b.close()


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("nsigma", [8, 16, 24, 32])
def test_level_vectors_do_not_change_results(precision, nsigma):
    """The level-vector kernels (each lane owns 4 (FP32) or 2 (FP64)
    consecutive levels and moves them with 16-byte accesses) give the
    one-level kernels' results bit for bit, at the ensemble's level counts;
    the column's unused last slot is left as it was."""
    a, b = _model(precision, nsigma), _model(precision, nsigma)
    W = 4 if precision else 2
    vec = ["k_velrhs", "k_tr_muscl", "k_el2nodes", "k_vertvel_div", "k_tr_gradn_muscl", "k_velgrad3d", "k_je",
           "k_update3d"]
    for k in vec:
        assert a.get_launch(k + ":levels") == W   # the default where nsigma is a multiple of W
        b.set_launch(k + ":levels", 1)
        assert b.get_launch(k + ":levels") == 1
    a.step(4)
    b.step(4)
    for f in ("eta_n", "U_n", "V_n", "U_rhsAB", "V_rhsAB", "TF", "SF", "Wvel", "Kv", "snu", "Unode", "Vnode"):
        assert np.array_equal(a.get(f), b.get(f)), f
    assert a.diagnostics()["status"] == 0 and b.diagnostics()["status"] == 0
    a.close()
    b.close()


def test_level_vector_option_checks_its_arguments():
    """Levels per lane other than 1 or the 16-byte vector width are refused
    before anything reaches a device."""
    from fesomc_amd import gpu
    L = gpu.lib()
    assert L.fcg_set_launch(None, b"k_velrhs:levels", 4) != 0
    assert L.fcg_get_launch(None, b"k_velrhs:levels") == 0


@pytest.mark.gpu
def test_tuner_picks_a_candidate_for_each_kernel():
    from fesomc_amd import autotune
    m = _model(1)
    with pytest.raises(RuntimeError):
        m.set_launch("k_velrhs", 96)
    with pytest.raises(RuntimeError):
        m.set_launch("k_bt_elem2", 64)
    with pytest.raises(RuntimeError):
        m.set_launch("k_tr_muscl", 128)
    kern = ["k_tr_ab", "k_jac_node", "k_velrhs", "k_tr_muscl", "k_el2nodes"]
    table = autotune.tune(m, kernels=kern, steps=1)
    assert set(table) == set(kern)
    for k, (bs, ms) in table.items():
        assert bs in autotune._cands(k, autotune.CANDIDATES) and ms > 0 and m.get_launch(k) == bs
    assert m.diagnostics()["status"] == 0
    m.close()


@pytest.mark.gpu
def test_whole_step_tuner_searches_levels_per_lane():
    """tune_set over a kernel with a block option and a level-vector variant
    and over one with only the latter: every chosen setting is a valid one and
    is left set."""
    from fesomc_amd import autotune
    m = _model(1, nsigma=16)
    W = m.get_launch("k_velrhs:levels")
    assert W == 4
    tables = autotune.tune_set(m, kernels=["k_velrhs", "k_je"], steps=1)
    t = tables[str(m.K)]
    assert t["k_velrhs:levels"][0] in (1, W) and t["k_je:levels"][0] in (1, W)
    for name, (v, ms) in t.items():
        assert ms > 0 and m.get_launch(name) == v
    if t["k_velrhs:levels"][0] == 1:
        assert t["k_velrhs"][0] in autotune.WIDE_CANDIDATES
    assert m.diagnostics()["status"] == 0
    m.close()
