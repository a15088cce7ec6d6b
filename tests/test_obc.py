"""Open-boundary T/S climatology (key_obc; src/fv_obc.F90): oracle pins, the
host reader, and the device against the oracle (bit for bit: the path has
only additions, multiplications and divisions).

Pins: a profile linear in depth is reproduced exactly at every layer's
mid-depth (the layer depths are the step's own Z, which the LE energy series
pins); a profile linear in time is blended linearly between the records;
below the deepest record the profile is constant.  Planted slip 17 (the
level depth for the layer mid-depth) must fail the first.
"""
import re

import numpy as np
import pytest

from oracle.pyoracle import Oracle, set_mutation


@pytest.fixture(autouse=True)
def _no_mutation():
    set_mutation(0)
    yield
    set_mutation(0)


def _oracle(steps=2):
    from fesomc_amd import meshgen
    txt, mesh, ex = meshgen.config("c2_small")
    o = Oracle(txt, mesh)
    o.set_tracers(ex["T"], ex["S"])
    o.step(steps)
    return o, txt, mesh, ex


def _linear_profiles(nobn, depth, t0, t1, c):
    """coef (nobn, ndep, 2) of T = c0 + c1 z + c2 (t - t0), S = c3 - c4 z."""
    nd = depth.size
    a = np.zeros((nobn, nd, 2))
    b = np.zeros((nobn, nd, 2))
    a[:, :, 0] = c[2]
    b[:, :, 0] = c[0] + c[1] * depth[None, :] - c[2] * t0
    b[:, :, 1] = c[3] - c[4] * depth[None, :]
    return a, b


def _check_linear(o, depth, c, t):
    N, L = o.size["N"], o.size["L"]
    ob = o.get_int("in_obn")
    a, b = _linear_profiles(ob.size, depth, 0.0, 0.0, c)
    o.obc_apply(depth, a, b, t)
    Z = o.get("Z").reshape(N, L)[ob]
    T = o.get("Tclim").reshape(N, L)[ob]
    S = o.get("Sclim").reshape(N, L)[ob]
    assert np.allclose(T, c[0] + c[1] * Z + c[2] * t, rtol=1e-12, atol=1e-9)
    assert np.allclose(S, c[3] - c[4] * Z, rtol=1e-12, atol=1e-12)


def test_profile_linear_in_depth_is_reproduced_at_the_layer_depths():
    o, *_ = _oracle()
    depth = np.array([0.0, 2.0, 5.0, 11.0, 20.0, 40.0])
    c = (10.0, -0.1, 0.0, 30.0, -0.05)
    _check_linear(o, depth, c, o.scalar("time_jd"))
    set_mutation(17)
    try:
        with pytest.raises(AssertionError):
            _check_linear(o, depth, c, o.scalar("time_jd"))
    finally:
        set_mutation(0)


def test_profile_linear_in_time_and_constant_below_the_last_depth():
    o, *_ = _oracle()
    N, L = o.size["N"], o.size["L"]
    ob = o.get_int("in_obn")
    t = o.scalar("time_jd")
    depth = np.array([0.0, 3.0])              # shallower than most OB columns
    c = (12.0, -0.2, 0.5, 31.0, 0.1)
    a, b = _linear_profiles(ob.size, depth, 0.0, 0.0, c)
    o.obc_apply(depth, a, b, t)
    Z = o.get("Z").reshape(N, L)[ob]
    T = o.get("Tclim").reshape(N, L)[ob]
    want = np.where(Z >= 3.0, c[0] + c[1] * 3.0 + c[2] * t, c[0] + c[1] * Z + c[2] * t)
    assert (Z > 3.0).any() and np.allclose(T, want, rtol=1e-12, atol=1e-9)


def test_layer_above_the_first_depth_is_refused():
    o, *_ = _oracle()
    ob = o.get_int("in_obn")
    depth = np.array([1e3, 2e3])
    a, b = _linear_profiles(ob.size, depth, 0.0, 0.0, (10.0, 0.0, 0.0, 30.0, 0.0))
    with pytest.raises(ValueError):
        o.obc_apply(depth, a, b, o.scalar("time_jd"))


def test_climatology_reader_records(tmp_path):
    """nc node count checked against the mesh; records bracketed as the
    reference's getcoeffld; past the last record the profiles freeze at the
    time of their last interpolation (one_field)."""
    from fesomc_amd import obc, sbc
    in_obn = np.array([3, 7, 9])
    depth = np.array([0.0, 10.0, 30.0])
    time = np.array([0.0, 1.0, 2.0])
    T = np.stack([np.full((3, 3), 10.0 + i) for i in range(3)])
    S = np.stack([np.full((3, 3), 30.0 - i) for i in range(3)])
    obc.write_climatology(str(tmp_path), in_obn, depth, time, T, S, secstep=1.0, iyear=2000)
    with pytest.raises(ValueError):
        obc.ObcClimatology(str(tmp_path), 4)
    c = obc.ObcClimatology(str(tmp_path), 3, in_obn=in_obn)
    assert not c.node_mismatch
    t0 = sbc.julday(2000, 1, 1) - 0.5
    d, a, b = c.update(t0 + 0.5)
    assert (c.t_indx, c.t_indx_p1) == (1, 2)
    assert np.allclose((t0 + 0.5) * a[..., 0] + b[..., 0], 10.5, atol=1e-6)
    assert c.update(t0 + 0.9) is None and c.eval_time(t0 + 0.9) == t0 + 0.9
    assert c.update(t0 + 1.5) is not None and (c.t_indx, c.t_indx_p1) == (2, 3)
    assert c.update(t0 + 2.5) is not None and c.one_field
    assert c.update(t0 + 3.0) is None and c.eval_time(t0 + 3.0) == t0 + 2.5
    with pytest.raises(ValueError):
        obc.ObcClimatology(str(tmp_path), 3).update(t0 - 1.0)


@pytest.mark.gpu
def test_obc_climatology_matches_oracle_bit_for_bit(tmp_path):
    """The climatology file read by obc.ObcClimatology, the device's
    Tclim/Sclim after obc_do against the oracle's apply_obc, and four steps of
    both with the open-boundary relaxation towards them: bit for bit."""
    from fesomc_amd import gpu, meshgen, obc, sbc
    from fesomc_amd.runconfig import RunConfig
    txt, mesh, ex = meshgen.config("c2_small")
    g = gpu.Model(RunConfig(txt), mesh)
    g.set_tracers(ex["T"], ex["S"])
    o = Oracle(txt, mesh)
    o.set_tracers(ex["T"], ex["S"])
    g.step(2)
    o.step(2)
    ob = o.get_int("in_obn")
    t_start = o.scalar("time_jd")
    t0 = sbc.julday(2000, 1, 1) - 0.5
    depth = np.array([0.0, 2.5, 6.0, 15.0, 40.0])
    times = np.array([t_start - t0 - 0.001, t_start - t0 + 0.2])
    rng = np.random.default_rng(2)
    T = 9.0 + rng.uniform(0, 3, (2, depth.size, ob.size))
    S = 30.0 + rng.uniform(0, 1, (2, depth.size, ob.size))
    obc.write_climatology(str(tmp_path), ob + 1, depth, times, T, S, secstep=1.0, iyear=2000)
    clim = obc.ObcClimatology(str(tmp_path), ob.size, in_obn=ob + 1)
    g.attach_obc(clim)
    coef = clim._coefficients(t_start)
    for _ in range(4):
        g.step(1)
        o.phase(0)
        o.obc_apply(*coef, o.scalar("time_jd"))
        for ph in range(1, 9):
            o.phase(ph)
        for f in ("Tclim", "Sclim", "TF", "SF", "eta_n"):
            assert np.array_equal(g.get(f), o.get(f)), f
    assert g.diagnostics()["status"] == 0
    g.close()
