"""River table of the host (fesom-c_amd/rivers.py, src/fv_rivers.F90):
list-directed reading, the record search, linear interpolation in time, the
ten-day ramp and the reference's end-of-table behaviour."""
import numpy as np
import pytest

from fesomc_amd.rivers import Rivers, binarysearch

T0 = 2450000.5


def _write(tmp_path, text):
    p = tmp_path / "rivers.dat"
    p.write_text(text)
    return str(p)


TABLE = f"""2
3
{T0 + 12.0}, {T0 + 32.0}
5 9
5
10.0 0.0 2*1.0d0
30.0, 40.0, 3.0
"""


def test_interpolates_linearly_between_records(tmp_path):
    """Between the records (t0+12, t0+32) the discharge is the straight line
    through them; the third river shares node 5 with the first."""
    r = Rivers(_write(tmp_path, TABLE), 12, T0 + 12.0, T0)
    assert np.array_equal(r.nodes, [4, 8, 4])
    assert np.array_equal(r.data, [[10.0, 0.0, 1.0], [30.0, 40.0, 3.0]])
    # rivers_ini starts with no discharge
    assert not r.per_node().any()
    q = r.do(T0 + 17.0)
    want = np.array([10.0 + 20.0 * 0.25, 40.0 * 0.25, 1.0 + 2.0 * 0.25])
    assert q[4] == pytest.approx(want[0] + want[2], rel=1e-12)
    assert q[8] == pytest.approx(want[1], rel=1e-12)
    assert np.count_nonzero(q) == 2


def test_first_ten_days_ramp(tmp_path):
    """Before time_jd0 + 10 the discharge is scaled by (time_jd - time_jd0)/10."""
    r = Rivers(_write(tmp_path, TABLE), 12, T0 + 12.0, T0 + 6.0)
    q = r.do(T0 + 14.0)
    assert q[8] == pytest.approx(40.0 * 0.1 * 0.8, rel=1e-12)


def test_past_the_last_record_holds_its_value(tmp_path):
    """Past the last record getcoeffld sets one_field: the discharge is the
    last record's and no longer changes."""
    r = Rivers(_write(tmp_path, TABLE), 12, T0 + 12.0, T0)
    r.do(T0 + 20.0)
    q = r.do(T0 + 40.0)
    assert q[8] == 40.0 and q[4] == 33.0 and r.one_field
    assert np.array_equal(r.do(T0 + 50.0), q)


def test_single_record_keeps_zero_discharge(tmp_path):
    """One record: one_field at start-up, so the discharge stays at the zero
    rivers_ini set (the reference's behaviour)."""
    r = Rivers(_write(tmp_path, f"1\n1\n{T0}\n3\n25.0\n"), 5, T0 + 1.0, T0)
    assert not r.do(T0 + 11.0).any()


def test_before_the_first_record_is_rejected(tmp_path):
    with pytest.raises(ValueError, match="before the first river record"):
        Rivers(_write(tmp_path, TABLE), 12, T0, T0)


def test_node_outside_mesh_is_rejected(tmp_path):
    with pytest.raises(ValueError, match="outside"):
        Rivers(_write(tmp_path, TABLE), 8, T0 + 12.0, T0)


def test_binarysearch_is_the_last_record_not_after_the_time():
    """Away from the tolerance the search returns the count of records at or
    before the time, np.searchsorted(side='right'); within 1e-9 of a record,
    that record."""
    rng = np.random.default_rng(3)
    for n in (1, 2, 3, 7, 16, 33):
        a = np.sort(rng.uniform(0.0, 100.0, n))
        for v in rng.uniform(-10.0, 110.0, 200):
            assert binarysearch(a, v) == np.searchsorted(a, v, side="right")
        for i in range(n):
            assert binarysearch(a, a[i] + 5e-10) == i + 1
