"""The rotation search that allocates nothing (dtc_rotation_search_noalloc,
fixed per-GPU work space) against dtc_rotation_search, through the C ABI: the
same kernels on the same inputs, so every output is bit-identical.  Two calls
in a row, the second with a smaller sweep, for both criteria; then the same
after the allocating entry point ran in between (the work space keeps the last
call's contents, which must not leak into the next result)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _tensor(seed):
    rng = np.random.default_rng(seed)
    A = 100.0 * rng.standard_normal((6, 6))
    return A @ A.T + np.diag([3000.0, 3100.0, 3300.0, 900.0, 950.0, 1000.0])


def _same(a, b):
    for k in ("alpha", "n", "aa", "EE_opt", "jj", "crit"):
        assert np.array_equal(np.asarray(a[k]), np.asarray(b[k])), k


@pytest.mark.parametrize("criterion", [1, 2])
def test_noalloc_rotation_search_equals_allocating_one(criterion):
    import dtc_amd
    EE1, EE2 = _tensor(1), _tensor(2)
    calls = [(EE1, 1.0), (EE2, 2.5)]          # second call: smaller sweep (fewer grid points)
    for EE, step in calls:
        a = dtc_amd.rotation_search(EE, criterion, step_deg=step, noalloc=True)
        b = dtc_amd.rotation_search(EE, criterion, step_deg=step)
        _same(a, b)
        assert a["crit"] >= 0.0 and np.isfinite(a["EE_opt"]).all()
    # interleaved: the allocating search between two no-alloc calls
    c = dtc_amd.rotation_search(EE2, criterion, step_deg=1.0)
    d = dtc_amd.rotation_search(EE2, criterion, step_deg=1.0, noalloc=True)
    _same(c, d)


def test_noalloc_rejects_a_sweep_beyond_the_work_space():
    import dtc_amd
    lib = dtc_amd._lib.load()
    mx = lib.dtc_rotation_noalloc_max_points()
    assert 8_000_000 < mx < 100_000_000
    with pytest.raises(dtc_amd.DTCError) as e:
        dtc_amd.rotation_search(_tensor(3), 1, step_deg=0.25, noalloc=True)
    assert e.value.rc == dtc_amd._lib.DTC_ERR_BAD_ARGUMENT and "work space" in str(e.value)
