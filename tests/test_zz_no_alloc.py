"""The allocation-free HLLC entry (a2d_hllc_flux_noalloc) against the
allocating one (a2d_hllc_flux): the same kernel on the same faces gives the
same bits, on two calls in a row (the second with fewer faces, over the
array's leftover contents), and sizes beyond the fixed array are refused."""
import numpy as np
import pytest

from aero2d_amd import _abi

pytestmark = pytest.mark.gpu


def _faces(n, seed):
    rng = np.random.default_rng(seed)
    WL = np.ascontiguousarray(np.stack([rng.uniform(0.3, 2.0, n), rng.normal(0, 0.8, n),
                                        rng.normal(0, 0.8, n), rng.uniform(0.2, 1.5, n)]))
    WR = np.ascontiguousarray(np.stack([rng.uniform(0.3, 2.0, n), rng.normal(0, 0.8, n),
                                        rng.normal(0, 0.8, n), rng.uniform(0.2, 1.5, n)]))
    S = np.ascontiguousarray(rng.normal(size=(2, n)))
    return WL, WR, S


def _flux(name, WL, WR, S):
    n = WL.shape[1]
    F = np.full((4, n), np.nan)
    _abi.lib()
    _abi.call(name, n, _abi.dptr(WL), _abi.dptr(WR), _abi.dptr(S), _abi.dptr(F))
    return F


def test_noalloc_matches_allocating_entry_twice_in_a_row():
    for n, seed in ((65536, 1), (1000, 2)):
        WL, WR, S = _faces(n, seed)
        a = _flux("a2d_hllc_flux", WL, WR, S)
        b = _flux("a2d_hllc_flux_noalloc", WL, WR, S)
        assert np.all(np.isfinite(b))
        assert np.array_equal(a, b), n


def test_noalloc_refuses_sizes_beyond_its_array():
    WL, WR, S = _faces(65537, 3)
    with pytest.raises(_abi.HipError) as e:
        _flux("a2d_hllc_flux_noalloc", WL, WR, S)
    assert e.value.code == -6
    WL, WR, S = _faces(16, 4)
    assert np.array_equal(_flux("a2d_hllc_flux_noalloc", WL, WR, S), _flux("a2d_hllc_flux", WL, WR, S))
