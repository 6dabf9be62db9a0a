"""The allocation-free sub-cell count entry point
(lbdem_psm_subcell_counts_static, static device scratch array) against the
allocating one, through the C ABI: bit-identical fractions, counts and
branches on two calls in a row, the second with fewer queries than the
first (its outputs must not carry the first call's data), and the
capacity limit."""
import ctypes

import numpy as np
import pytest

import lbdem_psm as L

pytestmark = pytest.mark.gpu


def _queries(rng, n):
    r = rng.uniform(2.0, 9.0, n)
    u = rng.normal(size=(n, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    d = u * (r + rng.uniform(-1.2, 1.2, n))[:, None]   # around the sphere surface
    return np.ascontiguousarray(d), np.ascontiguousarray(r)


def _static(lat, d, r):
    n = len(r)
    sf = np.full(n, -1.0)
    cnt = np.full(n, -1, dtype=np.int32)
    br = np.full(n, -1, dtype=np.int32)
    rc = lat.lib.lbdem_psm_subcell_counts_static(lat.h, n, L._ptr(d), L._ptr(r), L._ptr(sf), L._ptr(cnt),
                                                  L._ptr(br))
    assert rc == L.OK, lat.lib.lbdem_psm_last_error(lat.h)
    return sf, cnt, br


@pytest.mark.parametrize("N", [8, 5])
def test_static_subcell_counts_match_the_allocating_entry_point(N):
    rng = np.random.default_rng(100 + N)
    lat = L.Lattice(8, 8, 8, slices_per_dim=N)
    try:
        for n in (20000, 3001):
            d, r = _queries(rng, n)
            sf_a, cnt_a, br_a = lat.subcell_counts(d, r)
            sf_s, cnt_s, br_s = _static(lat, d, r)
            assert np.array_equal(sf_s, sf_a)
            assert np.array_equal(cnt_s, cnt_a)
            assert np.array_equal(br_s, br_a)
            assert (br_a == 2).sum() > n // 10      # counted cells are exercised
    finally:
        lat.close()


def test_static_subcell_counts_refuse_more_than_the_limit():
    lat = L.Lattice(8, 8, 8, slices_per_dim=8)
    try:
        n = 65537
        d = np.zeros((n, 3))
        r = np.ones(n)
        out = np.zeros(n)
        cnt = np.zeros(n, dtype=np.int32)
        rc = lat.lib.lbdem_psm_subcell_counts_static(lat.h, n, L._ptr(d), L._ptr(r), L._ptr(out), L._ptr(cnt),
                                                      L._ptr(cnt))
        assert rc == L.ERR_CAPACITY
        assert b"LBDEM_PSM_STATIC_MAX_QUERIES" in lat.lib.lbdem_psm_last_error(lat.h)
    finally:
        lat.close()
