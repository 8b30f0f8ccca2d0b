"""CPU-only checks of the C ABI library and the host-side logic.

No compute call is made here (there is no GPU in the build container):
the library must load, export every symbol declared in include/svh.h, and
the host code that prepares device calls must agree with the oracle.
"""

import os
import re

import numpy as np
import pytest

from oracle import vecchia_np as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    txt = open(os.path.join(ROOT, "include", "svh.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|void|const char\*)\s+(svh_\w+)\s*\(", txt, re.M)))


def test_header_lists_entry_points():
    names = _declared()
    assert "svh_vecchia_eval" in names and "svh_maximin" in names
    assert len(names) >= 12


def test_library_loads_and_exports_every_declared_symbol():
    from pyscaledvecchia_amd import _lib
    L = _lib._load()
    for name in _declared():
        assert hasattr(L, name), name
    assert set(_lib.EXPORTED) == set(_declared())
    assert b"gfx950" in L.svh_version()


def test_device_count_is_an_int():
    from pyscaledvecchia_amd import _lib
    assert _lib.device_count() >= 0


def test_create_without_device_fails_cleanly():
    from pyscaledvecchia_amd import _lib
    if _lib.device_count() > 0:
        pytest.skip("a device is present")
    with pytest.raises(RuntimeError):
        _lib.Context(0)


@pytest.mark.parametrize("P", [1, 6, 10])
def test_chunk_ids_follow_reference_slicing(P):
    from pyscaledvecchia_amd.likelihood import _chunk_ids
    rng = np.random.default_rng(0)
    X = rng.random((3000, 3))
    nn = O.find_ordered_nn(X, 30)
    groups = O.nn_groups(nn, 30)
    ids, nchunk = _chunk_ids(groups, P)
    # oracle: enumerate chunks exactly as likelihood.py:111-114 does
    ref = []
    c = 0
    for rows, nb in groups:
        K = nb.shape[1] + 1
        for a0, a1 in O.chunks(len(rows), K, P):
            ref += [c] * (a1 - a0)
            c += 1
    np.testing.assert_array_equal(ids, ref)
    assert nchunk == c


@pytest.mark.parametrize("n,m,P", [(1, 0, 3), (2, 1, 3), (10, 3, 5), (3000, 20, 7),
                                   (5000, 30, 1), (50000, 30, 10), (100, 99, 12),
                                   (200000, 140, 12)])
def test_row_chunk_ids_match_group_chunk_ids(n, m, P):
    """The closed form used by the device rebuilds (chunk i below m, then
    m + (i - m) // step) equals the reference slicing of the k = min(i, m)
    groups (likelihood.py:111-114, _covariance.py:255-259)."""
    from pyscaledvecchia_amd.likelihood import _chunk_ids_rows
    rows = np.arange(n)
    ks = np.minimum(rows, m)
    groups = [(rows[ks == k], np.zeros((int((ks == k).sum()), k), np.int64))
              for k in np.unique(ks)]
    ref = []
    c = 0
    for r, nb in groups:
        K = nb.shape[1] + 1
        for a0, a1 in O.chunks(len(r), K, P):
            ref += [c] * (a1 - a0)
            c += 1
    ids, nchunk = _chunk_ids_rows(n, m, P)
    np.testing.assert_array_equal(ids, ref)
    assert nchunk == c


def test_nn_groups_matches_oracle():
    from pyscaledvecchia_amd.ordering import _nn_groups
    nn = O.find_ordered_nn(np.random.default_rng(1).random((80, 2)), 6)
    for (r1, b1), (r2, b2) in zip(_nn_groups(nn, 6), O.nn_groups(nn, 6)):
        np.testing.assert_array_equal(r1, r2)
        np.testing.assert_array_equal(b1, b2)


def test_estimator_validates_before_device_use():
    from pyscaledvecchia_amd import ScaledVecchiaGP
    with pytest.raises(ValueError):
        ScaledVecchiaGP().fit(np.random.rand(10, 2), np.random.rand(9))
    with pytest.raises(ValueError):
        ScaledVecchiaGP(trend="bogus").fit(np.random.rand(20, 2), np.random.rand(20))
    with pytest.raises(RuntimeError):
        ScaledVecchiaGP().predict(np.zeros((1, 2)))
    with pytest.raises(ValueError):
        ScaledVecchiaGP(nu=-1.0).fit(np.random.rand(20, 2), np.random.rand(20))
    with pytest.raises(ValueError):
        ScaledVecchiaGP(nu=None, warm_start="learned").fit(np.random.rand(20, 2),
                                                           np.random.rand(20))


@pytest.mark.parametrize("n,d,rows", [(1, 2, 8192), (7, 3, 2), (20000, 8, 8192), (5000, 1, 777)])
def test_maximin_first_equals_one_pass_expression(n, d, rows):
    """ordering.py:72 evaluated on row blocks equals the one-pass expression,
    including the first index among equal distances (duplicated rows)."""
    from pyscaledvecchia_amd.ordering import maximin_first
    rng = np.random.default_rng(n + d)
    X = rng.random((n, d)) / 0.37
    if n > 10:
        X[n // 2:] = X[:n - n // 2]          # every distance occurs twice
    ref = int(np.argmin(((X - X.mean(0)) ** 2).sum(1)))
    assert maximin_first(X, rows) == ref
