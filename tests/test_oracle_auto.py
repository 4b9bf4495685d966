"""CPU tests pinning the oracle's auto pass (corels1, corr.cpp:103-216).

  * the oracle restates corels1's scans (start at j = i+1, or j = 0 for PBz > 0; stop at the first
    included galaxy outside the z window); tests/bruteforce.py states the same pairs as predicates
    (j after i for PBz <= 0, and whether the scan's first included galaxy is below the window);
  * the invariant that ties the auto pass to the cross pass over the same catalog: corels2(D, D)
    counts each unordered pair twice with opposite pi, corels1 once, so with symmetric pi bins the
    cross counts are the auto counts plus their pi mirror;
  * a one-sided pi window (p_min > 0), where the scan from j = i+1 stops at once and the auto pass
    finds only the pairs reached through PBz = +1.
"""
import numpy as np
import pytest

import bruteforce
from corr_pc_amd.engine import JK_TOUCH, KIND_DD, KIND_SHAPE_SHAPE, make_params, make_sky_params
from corr_pc_amd import synth
from helpers import assert_hist_match, sky_catalogs, small_catalog


def _auto(p):
    p.auto_pairs = 1
    return p


@pytest.mark.parametrize("kind,est,jk_mode", [(KIND_DD, 0, 0), (KIND_SHAPE_SHAPE, 0, 0), (KIND_SHAPE_SHAPE, 1, JK_TOUCH)])
def test_oracle_auto_vs_bruteforce(oracle_lib, kind, est, jk_mode):
    L = 40.0
    A = small_catalog(300, L, 81, clumps=10, inc=True)
    p = _auto(make_params(L, 6, 0.3, 9.0, 8, -11.0, 11.0, kind=kind, g_with_shape=1, estimator=est, do_jk=1,
                          n_jk=8, include_prob_max=0.8, jk_mode=jk_mode))
    h, st = oracle_lib.count(A, A, p)
    bf = bruteforce.count(A, A, p)
    assert_hist_match(h, bf, rtol=1e-12, what="auto oracle vs brute force")
    assert st["n_binned"] == bf["num"][0].sum() > 200


def test_oracle_auto_one_sided_window(oracle_lib):
    L = 30.0
    A = small_catalog(260, L, 82)
    p = _auto(make_params(L, 5, 0.5, 8.0, 6, 2.0, 12.0, kind=KIND_DD, do_jk=1, n_jk=8))
    h, _ = oracle_lib.count(A, A, p)
    bf = bruteforce.count(A, A, p)
    assert_hist_match(h, bf, rtol=1e-12, what="auto oracle vs brute force (p_min > 0)")
    # only the PBz = +1 scans contribute: pairs whose pi is measured through the upper image
    p.auto_pairs = 0
    full, _ = oracle_lib.count(A, A, p)
    assert 0 < h.num[0].sum() < full.num[0].sum() / 4


def test_cross_equals_auto_plus_mirror(oracle_lib):
    L = 50.0
    A = small_catalog(700, L, 83, clumps=20, weights=False)
    p = make_params(L, 6, 0.3, 10.0, 10, -12.0, 12.0, kind=KIND_DD)
    cross, _ = oracle_lib.count(A, A, p)
    auto, _ = oracle_lib.count(A, A, _auto(make_params(L, 6, 0.3, 10.0, 10, -12.0, 12.0, kind=KIND_DD)))
    assert auto.num[0].sum() > 1000
    assert np.array_equal(cross.num[0], auto.num[0] + auto.num[0][:, ::-1])


def test_oracle_auto_sky_vs_bruteforce(oracle_lib):
    table = synth.distance_table(0.3, 1e-4, 1.0)
    pt = synth.sky_params_table(2, table)
    A, _, n_jk, _ = sky_catalogs(300, 10, 9, table=table)
    p = _auto(make_sky_params(pt, 8, 0.5, 60.0, 10, -60.0, 60.0, kind=KIND_SHAPE_SHAPE, g_with_shape=1, do_jk=1,
                              n_jk=n_jk))
    h, st = oracle_lib.count(A, A, p)
    bf = bruteforce.count_sky(A, A, p)
    assert_hist_match(h, bf, rtol=1e-12, what="sky auto oracle vs brute force")
    assert st["n_binned"] > 100
