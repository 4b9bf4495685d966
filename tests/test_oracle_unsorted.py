"""CPU pins of the oracle's unsorted scans (data_sorted 0: do_pair_unsorted, corr.cpp:36-52, which
ends no scan; corels1 starting every image's scan at j = i+1, corr.cpp:167-182).

Pinned by the brute-force enumeration of tests/bruteforce.py (every (outer, inner, image) triple,
no sweep): the cross pass with a one-sided pi window, where the sorted scan of the PBz = +1 image
stops at its first galaxy and the unsorted one does not; the auto pass over a catalog in random
order, box and sky; and, on z-sorted input with a two-sided window, the unsorted auto pass of the
sky equal to the sorted one.
"""
import numpy as np

import bruteforce
from corr_pc_amd import synth
from corr_pc_amd.engine import KIND_DD, KIND_SHAPE_SHAPE, make_params, make_sky_params
from helpers import assert_hist_match, sky_catalogs, small_catalog


def _shuffled(c, seed):
    return c.subset(np.random.default_rng(seed).permutation(c.n))


def test_unsorted_cross_one_sided_window(oracle_lib):
    L = 30.0
    A = small_catalog(300, L, 84)
    B = small_catalog(250, L, 85)
    p = make_params(L, 5, 0.5, 8.0, 6, 2.0, 12.0, kind=KIND_SHAPE_SHAPE, estimator=1, g_with_shape=1, do_jk=1, n_jk=8)
    sorted_h, _ = oracle_lib.count(A, B, p)
    p.unsorted = 1
    h, st = oracle_lib.count(_shuffled(A, 1), _shuffled(B, 2), p)
    bf = bruteforce.count(A, B, p)
    assert_hist_match(h, bf, rtol=1e-12, what="unsorted cross vs brute force")
    # the PBz = +1 image's sorted scan stopped at its first galaxy for the outer galaxies near the top
    assert st["n_binned"] > sorted_h.num[0].sum()


def test_unsorted_auto_box(oracle_lib):
    L = 40.0
    A = _shuffled(small_catalog(300, L, 86, clumps=10, inc=True), 3)
    p = make_params(L, 6, 0.3, 9.0, 8, -11.0, 11.0, kind=KIND_SHAPE_SHAPE, g_with_shape=1, do_jk=1, n_jk=8,
                    include_prob_max=0.8)
    p.auto_pairs, p.unsorted = 1, 1
    h, st = oracle_lib.count(A, A, p)
    bf = bruteforce.count(A, A, p)
    assert_hist_match(h, bf, rtol=1e-12, what="unsorted auto vs brute force")
    # every unordered pair under each of its images once: with symmetric pi bins, cross = auto + mirror
    p.auto_pairs = 0
    cross, _ = oracle_lib.count(A, A, p)
    assert np.array_equal(cross.num[0], h.num[0] + h.num[0][:, ::-1])


def test_unsorted_auto_sky(oracle_lib):
    table = synth.distance_table(0.3, 1e-4, 1.0)
    A, _, n_jk, _ = sky_catalogs(700, 10, 87, table=table)
    pt = synth.sky_params_table(2, table)
    p = make_sky_params(pt, 6, 1.0, 40.0, 8, -40.0, 40.0, kind=KIND_DD, do_jk=1, n_jk=n_jk)
    p.auto_pairs = 1
    sorted_h, _ = oracle_lib.count(A, A, p)
    p.unsorted = 1
    same, _ = oracle_lib.count(A, A, p)
    assert np.array_equal(same.num, sorted_h.num)  # z order is catalog order, no scan stops at once
    S = _shuffled(A, 4)
    h, st = oracle_lib.count(S, S, p)
    bf = bruteforce.count_sky(S, S, p)
    assert_hist_match(h, bf, rtol=1e-12, what="unsorted sky auto vs brute force")
    assert st["n_binned"] > 500
