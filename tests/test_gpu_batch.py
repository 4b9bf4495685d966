"""GPU tests of cpc_count_batch: many passes in one call give, job by job, what cpc_count gives for
each alone (bit-identical counts, weighted sums within helpers.RTOL) and what the oracle gives.

The reference has no batched entry point (it runs its passes one after another, do_corr.cpp:63-131);
the batch must be a pure scheduling change, so single-pass results are its specification.
"""
import numpy as np
import pytest

from corr_pc_amd import synth
from corr_pc_amd.engine import (JK_LEAVE_ONE_OUT, JK_TOUCH, KIND_DD, KIND_SHAPE_DENSITY, KIND_SHAPE_SHAPE, PREC_FP64,
                                PREC_SPLIT, EngineError, make_params, make_sky_params)
from helpers import assert_hist_match, sky_catalogs, small_catalog

pytestmark = pytest.mark.gpu


def _mixed_jobs():
    """Box passes of different sizes, boxes, binnings, kinds, precisions and jackknife settings,
    an auto pass and a sky pass."""
    jobs = []
    L = 60.0
    A = small_catalog(3000, L, 201, clumps=40)
    B = small_catalog(2500, L, 202, clumps=30)
    C = small_catalog(400, L, 203)
    jobs.append((A, B, make_params(L, 8, 0.2, 12.0, 10, -15.0, 15.0, kind=KIND_SHAPE_SHAPE, g_with_shape=2,
                                   do_jk=1, n_jk=8, precision=PREC_SPLIT)))
    jobs.append((A, A, make_params(L, 8, 0.2, 12.0, 10, -15.0, 15.0, kind=KIND_DD, do_jk=1, n_jk=8,
                                   precision=PREC_SPLIT)))
    jobs.append((B, A, make_params(L, 6, 0.3, 10.0, 12, -12.0, 12.0, kind=KIND_SHAPE_DENSITY, g_with_shape=1,
                                   do_jk=1, n_jk=8, jk_mode=JK_TOUCH, precision=PREC_FP64)))
    jobs.append((C, A, make_params(L, 5, 0.0, 9.0, 6, -9.0, 9.0, kind=KIND_DD, lin_bin=1, do_jk=0,
                                   precision=PREC_SPLIT)))
    pa = make_params(L, 8, 0.2, 12.0, 10, -15.0, 15.0, kind=KIND_DD, do_jk=1, n_jk=8, precision=PREC_SPLIT)
    pa.auto_pairs = 1
    jobs.append((B, B, pa))
    L2 = 35.0
    D = small_catalog(1500, L2, 204, jk_side=3, clumps=10)
    jobs.append((D, D, make_params(L2, 6, 0.3, 8.0, 8, -10.0, 10.0, kind=KIND_SHAPE_SHAPE, do_jk=1, n_jk=27,
                                   precision=PREC_SPLIT, jk_mode=JK_LEAVE_ONE_OUT)))
    E = small_catalog(0, L2, 205)
    jobs.append((E, D, make_params(L2, 6, 0.3, 8.0, 8, -10.0, 10.0, kind=KIND_DD, do_jk=1, n_jk=27)))
    table = synth.distance_table(0.3, 1e-4, 1.0)
    pt = synth.sky_params_table(2, table)
    S, T, n_jk, _ = sky_catalogs(3000, 4000, 206, shapes_in=True, table=table)
    jobs.append((S, T, make_sky_params(pt, 20, 0.1, 100.0, 40, -100.0, 100.0, kind=KIND_SHAPE_DENSITY,
                                       g_with_shape=2, do_jk=1, n_jk=n_jk, precision=PREC_SPLIT)))
    return jobs


def test_batch_matches_single_passes_and_oracle(engine, oracle_lib):
    jobs = _mixed_jobs()
    print(f"[case] mixed batch of {len(jobs)} jobs", flush=True)
    results, tot = engine.count_batch(jobs)
    assert len(results) == len(jobs)
    nb = 0
    for k, ((outer, inner, p), (h, st)) in enumerate(zip(jobs, results)):
        single, sst = engine.count(outer, inner, p)
        assert_hist_match(h, single, what=f"job {k} batch vs single")
        ref, rst = oracle_lib.count(outer, inner, p)
        assert_hist_match(h, ref, what=f"job {k} batch vs oracle")
        assert st["n_binned"] == sst["n_binned"] == rst["n_binned"]
        assert st["n_candidates"] == sst["n_candidates"]
        nb += st["n_binned"]
    assert tot["n_binned"] == nb > 0
    assert tot["ms_kernel"] > 0


def test_batch_many_mocks_under_both_tunings(engine):
    """A small version of the batched-mocks workload: boxes of log-uniform sizes, each with its
    randoms (DD, DR, RR), every result equal to its single pass."""
    r = np.random.default_rng(207)
    L = 130.0
    mocks = []
    for m in range(12):
        n = int(np.exp(r.uniform(np.log(300), np.log(6000))))
        D = small_catalog(n, L, 300 + m, shapes=False, clumps=max(1, n // 60))
        R = small_catalog(2 * n, L, 400 + m, shapes=False, weights=False)
        mocks.append((D, R))
    p = make_params(L, 10, 0.1, 30.0, 20, -60.0, 60.0, kind=KIND_DD, do_jk=1, n_jk=8, precision=PREC_SPLIT)
    jobs = []
    for D, R in mocks:
        jobs += [(D, D, p), (D, R, p), (R, R, p)]
    for tun in ({}, {"col_frac": 0.5, "item_size": 200, "block_threads": 256}):
        print(f"[case] {len(jobs)} mock passes tuning={tun}", flush=True)
        engine.set_tuning(**tun)
        try:
            results, tot = engine.count_batch(jobs)
            for k, ((outer, inner, pp), (h, st)) in enumerate(zip(jobs, results)):
                single, sst = engine.count(outer, inner, pp)
                assert_hist_match(h, single, what=f"mock pass {k} tuning={tun}")
        finally:
            engine.set_tuning()
        assert tot["n_binned"] == sum(st["n_binned"] for _, st in results) > 0


def test_batch_rejects_bad_job(engine):
    L = 30.0
    A = small_catalog(500, L, 208)
    good = make_params(L, 5, 0.5, 8.0, 6, -10.0, 10.0, kind=KIND_DD, do_jk=1, n_jk=8)
    bad = make_params(L, 5, 0.5, 8.0, 6, -10.0, 10.0, kind=KIND_DD, do_jk=1, n_jk=8)
    bad.n_jk = 2  # jk regions of the catalog reach 7
    with pytest.raises(EngineError, match="job 1"):
        engine.count_batch([(A, A, good), (A, A, bad)])
    # the context stays usable
    res, _ = engine.count_batch([(A, A, good)])
    single, _ = engine.count(A, A, good)
    assert_hist_match(res[0][0], single, what="after a rejected batch")


def test_batch_empty(engine):
    res, tot = engine.count_batch([])
    assert res == [] and tot["n_binned"] == 0


def test_autotune_keeps_valid_results(engine):
    L = 130.0
    jobs = []
    for m in range(4):
        D = small_catalog(1500 * (m + 1), L, 500 + m, shapes=False, clumps=20)
        R = small_catalog(3000 * (m + 1), L, 600 + m, shapes=False, weights=False)
        p = make_params(L, 10, 0.1, 30.0, 20, -60.0, 60.0, kind=KIND_DD, do_jk=1, n_jk=8, precision=PREC_SPLIT)
        jobs += [(D, D, p), (D, R, p)]
    print(f"[case] autotune over {len(jobs)} passes", flush=True)
    try:
        tun, tried, results = engine.autotune(jobs)
        assert tried >= 6
        assert tun["col_frac"] in (0.0, 0.35, 0.5) and tun["item_size"] in (0, 512, 4096)
        assert tun["block_threads"] in (0, 256)
        for k, ((outer, inner, pp), (h, st)) in enumerate(zip(jobs, results)):
            single, _ = engine.count(outer, inner, pp)
            assert_hist_match(h, single, what=f"autotuned pass {k}")
    finally:
        engine.set_tuning()


def test_config3_mocks_batch_against_oracle(engine, oracle_lib):
    """The batched-mocks workload's shape at sizes the oracle checks in seconds: 16 independent
    periodic boxes of side 250 with data sizes log-uniform in [1e4, 3e4] (clustered like the wg+ box),
    2x randoms, 2x2x2 jackknife, 10 log rp bins in [0.1, 30) and 20 pi bins in [-60, 60); each mock's
    SD, SRd and RsRd passes (do_corr.cpp:110-131), all 48 in one cpc_count_batch call with the
    autotuned decomposition, every job against the oracle."""
    from corr_pc_amd.passes import config_datainfo, setup_pass
    sizes = synth.mock_sizes(16, n_min=10_000, n_max=30_000)
    di = config_datainfo(synth.MOCK_BINNING)
    jobs = []
    for k, n in enumerate(sizes):
        c = synth.mock_catalogs(k, int(n))
        for name, g1, g2 in (("SD", c["D"], c["S"]), ("SRd", c["Rd"], c["S"]), ("RsRd", c["Rd"], c["Rs"])):
            outer, inner, p, _ = setup_pass(name, g1, g2, di, precision=PREC_SPLIT)
            jobs.append((outer, inner, p))
    print(f"[case] config-3 mocks: {len(sizes)} boxes, {len(jobs)} passes, data sizes {sizes.min()}..{sizes.max()}",
          flush=True)
    try:
        tun, tried, _ = engine.autotune(jobs)
        print(f"[case] autotuned: {tun} after {tried} batches", flush=True)
        results, tot = engine.count_batch(jobs)
    finally:
        engine.set_tuning()
    nb = 0
    for k, ((outer, inner, p), (h, st)) in enumerate(zip(jobs, results)):
        ref, rst = oracle_lib.count(outer, inner, p)
        assert_hist_match(h, ref, what=f"mock {k // 3} pass {('SD', 'SRd', 'RsRd')[k % 3]} batch vs oracle")
        assert st["n_binned"] == rst["n_binned"] > 0
        nb += st["n_binned"]
    assert tot["n_binned"] == nb
