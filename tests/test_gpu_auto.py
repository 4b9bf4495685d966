"""GPU parity of the auto pass (corels1, corr.cpp:103-216): the HIP engine through the C ABI against
the CPU oracle on the same z-sorted catalogs, box and sky, both kernel precisions.

Counts must be bit-identical; weighted sums within helpers.RTOL.
"""
import numpy as np
import pytest

from corr_pc_amd import synth
from corr_pc_amd.engine import (JK_LEAVE_ONE_OUT, JK_TOUCH, KIND_DD, KIND_SHAPE_SHAPE, PREC_FP64, PREC_SPLIT,
                                EngineError, make_params, make_sky_params)
from helpers import assert_hist_match, sky_catalogs, small_catalog

pytestmark = pytest.mark.gpu

PRECS = [PREC_SPLIT, PREC_FP64]
TUNINGS = [{}, {"col_frac": 0.5, "item_size": 200, "block_threads": 256}]


def _auto(p):
    p.auto_pairs = 1
    return p


def _run(engine, oracle_lib, cat, p, what):
    ref, rst = oracle_lib.count(cat, cat, p)
    for tun in TUNINGS:
        print(f"[case] {what} tuning={tun}", flush=True)
        engine.set_tuning(**tun)
        try:
            got, st = engine.count(cat, cat, p)
        finally:
            engine.set_tuning()
        assert_hist_match(got, ref, what=f"{what} tuning={tun}")
        assert st["n_binned"] == rst["n_binned"] > 0


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("kind,est,jk_mode", [(KIND_DD, 0, JK_LEAVE_ONE_OUT), (KIND_SHAPE_SHAPE, 0, JK_TOUCH),
                                              (KIND_SHAPE_SHAPE, 1, JK_LEAVE_ONE_OUT)])
def test_auto_box(engine, oracle_lib, prec, kind, est, jk_mode):
    L = 60.0
    A = small_catalog(4000, L, 91, clumps=40, inc=True)
    p = _auto(make_params(L, 8, 0.2, 12.0, 10, -15.0, 15.0, kind=kind, g_with_shape=1, estimator=est, do_jk=1,
                          n_jk=8, include_prob_max=0.8, jk_mode=jk_mode, precision=prec))
    _run(engine, oracle_lib, A, p, f"auto box kind={kind} est={est} jk={jk_mode} prec={prec}")


@pytest.mark.parametrize("prec", PRECS)
def test_auto_box_one_sided_window(engine, oracle_lib, prec):
    """p_min > 0: the scans from j = i+1 stop at once; only PBz = +1 pairs remain."""
    L = 30.0
    A = small_catalog(3000, L, 92)
    p = _auto(make_params(L, 5, 0.5, 8.0, 6, 2.0, 12.0, kind=KIND_DD, do_jk=1, n_jk=8, precision=prec))
    _run(engine, oracle_lib, A, p, f"auto box one-sided prec={prec}")


@pytest.mark.parametrize("prec", PRECS)
def test_auto_sky(engine, oracle_lib, prec):
    table = synth.distance_table(0.3, 1e-4, 1.0)
    pt = synth.sky_params_table(2, table)
    A, _, n_jk, _ = sky_catalogs(6000, 10, 19, table=table)
    p = _auto(make_sky_params(pt, 20, 0.1, 100.0, 40, -100.0, 100.0, kind=KIND_SHAPE_SHAPE, g_with_shape=1,
                              do_jk=1, n_jk=n_jk, precision=prec))
    _run(engine, oracle_lib, A, p, f"auto sky prec={prec}")


def test_auto_cross_mirror_on_device(engine):
    """On the device alone, at a size the oracle would take long on: the cross pass of a catalog with
    itself equals its auto pass plus the pi mirror (symmetric pi bins, unit weights)."""
    L = 250.0
    A = small_catalog(60000, L, 93, clumps=300, weights=False, shapes=False)
    p = make_params(L, 10, 0.1, 30.0, 20, -60.0, 60.0, kind=KIND_DD)
    cross, _ = engine.count(A, A, p)
    auto, st = engine.count(A, A, _auto(make_params(L, 10, 0.1, 30.0, 20, -60.0, 60.0, kind=KIND_DD)))
    assert st["n_binned"] > 1e6
    assert np.array_equal(cross.num[0], auto.num[0] + auto.num[0][:, ::-1])


def test_auto_rejects_two_catalogs(engine):
    A = small_catalog(100, 20.0, 94)
    B = small_catalog(100, 20.0, 95)
    with pytest.raises(EngineError):
        engine.count(A, B, _auto(make_params(20.0, 4, 0.5, 4.0, 4, -5.0, 5.0)))
