"""GPU parity of the unsorted scans (cpc_params.unsorted, data_sorted 0): the HIP engine through
the C ABI against the CPU oracle, on catalogs in random order, both kernel precisions: the cross
pass with a one-sided pi window (where the sorted scan of the PBz = +1 image would stop at its first
galaxy), and the auto pass (corels1: j > i in catalog order under every image) in a box and on the
sky.  Counts bit-identical, weighted sums within helpers.RTOL."""
import numpy as np
import pytest

from corr_pc_amd import synth
from corr_pc_amd.engine import KIND_DD, KIND_SHAPE_SHAPE, PREC_FP64, PREC_SPLIT, make_params, make_sky_params
from helpers import assert_hist_match, sky_catalogs, small_catalog

pytestmark = pytest.mark.gpu

PRECS = [PREC_SPLIT, PREC_FP64]


def _shuffled(c, seed):
    return c.subset(np.random.default_rng(seed).permutation(c.n))


def _run(engine, oracle_lib, outer, inner, p, what):
    print(f"[case] {what}", flush=True)
    ref, rst = oracle_lib.count(outer, inner, p)
    got, st = engine.count(outer, inner, p)
    assert_hist_match(got, ref, what=what)
    assert st["n_binned"] == rst["n_binned"] > 0


@pytest.mark.parametrize("prec", PRECS)
def test_unsorted_box(engine, oracle_lib, prec):
    L = 50.0
    A = _shuffled(small_catalog(4000, L, 88, clumps=30), 5)
    B = _shuffled(small_catalog(3000, L, 89, clumps=30), 6)
    p = make_params(L, 8, 0.3, 10.0, 8, 2.0, 14.0, kind=KIND_SHAPE_SHAPE, estimator=0, g_with_shape=2, do_jk=1,
                    n_jk=8, precision=prec)
    p.unsorted = 1
    _run(engine, oracle_lib, A, B, p, f"unsorted cross, one-sided window prec={prec}")
    p = make_params(L, 8, 0.3, 10.0, 10, -12.0, 12.0, kind=KIND_SHAPE_SHAPE, estimator=1, g_with_shape=1, do_jk=1,
                    n_jk=8, precision=prec)
    p.unsorted, p.auto_pairs = 1, 1
    _run(engine, oracle_lib, A, A, p, f"unsorted auto box prec={prec}")


@pytest.mark.parametrize("prec", PRECS)
def test_unsorted_sky_auto(engine, oracle_lib, prec):
    table = synth.distance_table(0.3, 1e-4, 1.0)
    A, _, n_jk, _ = sky_catalogs(6000, 10, 90, table=table)
    A = _shuffled(A, 7)
    pt = synth.sky_params_table(2, table)
    p = make_sky_params(pt, 10, 0.5, 60.0, 12, -60.0, 60.0, kind=KIND_DD, do_jk=1, n_jk=n_jk, precision=prec)
    p.unsorted, p.auto_pairs = 1, 1
    _run(engine, oracle_lib, A, A, p, f"unsorted auto sky prec={prec}")
