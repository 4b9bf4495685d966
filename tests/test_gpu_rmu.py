"""GPU parity of r-mu binning in a periodic box (coordinates 7): the HIP engine through the C ABI
against the CPU oracle, for each pair kind, the auto pass and both kernel precisions.

Counts must be bit-identical; weighted sums within helpers.RTOL.  Every r-mu candidate takes the
engine's FP64 classification (r = sqrt(rp*rp + pi*pi) with rp rounded first, mu = pi / r).
"""
import numpy as np
import pytest

from corr_pc_amd.engine import KIND_DD, KIND_SHAPE_DENSITY, KIND_SHAPE_SHAPE, PREC_FP64, PREC_SPLIT, make_params
from helpers import assert_hist_match, small_catalog

pytestmark = pytest.mark.gpu

PRECS = [PREC_SPLIT, PREC_FP64]


def _run(engine, oracle_lib, outer, inner, p, what):
    print(f"[case] {what}", flush=True)
    ref, rst = oracle_lib.count(outer, inner, p)
    got, st = engine.count(outer, inner, p)
    assert_hist_match(got, ref, what=what)
    assert st["n_binned"] == rst["n_binned"] > 0
    return got


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("kind,est,gws", [(KIND_DD, 0, 2), (KIND_SHAPE_DENSITY, 0, 2), (KIND_SHAPE_SHAPE, 0, 2),
                                          (KIND_SHAPE_SHAPE, 1, 1)])
def test_rmu_kinds(engine, oracle_lib, prec, kind, est, gws):
    L = 60.0
    A = small_catalog(3000, L, 51, clumps=40)
    B = small_catalog(2500, L, 52, clumps=30)
    # 8 log r bins in [0.2, 15), 6 mu bins in [-1, 1): the z window is +- 15
    p = make_params(L, 8, 0.2, 15.0, 6, -1.0, 1.0, kind=kind, g_with_shape=gws, estimator=est, do_jk=1, n_jk=8,
                    precision=prec, r_mu=1)
    _run(engine, oracle_lib, A, B, p, f"r-mu kind={kind} est={est} gws={gws} prec={prec}")


@pytest.mark.parametrize("prec", PRECS)
def test_rmu_auto_and_linear(engine, oracle_lib, prec):
    L = 50.0
    A = small_catalog(3000, L, 53, clumps=30)
    p = make_params(L, 6, 0.5, 12.0, 1, -1.0, 1.0, kind=KIND_SHAPE_SHAPE, estimator=1, g_with_shape=1, do_jk=1,
                    n_jk=8, precision=prec, r_mu=1, lin_bin=1)
    p.auto_pairs = 1
    _run(engine, oracle_lib, A, A, p, f"r-mu auto linear prec={prec}")


def test_rmu_known_pair(engine):
    """One (3, 4, 12) pair: r = 13, mu = 12/13, whose |pi| lies far outside the mu range."""
    from corr_pc_amd.engine import Catalog
    a = Catalog(x=np.array([5.0]), y=np.array([5.0]), z=np.array([5.0]))
    b = Catalog(x=np.array([8.0]), y=np.array([9.0]), z=np.array([17.0]))
    p = make_params(50.0, 4, 1.0, 20.0, 4, -1.0, 1.0, kind=KIND_DD, r_mu=1)
    h, st = engine.count(a, b, p)
    assert st["n_binned"] == 1 and h.num[0, 3, 3] == 1
