"""GPU parity of the ED kind (which_corr 5, periodic box: ED_corel / ED_calc_PB, corels.cpp:246-262,
PB_calcs.cpp:101-134): the HIP engine through the C ABI against the CPU oracle, with rp-pi and r-mu
bins, the axis on either side, both kernel precisions.

Counts must be bit-identical; weighted sums within helpers.RTOL.
"""
import numpy as np
import pytest

from corr_pc_amd.engine import KIND_ED, PREC_FP64, PREC_SPLIT, make_params
from helpers import assert_hist_match, small_catalog

pytestmark = pytest.mark.gpu

PRECS = [PREC_SPLIT, PREC_FP64]


def _with_axes(cat, seed):
    c = cat.subset(np.arange(cat.n))
    v = np.random.default_rng(seed).normal(size=(cat.n, 3))
    v /= np.linalg.norm(v, axis=1)[:, None]
    c.e1, c.e2, c.e3 = (np.ascontiguousarray(v[:, k]) for k in range(3))
    c.phi = None
    return c


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("gws", [1, 2])
@pytest.mark.parametrize("r_mu", [0, 1])
def test_box_ed(engine, oracle_lib, prec, gws, r_mu):
    L = 60.0
    A = small_catalog(3000, L, 71, shapes=False, clumps=40)
    B = small_catalog(2500, L, 72, shapes=False, clumps=30)
    if gws == 2:
        B = _with_axes(B, 73)
    else:
        A = _with_axes(A, 73)
    pi_lo, pi_hi, n_pi = (-1.0, 1.0, 4) if r_mu else (-15.0, 15.0, 10)
    p = make_params(L, 8, 0.2, 12.0, n_pi, pi_lo, pi_hi, kind=KIND_ED, g_with_shape=gws, do_jk=1, n_jk=8,
                    precision=prec, r_mu=r_mu)
    what = f"ED gws={gws} r_mu={r_mu} prec={prec}"
    print(f"[case] {what}", flush=True)
    ref, rst = oracle_lib.count(A, B, p)
    got, st = engine.count(A, B, p)
    assert_hist_match(got, ref, what=what)
    assert st["n_binned"] == rst["n_binned"] > 0
    mean = got.val[0, ..., 6].sum() / got.num[0].sum()
    assert 0.25 < mean < 0.42  # isotropic axes: 1/3
