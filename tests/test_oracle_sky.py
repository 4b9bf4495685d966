"""CPU tests pinning the oracle's sky geometry (coordinates 0: RA, Dec in degrees, redshift).

  * an independent enumeration of every (outer, inner) pair through the per-galaxy windows of
    lim_calc_sky (tests/bruteforce.py, no z-sorted sweep);
  * closed forms for single pairs: an RA-only separation on the equator at equal redshift has
    rp = dRA[rad] * dA(z) and orientation angle 0, a Dec-only one has angle pi/2 (e+ = -e1); the
    sign of pi follows DC(inner) - DC(outer);
  * invariants: unit weights give wt_num == num; touch sums with every coin 1 or 0.
"""
import math

import numpy as np
import pytest

import bruteforce
from corr_pc_amd import synth
from corr_pc_amd.engine import (JK_TOUCH, KIND_DD, KIND_SHAPE_DENSITY, KIND_SHAPE_SHAPE, Catalog, make_sky_params)
from helpers import assert_hist_match, sky_catalogs

_TABLE = synth.distance_table(0.3, 1e-4, 1.0)
_PT = synth.sky_params_table(2, _TABLE)


def _sky_params(kind=KIND_DD, **kw):
    base = dict(n_rp=10, rp_min=0.5, rp_max=60.0, n_pi=12, pi_min=-60.0, pi_max=60.0)
    base.update({k: kw.pop(k) for k in list(kw) if k in base})
    return make_sky_params(_PT, base["n_rp"], base["rp_min"], base["rp_max"], base["n_pi"], base["pi_min"],
                           base["pi_max"], kind=kind, **kw)


@pytest.mark.parametrize("kind,est,gws", [(KIND_DD, 0, 2), (KIND_SHAPE_DENSITY, 0, 2), (KIND_SHAPE_DENSITY, 0, 1),
                                          (KIND_SHAPE_SHAPE, 0, 2), (KIND_SHAPE_SHAPE, 1, 1)])
def test_oracle_sky_vs_bruteforce(oracle_lib, kind, est, gws):
    A, B, n_jk, _ = sky_catalogs(240, 300, 5, shapes_in=True, table=_TABLE)
    p = _sky_params(kind, g_with_shape=gws, estimator=est, do_jk=1, n_jk=n_jk)
    h, st = oracle_lib.count(A, B, p)
    bf = bruteforce.count_sky(A, B, p)
    assert_hist_match(h, bf, rtol=1e-12, what="sky oracle vs brute force")
    assert st["n_binned"] == bf["num"][0].sum() > 300


@pytest.mark.parametrize("lin,jk_mode", [(1, JK_TOUCH), (0, 0)])
def test_oracle_sky_vs_bruteforce_windows(oracle_lib, lin, jk_mode):
    """Linear bins from 0, a positive-only pi window, include_prob subsampling."""
    A, B, n_jk, _ = sky_catalogs(260, 320, 6, inc=True, table=_TABLE)
    p = _sky_params(KIND_SHAPE_DENSITY, g_with_shape=1, rp_min=0.0 if lin else 1.0, pi_min=0.0, n_pi=8, lin_bin=lin,
                    do_jk=1, n_jk=n_jk, include_prob_max=0.6, jk_mode=jk_mode)
    h, _ = oracle_lib.count(A, B, p)
    bf = bruteforce.count_sky(A, B, p)
    assert_hist_match(h, bf, rtol=1e-12, what="sky oracle vs brute force (windows)")


def _one(ra, dec, z, e=(0.0, 0.0), phi=None, w=1.0):
    it = int(math.floor(z / _TABLE["dz"]))
    return Catalog(x=[ra], y=[dec], z=[z], w=[w], e1=[e[0]], e2=[e[1]], phi=None if phi is None else [phi],
                   jk=np.array([0], np.int32), jk_prob=[1.0], dc=[_TABLE["dC"][it]], da=[_TABLE["dA"][it]])


def test_kat_sky_ra_separation(oracle_lib):
    """Equator, equal redshift, dRA = 0.5 deg: rp = (0.5 / r2d) * dA(z), pi = 0, theta = 0 so
    e+ = e1 and ex = e2 of the shape side."""
    z = 0.5
    A = _one(30.0, 0.0, z, phi=0.3)
    B = _one(30.5, 0.0, z, e=(0.2, -0.1))
    da = _TABLE["dA"][int(math.floor(z / _TABLE["dz"]))]
    rp = math.radians(0.5) * da
    p = _sky_params(KIND_SHAPE_DENSITY, n_rp=10, rp_min=1.0, rp_max=100.0, g_with_shape=2)
    h, _ = oracle_lib.count(A, B, p)
    k = int(math.floor(math.log10(rp / 1.0) / 0.2))
    q = int(math.floor((0.0 + 60.0) / 10.0))
    assert h.num.sum() == 1 and h.num[0, k, q] == 1, (rp, k)
    assert h.val[0, k, q, 1] == pytest.approx(0.2, abs=1e-15)
    assert h.val[0, k, q, 2] == pytest.approx(-0.1, abs=1e-15)
    assert h.val[0, k, q, 6] == pytest.approx(0.3, abs=1e-15)  # |phi_outer - 0|


def test_kat_sky_dec_separation_and_pi_sign(oracle_lib):
    """Dec-only separation: theta = atan(dDec / 0) = pi/2, so e+ = -e1, ex = -e2; the inner galaxy
    further away (larger DC) gives pi = DC_in - DC_out > 0."""
    A = _one(100.0, 10.0, 0.4)
    B = _one(100.0, 10.3, 0.405, e=(0.25, 0.05))
    it_a, it_b = int(math.floor(0.4 / 1e-4)), int(math.floor(0.405 / 1e-4))
    pi = _TABLE["dC"][it_b] - _TABLE["dC"][it_a]
    assert 5 < pi < 60
    p = _sky_params(KIND_SHAPE_DENSITY, n_rp=10, rp_min=1.0, rp_max=100.0, g_with_shape=2)
    h, _ = oracle_lib.count(A, B, p)
    q = int(math.floor((pi + 60.0) / 10.0))
    assert h.num.sum() == 1 and h.num[0, :, q].sum() == 1
    k = int(np.argmax(h.num[0, :, q]))
    assert h.val[0, k, q, 1] == pytest.approx(-0.25, abs=1e-12)
    assert h.val[0, k, q, 2] == pytest.approx(-0.05, abs=1e-12)
    # swapped roles: pi changes sign
    h2, _ = oracle_lib.count(B, A, p)
    q2 = int(math.floor((-pi + 60.0) / 10.0))
    assert h2.num[0, :, q2].sum() == 1


def test_sky_unit_weights_and_touch_sums(oracle_lib):
    """jk_check_sky: a same-region pair goes to the region iff the outer coin is >= 0.5 (unlike the
    box, where it always does); a cross-region pair to each side whose coin is >= 0.5."""
    A, B, n_jk, _ = sky_catalogs(300, 300, 7, table=_TABLE)
    A.w = None
    B.w = None
    p = _sky_params(KIND_DD, do_jk=1, n_jk=n_jk, jk_mode=JK_TOUCH)
    A.jk_prob = np.ones(A.n)
    B.jk_prob = np.ones(B.n)
    A.jk = np.zeros(A.n, np.int32)
    B.jk = np.ones(B.n, np.int32)
    h, _ = oracle_lib.count(A, B, p)
    full = h.num[0].sum()
    assert full > 500
    assert np.array_equal(h.wt_num[0], h.num[0].astype(float))
    assert np.array_equal(h.num[1], h.num[0]) and np.array_equal(h.num[2], h.num[0]) and h.num[3:].sum() == 0
    B.jk = np.zeros(B.n, np.int32)
    h, _ = oracle_lib.count(A, B, p)
    assert np.array_equal(h.num[1], h.num[0]) and h.num[2:].sum() == 0
    A.jk_prob = np.zeros(A.n)
    h, _ = oracle_lib.count(A, B, p)
    assert h.num[1:].sum() == 0 and h.num[0].sum() == full
