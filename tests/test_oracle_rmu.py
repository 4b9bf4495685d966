"""CPU pins of the oracle's r-mu binning of a periodic box (coordinates 7: lim_calc_xyz's z +- rp_max
window, PB_calcs.cpp:14-23; r_mu_calc, calcs.cpp:135-140; RR_PB_calc, PB_calcs.cpp:63-85).

Pinned by:
  * known answers: a (3, 4, 12) offset is r = 13, mu = 12/13 (and -12/13 the other way round); a pair
    across the z boundary is found through its image; a pair whose |pi| is far outside the mu range
    but inside +- rp_max is counted (the window is rp_max, not the pi range);
  * a Poisson pin: two independent uniform catalogs give N1 N2 (4/3 pi (r2^3 - r1^3)) (dmu / 2) / L^3
    pairs per (r, mu) bin, so the cross estimator with the analytic coordinates-7 RR is 0 up to
    Poisson noise; a wrong r (2-D), mu or window fails it.
"""
import numpy as np

from corr_pc_amd.engine import (KIND_DD, Catalog, Histogram, Norm, WSum, bin_edges, box_rr, make_params, weight_sum,
                                write_cross)
from helpers import small_catalog

L = 50.0


def _params(**kw):
    # 4 log r bins in [1, 20), 4 mu bins in [-1, 1)
    return make_params(L, 4, 1.0, 20.0, 4, -1.0, 1.0, kind=KIND_DD, r_mu=1, **kw)


def _cat(pts, w=None):
    pts = np.asarray(pts, np.float64)
    pts = pts[np.argsort(pts[:, 2], kind="stable")]
    return Catalog(x=pts[:, 0], y=pts[:, 1], z=pts[:, 2], w=None if w is None else np.asarray(w, np.float64))


def _cell(r, mu, p):
    k = int(np.floor(np.log10(r / p.rp_min) / ((np.log10(p.rp_max) - np.log10(p.rp_min)) / p.n_rp)))
    q = int(np.floor((mu - p.pi_min) / ((p.pi_max - p.pi_min) / p.n_pi)))
    return k, q


def test_known_answers(oracle_lib):
    p = _params()
    a = _cat([[5.0, 5.0, 5.0]])
    b = _cat([[8.0, 9.0, 17.0]])
    h, st = oracle_lib.count(a, b, p)
    assert st["n_binned"] == 1
    assert h.num[0][_cell(13.0, 12.0 / 13.0, p)] == 1
    h, _ = oracle_lib.count(b, a, p)
    assert h.num[0][_cell(13.0, -12.0 / 13.0, p)] == 1
    # across the z boundary: dz = (48 - 50) - 1 = -3 through the PBz = -1 image, rp = 5
    c = _cat([[10.0, 10.0, 1.0]])
    d = _cat([[13.0, 14.0, 48.0]])
    h, st = oracle_lib.count(c, d, p)
    assert st["n_binned"] == 1
    assert h.num[0][_cell(np.sqrt(34.0), -3.0 / np.sqrt(34.0), p)] == 1
    # a pure line-of-sight pair 15 apart: mu = 1 lies outside [-1, 1), mu = -1 inside
    e = _cat([[20.0, 20.0, 10.0]])
    f = _cat([[20.0, 20.0, 25.0]])
    h, st = oracle_lib.count(f, e, p)
    assert st["n_binned"] == 1 and h.num[0][_cell(15.0, -1.0, p)] == 1
    h, st = oracle_lib.count(e, f, p)
    assert st["n_binned"] == 0


def test_poisson_uniform_box(oracle_lib, tmp_path):
    p = _params()
    S = small_catalog(4000, L, 91, shapes=False, weights=False, jk_side=1)
    D = small_catalog(4000, L, 92, shapes=False, weights=False, jk_side=1)
    h, st = oracle_lib.count(D, S, p)
    re, pe = bin_edges(p)
    expect = 4000 * 4000 * (4.0 / 3.0 * np.pi * (re[1:] ** 3 - re[:-1] ** 3))[:, None] * \
        (np.diff(pe) / 2.0)[None, :] / L ** 3
    n = h.num[0].astype(np.float64)
    ok = expect > 200
    assert ok.sum() >= 8
    z = (n[ok] - expect[ok]) / np.sqrt(expect[ok])
    assert np.all(np.abs(z) < 5), z
    # the analytic RR of coordinates 7 is that expectation per S*D
    np.testing.assert_allclose(box_rr(p) * 4000 * 4000, expect, rtol=1e-6)
    # and the estimator is 0 up to noise
    zero = Histogram.for_params(p)
    write_cross(p, Norm(weight_sum(S), weight_sum(D), WSum(1.0, 1.0), WSum(1.0, 1.0), 1.0), h, zero, zero, zero,
                str(tmp_path) + "/rm_")
    w = np.loadtxt(str(tmp_path) + "/rm_bins2D_cross_final.dat", comments="#")[:, 3].reshape(p.n_rp, p.n_pi)
    assert np.all(np.abs(w[ok]) * np.sqrt(expect[ok]) < 5)

