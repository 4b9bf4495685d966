"""CPU pins of the oracle's ED kind (which_corr 5, periodic box: ED_corel / ED_calc_PB,
corels.cpp:246-262, PB_calcs.cpp:101-134): data[6] += cos^2 of the angle between the pair's imaged
separation and the 3-D axis of g2 (g_with_shape 2) or g1, err += cos^4.

Pinned by:
  * known answers with the axis along z: separation along z gives 1, along x gives 0, at 45 degrees
    gives 1/2, through the periodic image too; the axis comes from the g_with_shape side;
  * every axis along z: cos^2 = mu^2 for every pair, so in r-mu bins each mu bin's sum lies between
    n mu_lo^2 and n mu_hi^2;
  * isotropic axes: the mean cos^2 in each well-populated cell is 1/3 up to noise (variance 4/45).
"""
import numpy as np
import pytest

from corr_pc_amd.engine import KIND_ED, Catalog, make_params
from helpers import small_catalog

L = 50.0


def _params(**kw):
    # r-mu bins: 4 log r bins in [1, 20), 4 mu bins in [-1, 1)
    kw.setdefault("r_mu", 1)
    return make_params(L, 4, 1.0, 20.0, 4, -1.0, 1.0, kind=KIND_ED, **kw)


def _axis(n, ax):
    """Unit axis components as separate contiguous arrays (ctypes reads the raw buffers)."""
    a = np.asarray(ax, np.float64) / np.linalg.norm(ax)
    return np.full(n, a[0]), np.full(n, a[1]), np.full(n, a[2])


def _pair(p1, p2, ax1=(0, 0, 1), ax2=(0, 0, 1)):
    def cat(p, ax):
        e1, e2, e3 = _axis(1, ax)
        return Catalog(x=np.array([p[0]]), y=np.array([p[1]]), z=np.array([p[2]]), e1=e1, e2=e2, e3=e3)
    return cat(p1, ax1), cat(p2, ax2)


@pytest.mark.parametrize("gws", [1, 2])
def test_known_answers(oracle_lib, gws):
    p = _params(g_with_shape=gws)
    # (mu = +1 lies outside [-1, 1): the line-of-sight pair is taken with pi < 0)
    for p1, p2, expect in (([10, 10, 15], [10, 10, 10], 1.0), ([10, 10, 10], [15, 10, 10], 0.0),
                           ([10, 10, 10], [14, 10, 14], 0.5), ([10, 10, 1], [10, 10, 47], 1.0)):
        a, b = _pair(p1, p2)
        h, st = oracle_lib.count(a, b, p)
        assert st["n_binned"] == 1
        np.testing.assert_allclose(h.val[0][..., 6].sum(), expect, atol=1e-15)
        np.testing.assert_allclose(h.err[0][..., 6].sum(), expect ** 2, atol=1e-15)
    # the axis is read from the g_with_shape side only: the other side's axis does not matter
    a, b = _pair([10, 10, 15], [10, 10, 10], ax1=(1, 0, 0) if gws == 2 else (0, 0, 1),
                 ax2=(0, 0, 1) if gws == 2 else (1, 0, 0))
    h, _ = oracle_lib.count(a, b, p)
    np.testing.assert_allclose(h.val[0][..., 6].sum(), 1.0, atol=1e-15)


def test_aligned_axes_give_mu_squared(oracle_lib):
    A = small_catalog(2000, L, 61, shapes=False, weights=False, jk_side=1)
    B = small_catalog(2000, L, 62, shapes=False, weights=False, jk_side=1)
    B.e1, B.e2, B.e3 = _axis(B.n, (0, 0, 1))
    p = _params(g_with_shape=2)
    h, st = oracle_lib.count(A, B, p)
    assert st["n_binned"] > 10000
    n = h.num[0].astype(np.float64)
    s = h.val[0][..., 6]
    mu_lo = np.array([-1.0, -0.5, 0.0, 0.5])
    mu_hi = mu_lo + 0.5
    lo = np.minimum(mu_lo ** 2, mu_hi ** 2) * (mu_lo * mu_hi > 0)
    hi = np.maximum(mu_lo ** 2, mu_hi ** 2)
    assert np.all(s >= n * lo[None, :] - 1e-9) and np.all(s <= n * hi[None, :] + 1e-9)
    assert np.all(s[:, [0, 3]] > 0.3 * n[:, [0, 3]])


def test_isotropic_axes_mean_one_third(oracle_lib):
    A = small_catalog(3000, L, 63, shapes=False, weights=False, jk_side=1)
    B = small_catalog(3000, L, 64, shapes=False, weights=False, jk_side=1)
    v = np.random.default_rng(65).normal(size=(B.n, 3))
    v /= np.linalg.norm(v, axis=1)[:, None]
    B.e1, B.e2, B.e3 = (np.ascontiguousarray(v[:, k]) for k in range(3))
    h, _ = oracle_lib.count(A, B, _params(g_with_shape=2, r_mu=0))
    n = h.num[0].astype(np.float64)
    ok = n > 500
    assert ok.sum() >= 5
    z = (h.val[0][..., 6][ok] / n[ok] - 1.0 / 3.0) / np.sqrt(4.0 / 45.0 / n[ok])
    assert np.all(np.abs(z) < 5), z
