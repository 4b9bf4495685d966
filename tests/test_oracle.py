"""CPU tests pinning the oracle (oracle/cpc_oracle.c) before it is trusted as the GPU checker.

Each check is chosen so that a plausible mistake in the restatement (a dropped image, a wrong sign
of pi, a transposed operand, a wrong jackknife slot) fails it:
  * an independent O(N^2) enumeration of every (outer, inner, image) triple (tests/bruteforce.py);
  * numpy minimum-image pair counts (no windows, no images: pure geometry);
  * hand-computed single-pair known answers (angles, rotations, ephi, the raw-coordinate angle of
    wrapped pairs, bin edges);
  * invariants of weights, jackknife slots and operand swaps.
"""
import math

import numpy as np
import pytest

import bruteforce
from corr_pc_amd.engine import (JK_TOUCH, KIND_DD, KIND_SHAPE_DENSITY, KIND_SHAPE_SHAPE, Catalog, make_params)
from helpers import assert_hist_match, small_catalog


def _pair(xyz_i, xyz_j, L=50.0, e_i=(0.0, 0.0), e_j=(0.0, 0.0), w=(1.0, 1.0), phi=(None, None), jk=(0, 0),
          jkp=(0.0, 0.0)):
    def one(p, e, wt, ph, r, jp):
        return Catalog(x=[p[0]], y=[p[1]], z=[p[2]], w=[wt], e1=[e[0]], e2=[e[1]],
                       phi=None if ph is None else [ph], jk=np.array([r], np.int32), jk_prob=[jp])
    return one(xyz_i, e_i, w[0], phi[0], jk[0], jkp[0]), one(xyz_j, e_j, w[1], phi[1], jk[1], jkp[1])


@pytest.mark.parametrize("kind,est,gws", [(KIND_DD, 0, 2), (KIND_SHAPE_DENSITY, 0, 2), (KIND_SHAPE_DENSITY, 0, 1),
                                          (KIND_SHAPE_SHAPE, 0, 2), (KIND_SHAPE_SHAPE, 0, 1),
                                          (KIND_SHAPE_SHAPE, 1, 2)])
def test_oracle_vs_bruteforce(oracle_lib, kind, est, gws):
    L = 40.0
    A = small_catalog(260, L, 1, clumps=10)
    B = small_catalog(230, L, 2, clumps=10)
    p = make_params(L, 6, 0.3, 9.0, 8, -11.0, 11.0, kind=kind, g_with_shape=gws, estimator=est, do_jk=1, n_jk=8)
    h, st = oracle_lib.count(A, B, p)
    bf = bruteforce.count(A, B, p)
    assert_hist_match(h, bf, rtol=1e-12, what="oracle vs brute force")
    assert st["n_binned"] == bf["num"][0].sum() > 500


@pytest.mark.parametrize("pmin,pmax,lin", [(0.0, 10.0, 0), (-7.0, 12.0, 1), (3.0, 12.0, 0)])
def test_oracle_vs_bruteforce_windows(oracle_lib, pmin, pmax, lin):
    """One-sided and positive-only pi windows, linear bins from 0, include_prob subsampling, touch mode."""
    L = 30.0
    A = small_catalog(220, L, 3, inc=True)
    B = small_catalog(240, L, 4, inc=True)
    p = make_params(L, 5, 0.0 if lin else 0.5, 8.0, 6, pmin, pmax, kind=KIND_SHAPE_SHAPE, lin_bin=lin, do_jk=1,
                    n_jk=8, include_prob_max=0.7, jk_mode=JK_TOUCH)
    h, _ = oracle_lib.count(A, B, p)
    bf = bruteforce.count(A, B, p)
    assert_hist_match(h, bf, rtol=1e-12, what="oracle vs brute force (windows)")


def test_minimum_image_counts(oracle_lib):
    """Counts against numpy minimum-image geometry: a pair is counted once iff its minimum-image
    separation has rp in [rmin, rmax) and pi = dz in [pmin, pmax)."""
    L, n = 60.0, 1500
    r = np.random.default_rng(9)
    P = r.uniform(0, L, (n, 3))
    P = P[np.argsort(P[:, 2], kind="stable")]
    Q = r.uniform(0, L, (n, 3))
    Q = Q[np.argsort(Q[:, 2], kind="stable")]
    rmin, rmax, nb, pmin, pmax, npi = 0.5, 12.0, 7, -14.0, 14.0, 7
    d = Q[None, :, :] - P[:, None, :]
    d -= L * np.round(d / L)
    rp = np.hypot(d[..., 0], d[..., 1])
    pi = d[..., 2]
    lg = np.log10(np.where(rp > 0, rp, 1e-300) / rmin) / ((math.log10(rmax) - math.log10(rmin)) / nb)
    pb = (pi - pmin) / ((pmax - pmin) / npi)
    # no pair within 1e-9 of any edge, so floating-point details cannot matter
    assert np.min(np.abs(lg - np.round(lg))) > 1e-9 and np.min(np.abs(pb - np.round(pb))) > 1e-9
    k = np.floor(lg).astype(int)
    q = np.floor(pb).astype(int)
    ok = (k >= 0) & (k < nb) & (q >= 0) & (q < npi)
    expect = np.zeros((nb, npi), np.int64)
    np.add.at(expect, (k[ok], q[ok]), 1)
    A = Catalog(x=P[:, 0], y=P[:, 1], z=P[:, 2])
    B = Catalog(x=Q[:, 0], y=Q[:, 1], z=Q[:, 2])
    p = make_params(L, nb, rmin, rmax, npi, pmin, pmax, kind=KIND_DD)
    h, _ = oracle_lib.count(A, B, p)
    assert np.array_equal(h.num[0], expect)


def test_kat_single_pair_shapes(oracle_lib):
    """i at (10,10,10), j at (13,14,12): rp = 5, pi = +2, theta = atan(4/3), cos2t = -7/25,
    sin2t = 24/25.  Shape of j (0.3, 0.1): e+ = -0.18, ex = 0.26.  Weights 2 and 0.5."""
    A, B = _pair((10, 10, 10), (13, 14, 12), e_j=(0.3, 0.1), e_i=(-0.2, 0.05), w=(2.0, 0.5), phi=(0.5, -0.4))
    p = make_params(50.0, 2, 1.0, 16.0, 4, -10.0, 10.0, kind=KIND_SHAPE_DENSITY, g_with_shape=2)
    h, _ = oracle_lib.count(A, B, p)
    assert h.num[0, 1, 2] == 1 and h.num.sum() == 1  # rp 5 in [4, 16); pi 2 in [0, 5)
    th = math.atan2(4, 3)
    assert h.wt_num[0, 1, 2] == pytest.approx(1.0, abs=1e-15)
    assert h.val[0, 1, 2, 1] == pytest.approx(-0.18, abs=1e-15)
    assert h.val[0, 1, 2, 2] == pytest.approx(0.26, abs=1e-15)
    assert h.err[0, 1, 2, 1] == pytest.approx(0.0324, abs=1e-15)
    assert h.val[0, 1, 2, 6] == pytest.approx(abs(0.5 - th), abs=1e-15)  # theta column uses the outer's phi
    # shape-shape, estimator 0: e+ of j only, e++ = w e+_i e+_j, exx, e+x
    p = make_params(50.0, 2, 1.0, 16.0, 4, -10.0, 10.0, kind=KIND_SHAPE_SHAPE, g_with_shape=2)
    h, _ = oracle_lib.count(A, B, p)
    c, s = -7 / 25, 24 / 25
    epi, exi = c * -0.2 - s * 0.05, s * -0.2 + c * 0.05
    epj, exj = -0.18, 0.26
    v = h.val[0, 1, 2]
    assert v[1] == pytest.approx(epj, abs=1e-15) and v[2] == 0.0
    assert v[3] == pytest.approx(epi * epj, abs=1e-15)
    assert v[4] == pytest.approx(exi * exj, abs=1e-15)
    assert v[5] == pytest.approx((exi * epj + epi * exj) / 2, abs=1e-15)
    assert v[6] == pytest.approx(abs(-0.4 - th), abs=1e-15)  # g_with_shape = 2: j's phi


def test_kat_wrapped_pair_uses_raw_angle(oracle_lib):
    """i at (1, 1, 25), j at (49, 2, 26) in L = 50: the pair is counted through the image
    x_j - L = -1 (rp = sqrt(5)), but the orientation angle is atan((2-1)/(49-1)) on raw coordinates
    (calcs.cpp:54-57), not atan(1/(-2))."""
    A, B = _pair((1, 1, 25), (49, 2, 26), e_j=(0.2, -0.1))
    p = make_params(50.0, 3, 1.0, 10.0, 2, -5.0, 5.0, kind=KIND_SHAPE_DENSITY, g_with_shape=2)
    h, _ = oracle_lib.count(A, B, p)
    assert h.num.sum() == 1
    k = int(math.floor(math.log10(math.sqrt(5.0)) / (1 / 3)))
    assert h.num[0, k, 1] == 1
    th = math.atan(1 / 48)
    ep = math.cos(2 * th) * 0.2 - math.sin(2 * th) * -0.1
    assert h.val[0, k, 1, 1] == pytest.approx(ep, abs=1e-15)
    th_img = math.atan(1 / -2)
    assert abs(h.val[0, k, 1, 1] - (math.cos(2 * th_img) * 0.2 - math.sin(2 * th_img) * -0.1)) > 0.1


def test_kat_phi_minus_ten(oracle_lib):
    """shape_density's theta column uses the outer galaxy's phi even when it has no shape: phi = -10
    gives pi - (10 + theta) (calcs.cpp:97-105, corels.cpp:229, read_dat.cpp:269)."""
    A, B = _pair((10, 10, 10), (13, 14, 12), e_j=(0.3, 0.1))
    A.phi = None
    p = make_params(50.0, 2, 1.0, 16.0, 4, -10.0, 10.0, kind=KIND_SHAPE_DENSITY, g_with_shape=2)
    h, _ = oracle_lib.count(A, B, p)
    th = math.atan2(4, 3)
    assert h.val[0, 1, 2, 6] == pytest.approx(math.pi - (10 + th), abs=1e-14)


def test_kat_bin_edges(oracle_lib):
    """pi = p_min is in bin 0, pi = p_max is out; rp = rp_min is in bin 0; rp = 0 is rejected."""
    L = 50.0
    pts = [(10, 10, 10), (11, 10, 5), (11, 10, 15), (10, 10, 10)]  # rp 1 & pi -5, rp 1 & pi +5, rp 0
    B = Catalog(x=[p[0] for p in pts[1:]], y=[p[1] for p in pts[1:]], z=[p[2] for p in pts[1:]])
    A = Catalog(x=[10.0], y=[10.0], z=[10.0])
    p = make_params(L, 2, 1.0, 4.0, 2, -5.0, 5.0, kind=KIND_DD)
    h, _ = oracle_lib.count(A, B, p)
    assert h.num[0].tolist() == [[1, 0], [0, 0]]


def test_kat_jackknife_slots(oracle_lib):
    """Different regions: each side's region receives the pair iff its jk_prob >= 0.5; same region:
    the outer's region once, whatever the coins (jk.cpp:64-71, 201-223)."""
    p = make_params(50.0, 1, 1.0, 10.0, 1, -5.0, 5.0, kind=KIND_DD, do_jk=1, n_jk=3, jk_mode=JK_TOUCH)
    for (ja, jb, pa, pb, expect) in [(0, 1, 0.7, 0.2, [1, 0, 0]), (0, 1, 0.2, 0.7, [0, 1, 0]),
                                     (0, 1, 0.5, 0.5, [1, 1, 0]), (0, 1, 0.1, 0.1, [0, 0, 0]),
                                     (2, 2, 0.1, 0.1, [0, 0, 1]), (2, 2, 0.9, 0.9, [0, 0, 1])]:
        A, B = _pair((10, 10, 10), (12, 10, 11), jk=(ja, jb), jkp=(pa, pb))
        h, _ = oracle_lib.count(A, B, p)
        assert h.num[0, 0, 0] == 1
        assert h.num[1:, 0, 0].tolist() == expect, (ja, jb, pa, pb)
    # leave-one-out: jk_r = full - touch_r, err untouched
    p.jk_mode = 0
    A, B = _pair((10, 10, 10), (12, 10, 11), jk=(0, 1), jkp=(0.9, 0.9))
    h, _ = oracle_lib.count(A, B, p)
    assert h.num[1:, 0, 0].tolist() == [0, 0, 1]


def test_invariants_weights_and_swap(oracle_lib):
    L = 50.0
    A = small_catalog(600, L, 61, weights=False)
    B = small_catalog(500, L, 62, weights=False)
    p = make_params(L, 6, 0.3, 9.0, 8, -12.0, 12.0, kind=KIND_SHAPE_SHAPE, do_jk=1, n_jk=8, g_with_shape=2)
    h, _ = oracle_lib.count(A, B, p)
    assert np.array_equal(h.wt_num[0], h.num[0].astype(float))  # unit weights: sum of w = count
    # swapping the operands flips pi: bin p <-> n_pi-1-p (symmetric bins, no pair on a pi edge)
    p2 = make_params(L, 6, 0.3, 9.0, 8, -12.0, 12.0, kind=KIND_SHAPE_SHAPE, do_jk=1, n_jk=8, g_with_shape=1)
    h2, _ = oracle_lib.count(B, A, p2)
    assert np.array_equal(h.num[0], h2.num[0][:, ::-1])
    # the shape-side e+ sum is the same galaxy's sum either way (g_with_shape follows the swap)
    assert np.allclose(h.val[0, :, :, 1], h2.val[0, :, ::-1, 1], rtol=1e-12, atol=1e-12)


def test_invariant_touch_sums(oracle_lib):
    """Touch mode: with every coin 1 the regions' counts sum to full + cross-region pairs; with every
    coin 0 to the same-region pairs only."""
    L = 50.0
    A = small_catalog(500, L, 71)
    B = small_catalog(500, L, 72)
    p = make_params(L, 5, 0.3, 9.0, 6, -10.0, 10.0, kind=KIND_DD, do_jk=1, n_jk=8, jk_mode=JK_TOUCH)
    A.jk_prob = np.ones(A.n)
    B.jk_prob = np.ones(B.n)
    h1, _ = oracle_lib.count(A, B, p)
    A.jk_prob = np.zeros(A.n)
    B.jk_prob = np.zeros(B.n)
    h0, _ = oracle_lib.count(A, B, p)
    full = h1.num[0].sum()
    same = h0.num[1:].sum()
    assert h0.num[0].sum() == full
    assert h1.num[1:].sum() == full + (full - same)
