"""CPU tests of the output side (cpc_write_cross): the cross branch's estimator, pi integration,
jackknife summary and .dat files (bins_calcs.cpp, PB_calcs.cpp:63-85, outp.cpp, do_corr.cpp:326-431).

Pinned by:
  * the Landau-Szalay combination's closed forms: (SD - SR - DR + RR)/RR with each term normalised by
    its catalogs' weight sums, on hand-made histograms (w = 0 when all four agree, w = 1 when SD doubles);
  * a Poisson pin of the periodic-box normalisation: two independent uniform catalogs, counted by the
    oracle, give w consistent with 0 in every well-populated (rp, pi) bin only if S*D, the analytic RR
    volume pi (b_max^2 - b_min^2) dpi / L^3 and the FP32-stepped edges are all right;
  * the pi integral of a constant (w * (pi_max - pi_min)) and the textbook delete-one jackknife
    variance sqrt((n-1)/n * sum (w_r - mean)^2);
  * the file set and column headers of do_bining's cross branch, and the six-significant-digit
    stream formatting (C's %g with precision 6).
"""
import os

import numpy as np
import pytest

from corr_pc_amd.engine import (GEOM_SKY, KIND_DD, KIND_SHAPE_SHAPE, Histogram, Norm, WSum, EngineError, bin_edges,
                                box_rr, make_params, weight_sum, write_cross)
from helpers import small_catalog

HEAD1 = ("# rp\tnpairs\twgg\twgg_err\twgp\twgp_err\twgx wgx_err\twpp \twpp_err\t\twxx\twxx_err\t\twpx  wpx_err \ttheta"
         "   theta_err   LS_err     wt_npair    rp_low   rp_high sig_crit")
HEAD2 = ("# rp     pi   npairs     wgg      wgg_err  wgp     wgp_err  wgx  wgx_err     wpp  wpp_err    wxx   wxx_err  "
         "wpx   wpx_err    theta    theta_err  wt_npairs ")


def _norm(s=10.0, d=10.0, rs=10.0, rd=10.0, inc=1.0):
    return Norm(WSum(s, s), WSum(d, d), WSum(rs, rs), WSum(rd, rd), inc)


def _hists(p, *wts):
    out = []
    for w in wts:
        h = Histogram.for_params(p)
        h.wt_num[...] = w
        h.num[...] = 1
        h.num_rp[...] = p.n_pi
        out.append(h)
    return out


def _read2d(path):
    return np.loadtxt(path, comments="#")


def test_weight_sum():
    c = small_catalog(100, 10.0, 1)
    ws = weight_sum(c)
    acc = 0.0
    acc2 = 0.0
    for x in c.w:
        acc += x
        acc2 += x * x
    assert ws.val == acc and ws.err == acc2
    c.w = None
    ws = weight_sum(c)
    assert ws.val == 100 and ws.err == 100


def test_file_set_and_headers(tmp_path):
    p = make_params(40.0, 4, 0.5, 8.0, 6, -6.0, 6.0, kind=KIND_DD, do_jk=1, n_jk=3)
    sd, sr, dr, rr = _hists(p, 2.0, 1.0, 1.0, 1.0)
    n = write_cross(p, _norm(), sd, sr, dr, rr, str(tmp_path) + "/run_")
    names = ["bins_cross_final.dat", "bins2D_cross_final.dat"]
    for t in ("SD", "SR", "DR", "RR"):
        names += [f"bins_cross_{t}.dat", f"bins2D_cross_{t}.dat"]
    for r in range(3):
        for t in ("SD", "RR", "DR", "SR", "final"):
            names += [f"bins_cross_jk{r}_{t}.dat", f"bins2D_cross_jk{r}_{t}.dat"]
    names += ["bins_cross_jk_final.dat", "bins2D_cross_jk_final.dat"]
    assert n == len(names)
    assert sorted(os.listdir(tmp_path)) == sorted("run_" + x for x in names)
    for x in names:
        first = open(tmp_path / ("run_" + x)).readline().rstrip("\n")
        assert first == (HEAD2 if x.startswith("bins2D") else HEAD1), x
        rows = _read2d(tmp_path / ("run_" + x))
        assert rows.shape == ((p.n_rp * p.n_pi, 18) if x.startswith("bins2D") else (p.n_rp, 21)), x


def test_landau_szalay_closed_forms(tmp_path):
    """Sky geometry (counted RR): w = SD/(S D) / RR*(Rs Rd) - SR.. - DR.. + 1."""
    p = make_params(40.0, 3, 1.0, 8.0, 4, -4.0, 4.0, kind=KIND_DD)
    p.geometry = GEOM_SKY
    rng = np.random.default_rng(3)
    X = rng.uniform(1, 3, (1, p.n_rp, p.n_pi))
    cases = [((X, X, X, X), _norm(), 0.0), ((2 * X, X, X, X), _norm(), 1.0), ((3 * X, 2 * X, X, X), _norm(), 1.0),
             # weights S = 4, D = 10, Rs = 25, Rd = 4: pair normalisations SD = S*D = 40, SR = S*Rd = 16,
             # DR = Rs*D = 250, RR = Rs*Rd = 100, so with equal counts w = 100/40 - 100/16 - 100/250 + 1
             ((X, X, X, X), _norm(s=4.0, rd=4.0, rs=25.0), 2.5 - 6.25 - 0.4 + 1.0)]
    for k, (w4, nm, expect) in enumerate(cases):
        hs = _hists(p, *w4)
        write_cross(p, nm, *hs, str(tmp_path) + f"/c{k}_")
        rows = _read2d(tmp_path / f"c{k}_bins2D_cross_final.dat")
        np.testing.assert_allclose(rows[:, 3], expect, rtol=1e-5, atol=1e-6, err_msg=f"case {k}")
        # pi integral of a constant: w * (pi_max - pi_min)
        rows1 = _read2d(tmp_path / f"c{k}_bins_cross_final.dat")
        np.testing.assert_allclose(rows1[:, 2], expect * (p.pi_max - p.pi_min), rtol=1e-5, atol=1e-5)
        assert np.all(rows1[:, 1] == p.n_pi)  # npairs: the SD count summed over pi


def test_periodic_box_poisson(tmp_path, oracle_lib):
    """Two independent uniform catalogs in a periodic box: w(rp, pi) = SD/(S D RR_box) - 1 is 0 up to
    Poisson noise in every bin, and SD's own normalised column is 1."""
    L = 50.0
    S = small_catalog(3000, L, 61, shapes=False, weights=False, jk_side=1)
    D = small_catalog(3000, L, 62, shapes=False, weights=False, jk_side=1)
    p = make_params(L, 8, 0.2, 12.0, 10, -15.0, 15.0, kind=KIND_DD)
    sd, _ = oracle_lib.count(D, S, p)
    zero = Histogram.for_params(p)
    nm = Norm(weight_sum(S), weight_sum(D), WSum(1.0, 1.0), WSum(1.0, 1.0), 1.0)
    write_cross(p, nm, sd, zero, zero, zero, str(tmp_path) + "/pb_")
    rows = _read2d(tmp_path / "pb_bins2D_cross_final.dat")
    n = rows[:, 2]
    sdn = _read2d(tmp_path / "pb_bins2D_cross_SD.dat")
    cnt = sdn[:, 2]
    ok = cnt > 400
    assert ok.sum() >= 40
    z = rows[ok, 3] / (1.0 / np.sqrt(cnt[ok]))
    assert np.all(np.abs(z) < 5.0), z
    assert abs(np.mean(z)) < 1.0
    np.testing.assert_allclose(sdn[ok, 3], 1.0 + rows[ok, 3], rtol=1e-5)
    # RR file: the analytic volume fraction, num = 1
    rr = _read2d(tmp_path / "pb_bins2D_cross_RR.dat")
    np.testing.assert_allclose(rr[:, -1], box_rr(p).ravel(), rtol=1e-5)
    assert np.all(rr[:, 2] == 1) and np.all(n == 0)
    # bin centres: the FP32-stepped edges
    re, pe = bin_edges(p)
    np.testing.assert_allclose(rows[:, 0], np.repeat((re[:-1] + re[1:]) / 2, p.n_pi), rtol=1e-5)
    np.testing.assert_allclose(rows[:, 1], np.tile((pe[:-1] + pe[1:]) / 2, p.n_rp), rtol=1e-5, atol=1e-6)


def test_jackknife_summary(tmp_path):
    """Box geometry: region r's SD is set to (c_r + 1) times the analytic RR times S*D, so its w is c_r;
    the summary holds the full sample's w with the delete-one jackknife error."""
    p = make_params(40.0, 3, 1.0, 8.0, 4, -4.0, 4.0, kind=KIND_DD, do_jk=1, n_jk=4)
    rr = box_rr(p)
    S = D = 10.0
    c_full = 0.3
    c_reg = np.array([0.1, 0.35, 0.2, 0.6])
    sd = Histogram.for_params(p)
    sd.wt_num[0] = (c_full + 1) * rr * S * D
    for r, c in enumerate(c_reg):
        sd.wt_num[r + 1] = (c + 1) * rr * S * D
    sd.num[...] = 1
    zero = Histogram.for_params(p)
    write_cross(p, _norm(s=S, d=D), sd, zero, zero, zero, str(tmp_path) + "/j_")
    jk2 = _read2d(tmp_path / "j_bins2D_cross_jk_final.dat")
    n = len(c_reg)
    err = np.sqrt((n - 1) / n * np.sum((c_reg - c_reg.mean()) ** 2))
    np.testing.assert_allclose(jk2[:, 3], c_full, rtol=1e-5)
    np.testing.assert_allclose(jk2[:, 4], err, rtol=1e-5)
    jk1 = _read2d(tmp_path / "j_bins_cross_jk_final.dat")
    np.testing.assert_allclose(jk1[:, 2], c_full * (p.pi_max - p.pi_min), rtol=1e-5)
    np.testing.assert_allclose(jk1[:, 3], err * (p.pi_max - p.pi_min), rtol=1e-5)
    for r, c in enumerate(c_reg):
        rows = _read2d(tmp_path / f"j_bins2D_cross_jk{r}_final.dat")
        np.testing.assert_allclose(rows[:, 3], c, rtol=1e-5)


def test_shape_fields_linear_in_e(tmp_path, oracle_lib):
    """The shape columns are linear in each side's shapes: negating the inner (shape-side) catalog's
    (e1, e2) flips e+, e++, exx and e+x of the final estimate; negating both sides flips e+ only;
    wgg and the theta column do not depend on the shapes."""
    L = 40.0
    A = small_catalog(1500, L, 71, clumps=20)
    B = small_catalog(1500, L, 72, clumps=20)
    p = make_params(L, 6, 0.3, 9.0, 8, -10.0, 10.0, kind=KIND_SHAPE_SHAPE, g_with_shape=2, estimator=0,
                    which_corr=2, do_jk=1, n_jk=8)

    def neg(c):
        n = c.subset(np.arange(c.n))
        n.e1, n.e2 = -c.e1, -c.e2
        return n

    zero = Histogram.for_params(p)
    nm = Norm(weight_sum(B), weight_sum(A), WSum(1.0, 1.0), WSum(1.0, 1.0), 1.0)
    out = {}
    for tag, (o, i) in {"a": (A, B), "b": (A, neg(B)), "c": (neg(A), neg(B))}.items():
        h, _ = oracle_lib.count(o, i, p)
        write_cross(p, nm, h, zero, zero, zero, str(tmp_path) + f"/{tag}_")
        out[tag] = _read2d(tmp_path / f"{tag}_bins2D_cross_final.dat")
    a, b, c = out["a"], out["b"], out["c"]
    same, flip = [3, 15], [5, 9, 11, 13]  # wgg, theta | e+, e++, exx, e+x (value columns)
    for col in same:
        np.testing.assert_allclose(b[:, col], a[:, col], rtol=1e-5, atol=1e-9)
        np.testing.assert_allclose(c[:, col], a[:, col], rtol=1e-5, atol=1e-9)
    for col in flip:
        np.testing.assert_allclose(b[:, col], -a[:, col], rtol=1e-5, atol=1e-9)
        assert np.abs(a[:, col]).max() > 0
    np.testing.assert_allclose(c[:, 5], -a[:, 5], rtol=1e-5, atol=1e-9)
    for col in (9, 11, 13):
        np.testing.assert_allclose(c[:, col], a[:, col], rtol=1e-5, atol=1e-9)
    assert np.all(a[:, 7] == 0)  # estimator 0 never assigns ex (shape_shape_corel, corels.cpp:140-160)


def test_formatting_six_significant_digits(tmp_path):
    p = make_params(40.0, 2, 1.0, 8.0, 2, -4.0, 4.0, kind=KIND_DD)
    p.geometry = GEOM_SKY
    vals = [1.0 / 3.0, 123456.789, 1.5e-7, 2.0]
    sd, sr, dr, rr = _hists(p, 1.0, 1.0, 1.0, 1.0)
    sd.wt_num[0, 0, 0] = 1 + vals[0]
    sd.wt_num[0, 0, 1] = 1 + vals[1]
    sd.wt_num[0, 1, 0] = 1 + vals[2]
    sd.wt_num[0, 1, 1] = 1 + vals[3]
    write_cross(p, _norm(), sd, sr, dr, rr, str(tmp_path) + "/f_")
    lines = open(tmp_path / "f_bins2D_cross_SD.dat").read().splitlines()[1:]
    for line, v in zip(lines, vals):
        tok = line.split()
        assert tok[3] == "%g" % (1 + v), (tok[3], v)  # SD normalised: sd_wt_num/rr_wt_num * RR_wt/SD_wt
        assert tok[-1] == "%g" % (1 + v)


def test_unsupported_estimator(tmp_path):
    """cpc_write_cross is the cross branch only: the auto estimators 1-3 go through cpc_do_bining."""
    p = make_params(40.0, 2, 1.0, 8.0, 2, -4.0, 4.0, kind=KIND_DD, estimator=1)
    h = _hists(p, 1.0, 1.0, 1.0, 1.0)
    with pytest.raises(EngineError):
        write_cross(p, _norm(), *h, str(tmp_path) + "/u_")
