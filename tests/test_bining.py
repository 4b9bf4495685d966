"""CPU tests of cpc_do_bining, the restatement of do_bining (do_corr.cpp:203-433): the auto branch
(final_calc_auto and re_bin, bins_calcs.cpp:54-91, 465-527), estimator 4 (final_calc_mean_projected,
bins_calcs.cpp:258-329), the .temp files, and the cross branch shared with cpc_write_cross.

Pinned by:
  * the auto Landau-Szalay form (SS - 2 SR + RR)/RR with the reference's pair normalisations S^2/2,
    S Rs and Rs^2/2 (Rs^2 when RR came from the ordered RsRd-same pass), on hand-made histograms;
  * a Poisson pin of the auto normalisation in a periodic box: one uniform catalog's unordered pairs,
    counted by the oracle's auto pass with pi = |dz|, give w = 2(N-1)/N - 1 on the positive pi side
    and, folded by re_bin (RR folded too), w = (N-1)/N - 1, only if SS is normalised by S^2/2 and RR
    is the analytic volume fraction;
  * re_bin's fold on symmetric pi bins (bin k of the negative side lands on bin n_pi - 1 - k, counts
    add exactly, the negative side empties) and on an asymmetric layout whose edges do not mirror
    (nothing moves);
  * estimator 4's per-row mean: with every cell's field equal to a * wt_num, the row value is
    a * p_size, minus the same of SR;
  * the file set of each branch.
"""
import os

import numpy as np

from corr_pc_amd.engine import (GEOM_SKY, KIND_DD, KIND_SHAPE_DENSITY, Histogram, Norm, WSum, box_rr, do_bining,
                                make_params, weight_sum, write_cross)
from helpers import small_catalog


def _norm(s=10.0, d=10.0, rs=10.0, rd=10.0, inc=1.0):
    return Norm(WSum(s, s), WSum(d, d), WSum(rs, rs), WSum(rd, rd), inc)


def _const(p, w, num=1):
    h = Histogram.for_params(p)
    h.wt_num[...] = w
    h.num[...] = num
    h.num_rp[...] = num * p.n_pi
    return h


def _read(path):
    return np.loadtxt(path, comments="#")


def _sky(p):
    p.geometry = GEOM_SKY
    return p


def test_file_sets(tmp_path):
    """Estimator 3 (auto and cross) with jackknife and re-binning: every file do_bining writes."""
    p = make_params(40.0, 3, 0.5, 8.0, 4, -4.0, 4.0, kind=KIND_DD, estimator=3, do_jk=1, n_jk=2)
    hs = {k: _const(p, 1.0) for k in ("SS", "SRs", "RsRs", "SD", "SRd", "DRs", "RsRd")}
    n, _ = do_bining(p, _norm(), hs, str(tmp_path) + "/o_", do_auto_corr=1, do_cross_corr=1, do_auto_rebin=1)
    names = ["bins_auto_SR.temp", "bins_cross_SR.temp", "bins_auto_SR2.temp"]
    for br, sets in (("auto", ("SS", "SR", "RR", "final")), ("cross", ("SD", "SR", "DR", "RR", "final"))):
        for t in sets:
            names += [f"bins_{br}_{t}.dat", f"bins2D_{br}_{t}.dat"]
            for r in range(2):
                names += [f"bins_{br}_jk{r}_{t}.dat", f"bins2D_{br}_jk{r}_{t}.dat"]
        names += [f"bins_{br}_jk_final.dat", f"bins2D_{br}_jk_final.dat"]
    assert n == len(names)
    assert sorted(os.listdir(tmp_path)) == sorted("o_" + x for x in names)
    # cross only: the two .temp files and the cross branch
    n, _ = do_bining(p, _norm(), hs, str(tmp_path) + "/c_")
    assert n == 2 + 10 + 10 * 2 + 2


def test_cross_branch_matches_write_cross(tmp_path):
    p = make_params(40.0, 3, 0.5, 8.0, 4, -4.0, 4.0, kind=KIND_DD, do_jk=1, n_jk=3)
    rng = np.random.default_rng(5)
    hs = {}
    for k in ("SD", "SRd", "DRs", "RsRd"):
        h = Histogram.for_params(p)
        h.num[...] = rng.integers(1, 50, h.num.shape)
        h.wt_num[...] = h.num * rng.uniform(0.5, 2.0, h.num.shape)
        h.num_rp[...] = h.num.sum(axis=2)
        hs[k] = h
    nm = _norm(s=7.0, d=9.0, rs=30.0, rd=40.0)
    do_bining(p, nm, hs, str(tmp_path) + "/a_")
    write_cross(p, nm, hs["SD"], hs["SRd"], hs["DRs"], hs["RsRd"], str(tmp_path) + "/b_")
    for x in sorted(os.listdir(tmp_path)):
        if x.startswith("b_"):
            assert open(tmp_path / x).read() == open(tmp_path / ("a_" + x[2:])).read(), x


def test_auto_closed_forms(tmp_path):
    """Counted RR (sky geometry): w = (SS/SS_wt - 2 SR/SR_wt + RR/RR_wt) / (RR/RR_wt) with SS_wt = S^2/2,
    SR_wt = S Rs and RR_wt = Rs^2/2, or Rs^2 when RsRd_same."""
    p = _sky(make_params(40.0, 3, 1.0, 8.0, 4, -4.0, 4.0, kind=KIND_DD, estimator=1))
    S, Rs = 4.0, 10.0
    nm = _norm(s=S, rs=Rs)
    for same, rr_wt in ((0, Rs * Rs / 2), (1, Rs * Rs)):
        for k, ss in enumerate((1.0, 2.0)):
            hs = {"SS": _const(p, ss), "SRs": _const(p, 1.0), "RsRs": _const(p, 1.0)}
            _, pn = do_bining(p, nm, hs, str(tmp_path) + f"/a{same}{k}_", do_auto_corr=1, do_cross_corr=0,
                              rsrd_same=same)
            expect = ss * rr_wt / (S * S / 2) - 2 * rr_wt / (S * Rs) + 1
            w = _read(tmp_path / f"a{same}{k}_bins2D_auto_final.dat")[:, 3]
            np.testing.assert_allclose(w, expect, rtol=1e-5)
            w1 = _read(tmp_path / f"a{same}{k}_bins_auto_final.dat")[:, 2]
            np.testing.assert_allclose(w1, expect * (p.pi_max - p.pi_min), rtol=1e-5)
            assert (pn.SD.val, pn.SRd.val, pn.DRs.val, pn.RsRd.val) == (S * S / 2, S * Rs, S * Rs, rr_wt)


def test_auto_periodic_poisson(tmp_path, oracle_lib):
    """One uniform catalog of N galaxies in a periodic box of side L, unordered pairs from the oracle's
    auto pass (corels1, corr.cpp:103-216).  Each pair is found once, from its lower galaxy, with
    pi = |dz| (periodic) >= 0: the PBz = -1 scan from j = i+1 meets z_{i+1} - L below the window first
    and stops at once (the sorted-mode break, corr.cpp:67-72).  So the positive pi side holds twice the
    density of a signed-pi count and the negative side none.  With SS_wt = N^2/2 and the analytic RR,
    up to Poisson noise:
      not folded:  w = 2 (N-1)/N - 1 for pi > 0 and w = -1 for pi < 0;
      folded:      w = (N-1)/N - 1 for pi > 0 (SS unchanged, RR doubled), the negative side empty."""
    L = 50.0
    N = 4000
    S = small_catalog(N, L, 81, shapes=False, weights=False, jk_side=1)
    p = make_params(L, 8, 0.2, 12.0, 10, -15.0, 15.0, kind=KIND_DD, estimator=1)
    p.auto_pairs = 1
    ss, _ = oracle_lib.count(S, S, p)
    p.auto_pairs = 0
    nm = Norm(weight_sum(S), WSum(0.0, 0.0), WSum(1.0, 1.0), WSum(1.0, 1.0), 1.0)
    f = (N - 1) / N
    for rebin in (0, 1):
        pre = str(tmp_path) + f"/r{rebin}_"
        do_bining(p, nm, {"SS": ss}, pre, do_auto_corr=1, do_cross_corr=0, do_auto_rebin=rebin)
        w = _read(pre + "bins2D_auto_final.dat")
        cnt = _read(pre + "bins2D_auto_SS.dat")[:, 2]
        pi_c = w[:, 1]
        expect = f - 1 if rebin else 2 * f - 1
        ok = (cnt > 300) & (pi_c > 0)
        assert ok.sum() >= 30
        assert np.all(cnt[pi_c < 0] == 0)
        if not rebin:
            assert np.all(w[pi_c < 0, 3] == -1)
        z = (w[ok, 3] - expect) * np.sqrt(cnt[ok]) / (1 + expect)
        assert np.all(np.abs(z) < 5.0), z
        assert abs(np.mean(z)) < 1.0
        rr = _read(pre + "bins2D_auto_RR.dat")
        vol = box_rr(p).ravel()
        if rebin:
            neg = pi_c < 0
            assert np.all(cnt[neg] == 0) and np.all(rr[neg, 2] == 0) and np.all(rr[~neg, 2] == 2)
            np.testing.assert_allclose(rr[~neg, -1], 2 * vol[~neg], rtol=1e-5)
            assert cnt.sum() == ss.num[0].sum()
        else:
            np.testing.assert_allclose(rr[:, -1], vol, rtol=1e-5)


def test_re_bin_symmetric(tmp_path):
    p = _sky(make_params(40.0, 3, 1.0, 8.0, 6, -6.0, 6.0, kind=KIND_DD, estimator=2))
    rng = np.random.default_rng(9)
    ss = Histogram.for_params(p)
    ss.num[...] = rng.integers(0, 1000, ss.num.shape)
    ss.wt_num[...] = rng.uniform(0, 10, ss.num.shape)
    ss.num_rp[...] = ss.num.sum(axis=2)
    do_bining(p, _norm(), {"SS": ss, "SRs": _const(p, 1.0), "RsRs": _const(p, 1.0)}, str(tmp_path) + "/s_",
              do_auto_corr=1, do_cross_corr=0, do_auto_rebin=1)
    rows = _read(tmp_path / "s_bins2D_auto_SS.dat").reshape(p.n_rp, p.n_pi, -1)
    n = ss.num[0]
    np.testing.assert_array_equal(rows[:, :3, 2], 0)
    np.testing.assert_array_equal(rows[:, 3:, 2], n[:, 3:] + n[:, 2::-1])
    np.testing.assert_allclose(rows[:, 3:, -1], ss.wt_num[0][:, 3:] + ss.wt_num[0][:, 2::-1], rtol=1e-5)
    # the raw pre-fold SR goes to bins_auto_SR.temp, the folded one to bins_auto_SR2.temp
    t1 = _read(tmp_path / "s_bins_auto_SR.temp")
    t2 = _read(tmp_path / "s_bins_auto_SR2.temp")
    assert np.all(t1[:, 2] == 1) and set(t2[:, 2]) == {0, 2}


def test_re_bin_unmirrored_edges_left_alone(tmp_path, capfd):
    """pi in [-2, 7) in three bins: the only negative bin [-2, 1] holds its own mirrored centre 0.5 but
    its edges do not mirror, so the reference reports it and moves nothing."""
    p = _sky(make_params(40.0, 2, 1.0, 8.0, 3, -2.0, 7.0, kind=KIND_DD, estimator=2))
    ss = Histogram.for_params(p)
    ss.num[...] = np.arange(p.n_rp * p.n_pi).reshape(1, p.n_rp, p.n_pi) + 1
    ss.wt_num[...] = ss.num
    do_bining(p, _norm(), {"SS": ss, "SRs": _const(p, 1.0), "RsRs": _const(p, 1.0)}, str(tmp_path) + "/u_",
              do_auto_corr=1, do_cross_corr=0, do_auto_rebin=1)
    rows = _read(tmp_path / "u_bins2D_auto_SS.dat")
    np.testing.assert_array_equal(rows[:, 2], ss.num[0].ravel())
    assert "rebinning problem" in capfd.readouterr().out


def test_mean_projected(tmp_path):
    """Estimator 4: each rp row's field j is sum_k val * p_size / sum_k wt_num over SD, minus the same
    over SR; its error column is 0, the (rp, pi) cells of the result stay 0, and npairs is SD's."""
    p = _sky(make_params(40.0, 4, 1.0, 8.0, 5, -5.0, 5.0, kind=KIND_SHAPE_DENSITY, estimator=4))
    rng = np.random.default_rng(11)
    a, b = 0.3, -0.1
    sd = Histogram.for_params(p)
    sr = Histogram.for_params(p)
    for h, c in ((sd, a), (sr, b)):
        h.num[...] = rng.integers(1, 40, h.num.shape)
        h.wt_num[...] = h.num * rng.uniform(0.5, 1.5, h.num.shape)
        h.val[..., 1] = c * h.wt_num
        h.val[..., 6] = 2 * c * h.wt_num
        h.num_rp[...] = h.num.sum(axis=2)
    sd.num[0, 3] = 0  # an empty row in both
    sd.wt_num[0, 3] = 0
    sd.val[0, 3] = 0
    sd.num_rp[0, 3] = 0
    sr.num[0, 3] = 0
    sr.wt_num[0, 3] = 0
    sr.val[0, 3] = 0
    sr.num_rp[0, 3] = 0
    zero = Histogram.for_params(p)
    do_bining(p, _norm(), {"SD": sd, "SRd": sr, "DRs": zero, "RsRd": zero}, str(tmp_path) + "/m_")
    f = _read(tmp_path / "m_bins_cross_final.dat")
    dC = (p.pi_max - p.pi_min) / p.n_pi
    full = np.arange(p.n_rp) != 3
    np.testing.assert_allclose(f[full, 4], (a - b) * dC, rtol=1e-5)
    np.testing.assert_allclose(f[full, 14], 2 * (a - b) * dC, rtol=1e-5)
    assert np.all(f[3, 2:16] == 0)
    assert np.all(f[:, 3::2][:, :7] == 0)  # error columns
    np.testing.assert_array_equal(f[:, 1], sd.num_rp[0])
    np.testing.assert_allclose(f[:, 17], sd.wt_num[0].sum(axis=1), rtol=1e-5)
    f2 = _read(tmp_path / "m_bins2D_cross_final.dat")
    assert np.all(f2[:, 2:] == 0)
    # write_cross takes estimator 4 the same way
    write_cross(p, _norm(), sd, sr, zero, zero, str(tmp_path) + "/w_")
    assert open(tmp_path / "w_bins_cross_final.dat").read() == open(tmp_path / "m_bins_cross_final.dat").read()
