"""GPU end-to-end test of the output side: the four pair sets of the plain wgg box (SD, SRd, DRs,
RsRd passes, do_corr.cpp:110-131) counted by the engine, then the cross branch's estimator and files
(cpc_write_cross), against the same files written from the oracle's histograms.

Integer columns must be identical; every other column agrees to the six printed digits (the
histograms' weighted sums differ by at most 1e-10 relative)."""
import os

import numpy as np
import pytest

from corr_pc_amd import synth
from corr_pc_amd.engine import PREC_FP64, PREC_SPLIT, Norm, weight_sum, write_cross
from corr_pc_amd.passes import config_datainfo, setup_pass

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec", [PREC_SPLIT, PREC_FP64])
def test_config0_files_match_oracle(engine, oracle_lib, tmp_path, prec):
    cats = synth.catalogs(0)
    di = config_datainfo(cats["config"])
    S, Rs, D, Rd = cats["S"], cats["Rs"], cats["D"], cats["Rd"]
    sets = {"SD": (D, S), "SRd": (Rd, S), "DRs": (D, Rs), "RsRd": (Rd, Rs)}
    got, ref = {}, {}
    p0 = None
    for name, (g1, g2) in sets.items():
        outer, inner, p, _ = setup_pass(name, g1, g2, di, precision=prec)
        print(f"[case] config0 {name} prec={prec}", flush=True)
        got[name], _ = engine.count(outer, inner, p)
        ref[name], _ = oracle_lib.count(outer, inner, p)
        if name == "SD":
            p0 = p
    norm = Norm(weight_sum(S), weight_sum(D), weight_sum(Rs), weight_sum(Rd), di.include_prob_RR)
    os.makedirs(tmp_path / "gpu")
    os.makedirs(tmp_path / "cpu")
    n_g = write_cross(p0, norm, got["SD"], got["SRd"], got["DRs"], got["RsRd"], str(tmp_path / "gpu") + "/")
    n_c = write_cross(p0, norm, ref["SD"], ref["SRd"], ref["DRs"], ref["RsRd"], str(tmp_path / "cpu") + "/")
    assert n_g == n_c == 2 + 8 + 10 * di.n_jk + 2
    names = sorted(os.listdir(tmp_path / "cpu"))
    assert names == sorted(os.listdir(tmp_path / "gpu"))
    for x in names:
        a = np.loadtxt(tmp_path / "gpu" / x, comments="#")
        b = np.loadtxt(tmp_path / "cpu" / x, comments="#")
        assert a.shape == b.shape, x
        np.testing.assert_allclose(a, b, rtol=2e-6, atol=1e-12, err_msg=x)
        count_col = 2 if x.startswith("bins2D") else 1
        assert np.array_equal(a[:, count_col], b[:, count_col]), x
    final = np.loadtxt(tmp_path / "gpu" / "bins_cross_final.dat", comments="#")
    assert np.all(np.isfinite(final)) and np.abs(final[:, 2]).max() > 0
