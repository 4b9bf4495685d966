"""GPU end-to-end tests of the corr_pc program: an unchanged-format .inp file and catalog files go in,
and every output file the program writes is compared with the same file produced by the test-side
restatement of the program (tests/program_ref.py) with the oracle counting the pairs.

Integer columns must be identical; every other column agrees to the six printed significant digits
(the histograms' weighted sums differ by at most 1e-10 relative)."""
import os
import subprocess

import numpy as np
import pytest

from corr_pc_amd.engine import PREC_FP64, PREC_SPLIT
import program_ref
from test_program import PROG, box_case, sky_case

pytestmark = pytest.mark.gpu


def _compare_dirs(got_prefix, ref_prefix):
    gdir, gpre = os.path.split(got_prefix)
    rdir, rpre = os.path.split(ref_prefix)
    got = sorted(x[len(gpre):] for x in os.listdir(gdir) if x.startswith(gpre) and not x.endswith(".info"))
    ref = sorted(x[len(rpre):] for x in os.listdir(rdir) if x.startswith(rpre))
    assert got == ref, (set(got) ^ set(ref))
    for x in got:
        a = np.loadtxt(got_prefix + x, comments="#", ndmin=2)
        b = np.loadtxt(ref_prefix + x, comments="#", ndmin=2)
        assert a.shape == b.shape, x
        count_col = 2 if (x.startswith("bins2D") or x.endswith(".temp")) else 1
        assert np.array_equal(a[:, count_col], b[:, count_col]), x
        np.testing.assert_allclose(a, b, rtol=2e-6, atol=1e-12, equal_nan=True, err_msg=x)
    return len(got)


def _info(path):
    out = {}
    for line in open(path):
        tok = line.split()
        if len(tok) == 2:
            out[tok[0]] = tok[1]
    return out


def _run_program(inp_path, precision):
    print(f"[case] corr_pc {inp_path} precision={precision}", flush=True)
    r = subprocess.run([PROG, inp_path, "--precision", "fp64" if precision == PREC_FP64 else "split"],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    return r.stdout


CASES = [
    # (geometry, which_corr, estimator, precision, extra .inp settings)
    ("box", 2, 0, PREC_SPLIT, {}),
    ("box", 1, 3, PREC_SPLIT, {}),
    ("box", 0, 0, PREC_FP64, {}),
    ("box", 1, 0, PREC_SPLIT, {"data_sorted": 0, "lin_bin": 1, "bin_r_min": 1.0}),
    # unsorted input with the auto estimators (corels1's unsorted scans), box and sky
    ("box", 2, 3, PREC_SPLIT, {"data_sorted": 0}),
    ("sky", 0, 2, PREC_FP64, {"data_sorted": 0}),
    # Dec-sorted sky input (data_sorted 2) with the auto and cross estimators
    ("sky", 1, 3, PREC_SPLIT, {"data_sorted": 2}),
    # the generator notebook's box run: shape and density samples only, no randoms
    ("box", 2, 0, PREC_SPLIT, {"n_shape_rand": 0, "n_density_rand": 0}),
    ("sky", 1, 0, PREC_SPLIT, {}),
    ("sky", 0, 2, PREC_SPLIT, {}),
    ("sky", 2, 4, PREC_SPLIT, {}),
    ("sky", 0, 0, PREC_SPLIT, {"rand_subsample": 1}),
    # convergence kinds: which_corr 3 (kappa on the shape sample), 4 (kappa on the density sample)
    ("box", 3, 0, PREC_SPLIT, {}),
    ("box", 4, 0, PREC_SPLIT, {"n_shape_rand": 0, "n_density_rand": 0}),
    ("sky", 3, 0, PREC_SPLIT, {}),
    ("sky", 4, 0, PREC_FP64, {}),
    # the ED notebook's box run: which_corr 5 with r-mu bins (coordinates 7), one mu bin in [-1, 1)
    ("box", 5, 0, PREC_SPLIT, {"coordinates": 7, "n_p_bin": 1, "p_min": -1.0, "p_max": 1.0, "bin_r_max": 10.0}),
    # the generator notebook's box: L = 75, |pi| < 75 (a z window twice the box), 20 pi bins
    ("box", 2, 0, PREC_SPLIT, {"L": 75.0, "p_min": -75.0, "p_max": 75.0, "n_p_bin": 20, "bin_r_min": 0.1,
                               "bin_r_max": 10.0, "n_bins": 10}),
    # r-mu bins of the density and IA kinds
    ("box", 2, 0, PREC_SPLIT, {"coordinates": 7, "n_p_bin": 4, "p_min": -1.0, "p_max": 1.0}),
    # the other sky coordinates: r-mu, rp-ephi (with a z window), projected, angular-ephi, angular
    ("sky", 1, 0, PREC_SPLIT, {"coordinates": 1, "n_p_bin": 4, "p_min": -1.0, "p_max": 1.0}),
    ("sky", 2, 0, PREC_SPLIT, {"coordinates": 2, "n_p_bin": 4, "p_min": 0.0, "p_max": 1.6,
                               "z_sep_min": -0.05, "z_sep_max": 0.05}),
    ("sky", 0, 3, PREC_SPLIT, {"coordinates": 3, "n_p_bin": 1, "p_min": 0.0, "p_max": 1.0,
                               "z_sep_max": 0.1}),
    ("sky", 1, 0, PREC_FP64, {"coordinates": 4, "n_p_bin": 4, "p_min": 0.0, "p_max": 1.6, "bin_r_min": 0.02,
                              "bin_r_max": 1.5}),
    ("sky", 2, 4, PREC_SPLIT, {"coordinates": 5, "n_p_bin": 1, "p_min": 0.0, "p_max": 1.0, "bin_r_min": 0.02,
                               "bin_r_max": 1.5}),
    # the HSC notebook's run: projected coordinates 3, estimator 4, no randoms, pmin = pmax = 0 (pi
    # bins of width 1), z_sep -100..100, 10 log rp bins in [0.1, 10)
    ("sky", 1, 4, PREC_SPLIT, {"coordinates": 3, "n_p_bin": 1, "p_min": 0, "p_max": 0, "z_sep_min": -100,
                               "z_sep_max": 100, "bin_r_min": 0.1, "bin_r_max": 10.0, "n_bins": 10,
                               "n_shape_rand": 0, "n_density_rand": 0}),
]


@pytest.mark.parametrize("geom,which,estimator,precision,extra", CASES)
def test_program_matches_restatement(tmp_path, oracle_lib, geom, which, estimator, precision, extra):
    table = None
    if geom == "box":
        path, inp, cols = box_case(tmp_path, which=which, estimator=estimator,
                                   sorted_z=extra.get("data_sorted", 1) == 1, **extra)
    else:
        # which_corr 3-4 never share the shape randoms as density randoms (RsRd_same needs
        # which_corr <= 2, initialization.cpp:189-192): give them density randoms
        n_rd = 1800 if (estimator == 4 or extra.get("rand_subsample") or which >= 3) else 0
        path, inp, cols, table = sky_case(tmp_path, which=which, estimator=estimator, n_rd=n_rd,
                                          sorted_z=extra.get("data_sorted", 1) == 1,
                                          sort_dec=extra.get("data_sorted", 1) == 2, **extra)
    _run_program(path, precision)
    ref_prefix = str(tmp_path / "ref" / "out_")
    os.makedirs(tmp_path / "ref")
    hists, n_files, pn, ws = program_ref.run(inp, cols, oracle_lib.count, ref_prefix, table=table,
                                             precision=precision)
    assert n_files > 0
    n = _compare_dirs(inp["out_file"], ref_prefix)
    assert n == n_files
    info = _info(inp["out_file"] + "inp2.info")
    assert float(info["S_wt"]) == pytest.approx(ws["S"].val, rel=1e-5)
    assert float(info["SD_wt"]) == pytest.approx(pn.SD.val, rel=1e-5, abs=1e-300)
    assert float(info["RsRd_wt"]) == pytest.approx(pn.RsRd.val, rel=1e-5, abs=1e-300)
    final = np.loadtxt(inp["out_file"] + ("bins_cross_final.dat" if estimator != 2 else "bins_auto_final.dat"),
                       comments="#")
    # estimator 4 reports per-pair means of the fields; field 0 (W) is never accumulated by the pair
    # loop, so its e+ column (4) is the one that carries the signal
    col = 4 if estimator == 4 else 2
    assert np.all(np.isfinite(final[:, col])) and np.abs(final[:, col]).max() > 0
