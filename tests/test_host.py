"""CPU tests of the host side: the C ABI exports, the rp thresholds against the host libm, the coin
stream against glibc's published generator, the corels2 pass mirror and the synthetic workloads.
No test here launches a kernel."""
import ctypes
import json
import math
import os
import re

import numpy as np
import pytest

from corr_pc_amd import synth
from corr_pc_amd.engine import (EXPORTED_SYMBOLS, KIND_DD, KIND_SHAPE_DENSITY, KIND_SHAPE_SHAPE, LIB_PATH, Catalog,
                                bin_edges, box_rr, make_params, rand_stream, rp_thresholds)
from corr_pc_amd.passes import DataInfo, config_datainfo, setup_pass

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def test_library_exports_every_declared_symbol():
    header = open(os.path.join(ROOT, "include", "cpc.h")).read()
    declared = set(re.findall(r"^\w[\w\s\*]*?\b(cpc_\w+)\(", header, re.M))
    assert declared == set(EXPORTED_SYMBOLS), declared ^ set(EXPORTED_SYMBOLS)
    lib = ctypes.CDLL(LIB_PATH)
    for name in declared:
        assert hasattr(lib, name), name


def test_rand_stream_matches_glibc_generator():
    """cpc_rand_stream (the host rand()) against glibc's TYPE_3 generator restated in make_golden.py,
    through the committed fixture."""
    fx = json.load(open(os.path.join(GOLDEN, "glibc_rand_124693517.json")))
    expect = np.array(fx["rand"], np.float64) / (fx["RAND_MAX"] + 1.0)
    assert np.array_equal(rand_stream(len(expect)), expect)
    # skip is the same stream
    assert np.array_equal(rand_stream(10, skip=20), expect[20:30])


@pytest.mark.parametrize("name", ["config0", "config1", "config4", "linear"])
def test_rp_thresholds(oracle_lib, name):
    """thr[k] is the first double at which the reference's rp bin index (the oracle's log10/sqrt
    restatement) reaches k, the index is monotone for 4096 doubles either side, the values agree
    with the closed form rp_k = rmin*10^(k*rsize), and they equal the committed fixture."""
    fx = json.load(open(os.path.join(GOLDEN, "rp_thresholds.json")))[name]
    b = fx["binning"]
    p = make_params(1000.0, b["n_rp"], b["rp_min"], b["rp_max"], 10, -10.0, 10.0, lin_bin=b["lin_bin"])
    thr = rp_thresholds(p)
    assert [float(t).hex() for t in thr] == fx["thresholds_hex"]
    if b["lin_bin"]:
        rsize = (b["rp_max"] - b["rp_min"]) / b["n_rp"]
    else:
        rsize = (math.log10(b["rp_max"]) - math.log10(b["rp_min"])) / b["n_rp"]
    for k, t in enumerate(thr):
        rp_k = b["rp_min"] + k * rsize if b["lin_bin"] else b["rp_min"] * 10 ** (k * rsize)
        assert math.sqrt(t) == pytest.approx(rp_k, rel=1e-13, abs=1e-300)
        if t > 0:
            assert oracle_lib.rp_bin(p, t) >= k
            assert oracle_lib.rp_bin(p, np.nextafter(t, 0)) < k
        u = np.float64(t)
        lo, hi = u, u
        for _ in range(4096):
            lo, hi = np.nextafter(lo, -1.0), np.nextafter(hi, np.inf)
            if lo >= 0:
                assert oracle_lib.rp_bin(p, lo) < k
            assert oracle_lib.rp_bin(p, hi) >= k


def test_pi_bin_edges(oracle_lib):
    p = make_params(500.0, 20, 0.1, 100.0, 40, -100.0, 100.0)
    assert oracle_lib.pi_bin(p, -100.0) == 0
    assert oracle_lib.pi_bin(p, np.nextafter(-100.0, -200)) == -1
    assert oracle_lib.pi_bin(p, 100.0) == 40
    # the double just below p_max: pi - p_min = 200 - 2^-46 exactly halfway between two doubles,
    # rounds to even (200), so the reference's formula puts it out of range as well
    assert oracle_lib.pi_bin(p, np.nextafter(100.0, 0)) == 40
    assert oracle_lib.pi_bin(p, 100.0 - 1e-13) == 39


def _cat(n, shapes=False):
    z = np.sort(np.random.default_rng(n).uniform(0, 10, n))
    e = np.zeros(n) if shapes else None
    return Catalog(x=np.zeros(n) + 1, y=np.zeros(n) + 1, z=z, e1=e, e2=e)


def test_pass_roles_and_swap():
    """corels2's role table and larger-catalog-outer rule (corr.cpp:226-296), choose_corel_func
    (initialization.cpp:490-543)."""
    di = DataInfo(which_corr=2, estimator=0, box_size=100.0, n_rp=5, rp_min=0.1, rp_max=10.0, n_pi=4, pi_min=-8.0,
                  pi_max=8.0, do_jk=1, n_jk=8, include_prob_RR=0.3)
    D, S, R = _cat(100, shapes=True), _cat(100, shapes=True), _cat(300)
    o, i, p, sw = setup_pass("SD", D, S, di)  # corels2(gdensity, gshape), equal sizes: no swap
    assert (o is D, i is S, sw, p.g_with_shape, p.kind) == (True, True, False, 2, KIND_SHAPE_SHAPE)
    # a shape kind on a catalog without shapes is a reference read of an unset pointer: density pairs
    assert setup_pass("SD", _cat(100), S, di)[2].kind == KIND_DD
    o, i, p, sw = setup_pass("SRd", R, S, di)  # randoms larger: stay outer
    assert (o is R, i is S, sw, p.g_with_shape, p.kind) == (True, True, False, 2, KIND_SHAPE_DENSITY)
    R2 = _cat(50)
    o, i, p, sw = setup_pass("SRd", R2, S, di)  # shapes larger: swapped, shape now on the outer side
    assert (o is S, i is R2, sw, p.g_with_shape) == (True, True, True, 1)
    o, i, p, sw = setup_pass("RsRd", R, R2, di)
    assert p.kind == KIND_DD and p.include_prob_max == 0.3
    assert setup_pass("SD", _cat(0), S, di) is None  # corels2 returns 0 for an empty catalog
    di.which_corr = 0
    assert setup_pass("SD", D, S, di)[2].kind == KIND_DD
    di.which_corr = 1
    assert setup_pass("SD", D, S, di)[2].kind == KIND_SHAPE_DENSITY


def test_synthetic_config0_and_coins():
    cats = synth.catalogs(0)
    c = cats["config"]
    S, D, Rs, Rd = cats["S"], cats["D"], cats["Rs"], cats["Rd"]
    assert S.n == D.n == 20_000 and Rs.n == Rd.n == 40_000
    for cat in (S, D, Rs, Rd):
        assert np.all(np.diff(cat.z) >= 0)
        assert cat.x.min() >= 0 and cat.x.max() < c.box
        assert set(np.unique(cat.jk)) == set(range(8))
    # the coin stream in read order S, Rs (jk_prob, include_prob), D, Rd
    stream = rand_stream(2 * 20_000 + 4 * 40_000)
    assert np.array_equal(S.jk_prob, stream[:20_000])
    assert np.array_equal(Rs.jk_prob, stream[20_000:100_000:2])
    assert np.array_equal(Rs.include_prob, stream[20_001:100_000:2])
    assert np.array_equal(D.jk_prob, stream[100_000:120_000])
    assert np.array_equal(S.x, D.x) and not np.array_equal(S.jk_prob, D.jk_prob)


def test_synthetic_config1_shapes_small():
    data = synth.box_data(1, n=100_000)
    e = np.hypot(data["e1"], data["e2"])
    assert e.max() < 1.0 and 0.3 < e.mean() < 0.33
    assert data["w"].mean() == pytest.approx(1.0, rel=1e-12)
    assert np.all(np.abs(data["phi"]) <= math.pi / 2)
    # clumped galaxies: 30% sit within a few sigma of one of 5000 centres -> excess small-scale pairs
    assert np.all(np.diff(data["z"]) >= 0)
    di = config_datainfo(synth.CONFIGS[1])
    assert (di.n_jk, di.pi_min, di.pi_max, di.n_rp, di.n_pi) == (27, -100.0, 100.0, 20, 40)


@pytest.mark.parametrize("lin", [0, 1])
def test_bin_edges_fp32_stepping(lin):
    """Output edges as bin_log / bin_lin / P_bin_ini step them (initialization.cpp:357-444): FP32 running
    sums, log edges as pow(10, running FP32 log10); restated here with numpy float32."""
    p = make_params(250.0, 10, 0.0 if lin else 0.1, 30.0, 20, -60.0, 60.0, lin_bin=lin)
    re_, pe = bin_edges(p)
    f32 = np.float32
    if lin:
        b = f32((30.0 - 0.0) / float(f32(10)))
        t = f32(0.0)
        expect = [float(t)]
        for _ in range(10):
            t = f32(t + b)
            expect.append(float(t))
    else:
        d = f32(math.log10(0.1))
        b = f32((f32(math.log10(30.0)) - d) / f32(10))
        expect = [10.0 ** float(d)]
        for _ in range(10):
            d = f32(d + b)
            expect.append(10.0 ** float(d))
    assert re_.tolist() == expect
    ps = f32(120.0 / 20)
    t = f32(-60.0)
    pexp = [float(t)]
    for _ in range(20):
        t = f32(t + ps)
        pexp.append(float(t))
    assert pe.tolist() == pexp
    # the edges are close to the exact ones: log edges within FP32 rounding of the exponent
    assert np.allclose(re_, np.linspace(0, 30, 11) if lin else np.logspace(-1, math.log10(30), 11), rtol=1e-5)


def test_box_rr_analytic():
    """RR_PB_calc (PB_calcs.cpp:63-85): cylinder-shell volume over the box volume; the sum over all cells
    is pi (rmax^2 - rmin^2)(pmax - pmin)/L^3 on the same edges."""
    L = 250.0
    p = make_params(L, 10, 0.1, 30.0, 20, -60.0, 60.0)
    re_, pe = bin_edges(p)
    rr = box_rr(p)
    assert rr.shape == (10, 20) and np.all(rr > 0)
    k, q = 7, 3
    assert rr[k, q] == math.pi * (re_[k + 1] ** 2 - re_[k] ** 2) * (pe[q + 1] - pe[q]) / L ** 3
    total = math.pi * (re_[-1] ** 2 - re_[0] ** 2) * (pe[-1] - pe[0]) / L ** 3
    assert rr.sum() == pytest.approx(total, rel=1e-12)


def test_batch_histograms_share_one_buffer_without_overlap():
    """Histogram.many: per pass the shapes of Histogram.for_params, C-contiguous views whose address
    ranges (the pointers c_struct hands to the library) are disjoint and zeroed."""
    from corr_pc_amd.engine import Histogram
    ps = [make_params(100.0, 5, 0.1, 10.0, 4, -4.0, 4.0, do_jk=1, n_jk=8),
          make_params(100.0, 3, 0.1, 10.0, 6, -6.0, 6.0),
          make_params(100.0, 7, 0.1, 10.0, 2, -2.0, 2.0, do_jk=1, n_jk=27)]
    hs = Histogram.many(ps)
    spans = []
    for p, h in zip(ps, hs):
        ref = Histogram.for_params(p)
        for name in ("num", "wt_num", "val", "err", "num_rp"):
            a, b = getattr(h, name), getattr(ref, name)
            assert a.shape == b.shape and a.dtype == b.dtype and a.flags["C_CONTIGUOUS"]
            assert not a.any()
            spans.append((a.ctypes.data, a.ctypes.data + a.nbytes))
        c = h.c_struct()
        assert ctypes.cast(c.num, ctypes.c_void_p).value == h.num.ctypes.data
        assert ctypes.cast(c.err, ctypes.c_void_p).value == h.err.ctypes.data
        assert ctypes.cast(c.num_rp, ctypes.c_void_p).value == h.num_rp.ctypes.data
    spans.sort()
    assert all(e <= s for (_, e), (s, _) in zip(spans, spans[1:]))
