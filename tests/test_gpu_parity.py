"""GPU parity: the HIP engine (through the C ABI) against the CPU oracle on the same inputs.

Counts must be bit-identical; weighted sums within helpers.RTOL (1e-10 of the bin's accumulated
magnitude).  Every case runs both the split-precision and the plain FP64 kernel.
"""
import numpy as np
import pytest

from corr_pc_amd import synth
from corr_pc_amd.compare import field_errors, kind_fields
from corr_pc_amd.engine import (JK_LEAVE_ONE_OUT, JK_TOUCH, KIND_DD, KIND_SHAPE_DENSITY, KIND_SHAPE_SHAPE, PREC_FP64,
                                PREC_SPLIT, make_params)
from corr_pc_amd.passes import config_datainfo, setup_pass
from helpers import assert_hist_match, small_catalog

pytestmark = pytest.mark.gpu

PRECS = [PREC_SPLIT, PREC_FP64]


# the automatic decomposition and a different one (narrower work groups, short items, wider columns):
# results must not depend on how the pairs are scheduled
TUNINGS = [{}, {"col_frac": 0.5, "item_size": 200, "block_threads": 256}]


def _run(engine, oracle_lib, outer, inner, p, what):
    """The engine under each tuning against the oracle."""
    ref, rst = oracle_lib.count(outer, inner, p)
    st = None
    for tun in TUNINGS:
        print(f"[case] {what} tuning={tun}", flush=True)
        engine.set_tuning(**tun)
        try:
            got, st = engine.count(outer, inner, p)
        finally:
            engine.set_tuning()
        assert_hist_match(got, ref, what=f"{what} tuning={tun}")
        assert st["n_binned"] == rst["n_binned"]
        # the stated error of the weighted bins against the oracle (corr_pc's CPU arithmetic)
        fe = field_errors(got, ref, kind_fields(p.kind, p.estimator))
        rel = max((v["max_rel"] for v in fe.values() if isinstance(v, dict)), default=0.0)
        print(f"[err] {what} tuning={tun}: max |d|/|ref| {rel:.2e}, max |d|/magnitude {fe['max_rel_mag_all']:.2e}",
              flush=True)
    return st


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("kind,est,gws", [(KIND_DD, 0, 2), (KIND_SHAPE_DENSITY, 0, 2), (KIND_SHAPE_DENSITY, 0, 1),
                                          (KIND_SHAPE_SHAPE, 0, 2), (KIND_SHAPE_SHAPE, 0, 1),
                                          (KIND_SHAPE_SHAPE, 1, 2)])
def test_small_box_kinds(engine, oracle_lib, prec, kind, est, gws):
    L = 60.0
    A = small_catalog(3000, L, 11, clumps=40)
    B = small_catalog(2500, L, 12, clumps=30)
    p = make_params(L, 8, 0.2, 12.0, 10, -15.0, 15.0, kind=kind, g_with_shape=gws, estimator=est, do_jk=1, n_jk=8,
                    precision=prec)
    _run(engine, oracle_lib, A, B, p, f"kind={kind} est={est} gws={gws} prec={prec}")


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("jk_mode", [JK_LEAVE_ONE_OUT, JK_TOUCH])
def test_jackknife_modes(engine, oracle_lib, prec, jk_mode):
    L = 40.0
    A = small_catalog(2000, L, 21, jk_side=3)
    p = make_params(L, 6, 0.3, 8.0, 8, -10.0, 10.0, kind=KIND_SHAPE_SHAPE, do_jk=1, n_jk=27, precision=prec,
                    jk_mode=jk_mode)
    # same handle on both sides: self pairs are rejected by rp = 0
    _run(engine, oracle_lib, A, A, p, f"jk_mode={jk_mode} prec={prec}")


@pytest.mark.parametrize("prec", PRECS)
def test_no_jackknife_linear_bins(engine, oracle_lib, prec):
    L = 50.0
    A = small_catalog(2500, L, 31, jk_side=1)
    B = small_catalog(1500, L, 32, jk_side=1, shapes=False)
    p = make_params(L, 7, 0.0, 9.0, 9, -9.0, 9.0, kind=KIND_SHAPE_DENSITY, g_with_shape=1, lin_bin=1, do_jk=0,
                    precision=prec)
    _run(engine, oracle_lib, A, B, p, f"lin prec={prec}")


@pytest.mark.parametrize("prec", PRECS)
def test_include_prob_subsample(engine, oracle_lib, prec):
    L = 50.0
    A = small_catalog(2500, L, 41, shapes=False, inc=True)
    B = small_catalog(2000, L, 42, shapes=False, inc=True)
    p = make_params(L, 6, 0.2, 10.0, 8, -12.0, 12.0, kind=KIND_DD, do_jk=1, n_jk=8, include_prob_max=0.4,
                    precision=prec)
    _run(engine, oracle_lib, A, B, p, f"include prec={prec}")


@pytest.mark.parametrize("prec", PRECS)
def test_edge_cases(engine, oracle_lib, prec):
    """Galaxies on the box faces and at exact bin edges, duplicates, one-sided pi range."""
    L = 32.0
    r = np.random.default_rng(5)
    n = 1200
    pos = r.uniform(0, L, (n, 3))
    pos[:50, 0] = 0.0
    pos[50:100, 1] = 0.0
    pos[100:150, 2] = 0.0
    pos[150:200, 0] = np.nextafter(L, 0)
    # exact separations on bin edges: partners at rp = 1, 2, 4 and pi = 0, +-2
    for k, (dx, dz) in enumerate([(1.0, 0.0), (2.0, 2.0), (4.0, -2.0), (0.5, 4.0)]):
        i = 200 + 2 * k
        pos[i] = [10.0, 10.0, 10.0]
        pos[i + 1] = [10.0 + dx, 10.0, 10.0 + dz]
    pos[300:320] = pos[320:340]  # duplicates: rp = 0
    pos = pos[np.argsort(pos[:, 2], kind="stable")]
    from corr_pc_amd.engine import Catalog
    A = Catalog(x=pos[:, 0], y=pos[:, 1], z=pos[:, 2], jk=np.zeros(n, np.int32), jk_prob=r.uniform(0, 1, n))
    for pmin, pmax in [(-8.0, 8.0), (0.0, 8.0), (-6.0, 10.0)]:
        p = make_params(L, 5, 1.0, 8.0, 8, pmin, pmax, kind=KIND_DD, lin_bin=0, do_jk=1, n_jk=1, precision=prec)
        _run(engine, oracle_lib, A, A, p, f"edges pi=[{pmin},{pmax}) prec={prec}")
        p = make_params(L, 8, 0.0, 8.0, 8, pmin, pmax, kind=KIND_DD, lin_bin=1, do_jk=0, precision=prec)
        _run(engine, oracle_lib, A, A, p, f"edges lin pi=[{pmin},{pmax}) prec={prec}")


@pytest.mark.parametrize("prec", PRECS)
def test_empty_and_tiny(engine, oracle_lib, prec):
    L = 20.0
    A = small_catalog(1, L, 51)
    B = small_catalog(3, L, 52)
    p = make_params(L, 4, 0.1, 5.0, 4, -5.0, 5.0, kind=KIND_SHAPE_SHAPE, do_jk=1, n_jk=8, precision=prec)
    _run(engine, oracle_lib, A, B, p, "tiny")
    E = A.subset(np.arange(0))
    got, st = engine.count(E, B, p)
    assert got.num.sum() == 0 and st["n_binned"] == 0


@pytest.mark.parametrize("prec", PRECS)
def test_config0_passes(engine, oracle_lib, prec):
    """The CPU-runnable wgg box at full size: SD (DD), SRd (DR) and RsRd (RR) passes."""
    cats = synth.catalogs(0)
    di = config_datainfo(cats["config"])
    for name, g1, g2 in (("SD", cats["D"], cats["S"]), ("SRd", cats["Rd"], cats["S"]),
                         ("RsRd", cats["Rd"], cats["Rs"])):
        outer, inner, p, _ = setup_pass(name, g1, g2, di, precision=prec)
        _run(engine, oracle_lib, outer, inner, p, f"config0 {name} prec={prec}")


@pytest.mark.parametrize("prec", PRECS)
def test_config4_binning(engine, oracle_lib, prec):
    """The large box's binning (25 log rp bins in [0.1, 200), 60 pi bins in [-150, 150), 4x4x4
    jackknife: 1500 cells, a shape-shape histogram of 138 KB that takes most of a CU's LDS) on a
    40k-galaxy sample of that L = 1000 box."""
    cats = synth.catalogs(4, n_data=40_000, n_rand=10)
    di = config_datainfo(cats["config"])
    outer, inner, p, _ = setup_pass("SD", cats["D"], cats["S"], di, precision=prec)
    st = _run(engine, oracle_lib, outer, inner, p, f"config4 binning prec={prec}")
    assert st["n_binned"] > 1e7


@pytest.mark.parametrize("prec", PRECS)
def test_config1_subsample(engine, oracle_lib, prec):
    """The 1M-galaxy wg+/w++ box: the full inner catalog against every 500th outer galaxy (SD pass,
    shape_shape) and every 1500th random (SRd pass, shape_density)."""
    cats = synth.catalogs(1)
    di = config_datainfo(cats["config"])
    D, S, Rd = cats["D"], cats["S"], cats["Rd"]
    outer, inner, p, _ = setup_pass("SD", D, S, di, precision=prec)
    sub = outer.subset(np.arange(0, outer.n, 500))
    _run(engine, oracle_lib, sub, inner, p, f"config1 SD/500 prec={prec}")
    outer, inner, p, _ = setup_pass("SRd", Rd, S, di, precision=prec)
    sub = outer.subset(np.arange(0, outer.n, 1500))
    _run(engine, oracle_lib, sub, inner, p, f"config1 SRd/1500 prec={prec}")


def test_split_matches_fp64_full_config1(engine):
    """Full 1M x 1M SD pass: the split kernel against the plain FP64 kernel (size-independent check:
    identical counts everywhere, weighted sums within tolerance)."""
    cats = synth.catalogs(1)
    di = config_datainfo(cats["config"])
    outer, inner, p, _ = setup_pass("SD", cats["D"], cats["S"], di, precision=PREC_SPLIT)
    h_split, st_s = engine.count(outer, inner, p)
    p.precision = PREC_FP64
    h_64, st_d = engine.count(outer, inner, p)
    assert_hist_match(h_split, h_64, what="config1 full SD split vs fp64")
    fe = field_errors(h_split, h_64, kind_fields(p.kind, p.estimator))
    print("[err] config1 full SD split vs fp64: " + ", ".join(
        f"{k} {v['max_rel']:.2e}/{v['max_rel_mag']:.2e}" for k, v in fe.items() if isinstance(v, dict)), flush=True)
    assert st_s["n_binned"] == st_d["n_binned"] > 4e10
    # invariants of the jackknife layout: leave-one-out counts never exceed the full sample and the
    # rp-level counts are the pi sums
    assert np.all(h_split.num[1:] <= h_split.num[0][None])
    assert np.array_equal(h_split.num_rp, h_split.num.sum(axis=2))


@pytest.mark.parametrize("prec", PRECS)
def test_pi_window_wider_than_the_box(engine, oracle_lib, prec):
    """The generator notebook's box (L = 75, |pi| < 75, 20 pi bins, rp in [0.1, 10)): the z window
    is twice the box, so every pair is met under two z images; cross and auto passes."""
    L = 75.0
    A = small_catalog(3000, L, 94, clumps=30)
    B = small_catalog(2500, L, 95, clumps=30)
    p = make_params(L, 10, 0.1, 10.0, 20, -75.0, 75.0, kind=KIND_SHAPE_SHAPE, estimator=0, g_with_shape=2, do_jk=1,
                    n_jk=8, precision=prec)
    _run(engine, oracle_lib, A, B, p, f"wide pi window cross prec={prec}")
    p.auto_pairs = 1
    _run(engine, oracle_lib, A, A, p, f"wide pi window auto prec={prec}")


def test_candidate_counters_do_not_depend_on_the_kind(engine):
    """The split classification (candidate test, guard bands) is the same for every pair kind, so one
    pass counted as density, shape-density and shape-shape pairs must report the same candidates and
    binned pairs, and the same FP64 re-evaluations up to the rounding of the log-bin estimate (one
    FMA in the density kernel, a product and a sum in the shape kinds: a candidate at the guard
    band's edge may fall on either side); the FP64 kernel re-evaluates nothing it counts there.  Also
    a regression test for a re-evaluation count that was kept per wave inside a divergent branch
    (it read a small fraction of the true count)."""
    L = 120.0
    A = small_catalog(20000, L, 21, clumps=80)
    B = small_catalog(20000, L, 22, clumps=80)
    seen = {}
    for kind in (KIND_DD, KIND_SHAPE_DENSITY, KIND_SHAPE_SHAPE):
        p = make_params(L, 12, 0.05, 20.0, 16, -20.0, 20.0, kind=kind, g_with_shape=2, do_jk=1, n_jk=8,
                        precision=PREC_SPLIT)
        _, st = engine.count(A, B, p)
        assert st["ms_cells"] > 0 and st["ms_kernel"] > 0  # the cell lists and the kernels are timed
        seen[kind] = (st["n_candidates"], st["n_fp64_reeval"], st["n_binned"])
        print(f"[counters] kind={kind}: {seen[kind]}", flush=True)
    assert len({(c, b) for c, _, b in seen.values()}) == 1, seen
    assert seen[KIND_SHAPE_DENSITY][1] == seen[KIND_SHAPE_SHAPE][1], seen  # the same arithmetic
    re_dd, re_ss = seen[KIND_DD][1], seen[KIND_SHAPE_SHAPE][1]
    assert re_ss > 0  # the guard band is met at this size
    assert abs(re_dd - re_ss) <= max(2, 2e-3 * re_ss), seen
    p = make_params(L, 12, 0.05, 20.0, 16, -20.0, 20.0, kind=KIND_DD, do_jk=1, n_jk=8, precision=PREC_FP64)
    _, st = engine.count(A, B, p)
    assert st["n_candidates"] == seen[KIND_DD][0] and st["n_fp64_reeval"] == 0
