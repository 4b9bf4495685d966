"""GPU parity of the sky geometry (RA, Dec, redshift; coordinates 0): the HIP engine through the
C ABI against the CPU oracle on the same survey subsamples.

Counts must be bit-identical; weighted sums within helpers.RTOL.  Pairs whose rp lies within the
device's error band of a bin edge are binned on the host with the host libm (stats n_host_pairs);
the comparison covers them like every other pair.
"""
import numpy as np
import pytest

from corr_pc_amd import synth
from corr_pc_amd.engine import (JK_LEAVE_ONE_OUT, JK_TOUCH, KIND_DD, KIND_SHAPE_DENSITY, KIND_SHAPE_SHAPE, PREC_FP64,
                                PREC_SPLIT, make_sky_params)
from helpers import assert_hist_match, sky_catalogs

pytestmark = pytest.mark.gpu

PRECS = [PREC_SPLIT, PREC_FP64]
_TABLE = synth.distance_table(0.3, 1e-4, 1.0)
_PT = synth.sky_params_table(2, _TABLE)


def _params(kind, n_jk, **kw):
    """The wg+ survey binning: 20 log rp bins in [0.1, 100), 40 pi bins in [-100, 100)."""
    return make_sky_params(_PT, 20, 0.1, 100.0, 40, -100.0, 100.0, kind=kind, do_jk=1 if n_jk else 0, n_jk=n_jk,
                           **kw)


def _run(engine, oracle_lib, outer, inner, p, what):
    print(f"[case] {what}", flush=True)
    ref, rst = oracle_lib.count(outer, inner, p)
    got, st = engine.count(outer, inner, p)
    assert_hist_match(got, ref, what=what)
    assert st["n_binned"] == rst["n_binned"] > 0
    print(f"  binned {st['n_binned']} host-binned {st['n_host_pairs']} candidates {st['n_candidates']}", flush=True)
    return st


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("kind,est,gws", [(KIND_DD, 0, 2), (KIND_SHAPE_DENSITY, 0, 1), (KIND_SHAPE_DENSITY, 0, 2),
                                          (KIND_SHAPE_SHAPE, 0, 2), (KIND_SHAPE_SHAPE, 1, 1)])
def test_sky_kinds(engine, oracle_lib, prec, kind, est, gws):
    A, B, n_jk, _ = sky_catalogs(6000, 8000, 15, shapes_in=True, table=_TABLE)
    p = _params(kind, n_jk, g_with_shape=gws, estimator=est, precision=prec)
    _run(engine, oracle_lib, A, B, p, f"sky kind={kind} est={est} gws={gws} prec={prec}")


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("jk_mode", [JK_LEAVE_ONE_OUT, JK_TOUCH])
def test_sky_jackknife_modes_and_include(engine, oracle_lib, prec, jk_mode):
    A, B, n_jk, _ = sky_catalogs(5000, 7000, 16, inc=True, table=_TABLE)
    p = _params(KIND_SHAPE_DENSITY, n_jk, g_with_shape=1, include_prob_max=0.7, jk_mode=jk_mode, precision=prec)
    _run(engine, oracle_lib, A, B, p, f"sky jk_mode={jk_mode} prec={prec}")


@pytest.mark.parametrize("prec", PRECS)
def test_sky_self_pass_linear_bins(engine, oracle_lib, prec):
    """One catalog on both sides (the randoms' RR-like pass), linear rp bins from 0, no jackknife."""
    A, _, _, _ = sky_catalogs(6000, 10, 17, table=_TABLE)
    p = make_sky_params(_PT, 10, 0.0, 50.0, 10, -50.0, 50.0, kind=KIND_DD, lin_bin=1, precision=prec)
    _run(engine, oracle_lib, A, A, p, f"sky self lin prec={prec}")


def test_sky_large_split(engine, oracle_lib):
    """60k data x 120k randoms, wg+ pass with 100 jackknife regions."""
    A, B, n_jk, _ = sky_catalogs(60000, 120000, 18, table=_TABLE)
    p = _params(KIND_SHAPE_DENSITY, n_jk, g_with_shape=1, which_corr=1, precision=PREC_SPLIT)
    st = _run(engine, oracle_lib, A, B, p, "sky large split")
    assert st["n_host_pairs"] < 1e-4 * st["n_binned"] + 100


def _dec_sorted(cat):
    o = np.argsort(cat.y, kind="stable")
    return cat.subset(o)


def test_sky_dec_sorted_input_counts_every_window_pair_and_flags_the_pointer(engine, oracle_lib):
    """data_sorted 2 (unsorted = 2): catalogs in Dec order.  The counts are every pair of each
    outer galaxy's window (the oracle's unsorted scans).  corels2's forward-only Dec start pointer
    (corr.cpp:356-362) would cut windows short where a later outer galaxy's window begins lower
    (its redshift, hence its Dec half-width, differs): with redshifts spread over 0.2-1 that happens
    and n_sparse_windows says so; with every outer galaxy at one redshift the windows are ordered
    like the galaxies and nothing is flagged."""
    A, B, n_jk, _ = sky_catalogs(3000, 5000, 18, table=_TABLE)
    A, B = _dec_sorted(A), _dec_sorted(B)
    p = _params(KIND_DD, n_jk, precision=PREC_SPLIT)
    p.unsorted = 2
    st = _run(engine, oracle_lib, A, B, p, "sky dec-sorted mixed redshifts")
    assert st["n_sparse_windows"] > 0
    # the same outer galaxies all at the redshift of the first one: ordered windows
    from corr_pc_amd.engine import Catalog
    z0 = np.full(A.n, A.z[0])
    C = Catalog(x=A.x, y=A.y, z=z0, jk=A.jk, jk_prob=A.jk_prob, dc=np.full(A.n, A.dc[0]), da=np.full(A.n, A.da[0]))
    st = _run(engine, oracle_lib, C, B, p, "sky dec-sorted one redshift")
    assert st["n_sparse_windows"] == 0
