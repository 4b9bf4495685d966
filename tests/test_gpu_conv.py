"""GPU parity of the convergence kinds (which_corr 3 and 4: density_convergence_corel and
shape_convergence_corel, corels.cpp:79-120): the HIP engine through the C ABI against the CPU
oracle, box and sky, every g_with_shape, both kernel precisions.

Counts must be bit-identical; weighted sums within helpers.RTOL of the bin's accumulated
magnitude, here the sum of |terms| over the full sample (these kinds accumulate no err sums to take
it from).  Kappa travels in e1.
"""
import numpy as np
import pytest

from corr_pc_amd import synth
from corr_pc_amd.engine import (KIND_DENSITY_CONVERGENCE, KIND_SHAPE_CONVERGENCE, PREC_FP64, PREC_SPLIT, make_params,
                                make_sky_params)
from helpers import assert_hist_match, sky_catalogs, small_catalog

pytestmark = pytest.mark.gpu

PRECS = [PREC_SPLIT, PREC_FP64]
_TABLE = synth.distance_table(0.3, 1e-4, 1.0)
_PT = synth.sky_params_table(2, _TABLE)


def _kappa(cat, seed):
    c = cat.subset(np.arange(cat.n))
    c.e1 = np.random.default_rng(seed).normal(0.0, 0.05, cat.n)
    c.e2 = np.zeros(cat.n)
    c.phi = None
    return c


def _sides(kind, gws, shaped_a, shaped_b, seed):
    """(outer, inner): kappa where the kind reads it, shapes on the side shape-convergence rotates."""
    if kind == KIND_DENSITY_CONVERGENCE:
        return (_kappa(shaped_a, seed), shaped_b) if gws == 1 else (shaped_a, _kappa(shaped_b, seed))
    return (shaped_a, _kappa(shaped_b, seed)) if gws == 1 else (_kappa(shaped_a, seed), shaped_b)


def _abs_kappa(cat):
    c = cat.subset(np.arange(cat.n))
    if c.e1 is not None and c.phi is None:  # a kappa catalog
        c.e1 = np.abs(c.e1)
    return c


def _magnitude(oracle_lib, outer, inner, p):
    """Sum over each full-sample cell of |term|: these kinds carry no err sums, whose square root
    gives the other kinds' scale.  |w kappa| from an oracle run on |kappa|; the rotated fields are
    at most that times the largest |e|."""
    a, _ = oracle_lib.count(_abs_kappa(outer), _abs_kappa(inner), p)
    shaped = [c for c in (outer, inner) if c.e1 is not None and c.phi is not None]
    emax = max(float(np.max(np.hypot(c.e1, c.e2))) for c in shaped) if shaped else 1.0
    mag = np.zeros(a.val.shape[1:])
    mag[..., 1] = np.abs(a.val[0, ..., 1])
    mag[..., 3] = mag[..., 5] = mag[..., 1] * emax
    return mag


def _run(engine, oracle_lib, outer, inner, p, what):
    print(f"[case] {what}", flush=True)
    ref, rst = oracle_lib.count(outer, inner, p)
    got, st = engine.count(outer, inner, p)
    assert_hist_match(got, ref, what=what, mag=_magnitude(oracle_lib, outer, inner, p))
    assert st["n_binned"] == rst["n_binned"] > 0
    return got


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("kind", [KIND_DENSITY_CONVERGENCE, KIND_SHAPE_CONVERGENCE])
@pytest.mark.parametrize("gws", [0, 1, 2])
def test_box_convergence(engine, oracle_lib, prec, kind, gws):
    L = 60.0
    A = small_catalog(3000, L, 41, clumps=40)
    B = small_catalog(2500, L, 42, clumps=30)
    outer, inner = _sides(kind, gws, A, B, 43)
    p = make_params(L, 8, 0.2, 12.0, 10, -15.0, 15.0, kind=kind, g_with_shape=gws, do_jk=1, n_jk=8, precision=prec)
    got = _run(engine, oracle_lib, outer, inner, p, f"box kind={kind} gws={gws} prec={prec}")
    if gws == 0:
        assert np.all(got.val == 0)
    else:
        assert np.abs(got.val[0, ..., 1]).max() > 0


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("kind", [KIND_DENSITY_CONVERGENCE, KIND_SHAPE_CONVERGENCE])
def test_sky_convergence(engine, oracle_lib, prec, kind):
    A, B, n_jk, _ = sky_catalogs(6000, 8000, 44, shapes_in=True, table=_TABLE)
    outer, inner = _sides(kind, 2, A, B, 45)
    p = make_sky_params(_PT, 20, 0.1, 100.0, 40, -100.0, 100.0, kind=kind, g_with_shape=2, do_jk=1, n_jk=n_jk,
                        precision=prec)
    _run(engine, oracle_lib, outer, inner, p, f"sky kind={kind} prec={prec}")


def test_convergence_needs_kappa(engine):
    """A kappa side without e1 is refused (the reference would read an unset pointer)."""
    from corr_pc_amd.engine import EngineError
    L = 40.0
    A = small_catalog(500, L, 46, shapes=False)
    B = small_catalog(500, L, 47)
    p = make_params(L, 4, 0.5, 8.0, 4, -8.0, 8.0, kind=KIND_DENSITY_CONVERGENCE, g_with_shape=1)
    with pytest.raises(EngineError):
        engine.count(A, B, p)
