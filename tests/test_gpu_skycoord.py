"""GPU parity of the other sky coordinates (1 r-mu, 2 rp-ephi, 3 projected, 4 angular-ephi,
5 angular projected, and p_max > 600 with 0): the HIP engine through the C ABI against the CPU
oracle, each pair kind, both kernel precisions, the auto pass and jackknife.

Counts must be bit-identical; weighted sums within helpers.RTOL.  These coordinates take the
engine's FP64 classification for every candidate; pairs whose rp, r, mu or ephi lies within the
device/host difference bound of a bin edge are binned on the host with its libm.
"""
import numpy as np
import pytest

from corr_pc_amd import synth
from corr_pc_amd.engine import KIND_DD, KIND_SHAPE_DENSITY, KIND_SHAPE_SHAPE, PREC_FP64, PREC_SPLIT, make_sky_params
from helpers import assert_hist_match, sky_catalogs

pytestmark = pytest.mark.gpu

PRECS = [PREC_SPLIT, PREC_FP64]
_TABLE = synth.distance_table(0.3, 1e-4, 1.0)
_H_ZMAX = _TABLE["H"][int(np.ceil(1.0 / _TABLE["dz"]))]


def _table(z_sep_max):
    return _TABLE["H"], _TABLE["dA"], _TABLE["dz"], 1.0, z_sep_max


def _zero_z(c):
    """Coordinates 4 and 5 read no redshift (read_dat.cpp:27): z = 0 and the table's first row."""
    c = c.subset(np.arange(c.n))
    c.z = np.zeros(c.n)
    c.dc = np.full(c.n, _TABLE["dC"][0])
    c.da = np.full(c.n, _TABLE["dA"][0])
    return c


def _params(coordinates, kind, gws, est, n_jk, prec):
    kw = dict(kind=kind, g_with_shape=gws, estimator=est, which_corr={KIND_DD: 0, KIND_SHAPE_DENSITY: 1,
                                                                          KIND_SHAPE_SHAPE: 2}[kind],
              do_jk=1, n_jk=n_jk, precision=prec, coordinates=coordinates)
    half_pi = np.pi / 2 + 1e-9
    if coordinates == 1:
        return make_sky_params(_table(60.0 * _H_ZMAX / 300000.0 + 0.001), 12, 0.5, 60.0, 10, -1.0, 1.0, **kw)
    if coordinates == 2:
        return make_sky_params(_table(0.05), 12, 0.5, 60.0, 8, 0.0, half_pi, z_min=0.25, z_sep_min=-0.05, **kw)
    if coordinates == 3:
        return make_sky_params(_table(0.04), 12, 0.5, 60.0, 2, -1.0, 5.0, z_min=0.2, **kw)
    if coordinates == 4:
        return make_sky_params(_table(9.0), 12, 0.02, 3.0, 8, 0.0, half_pi, **kw)
    if coordinates == 5:
        return make_sky_params(_table(9.0), 12, 0.02, 3.0, 1, 0.0, 1.0, **kw)
    # coordinates 0 with p_max > 600: projected rp, pi = 0
    return make_sky_params(_table(800.0 * _H_ZMAX / 300000.0 + 0.001), 12, 0.5, 60.0, 4, -800.0, 800.0, **kw)


def _run(engine, oracle_lib, outer, inner, p, what):
    print(f"[case] {what}", flush=True)
    ref, rst = oracle_lib.count(outer, inner, p)
    got, st = engine.count(outer, inner, p)
    assert_hist_match(got, ref, what=what)
    assert st["n_binned"] == rst["n_binned"] > 0
    return got, st


CASES = [(1, KIND_DD, 2, 0), (1, KIND_SHAPE_DENSITY, 2, 0), (2, KIND_DD, 2, 0), (2, KIND_SHAPE_SHAPE, 2, 1),
         (2, KIND_SHAPE_DENSITY, 1, 0), (3, KIND_SHAPE_DENSITY, 2, 0), (4, KIND_DD, 2, 0),
         (4, KIND_SHAPE_SHAPE, 2, 1), (5, KIND_SHAPE_SHAPE, 2, 1), ("600", KIND_DD, 2, 0)]


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("coordinates,kind,gws,est", CASES)
def test_sky_coordinates(engine, oracle_lib, prec, coordinates, kind, gws, est):
    A, B, n_jk, _ = sky_catalogs(4000, 6000, 71, shapes_in=True, table=_TABLE)
    if kind == KIND_SHAPE_SHAPE or gws == 1:
        A.e1, A.e2, A.phi = B.e1[: A.n].copy(), B.e2[: A.n].copy(), B.phi[: A.n].copy()
    if coordinates in (4, 5):
        A, B = _zero_z(A), _zero_z(B)
    c = 0 if coordinates == "600" else coordinates
    p = _params(c, kind, gws, est, n_jk, prec)
    _run(engine, oracle_lib, A, B, p, f"sky coordinates={coordinates} kind={kind} gws={gws} est={est} prec={prec}")


@pytest.mark.parametrize("coordinates", [1, 2, 4])
def test_sky_coordinates_auto(engine, oracle_lib, coordinates):
    A, _, n_jk, _ = sky_catalogs(5000, 10, 72, table=_TABLE)
    if coordinates == 4:
        A = _zero_z(A)
    A.phi = np.random.default_rng(73).uniform(-1.5, 1.5, A.n)
    p = _params(coordinates, KIND_DD, 1, 0, n_jk, PREC_SPLIT)
    p.auto_pairs = 1
    _run(engine, oracle_lib, A, A, p, f"sky auto coordinates={coordinates}")


def test_sky_coordinates_rejected(engine):
    from corr_pc_amd.engine import EngineError
    A, B, n_jk, _ = sky_catalogs(200, 200, 74, table=_TABLE)
    for bad in (6, 7, -1):
        p = _params(0, KIND_DD, 2, 0, n_jk, PREC_SPLIT)
        p.coordinates = bad
        with pytest.raises(EngineError):
            engine.count(A, B, p)
    p = _params(0, KIND_DD, 2, 0, n_jk, PREC_SPLIT)
    p.r_mu = 1
    with pytest.raises(EngineError):
        engine.count(A, B, p)


@pytest.mark.parametrize("prec", PRECS)
def test_sky_coordinates_variants(engine, oracle_lib, prec):
    """Linear bins without jackknife and with subsampled randoms (coordinates 1); the projected rp
    with ephi bins (coordinates 2 with p_max > 600); shape_shape's estimator-1 branch with the
    projected coordinates 3."""
    A, B, n_jk, _ = sky_catalogs(3000, 5000, 75, shapes_in=True, inc=True, table=_TABLE)
    A.e1, A.e2, A.phi = B.e1[: A.n].copy(), B.e2[: A.n].copy(), B.phi[: A.n].copy()
    p = make_sky_params(_table(50.0 * _H_ZMAX / 300000.0 + 0.001), 10, 0.0, 50.0, 8, -1.0, 1.0, coordinates=1,
                        kind=KIND_DD, lin_bin=1, do_jk=0, include_prob_max=0.6, precision=prec)
    _run(engine, oracle_lib, A, B, p, f"sky r-mu linear, no jk, include_prob prec={prec}")
    p = make_sky_params(_table(0.05), 12, 0.5, 60.0, 4, 0.0, 700.0, coordinates=2, kind=KIND_SHAPE_DENSITY,
                        g_with_shape=2, which_corr=1, do_jk=1, n_jk=n_jk, precision=prec)
    _run(engine, oracle_lib, A, B, p, f"sky rp-ephi projected prec={prec}")
    p = make_sky_params(_table(0.04), 12, 0.5, 60.0, 2, -1.0, 5.0, coordinates=3, kind=KIND_SHAPE_SHAPE,
                        g_with_shape=2, estimator=1, which_corr=2, do_jk=1, n_jk=n_jk, z_min=0.2, precision=prec)
    _run(engine, oracle_lib, A, B, p, f"sky projected shape_shape est 1 prec={prec}")


def test_angular_ephi_sums_to_angular_projected_at_scale(engine):
    """A size-independent check on 2e5 x 3e5 galaxies: coordinates 4 bins the same pairs as
    coordinates 5 (same rp in degrees, no z test), spread over ephi bins covering [0, pi/2]; summed
    over pi, its counts equal coordinates 5's one pi bin exactly, whichever pairs the guard bands
    sent to the host."""
    A, B, n_jk, _ = sky_catalogs(200_000, 300_000, 76, shapes_in=True, table=_TABLE)
    A, B = _zero_z(A), _zero_z(B)
    p4 = _params(4, KIND_DD, 2, 0, n_jk, PREC_SPLIT)
    p5 = _params(5, KIND_DD, 2, 0, n_jk, PREC_SPLIT)
    print("[case] sky coordinates 4 vs 5 at 2e5 x 3e5", flush=True)
    h4, st4 = engine.count(A, B, p4)
    h5, st5 = engine.count(A, B, p5)
    assert st4["n_binned"] == st5["n_binned"] > 1e8
    assert np.array_equal(h4.num.sum(axis=2), h5.num[..., 0])
    np.testing.assert_allclose(h4.wt_num.sum(axis=2), h5.wt_num[..., 0], rtol=1e-10)
