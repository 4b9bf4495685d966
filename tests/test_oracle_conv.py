"""CPU pins of the oracle's convergence kinds (which_corr 3 and 4): density_convergence_corel and
shape_convergence_corel (corels.cpp:79-120).  The convergence kappa travels in e1.

Pinned by:
  * density-convergence with kappa = c on the g_with_shape side: data[1] = c * wt_num in every cell
    and region (the pair value is w c), nothing else accumulates, err stays 0; g_with_shape 0 adds 0;
  * shape-convergence with kappa = 1 on the other side: data[1] = wt_num, and data[3], data[5] equal
    the shape-density kind's w e+ and w ex of the same pairs (its own pins fix the rotation); doubling
    kappa doubles every column;
both in the box and on the sky.
"""
import numpy as np
import pytest

from corr_pc_amd import synth
from corr_pc_amd.engine import (KIND_DENSITY_CONVERGENCE, KIND_SHAPE_CONVERGENCE, KIND_SHAPE_DENSITY, make_params,
                                make_sky_params)
from helpers import sky_catalogs, small_catalog

_TABLE = synth.distance_table(0.3, 1e-4, 1.0)
_PT = synth.sky_params_table(2, _TABLE)


def _box(kind, gws, **kw):
    return make_params(40.0, 6, 0.3, 8.0, 8, -10.0, 10.0, kind=kind, g_with_shape=gws, do_jk=1, n_jk=8, **kw)


def _sky(kind, gws, n_jk, **kw):
    return make_sky_params(_PT, 8, 0.5, 40.0, 10, -50.0, 50.0, kind=kind, g_with_shape=gws, do_jk=1, n_jk=n_jk, **kw)


def _with_kappa(cat, kappa):
    c = cat.subset(np.arange(cat.n))
    c.e1 = np.full(cat.n, kappa, np.float64) if np.isscalar(kappa) else np.asarray(kappa, np.float64)
    c.e2 = np.zeros(cat.n)
    c.phi = None
    return c


def _cases():
    A = small_catalog(1500, 40.0, 31, clumps=20)
    B = small_catalog(1200, 40.0, 32, clumps=20)
    yield "box", A, B, (lambda kind, gws: _box(kind, gws))
    SA, SB, n_jk, _ = sky_catalogs(2500, 3000, 33, shapes_in=True, table=_TABLE)
    yield "sky", SA, SB, (lambda kind, gws: _sky(kind, gws, n_jk))


@pytest.mark.parametrize("gws", [1, 2])
def test_density_convergence_constant_kappa(oracle_lib, gws):
    for geom, A, B, params in _cases():
        c = 0.37
        outer, inner = (_with_kappa(A, c), B) if gws == 1 else (A, _with_kappa(B, c))
        h, st = oracle_lib.count(outer, inner, params(KIND_DENSITY_CONVERGENCE, gws))
        assert st["n_binned"] > 0, geom
        np.testing.assert_allclose(h.val[..., 1], c * h.wt_num, rtol=1e-12, atol=1e-300, err_msg=geom)
        assert np.all(h.val[..., [0, 2, 3, 4, 5, 6]] == 0) and np.all(h.err == 0), geom


def test_density_convergence_gws0_adds_zero(oracle_lib):
    for geom, A, B, params in _cases():
        h, st = oracle_lib.count(_with_kappa(A, 1.0), _with_kappa(B, 1.0), params(KIND_DENSITY_CONVERGENCE, 0))
        assert st["n_binned"] > 0 and np.all(h.val == 0) and np.all(h.err == 0), geom


@pytest.mark.parametrize("gws", [1, 2])
def test_shape_convergence_against_shape_density(oracle_lib, gws):
    for geom, A, B, params in _cases():
        # the g_with_shape side keeps its shapes; the other side carries kappa = 1
        outer, inner = (A, _with_kappa(B, 1.0)) if gws == 1 else (_with_kappa(A, 1.0), B)
        if geom == "sky" and gws == 1:
            continue  # the sky helper gives shapes to the inner catalog only
        sk, st = oracle_lib.count(outer, inner, params(KIND_SHAPE_CONVERGENCE, gws))
        sd, _ = oracle_lib.count(A if gws == 1 else outer, inner if gws == 1 else B, params(KIND_SHAPE_DENSITY, gws))
        assert st["n_binned"] > 0, geom
        assert np.array_equal(sk.num, sd.num), geom
        np.testing.assert_allclose(sk.val[..., 1], sk.wt_num, rtol=1e-12, atol=1e-300, err_msg=geom)
        # the same per-pair products; the oracle's threads take pairs dynamically, so two runs sum
        # them in different orders
        np.testing.assert_allclose(sk.val[..., 3], sd.val[..., 1], rtol=1e-12, atol=1e-12, err_msg=geom)  # w e+
        np.testing.assert_allclose(sk.val[..., 5], sd.val[..., 2], rtol=1e-12, atol=1e-12, err_msg=geom)  # w ex
        assert np.all(sk.val[..., [2, 4, 6]] == 0) and np.all(sk.err == 0), geom
        # linear in kappa
        outer2, inner2 = (A, _with_kappa(B, 2.0)) if gws == 1 else (_with_kappa(A, 2.0), B)
        sk2, _ = oracle_lib.count(outer2, inner2, params(KIND_SHAPE_CONVERGENCE, gws))
        np.testing.assert_allclose(sk2.val, 2 * sk.val, rtol=1e-12, atol=1e-12, err_msg=geom)
