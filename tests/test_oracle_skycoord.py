"""CPU pins of the oracle's other sky coordinates: choose_rp_pi_func (initialization.cpp:545-571),
lim_calc_sky (sky_calcs.cpp:105-190), pi_calc_projected and pi_calc_phi (calcs.cpp:108-123),
rp_calc_sky_projected and rp_calc_sky_angular (sky_calcs.cpp:84-102), r_mu_calc (calcs.cpp:135-140)
and the pi bin width of 1 for coordinates 3 and 5 (initialization.cpp:307-308).

Pinned by brute force over every pair of small catalogs, with formulas other than the oracle's:
the angle from the chord between 3-D unit vectors, 2 asin(|u_i - u_j| / 2), in place of
acos(sin sin + cos cos cos); ephi as arccos(|cos(phi - theta)|) with theta from atan2, in place of
the folded |phi - theta|.  Every pair the reference's formulas bin must be found (the RA/Dec/z window
is only a search bound) and put in the same cell:
  * coordinates 5: rp = the angle in degrees, one pi bin (pi = 0, bins of width 1), no z test;
  * coordinates 4: the same rp, pi = ephi of the g_with_shape galaxy; summed over pi it is 5;
  * coordinates 3: rp = angle * DA of the outer galaxy (the two orders differ), pi = 0;
  * coordinates 0 with p_max > 600: the projected rp and pi = 0;
  * coordinates 1: r = sqrt(rp^2 + (DC_j - DC_i)^2) with rp = angle (DA_i + DA_j)/2 and mu = pi / r,
    and a line-of-sight pair at r = |DC_j - DC_i|, mu = -1;
  * coordinates 2: pi = ephi, the z window [z_min, z_i + z_sep_max];
  * shape_shape with coordinates 2 or 4 takes the estimator-0 branch (corels.cpp:139).
"""
import numpy as np
import pytest

from corr_pc_amd import synth
from corr_pc_amd.engine import KIND_DD, KIND_SHAPE_SHAPE, Catalog, make_sky_params

_TABLE = synth.distance_table(0.3, 1e-4, 1.0)
C = 300000.0
Z_MAX = 1.0


def _table(z_sep_max):
    return _TABLE["H"], _TABLE["dA"], _TABLE["dz"], Z_MAX, z_sep_max


def _cat(n, seed, zlo=0.30, zhi=0.36, patch=(10.0, 13.0, -1.5, 1.5), zero_z=False):
    r = np.random.default_rng(seed)
    ra = r.uniform(patch[0], patch[1], n)
    dec = np.degrees(np.arcsin(r.uniform(np.sin(np.radians(patch[2])), np.sin(np.radians(patch[3])), n)))
    z = np.zeros(n) if zero_z else r.uniform(zlo, zhi, n)
    o = np.argsort(z, kind="stable")
    ra, dec, z = ra[o], dec[o], z[o]
    it = np.floor(z / _TABLE["dz"]).astype(np.int64)
    return Catalog(x=ra, y=dec, z=z, w=r.lognormal(0, 0.3, n), e1=r.normal(0, 0.3, n), e2=r.normal(0, 0.3, n),
                   phi=r.uniform(-1.5, 1.5, n), dc=_TABLE["dC"][it].copy(), da=_TABLE["dA"][it].copy())


def _unit(c):
    ra, dec = np.radians(c.x), np.radians(c.y)
    return np.stack([np.cos(dec) * np.cos(ra), np.cos(dec) * np.sin(ra), np.sin(dec)], axis=-1)


def _angle(a, b):
    """[n_a, n_b] great-circle angle (radians) from the chord."""
    d = _unit(a)[:, None, :] - _unit(b)[None, :, :]
    return 2 * np.arcsin(np.minimum(np.linalg.norm(d, axis=-1) / 2, 1.0))


def _ephi(a, b, phi_side):
    """[n_a, n_b] ephi of the galaxy on `phi_side` (1 = a, 2 = b) against the pair's orientation."""
    dra = b.x[None, :] - a.x[:, None]
    ddec = b.y[None, :] - a.y[:, None]
    th = np.arctan2(ddec, dra * np.cos(np.radians(a.y[:, None] + b.y[None, :]) / 2))
    phi = a.phi[:, None] if phi_side == 1 else b.phi[None, :]
    return np.arccos(np.minimum(np.abs(np.cos(phi - th)), 1.0))


def _log_bin(r, p, n):
    return np.floor(np.log10(r / p.rp_min) / ((np.log10(p.rp_max) - np.log10(p.rp_min)) / n))


def _brute(p, rp, pi, width, mask=None):
    """[n_rp, n_pi] counts of the pairs (rp, pi) with pi bins of `width`, plus the smallest distance
    of any pair to a bin edge (relative in rp, absolute in pi)."""
    k = _log_bin(rp, p, p.n_rp)
    q = np.floor((pi - p.pi_min) / width)
    ok = (k >= 0) & (k < p.n_rp) & (q >= 0) & (q < p.n_pi)
    if mask is not None:
        ok &= mask
    h = np.zeros((p.n_rp, p.n_pi), np.int64)
    np.add.at(h, (k[ok].astype(int), q[ok].astype(int)), 1)
    kf = np.log10(rp / p.rp_min) / ((np.log10(p.rp_max) - np.log10(p.rp_min)) / p.n_rp)
    qf = (pi - p.pi_min) / width
    fin = np.isfinite(kf) & np.isfinite(qf)
    m_r = np.abs(kf[fin] - np.round(kf[fin])).min() if fin.any() else 1.0
    fq = fin & (pi != 0)  # pi = 0 exactly (one table row, or projected) is exact in both
    m_q = np.abs(qf[fq] - np.round(qf[fq])).min() if fq.any() and width < 1 else 1.0
    return h, min(m_r, m_q)


def _check(oracle_lib, a, b, p, rp, pi, width, mask=None):
    h, st = oracle_lib.count(a, b, p)
    want, margin = _brute(p, rp, pi, width, mask)
    assert margin > 1e-9, "a pair lies on a bin edge: pick another seed"
    assert want.sum() > 200
    np.testing.assert_array_equal(h.num[0], want)
    return h


def test_angular_projected_and_angular_ephi(oracle_lib):
    a, b = _cat(500, 1, zero_z=True), _cat(600, 2, zero_z=True)
    th = np.degrees(_angle(a, b))
    p5 = make_sky_params(_table(9.0), 6, 0.05, 1.5, 1, 0.0, 1.0, coordinates=5)
    h5 = _check(oracle_lib, a, b, p5, th, np.zeros_like(th), 1.0)
    for gws in (1, 2):
        p4 = make_sky_params(_table(9.0), 6, 0.05, 1.5, 5, 0.0, np.pi / 2 + 1e-9, coordinates=4, g_with_shape=gws)
        h4 = _check(oracle_lib, a, b, p4, th, _ephi(a, b, gws), (np.pi / 2 + 1e-9) / 5)
        np.testing.assert_array_equal(h4.num[0].sum(axis=1), h5.num[0][:, 0])
    # pi = 0 lands in bin floor(-p_min) of width 1, not (p_max - p_min) / n_pi
    p5b = make_sky_params(_table(9.0), 6, 0.05, 1.5, 3, -1.5, 10.0, coordinates=5)
    h, _ = oracle_lib.count(a, b, p5b)
    np.testing.assert_array_equal(h.num[0][:, 1], h5.num[0][:, 0])
    assert h.num[0][:, [0, 2]].sum() == 0


@pytest.mark.parametrize("coordinates,pmin,pmax,npi", [(3, 0.0, 1.0, 1), (0, -1000.0, 1000.0, 4)])
def test_projected(oracle_lib, coordinates, pmin, pmax, npi):
    a, b = _cat(500, 3), _cat(600, 4)
    th = _angle(a, b)
    zsep = 9.0 if coordinates == 3 else pmax * _TABLE["H"][int(np.ceil(Z_MAX / _TABLE["dz"]))] / C + 0.001
    p = make_sky_params(_table(zsep), 6, 1.0, 30.0, npi, pmin, pmax, coordinates=coordinates)
    width = 1.0 if coordinates == 3 else (pmax - pmin) / npi
    # rp from the outer galaxy's DA: the two orders differ
    _check(oracle_lib, a, b, p, th * a.da[:, None], np.zeros_like(th), width)
    _check(oracle_lib, b, a, p, th.T * b.da[:, None], np.zeros_like(th.T), width)


def test_r_mu(oracle_lib):
    a, b = _cat(500, 5, 0.30, 0.32), _cat(600, 6, 0.30, 0.32)
    rmax = 30.0
    zsep = rmax * _TABLE["H"][int(np.ceil(Z_MAX / _TABLE["dz"]))] / C + 0.001
    p = make_sky_params(_table(zsep), 6, 1.0, rmax, 4, -1.0, 1.0, coordinates=1)
    rp = _angle(a, b) * (a.da[:, None] + b.da[None, :]) / 2
    dpi = b.dc[None, :] - a.dc[:, None]
    r = np.hypot(rp, dpi)
    _check(oracle_lib, a, b, p, r, dpi / r, 0.5)
    # a line-of-sight pair: rp = 0, r = |DC_j - DC_i|, mu = -1
    it = np.floor(np.array([0.31, 0.305]) / _TABLE["dz"]).astype(int)
    one = [Catalog(x=np.array([11.0]), y=np.array([0.5]), z=np.array([z]), dc=_TABLE["dC"][[k]].copy(),
                   da=_TABLE["dA"][[k]].copy()) for z, k in zip((0.31, 0.305), it)]
    h, st = oracle_lib.count(one[0], one[1], p)
    r1 = _TABLE["dC"][it[0]] - _TABLE["dC"][it[1]]
    assert st["n_binned"] == 1 and h.num[0][int(_log_bin(r1, p, 6)), 0] == 1


def test_rp_ephi_window(oracle_lib):
    # a narrow z range: DA varies by under 5%, so the Dec bound (the outer galaxy's da_max times
    # 1.05) cuts no pair with rp < rp_max
    a, b = _cat(500, 7, 0.30, 0.31), _cat(600, 8, 0.30, 0.31)
    zsep = 0.004
    p = make_sky_params(_table(zsep), 6, 1.0, 30.0, 5, 0.0, np.pi / 2 + 1e-9, coordinates=2, g_with_shape=2,
                        z_min=0.302)
    rp = _angle(a, b) * (a.da[:, None] + b.da[None, :]) / 2
    zw = (b.z[None, :] >= 0.302) & (b.z[None, :] <= a.z[:, None] + zsep)
    _check(oracle_lib, a, b, p, rp, _ephi(a, b, 2), (np.pi / 2 + 1e-9) / 5, zw)
    # a z_sep_min moves the lower edge to z_i + z_sep_min
    p.z_sep_min = -0.002
    zw = (b.z[None, :] >= a.z[:, None] - 0.002) & (b.z[None, :] <= a.z[:, None] + zsep)
    _check(oracle_lib, a, b, p, rp, _ephi(a, b, 2), (np.pi / 2 + 1e-9) / 5, zw)


def test_rp_ephi_window_uses_the_outer_redshift(oracle_lib):
    """lim_calc_ra_dec takes da_max at the outer galaxy's z for coordinates 2: an inner galaxy much
    nearer has a smaller DA, so a pair with rp < rp_max can lie outside the Dec bound."""
    def one(dec, z):
        k = int(np.floor(z / _TABLE["dz"]))
        return Catalog(x=np.array([11.0]), y=np.array([dec]), z=np.array([z]), phi=np.array([0.2]),
                       dc=_TABLE["dC"][[k]].copy(), da=_TABLE["dA"][[k]].copy())
    o, i = one(0.0, 0.6), one(0.0, 0.2)
    da_max = 20.0 / np.interp(0.6, _TABLE["z"], _TABLE["dA"])  # radians, at the outer z
    ddec = np.degrees(da_max) * 1.05 * 1.2                       # just outside the Dec bound
    i.y[:] = ddec
    rp = np.radians(ddec) * (o.da[0] + i.da[0]) / 2
    assert rp < 20.0
    p = make_sky_params(_table(9.0), 4, 1.0, 20.0, 4, 0.0, np.pi / 2 + 1e-9, coordinates=2, g_with_shape=2)
    _, st = oracle_lib.count(o, i, p)
    assert st["n_binned"] == 0
    i.y[:] = ddec / 1.2 / 1.05 * 0.99                            # inside: counted
    _, st = oracle_lib.count(o, i, p)
    assert st["n_binned"] == 1


@pytest.mark.parametrize("coordinates", [2, 4])
def test_shape_shape_takes_the_shape_side(oracle_lib, coordinates):
    """One pair along RA (theta = 0): e+ = e1 per galaxy; estimator 1 would average both sides."""
    def one(ra, e1, phi):
        z = np.array([0.0 if coordinates == 4 else 0.31])
        k = int(np.floor(z[0] / _TABLE["dz"]))
        return Catalog(x=np.array([ra]), y=np.array([0.0]), z=z, e1=np.array([e1]), e2=np.array([0.0]),
                       phi=np.array([phi]), dc=_TABLE["dC"][[k]].copy(), da=_TABLE["dA"][[k]].copy())
    a, b = one(11.0, 0.2, 0.3), one(11.01, 0.4, 0.6)
    for est in (0, 1):
        p = make_sky_params(_table(9.0), 4, 0.001, 20.0, 4, 0.0, np.pi / 2 + 1e-9, coordinates=coordinates,
                            kind=KIND_SHAPE_SHAPE, g_with_shape=2, estimator=est, which_corr=2)
        h, st = oracle_lib.count(a, b, p)
        assert st["n_binned"] == 1
        v = h.val[0].sum(axis=(0, 1))
        np.testing.assert_allclose(v[1], 0.4, rtol=1e-12)    # w e+ of the shape (inner) galaxy
        np.testing.assert_allclose(v[6], 0.6, rtol=1e-12)    # ephi of the shape galaxy
        assert v[2] == 0
    # ...while coordinates 0 with estimator 1 averages both sides (e+ = (0.2 + 0.4)/2)
    z = np.array([0.31])
    k = int(np.floor(0.31 / _TABLE["dz"]))
    a0 = Catalog(x=a.x, y=a.y, z=z, e1=a.e1, e2=a.e2, phi=a.phi, dc=_TABLE["dC"][[k]].copy(), da=_TABLE["dA"][[k]].copy())
    b0 = Catalog(x=b.x, y=b.y, z=z, e1=b.e1, e2=b.e2, phi=b.phi, dc=_TABLE["dC"][[k]].copy(), da=_TABLE["dA"][[k]].copy())
    p = make_sky_params(_table(0.1), 4, 0.01, 20.0, 4, -10.0, 10.0, kind=KIND_SHAPE_SHAPE, g_with_shape=2,
                        estimator=1, which_corr=2)
    h, st = oracle_lib.count(a0, b0, p)
    assert st["n_binned"] == 1
    np.testing.assert_allclose(h.val[0].sum(axis=(0, 1))[1], 0.3, rtol=1e-12)


def test_invalid_coordinates(oracle_lib):
    a = _cat(10, 9)
    p = make_sky_params(_table(9.0), 4, 1.0, 20.0, 4, 0.0, 1.0, coordinates=6)
    with pytest.raises(Exception):
        oracle_lib.count(a, a, p)
    p = make_sky_params(_table(9.0), 4, 1.0, 20.0, 4, -1.0, 1.0, r_mu=1)
    with pytest.raises(Exception):
        oracle_lib.count(a, a, p)
