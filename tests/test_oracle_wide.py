"""CPU pin of the oracle's box windows wider than the box in z (lim_calc_PB, PB_calcs.cpp:30-61):
every outer galaxy then takes the z images +1, 0 and -1 and each pair is found under two of them."""
import numpy as np

from corr_pc_amd.engine import KIND_DD, Catalog, make_params


def test_pi_window_wider_than_the_box(oracle_lib):
    """The generator notebook's box: L = 75 with |pi| < 75.  lim_calc_PB then gives every outer
    galaxy the z images +1, 0 and -1 (PB_calcs.cpp:30-61), and each pair is found twice: at image 0
    with pi = dz, and through the image on the other side with pi = dz -+ L.  Pinned by brute force
    over all pairs with minimum-image rp: summed over pi, every rp bin holds twice the pairs."""
    L2 = 75.0
    r = np.random.default_rng(93)
    pts = r.uniform(0, L2, (1500, 3))
    pts = pts[np.argsort(pts[:, 2], kind="stable")]
    A = Catalog(x=pts[:, 0], y=pts[:, 1], z=pts[:, 2])
    p = make_params(L2, 8, 0.5, 10.0, 4, -75.0, 75.0, kind=KIND_DD)
    h, st = oracle_lib.count(A, A, p)
    d = pts[:, None, :2] - pts[None, :, :2]
    d -= L2 * np.round(d / L2)
    rp = np.hypot(d[..., 0], d[..., 1])
    np.fill_diagonal(rp, np.inf)
    k = np.floor(np.log10(rp / 0.5) / ((np.log10(10.0) - np.log10(0.5)) / 8))
    want = np.array([np.sum(k == b) for b in range(8)])
    np.testing.assert_array_equal(h.num[0].sum(axis=1), 2 * want)
    # the self pairs (rp = 0) stay out; the pi = 0 half holds the image-0 pairs with dz >= 0 and
    # the image +-1 pairs with dz <= 0 alike: the two pi halves are mirror images
    assert np.array_equal(h.num[0][:, :2].sum(axis=1), h.num[0][:, 2:].sum(axis=1))
