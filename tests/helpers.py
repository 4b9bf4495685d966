"""Shared test helpers: small synthetic catalogs and the parity comparison."""
import numpy as np

from corr_pc_amd.engine import Catalog

# Weighted-bin tolerance: |a - b| <= RTOL * (|ref| + sqrt(num * err)) per (region, rp, pi, field).
# sqrt(num * err) bounds the sum of |terms| of the bin (Cauchy-Schwarz with err = sum of squares),
# so this is a 1e-10 relative bound on the accumulated magnitude, as stated for every wgg/wg+/w++ bin.
RTOL = 1e-10


def small_catalog(n, L, seed, shapes=True, weights=True, jk_side=2, clumps=0, coin=True, inc=False):
    r = np.random.default_rng(seed)
    pos = r.uniform(0, L, (n, 3))
    if clumps:
        c = r.uniform(0, L, (clumps, 3))
        k = r.integers(0, clumps, n // 2)
        pos[: n // 2] = np.mod(c[k] + r.normal(0, 0.5, (n // 2, 3)), L)
        pos[pos >= L] = 0.0
    pos = pos[np.argsort(pos[:, 2], kind="stable")]
    jk = (np.minimum((pos[:, 0] * jk_side / L).astype(int), jk_side - 1) * jk_side * jk_side
          + np.minimum((pos[:, 1] * jk_side / L).astype(int), jk_side - 1) * jk_side
          + np.minimum((pos[:, 2] * jk_side / L).astype(int), jk_side - 1)).astype(np.int32)
    return Catalog(x=pos[:, 0], y=pos[:, 1], z=pos[:, 2], w=r.lognormal(0, 0.3, n) if weights else None,
                   e1=r.normal(0, 0.3, n) if shapes else None, e2=r.normal(0, 0.3, n) if shapes else None,
                   phi=r.uniform(-1.5, 1.5, n) if shapes else None, jk=jk,
                   jk_prob=r.uniform(0, 1, n) if coin else None, include_prob=r.uniform(0, 1, n) if inc else None)


def _get(h, name):
    return h[name] if isinstance(h, dict) else getattr(h, name)


def assert_hist_match(got, ref, rtol=RTOL, what="", mag=None):
    """Bit-identical integer counts; weighted sums within rtol of the bin's accumulated magnitude.
    That magnitude is sqrt(num * err) of the full sample for the fields that carry err sums; kinds
    without them pass `mag` ([rp, pi, field], the full sample's sum of |terms|) instead."""
    gn, rn = _get(got, "num"), _get(ref, "num")
    if not np.array_equal(gn, rn):
        bad = np.argwhere(gn != rn)
        r, k, p = bad[0]
        raise AssertionError(f"{what}: {len(bad)} count cells differ, first (region {r}, rp {k}, pi {p}): "
                             f"{gn[r, k, p]} vs {rn[r, k, p]}")
    assert np.array_equal(_get(got, "num_rp"), _get(ref, "num_rp")), f"{what}: rp-level counts differ"
    num0 = rn[0].astype(np.float64)
    gw, rw = _get(got, "wt_num"), _get(ref, "wt_num")
    scale_w = np.abs(rw) + np.abs(rw[0])[None]
    assert np.all(np.abs(gw - rw) <= rtol * scale_w + 1e-300), f"{what}: wt_num differs beyond {rtol}"
    gv, rv = _get(got, "val"), _get(ref, "val")
    ge, re_ = _get(got, "err"), _get(ref, "err")
    if mag is None:
        mag = np.sqrt(num0[..., None] * np.abs(re_[0]))  # [rp, pi, field]
    scale_v = np.abs(rv) + mag[None]
    dv = np.abs(gv - rv)
    if not np.all(dv <= rtol * scale_v + 1e-300):
        idx = np.unravel_index(np.argmax(dv / (scale_v + 1e-300)), dv.shape)
        raise AssertionError(f"{what}: val differs at {idx}: {gv[idx]!r} vs {rv[idx]!r} (scale {scale_v[idx]:.3e})")
    scale_e = np.abs(re_) + np.abs(re_[0])[None]
    de = np.abs(ge - re_)
    if not np.all(de <= rtol * scale_e + 1e-300):
        idx = np.unravel_index(np.argmax(de / (scale_e + 1e-300)), de.shape)
        raise AssertionError(f"{what}: err differs at {idx}: {ge[idx]!r} vs {re_[idx]!r}")


def sky_catalogs(n_out, n_in, seed, shapes_in=False, coins=True, inc=False, table=None):
    """Sky survey subsamples (six 5x5 degree fields, 0.2 < z < 1): data as the outer catalog, randoms
    as the inner one.  Returns (outer, inner, n_jk, table).  Coins (jk_prob) are drawn uniformly so
    that both jackknife slot rules are exercised; the reader itself would set jk_prob = 1."""
    from corr_pc_amd import synth
    A, n_jk, table = synth.sky_catalog(2, "data", n=n_out, table=table)
    B, _, _ = synth.sky_catalog(2, "rand", n=n_in, table=table)
    r = np.random.default_rng(seed)
    if coins:
        A.jk_prob = r.uniform(0, 1, A.n)
        B.jk_prob = r.uniform(0, 1, B.n)
    if inc:
        A.include_prob = r.uniform(0, 1, A.n)
        B.include_prob = r.uniform(0, 1, B.n)
    if shapes_in:
        B.e1 = r.normal(0, 0.3, B.n)
        B.e2 = r.normal(0, 0.3, B.n)
        B.phi = r.uniform(-1.5, 1.5, B.n)
    return A, B, n_jk, table
