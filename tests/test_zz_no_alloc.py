"""cpc_count_noalloc (the entry point that allocates no device memory) against cpc_count, through the
C ABI: two calls in a row, the second on smaller catalogs, each bit-identical in its counts and
within the weighted tolerance; a pass larger than the scratch fails with CPC_ERR_NOMEM and leaves
the context usable.  No counterpart in the reference: the callers it serves are those that cannot
allocate (graph capture, serving loops)."""
import numpy as np
import pytest

from corr_pc_amd.engine import (CPC_ERR_NOMEM, KIND_DD, KIND_SHAPE_SHAPE, PREC_FP64, PREC_SPLIT, EngineError,
                                make_params)
from helpers import assert_hist_match, small_catalog

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec", [PREC_SPLIT, PREC_FP64])
def test_noalloc_matches_count_twice(engine, prec):
    L = 60.0
    big_o = small_catalog(6000, L, 901, clumps=40)
    big_i = small_catalog(5000, L, 902, clumps=30)
    small_o = small_catalog(1500, L, 903, clumps=10)
    small_i = small_catalog(900, L, 904)
    p = make_params(L, 8, 0.2, 12.0, 10, -15.0, 15.0, kind=KIND_SHAPE_SHAPE, g_with_shape=2, do_jk=1, n_jk=8,
                    precision=prec)
    for what, (o, i) in (("first call", (big_o, big_i)), ("second call, smaller", (small_o, small_i))):
        print(f"[case] noalloc {what} prec={prec} outer={o.n} inner={i.n}", flush=True)
        got, st = engine.count_noalloc(o, i, p)
        ref, rst = engine.count(o, i, p)
        assert_hist_match(got, ref, what=f"noalloc {what}")
        assert st["n_binned"] == rst["n_binned"] > 0


def test_noalloc_auto_and_plain_dd(engine):
    L = 50.0
    A = small_catalog(4000, L, 905, shapes=False)
    p = make_params(L, 6, 0.2, 10.0, 8, -12.0, 12.0, kind=KIND_DD, do_jk=0, precision=PREC_SPLIT)
    print("[case] noalloc dd no-jackknife", flush=True)
    got, _ = engine.count_noalloc(A, A, p)
    ref, _ = engine.count(A, A, p)
    assert_hist_match(got, ref, what="noalloc dd")
    p.auto_pairs = 1
    print("[case] noalloc auto", flush=True)
    got, _ = engine.count_noalloc(A, A, p)
    ref, _ = engine.count(A, A, p)
    assert_hist_match(got, ref, what="noalloc auto")


def test_noalloc_rejects_a_pass_beyond_the_scratch(engine):
    """1.2e7 galaxies on each side need more than the scratch's 256 MiB of records alone."""
    L = 500.0
    r = np.random.default_rng(906)
    n = 12_000_000
    z = np.sort(r.uniform(0, L, n))
    from corr_pc_amd.engine import Catalog
    big = Catalog(x=r.uniform(0, L, n), y=r.uniform(0, L, n), z=z, jk=np.zeros(n, np.int32))
    p = make_params(L, 4, 0.5, 2.0, 4, -2.0, 2.0, kind=KIND_DD, do_jk=0, precision=PREC_SPLIT)
    print("[case] noalloc beyond the scratch", flush=True)
    with pytest.raises(EngineError) as e:
        engine.count_noalloc(big, big, p)
    assert e.value.status == CPC_ERR_NOMEM and "no-alloc scratch" in str(e.value)
    engine.free(big)
    # the context is still usable afterwards, through both entry points
    A = small_catalog(2000, 40.0, 907)
    q = make_params(40.0, 6, 0.3, 8.0, 8, -10.0, 10.0, kind=KIND_SHAPE_SHAPE, do_jk=1, n_jk=8, precision=PREC_SPLIT)
    got, _ = engine.count_noalloc(A, A, q)
    ref, _ = engine.count(A, A, q)
    assert_hist_match(got, ref, what="after a rejected pass")
