"""GPU test of vkhel_ext_polymul_noalloc (include/vkhel/vkhel_ext.h): the
polynomial product whose scratch is the library's static per-GPU array.
Its words must equal the batch path's VKHEL_JOB_POLYMUL (which allocates its
own scratch) bit for bit, over consecutive calls of different sizes sharing
that array (a later, smaller call must not see the earlier one's data), and
it must reject what it cannot do without allocating before enqueuing
anything.
"""
import numpy as np
import pytest

from oracle import oracle as O
import vkhel_amd as V
from vkhel_amd import VkhelError

pytestmark = pytest.mark.gpu

MAX_WORDS = 1 << 22  # VKHEL_POLYMUL_NOALLOC_MAX_WORDS


@pytest.fixture(scope="module")
def ctx():
    c = V.Context()
    yield c
    c.close()


def _tables(logn, nlimbs, bits):
    n = 1 << logn
    qs = O.primes_ntt(1 << bits, 2 * n, nlimbs)
    return [V.NttTables(n, q, O.root_2n(q, n)) for q in qs]


def _operands(ctx, tabs, seed):
    n, L = tabs[0].n, len(tabs)
    a, b = ctx.vector(L * n, False), ctx.vector(L * n, False)
    for i, t in enumerate(tabs):
        a.fill_uniform(t.q, seed + i, i * n, n)
        b.fill_uniform(t.q, seed + 100 + i, i * n, n)
    return a, b


def _batch_product(ctx, a, b, tabs):
    res = ctx.vector(a.length)
    bt = V.Batch(ctx)
    bt.add(V.JOB_POLYMUL, a, b, res, tabs)
    bt.run()
    out = res.read()
    bt.close()
    res.close()
    return out


# (log2 n, limbs, modulus bits): two-pass fused product, one-tile fused
# product (smaller than the first, so it runs on what the first left in the
# scratch), fused transforms without the fused product (2^19), the radix-2
# path (n = 4), and a call at the scratch maximum
CALLS = [(15, 4, 55), (12, 2, 60), (19, 1, 50), (2, 3, 61), (17, 32, 59),
         (10, 1, 50)]


def test_polymul_noalloc_matches_batch_on_consecutive_calls(ctx):
    for k, (logn, L, bits) in enumerate(CALLS):
        print(f"noalloc call {k}: n=2^{logn} limbs={L} q<2^{bits}", flush=True)
        tabs = _tables(logn, L, bits)
        for t in tabs:
            t.upload(ctx)
        a, b = _operands(ctx, tabs, 1000 * k + 1)
        want = _batch_product(ctx, a, b, tabs)
        res = ctx.vector(a.length)
        ctx.polymul_noalloc(a, b, res, tabs)
        ctx.synchronize()
        got = res.read()
        assert np.array_equal(got, want), f"call {k} (n=2^{logn}, L={L}) differs"
        if logn <= 12:
            # one limb against the oracle's negacyclic product as well
            n = 1 << logn
            w = O.polymul_ntt(a.read(0, n), b.read(0, n), tabs[0].q, tabs[0].w)
            assert np.array_equal(got[:n], w)
        for v in (a, b, res):
            v.close()
        for t in tabs:
            t.close()


def test_polymul_noalloc_in_place_on_a(ctx):
    tabs = _tables(14, 2, 55)
    for t in tabs:
        t.upload(ctx)
    a, b = _operands(ctx, tabs, 77)
    want = _batch_product(ctx, a, b, tabs)
    ctx.polymul_noalloc(a, b, a, tabs)
    ctx.synchronize()
    assert np.array_equal(a.read(), want)
    for v in (a, b):
        v.close()
    for t in tabs:
        t.close()


def test_polymul_noalloc_rejects_before_enqueuing(ctx):
    # beyond the scratch maximum: 2^17 x 33 limbs
    tabs = _tables(17, 33, 59)
    for t in tabs:
        t.upload(ctx)
    words = 33 << 17
    assert words > MAX_WORDS
    a, b = ctx.vector(words), ctx.vector(words)
    res = ctx.vector(words)
    res.fill_uniform(5, 3)
    before = res.read(0, 64)
    with pytest.raises(VkhelError) as e:
        ctx.polymul_noalloc(a, b, res, tabs)
    assert e.value.code == V.vkhel.ERR_INVALID
    assert "scratch maximum" in str(e.value)
    # tables not resident on this context
    fresh = _tables(12, 1, 50)
    a2, b2, r2 = ctx.vector(4096), ctx.vector(4096), ctx.vector(4096)
    with pytest.raises(VkhelError) as e:
        ctx.polymul_noalloc(a2, b2, r2, fresh)
    assert "not resident" in str(e.value)
    # result aliasing b
    with pytest.raises(VkhelError):
        ctx.polymul_noalloc(a2, b2, b2, fresh)
    ctx.synchronize()
    assert np.array_equal(res.read(0, 64), before)
    for v in (a, b, res, a2, b2, r2):
        v.close()
    for t in tabs + fresh:
        t.close()
