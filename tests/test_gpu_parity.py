"""GPU parity of libvkhel.so against the reference's known answers and the
CPU oracle.  Every call goes through the C ABI (vkhel_amd mirrors it with
ctypes); integer results must be bit-exact.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests import mathref as R
import vkhel_amd as V

pytestmark = pytest.mark.gpu

U64MAX = (1 << 64) - 1


@pytest.fixture(scope="module")
def ctx():
    c = V.Context()
    assert c.device_name.startswith("gfx950"), c.device_name
    yield c
    c.close()


def vec(ctx, data):
    return ctx.from_array(np.asarray(data, dtype=np.uint64))


# ---------------- test/vector.c known answers ----------------

def test_ref_copy_from_host_and_dup(ctx, golden):
    v = golden["vectors"]["vector"]
    a = vec(ctx, v["copy_from_host"]["data"])
    assert a.map_read().tolist() == v["copy_from_host"]["data"]
    d = vec(ctx, v["dup"]["data"])
    e = d.dup()
    assert e.map_read().tolist() == v["dup"]["data"]
    for x in (a, d, e):
        x.close()


def test_ref_create_zeroes(ctx):
    z = ctx.vector(1000, zero=True)
    assert not z.map_read().any()
    z.close()


def test_ref_elementwise(ctx, golden):
    v = golden["vectors"]["vector"]
    f = v["elemfma"]
    a, b = vec(ctx, f["a"]), vec(ctx, f["b"])
    for c in f["cases"]:
        out = ctx.vector(len(f["a"]))
        a.elemfma(b, out, c["multiplier"], f["mod"])
        assert out.map_read().tolist() == c["expected"]
    for c in v["elemmul"]["cases"]:
        a, b = vec(ctx, c["a"]), vec(ctx, c["b"])
        out = ctx.vector(len(c["a"]))
        a.elemmul(b, out, c["mod"])
        assert out.map_read().tolist() == c["expected"]
    g = v["elemgtadd"]
    a = vec(ctx, g["a"])
    for c in g["cases"]:
        out = ctx.vector(len(g["a"]))
        a.elemgtadd(out, c["bound"], c["diff"])
        assert out.map_read().tolist() == c["expected"]
    g = v["elemgtsub"]
    a = vec(ctx, g["a"])
    for c in g["cases"]:
        out = ctx.vector(len(g["a"]))
        a.elemgtsub(out, c["bound"], c["diff"], c["mod"])
        assert out.map_read().tolist() == c["expected"]
    for c in v["elemmod"]["cases"]:
        a = vec(ctx, c["a"])
        out = ctx.vector(len(c["a"]))
        a.elemmod(out, c["mod"], c["q"])
        assert out.map_read().tolist() == c["expected"]


@pytest.mark.parametrize("algo", [V.NTT_FUSED, V.NTT_RADIX2])
def test_ref_transforms(ctx, golden, algo):
    ctx.set_ntt_algo(algo)
    try:
        for c in golden["vectors"]["vector"]["transform"]["cases"]:
            t = V.NttTables(c["n"], c["q"], c["w"])
            if c["in_place"]:
                x = vec(ctx, c["coeffs"])
                x.forward_transform(x, t)
                assert x.map_read().tolist() == c["ntt"]
                x.inverse_transform(x, t)
                assert x.map_read().tolist() == c["coeffs"]
            else:
                a = vec(ctx, c["coeffs"])
                b = ctx.vector(c["n"])
                a.forward_transform(b, t)
                assert b.map_read().tolist() == c["ntt"]
                d = ctx.vector(c["n"])
                b.inverse_transform(d, t)
                assert d.map_read().tolist() == c["coeffs"]
            t.close()
    finally:
        ctx.set_ntt_algo(V.NTT_AUTO)


# ---------------- elementwise vs oracle ----------------

MODS = [3, 5, 769, 65537, 1125891450734593, 1152921504606830593,
        (1 << 61) - 1, (1 << 63) + 29]


def test_elemfma_wrap_branches_match_case_analysis(ctx):
    """elemfma's `q - x` corrections (elemfma.comp:47-48, 51-52) on the GPU
    against the case analysis of tests/mathref.py, every case reached."""
    rng = np.random.default_rng(21)
    seen = [0, 0, 0, 0]
    for q in [3, 97, 1125891450734593, 1152921504606830593,
              O.primes_ntt(1 << 61, 2, 1)[0]]:
        count = 4000
        a = np.concatenate([
            rng.integers(0, U64MAX, count, dtype=np.uint64, endpoint=True),
            np.uint64(U64MAX) - rng.integers(0, q, count, dtype=np.uint64),
            rng.integers(0, q, count, dtype=np.uint64)])
        b = rng.integers(0, q, a.size, dtype=np.uint64)
        va, vb, out = vec(ctx, a), vec(ctx, b), ctx.vector(a.size)
        for m in [1, 2, int(rng.integers(1, q)), q - 1]:
            want, cases = R.elemfma_closed_form(a.tolist(), b.tolist(), m, q)
            seen = [s + c for s, c in zip(seen, cases)]
            va.elemfma(vb, out, m, q)
            assert out.read().tolist() == want, (q, m)
        for x in (va, vb, out):
            x.close()
    assert all(seen), seen


@pytest.mark.parametrize("offset", [1, 3])
def test_elementwise_on_odd_offset_views(ctx, offset):
    """Views at odd word offsets are 8-byte aligned only; the pointwise
    kernels then move one word per access.  Single calls and batch jobs on
    such views against the oracle."""
    q = 1152921504606830593
    rng = np.random.default_rng(offset)
    length = 10007
    base_a = vec(ctx, rng.integers(0, U64MAX, length + 8, dtype=np.uint64,
                                   endpoint=True))
    base_b = vec(ctx, rng.integers(0, q, length + 8, dtype=np.uint64))
    base_r = ctx.vector(length + 8)
    a, b = base_a.view(offset, length), base_b.view(offset + 1, length)
    r = base_r.view(offset + 2, length)
    xa, xb = a.read(), b.read()
    a.elemmul(b, r, q)
    assert np.array_equal(r.read(), O.elemmul(xa, xb, q))
    a.elemfma(b, r, 3, q)
    assert np.array_equal(r.read(), O.elemfma(xa, xb, 3, q))
    a.elemgtadd(r, q // 2, 12345)
    assert np.array_equal(r.read(), O.elemgtadd(xa, q // 2, 12345))
    a.elemgtsub(r, q // 2, 12345, 65537)
    assert np.array_equal(r.read(), O.elemgtsub(xa, q // 2, 12345, 65537))
    a.elemmod(r, 2, q)
    assert np.array_equal(r.read(), O.elemmod(xa, 2, q))
    # batch pointwise jobs on the same views (tables only name n and q)
    n = 1 << 13
    t = V.NttTables(n, q, O.root_2n(q, n))
    ca = base_a.view(offset, n)
    cb, cr = base_b.view(offset + 1, n), base_r.view(offset + 2, n)
    ya, yb = ca.read() % np.uint64(q), cb.read()
    ca.write(ya)
    for op, want in [(V.JOB_MUL, O.elemmul(ya, yb, q)),
                     (V.JOB_ADD, O.addmul_mod(ya, yb, 1, q)),
                     (V.JOB_ADD_REF, O.elemfma(ya, yb, 1, q))]:
        bt = V.Batch(ctx)
        bt.add(op, ca, cb, cr, [t])
        bt.run()
        bt.close()
        assert np.array_equal(cr.read(), want), op
    for v in (a, b, r, ca, cb, cr, base_a, base_b, base_r):
        v.close()
    t.close()


@pytest.mark.parametrize("length", [1, 2, 7, 64, 1000, 100003])
def test_elementwise_vs_oracle(ctx, length):
    rng = np.random.default_rng(length)
    raw_a = rng.integers(0, U64MAX, length, dtype=np.uint64, endpoint=True)
    raw_b = rng.integers(0, U64MAX, length, dtype=np.uint64, endpoint=True)
    a, b = vec(ctx, raw_a), vec(ctx, raw_b)
    out = ctx.vector(length)
    for q in MODS:
        a.elemmul(b, out, q)
        assert np.array_equal(out.read(), O.elemmul(raw_a, raw_b, q)), q
        ca = raw_a % np.uint64(q)
        cb = raw_b % np.uint64(q)
        ga, gb = vec(ctx, ca), vec(ctx, cb)
        for m in [1, 2, int(rng.integers(0, q, dtype=np.uint64))]:
            ga.elemfma(gb, out, m, q)
            assert np.array_equal(out.read(), O.elemfma(ca, cb, m, q)), (q, m)
        # raw operands exercise the wrap branches of elemfma.comp
        a.elemfma(b, out, 7, q)
        assert np.array_equal(out.read(), O.elemfma(raw_a, raw_b, 7, q)), q
        bound, diff = (int(x) for x in rng.integers(0, U64MAX, 2, dtype=np.uint64, endpoint=True))
        a.elemgtsub(out, bound, diff, q)
        assert np.array_equal(out.read(), O.elemgtsub(raw_a, bound, diff, q)), q
        a.elemgtadd(out, bound, diff)
        assert np.array_equal(out.read(), O.elemgtadd(raw_a, bound, diff))
        for p in [2, 3, 17, q]:
            a.elemmod(out, p, q)
            assert np.array_equal(out.read(), O.elemmod(raw_a, p, q)), (p, q)
        ga.close()
        gb.close()
    for x in (a, b, out):
        x.close()


def test_elementwise_in_place_and_length_rules(ctx):
    rng = np.random.default_rng(3)
    q = 1125891450734593
    x = rng.integers(0, q, 999, dtype=np.uint64)
    a = vec(ctx, x)
    a.elemmul(a, a, q)
    assert np.array_equal(a.read(), O.elemmul(x, x, q))
    short = ctx.vector(10)
    with pytest.raises(V.VkhelError):
        short.elemmul(a, a, q)  # input shorter than the result


# ---------------- transforms vs oracle ----------------

def _prime(logn, bits=60):
    return O.primes_ntt(1 << bits, 2 << logn, 1)[0]


@pytest.mark.parametrize("algo", [V.NTT_FUSED, V.NTT_RADIX2])
@pytest.mark.parametrize("logn", list(range(1, 18)))
def test_ntt_vs_oracle(ctx, logn, algo):
    n = 1 << logn
    ctx.set_ntt_algo(algo)
    try:
        # 61 bits: the largest moduli accepted, where the lazy bound 4q
        # comes closest to 2^63; 56 bits: the largest moduli of the
        # small-modulus kernels, where their bounds come closest; 57 bits:
        # the smallest of the others
        for bits in (50, 56, 57, 60, 61):
            q = _prime(logn, bits)
            w = O.root_2n(q, n)
            t = V.NttTables(n, q, w)
            x = O.fill_uniform(n, q, seed=logn * 7 + bits)
            a = vec(ctx, x)
            f = ctx.vector(n)
            a.forward_transform(f, t)
            want = O.ntt_forward(x, q, w)
            assert np.array_equal(f.read(), want), (logn, bits)
            g = ctx.vector(n)
            f.inverse_transform(g, t)
            assert np.array_equal(g.read(), x), (logn, bits)
            assert np.array_equal(O.ntt_inverse(want, q, w), x)
            for v in (a, f, g):
                v.close()
            t.close()
    finally:
        ctx.set_ntt_algo(V.NTT_AUTO)


@pytest.mark.parametrize("q,max_logn", [(17, 3), (97, 4), (7681, 8),
                                         (12289, 11), (65537, 15)])
def test_ntt_small_moduli(ctx, q, max_logn):
    """Tiny moduli: the reductions' shift and quotient constants at their
    smallest, every size the modulus admits (2n | q - 1, n < q)."""
    for logn in range(1, max_logn + 1):
        n = 1 << logn
        w = O.root_2n(q, n)
        t = V.NttTables(n, q, w)
        for seed in range(3):
            x = O.fill_uniform(n, q, seed=seed * 100 + logn)
            if seed == 2:
                x[:] = q - 1  # every word at the largest residue
            a = vec(ctx, x)
            f = ctx.vector(n)
            a.forward_transform(f, t)
            want = O.ntt_forward(x, q, w)
            assert np.array_equal(f.read(), want), (q, logn, seed)
            f.inverse_transform(f, t)
            assert np.array_equal(f.read(), x), (q, logn, seed)
            a.close()
            f.close()
        t.close()


@pytest.mark.parametrize("bits", [50, 61])
@pytest.mark.parametrize("logn", [1, 4, 10, 13, 14, 16])
def test_ntt_noncanonical_inputs_match_reference(ctx, logn, bits):
    # nttfwdbutterfly.comp reduces arbitrary words in its first stage,
    # including sums that wrap past 2^64.
    n = 1 << logn
    q = _prime(logn, bits)
    w = O.root_2n(q, n)
    t = V.NttTables(n, q, w)
    rng = np.random.default_rng(logn)
    x = rng.integers(0, U64MAX, n, dtype=np.uint64, endpoint=True)
    x[: n // 4] = np.uint64(U64MAX) - rng.integers(0, q, n // 4, dtype=np.uint64)
    a = vec(ctx, x)
    a.forward_transform(a, t)
    assert np.array_equal(a.read(), O.ntt_forward(x, q, w))
    # canonical words with a few large ones: waves of both kinds in a launch
    y = O.fill_uniform(n, q, seed=logn + bits)
    pos = rng.integers(0, n, 3)
    y[pos] = np.uint64(U64MAX) - rng.integers(0, 1 << 40, 3, dtype=np.uint64)
    b = vec(ctx, y)
    b.forward_transform(b, t)
    assert np.array_equal(b.read(), O.ntt_forward(y, q, w))
    t.close()
    a.close()
    b.close()


@pytest.mark.parametrize("algo", [V.NTT_FUSED, V.NTT_RADIX2])
@pytest.mark.parametrize("bits", [50, 60, 61])
@pytest.mark.parametrize("logn", [1, 2, 3, 4, 5, 6])
def test_ntt_first_stage_wrap_matches_closed_form(ctx, logn, bits, algo):
    """Words near 2^64 whose first-stage sum X + WY wraps (nttfwdbutterfly
    .comp:56) against the closed form of tests/mathref.py (the DFT of the
    corrected canonical vector), with both wrapped and unwrapped sums."""
    n = 1 << logn
    q = O.primes_ntt(1 << bits, 2 * n, 1)[0]
    psi = O.root_2n(q, n)
    t = V.NttTables(n, q, psi)
    ctx.set_ntt_algo(algo)
    rng = np.random.default_rng(logn * 100 + bits + 7)
    wraps = 0
    try:
        for trial in range(8):
            near = np.uint64(U64MAX) - rng.integers(0, q, n, dtype=np.uint64)
            raw = rng.integers(0, U64MAX, n, dtype=np.uint64, endpoint=True)
            x = np.where(rng.random(n) < 0.6, near, raw)
            want, w = R.ntt_forward_noncanonical(x.tolist(), q, psi)
            wraps += w
            a = vec(ctx, x)
            a.forward_transform(a, t)
            assert a.read().tolist() == want, trial
            a.close()
    finally:
        ctx.set_ntt_algo(V.NTT_AUTO)
        t.close()
    assert wraps > 0


@pytest.mark.parametrize("logn", [2, 3, 5, 7])
def test_ntt_closed_form(ctx, logn):
    n = 1 << logn
    q = 1125899903827969
    psi = O.root_2n(q, n)
    t = V.NttTables(n, q, psi)
    x = O.fill_uniform(n, q, seed=logn)
    a = vec(ctx, x)
    a.forward_transform(a, t)
    assert a.read().tolist() == R.dft_negacyclic_bitrev(x.tolist(), q, psi)
    t.close()


def test_ntt_length_checks(ctx):
    t = V.NttTables(16, 1125899903827969, O.root_2n(1125899903827969, 16))
    a = ctx.vector(8)
    with pytest.raises(V.VkhelError):
        a.forward_transform(a, t)
    t.close()


# ---------------- batch API ----------------

def test_batch_mixed_jobs_vs_oracle(ctx):
    rng = np.random.default_rng(5)
    jobs = []
    batch = V.Batch(ctx)
    keep = []
    for k in range(24):
        logn = int(rng.integers(4, 15))
        n = 1 << logn
        L = int(rng.integers(1, 4))
        qs = O.primes_ntt(1 << 60, 2 * n, L)
        ws = [O.root_2n(q, n) for q in qs]
        tabs = [V.NttTables(n, q, w) for q, w in zip(qs, ws)]
        op = int(rng.integers(0, 5))
        xa = np.concatenate([O.fill_uniform(n, q, 100 * k + i) for i, q in enumerate(qs)])
        xb = np.concatenate([O.fill_uniform(n, q, 100 * k + 50 + i) for i, q in enumerate(qs)])
        a, b, r = vec(ctx, xa), vec(ctx, xb), ctx.vector(L * n)
        batch.add(op, a, b, r, tabs)
        keep += [a, b, r] + tabs
        jobs.append((op, n, qs, ws, xa, xb, r))
    batch.run()
    for op, n, qs, ws, xa, xb, r in jobs:
        got = r.read()
        for i, (q, w) in enumerate(zip(qs, ws)):
            sa, sb = xa[i * n:(i + 1) * n], xb[i * n:(i + 1) * n]
            if op == V.JOB_NTT_FWD:
                want = O.ntt_forward(sa, q, w)
            elif op == V.JOB_NTT_INV:
                want = O.ntt_inverse(sa, q, w)
            elif op == V.JOB_POLYMUL:
                want = O.polymul_ntt(sa, sb, q, w)
            elif op == V.JOB_MUL:
                want = O.elemmul(sa, sb, q)
            else:
                want = O.addmul_mod(sa, sb, 1, q)
            assert np.array_equal(got[i * n:(i + 1) * n], want), (op, n, q)
    # a replay of the captured graph gives the same results (no job's
    # result vector is one of its inputs)
    snap = [r.read() for *_, r in jobs]
    batch.run()
    for (*_, r), s in zip(jobs, snap):
        assert np.array_equal(r.read(), s)
    batch.close()
    for x in keep:
        x.close()


def test_polymul_matches_schoolbook(ctx):
    # n = 2, 4: separate pointwise phase; larger n: the fused inverse forms
    # the product of the two spectra as it loads them
    for logn in (1, 2, 4, 8, 11):
        n = 1 << logn
        q = O.primes_ntt(1 << 55, 2 * n, 1)[0]
        w = O.root_2n(q, n)
        t = V.NttTables(n, q, w)
        xa, xb = O.fill_uniform(n, q, 1), O.fill_uniform(n, q, 2)
        a, b, r = vec(ctx, xa), vec(ctx, xb), ctx.vector(n)
        batch = V.Batch(ctx)
        batch.add(V.JOB_POLYMUL, a, b, r, [t])
        batch.run()
        assert np.array_equal(r.read(), O.negacyclic_schoolbook(xa, xb, q))
        batch.close()
        t.close()


@pytest.mark.parametrize("mode", [1, 2])
@pytest.mark.parametrize("logn", [12, 13, 14, 16])
def test_polymul_rns_vs_oracle(ctx, logn, mode):
    """Three-limb products (50-, 60- and 61-bit primes) in both pass modes,
    against the oracle's NTT-based product (itself pinned to the schoolbook
    product in the oracle tests)."""
    n = 1 << logn
    qs = [O.primes_ntt(1 << bits, 2 * n, 1)[0] for bits in (50, 60, 61)]
    tabs = [V.NttTables(n, q, O.root_2n(q, n)) for q in qs]
    xa = np.concatenate([O.fill_uniform(n, q, 10 + i) for i, q in enumerate(qs)])
    xb = np.concatenate([O.fill_uniform(n, q, 20 + i) for i, q in enumerate(qs)])
    a, b, r = vec(ctx, xa), vec(ctx, xb), ctx.vector(3 * n)
    ctx.set_pass_mode(logn, mode)
    try:
        batch = V.Batch(ctx)
        batch.add(V.JOB_POLYMUL, a, b, r, tabs)
        batch.run()
        got = r.read()
        for i, q in enumerate(qs):
            want = O.polymul_ntt(xa[i * n:(i + 1) * n], xb[i * n:(i + 1) * n], q,
                                 tabs[i].w)
            assert np.array_equal(got[i * n:(i + 1) * n], want), (logn, i)
        batch.close()
    finally:
        ctx.set_pass_mode(logn, 0)
        for t in tabs:
            t.close()


def test_batch_direct_launch_replays(ctx):
    """A batch of few launches (forward transforms of one size: two pass
    kernels) is enqueued directly rather than replayed from a graph: each
    launch reads the inputs as they are at that launch."""
    n = 1 << 16
    qs = O.primes_ntt(1 << 50, 2 * n, 2)
    tabs = [V.NttTables(n, q, O.root_2n(q, n)) for q in qs]
    xs = [O.fill_uniform(n, qs[i % 2], 70 + i) for i in range(4)]
    ins = [vec(ctx, x) for x in xs]
    outs = [ctx.vector(n) for _ in xs]
    batch = V.Batch(ctx)
    for i, (a, r) in enumerate(zip(ins, outs)):
        batch.add(V.JOB_NTT_FWD, a, None, r, [tabs[i % 2]])
    for rep in range(3):
        if rep == 2:
            # new inputs, same job list: the next launch transforms them
            xs = [O.fill_uniform(n, qs[i % 2], 170 + i) for i in range(4)]
            for a, x in zip(ins, xs):
                a.write(x)
        for r in outs:
            r.write(np.zeros(n, dtype=np.uint64))
        batch.run()
        for i, r in enumerate(outs):
            assert np.array_equal(r.read(), O.ntt_forward(xs[i], qs[i % 2],
                                                          tabs[i % 2].w)), (rep, i)
    batch.close()
    for x in ins + outs + tabs:
        x.close()


@pytest.mark.parametrize("tb", [11, 12, 13, 14])
def test_fused_polymul_tiles(ctx, tb):
    """Polymul jobs run a's forward, the product and the inverse in one
    kernel (after a's column pass for two-pass sizes): every tile size, one-
    and two-pass sizes, the three modulus classes (50/55-, 59- and 61-bit
    primes) and a 24-bit prime, in one batch, against the oracle's product."""
    lognes = (3, 5, 9, 11, 12, 13, 14, 15, 17, 19)
    jobs = []
    try:
        for logn in lognes:
            ctx.set_tile_bits(logn, tb)
        batch = V.Batch(ctx)
        for logn in lognes:
            n = 1 << logn
            bits = (50, 55, 59, 61) + ((24,) if logn <= 14 else ())
            qs = [O.primes_ntt(1 << b, 2 * n, 1)[0] for b in bits]
            tabs = [V.NttTables(n, q, O.root_2n(q, n)) for q in qs]
            xa = np.concatenate([O.fill_uniform(n, q, 30 + i) for i, q in enumerate(qs)])
            xb = np.concatenate([O.fill_uniform(n, q, 40 + i) for i, q in enumerate(qs)])
            a, b, r = vec(ctx, xa), vec(ctx, xb), ctx.vector(len(qs) * n)
            batch.add(V.JOB_POLYMUL, a, b, r, tabs)
            jobs.append((logn, qs, tabs, xa, xb, a, b, r))
        batch.run()
        for logn, qs, tabs, xa, xb, a, b, r in jobs:
            n = 1 << logn
            got = r.read()
            for i, q in enumerate(qs):
                want = O.polymul_ntt(xa[i * n:(i + 1) * n], xb[i * n:(i + 1) * n],
                                     q, tabs[i].w)
                assert np.array_equal(got[i * n:(i + 1) * n], want), (tb, logn, q)
            # the operands are left as they were
            assert np.array_equal(a.read(), xa) and np.array_equal(b.read(), xb)
        batch.close()
    finally:
        for logn in lognes:
            ctx.set_tile_bits(logn, 0)
        for job in jobs:
            for t in job[2]:
                t.close()


def test_batch_roundtrip_full_size(ctx):
    # 2^16 x 32 polynomials, 4 primes: NTT then INTT is the identity.
    n, per, L = 1 << 16, 8, 4
    qs = O.primes_ntt(1 << 50, 2 * n, L)
    tabs = [V.NttTables(n, q, O.root_2n(q, n)) for q in qs]
    rng = np.random.default_rng(9)
    fwd, inv = V.Batch(ctx), V.Batch(ctx)
    data = []
    for i in range(per * L):
        q = qs[i % L]
        x = rng.integers(0, q, n, dtype=np.uint64)
        v = vec(ctx, x)
        fwd.add(V.JOB_NTT_FWD, v, None, v, [tabs[i % L]])
        inv.add(V.JOB_NTT_INV, v, None, v, [tabs[i % L]])
        data.append((x, v))
    fwd.run()
    x0, v0 = data[0]
    assert np.array_equal(v0.read(), O.ntt_forward(x0, qs[0], tabs[0].w))
    inv.run()
    for x, v in data:
        assert np.array_equal(v.read(), x)


def _eval_at(a, z, q):
    # a(z) mod q by Horner's rule in exact integers
    acc = 0
    for x in reversed(a.tolist()):
        acc = (acc * z + x) % q
    return acc


def test_headline_batch_properties(ctx):
    """The headline workload at its full size: 512 polynomials of n = 2^16
    over 8 primes < 2^50, interleaved polynomial by polynomial, one batch
    per direction.  Properties that hold at any size: the forward transform
    is linear (NTT(a) + NTT(b) = NTT(a + b) mod q on every polynomial);
    sampled outputs equal the polynomial evaluated at psi^(2 bitrev(i) + 1),
    in exact integers and independent of the oracle; and the inverse gives
    every input back."""
    n, primes, per = 1 << 16, 8, 64
    L = primes * per
    qs = O.primes_ntt(1 << 50, 2 * n, primes)
    ws = [O.root_2n(q, n) for q in qs]
    tabs = [V.NttTables(n, q, w) for q, w in zip(qs, ws)]
    qcol = np.array([qs[i % primes] for i in range(L)], dtype=np.uint64)[:, None]
    rng = np.random.default_rng(1 << 16)
    a = rng.integers(0, 1 << 62, (L, n), dtype=np.uint64) % qcol
    b = rng.integers(0, 1 << 62, (L, n), dtype=np.uint64) % qcol
    c = (a + b) % qcol
    slabs = [vec(ctx, m.ravel()) for m in (a, b, c)]
    views = [[s.view(i * n, n) for i in range(L)] for s in slabs]
    fwd, inv = V.Batch(ctx), V.Batch(ctx)
    for vs in views:
        for i, v in enumerate(vs):
            fwd.add(V.JOB_NTT_FWD, v, None, v, [tabs[i % primes]])
    for i, v in enumerate(views[0]):
        inv.add(V.JOB_NTT_INV, v, None, v, [tabs[i % primes]])
    fwd.run()
    fa, fb, fc = (s.read().reshape(L, n) for s in slabs)
    assert np.array_equal((fa + fb) % qcol, fc)
    rng2 = np.random.default_rng(7)
    sample = [(0, 0), (1, 1), (L - 1, n - 1), (L // 2, 12345)] + [
        (int(p), int(i)) for p, i in zip(rng2.integers(0, L, 6),
                                         rng2.integers(0, n, 6))]
    for p, i in sample:
        q, w = qs[p % primes], ws[p % primes]
        z = pow(w, 2 * R.bitrev(i, 16) + 1, q)
        assert int(fa[p, i]) == _eval_at(a[p], z, q), (p, i)
    inv.run()
    assert np.array_equal(slabs[0].read().reshape(L, n), a)
    fwd.close()
    inv.close()
    for vs in views:
        for v in vs:
            v.close()
    for s in slabs:
        s.close()
    for t in tabs:
        t.close()


def test_polymul_batch_properties(ctx):
    """The polynomial-product workload at its full size: 256 products of
    16-limb RNS polynomials (16 primes < 2^55, n = 2^15) in one batch.  For
    the negacyclic product c = a b mod (X^n + 1, q), c(z) = a(z) b(z) mod q
    at every root z of X^n + 1, i.e. at the odd powers of psi: sampled limbs
    are checked there by Horner's rule in exact integers, independent of
    any transform, and one product's every limb against the oracle."""
    n, primes, pairs = 1 << 15, 16, 256
    qs = O.primes_ntt(1 << 55, 2 * n, primes)
    ws = [O.root_2n(q, n) for q in qs]
    tabs = [V.NttTables(n, q, w) for q, w in zip(qs, ws)]
    qcol = np.array(qs, dtype=np.uint64)[:, None]
    rng = np.random.default_rng(1 << 15)
    xs, vs = [], []
    batch = V.Batch(ctx)
    for j in range(pairs):
        xa = rng.integers(0, 1 << 62, (primes, n), dtype=np.uint64) % qcol
        xb = rng.integers(0, 1 << 62, (primes, n), dtype=np.uint64) % qcol
        a, b, r = vec(ctx, xa.ravel()), vec(ctx, xb.ravel()), ctx.vector(primes * n)
        batch.add(V.JOB_POLYMUL, a, b, r, tabs)
        xs.append((xa, xb))
        vs.append((a, b, r))
    batch.run()
    rng2 = np.random.default_rng(15)
    for j, l, k in zip(rng2.integers(0, pairs, 12), rng2.integers(0, primes, 12),
                       rng2.integers(0, n, 12)):
        j, l = int(j), int(l)
        q = qs[l]
        z = pow(ws[l], 2 * int(k) + 1, q)
        xa, xb = xs[j]
        c = vs[j][2].read()[l * n:(l + 1) * n]
        assert _eval_at(c, z, q) == _eval_at(xa[l], z, q) * _eval_at(xb[l], z, q) % q, \
            (j, l, int(k))
    xa, xb = xs[pairs - 1]
    got = vs[pairs - 1][2].read()
    for l, q in enumerate(qs):
        want = O.polymul_ntt(xa[l], xb[l], q, ws[l])
        assert np.array_equal(got[l * n:(l + 1) * n], want), l
    batch.close()
    for trio in vs:
        for v in trio:
            v.close()
    for t in tabs:
        t.close()


# ---------------- tile sizes, split vs whole-polynomial kernels ----------------

@pytest.mark.parametrize("bits", [55, 59])
@pytest.mark.parametrize("mode", [1, 2])
@pytest.mark.parametrize("tb", [11, 12, 13, 14])
@pytest.mark.parametrize("logn", [12, 13, 14, 15, 16, 17, 18, 19, 20])
def test_tile_sizes_per_call_and_batched(ctx, logn, tb, mode, bits):
    """Every compiled tile size in both pass modes (1 = one workgroup per
    polynomial running both passes, 2 = one grid per pass, one workgroup per
    tile) for moduli of the small (55-bit) and mid (59-bit) classes."""
    n = 1 << logn
    q = _prime(logn, bits)
    w = O.root_2n(q, n)
    t = V.NttTables(n, q, w)
    ctx.set_tile_bits(logn, tb)
    ctx.set_pass_mode(logn, mode)
    try:
        # per-call path: one polynomial
        x = O.fill_uniform(n, q, seed=logn + tb)
        a = vec(ctx, x)
        f = ctx.vector(n)
        a.forward_transform(f, t)
        want = O.ntt_forward(x, q, w)
        assert np.array_equal(f.read(), want)
        f.inverse_transform(f, t)
        assert np.array_equal(f.read(), x)
        # a batch through the task array (80 polynomials; 6 at 2^19 and up)
        npoly = 80 if logn <= 18 else 6
        rng = np.random.default_rng(logn * 100 + tb + bits)
        xs = [rng.integers(0, q, n, dtype=np.uint64) for _ in range(npoly)]
        vs = [vec(ctx, z) for z in xs]
        fb, ib = V.Batch(ctx), V.Batch(ctx)
        for v in vs:
            fb.add(V.JOB_NTT_FWD, v, None, v, [t])
            ib.add(V.JOB_NTT_INV, v, None, v, [t])
        fb.run()
        for i in (0, npoly // 2, npoly - 1):
            assert np.array_equal(vs[i].read(), O.ntt_forward(xs[i], q, w)), i
        ib.run()
        for z, v in zip(xs, vs):
            assert np.array_equal(v.read(), z)
        for v in vs + [a, f]:
            v.close()
        fb.close()
        ib.close()
    finally:
        ctx.set_tile_bits(logn, 0)
        ctx.set_pass_mode(logn, 0)
        t.close()


def _tuning(ctx, l):
    return (ctx.tile_bits(l), ctx.pass_mode(l), ctx.split(l), ctx.tile_bits2(l))


def _untune(ctx, l):
    ctx.set_tile_bits(l, 0)
    ctx.set_pass_mode(l, 0)
    ctx.set_split(l, 0)
    ctx.set_tile_bits2(l, 0)


def test_autotune_choices_and_cache(ctx, tmp_path):
    ctx.autotune(12, 16, 64)
    chosen = {l: _tuning(ctx, l) for l in (12, 13, 14, 15, 16)}
    assert all(tb in (11, 12, 13, 14) and m in (1, 2, 3) and sp in (0, 1) and
               tb2 in (0, 11, 12, 13, 14) for tb, m, sp, tb2 in chosen.values()), chosen
    path = tmp_path / "tune.txt"
    ctx.autotune_save(path)
    for l in chosen:
        _untune(ctx, l)
    ctx.autotune_load(path)
    assert {l: _tuning(ctx, l) for l in chosen} == chosen
    # files of three fields per line (tile, pass mode) still load
    old = tmp_path / "old.txt"
    old.write_text("16 12 2\n14 11 1\n")
    ctx.autotune_load(old)
    assert _tuning(ctx, 16) == (12, 2, 0, 0) and _tuning(ctx, 14) == (11, 1, 0, 0)
    for l in chosen:
        _untune(ctx, l)


@pytest.mark.parametrize("logn,tb,sp,tb2", [
    (14, 12, 1, 0), (14, 11, 0, 13), (14, 13, 1, 12), (15, 12, 1, 14),
    (16, 12, 1, 0), (16, 13, 0, 11), (16, 11, 0, 14), (16, 12, 1, 13),
    (17, 12, 1, 0), (17, 13, 1, 14), (17, 12, 0, 13), (19, 14, 1, 0),
    (19, 13, 0, 12)])
def test_split_and_row_pass_tile_choices(ctx, logn, tb, sp, tb2):
    """The alternative stage split and a row-pass tile of its own give the
    same words as the oracle, single calls and batches, both directions."""
    n = 1 << logn
    q = _prime(logn, 55)
    w = O.root_2n(q, n)
    t = V.NttTables(n, q, w)
    ctx.set_tile_bits(logn, tb)
    ctx.set_pass_mode(logn, 2)
    ctx.set_split(logn, sp)
    ctx.set_tile_bits2(logn, tb2)
    try:
        rng = np.random.default_rng(logn * 1000 + tb * 10 + sp + tb2)
        npoly = 24 if logn <= 17 else 3
        xs = [rng.integers(0, q, n, dtype=np.uint64) for _ in range(npoly)]
        vs = [vec(ctx, z) for z in xs]
        fb, ib = V.Batch(ctx), V.Batch(ctx)
        for v in vs:
            fb.add(V.JOB_NTT_FWD, v, None, v, [t])
            ib.add(V.JOB_NTT_INV, v, None, v, [t])
        fb.run()
        for i in (0, npoly - 1):
            assert np.array_equal(vs[i].read(), O.ntt_forward(xs[i], q, w)), i
        ib.run()
        for z, v in zip(xs, vs):
            assert np.array_equal(v.read(), z)
        a = vec(ctx, xs[0])
        a.forward_transform(a, t)
        assert np.array_equal(a.read(), O.ntt_forward(xs[0], q, w))
        a.inverse_transform(a, t)
        assert np.array_equal(a.read(), xs[0])
        for v in vs + [a]:
            v.close()
        fb.close()
        ib.close()
    finally:
        _untune(ctx, logn)
        t.close()


# ---------------- asynchronous transfers ----------------

def test_async_transfers_on_a_copy_stream(ctx):
    """Pinned host buffers, uploads and downloads on a second stream ordered
    against the library's stream by events: results equal the oracle's."""
    from vkhel_amd import hipstream as H
    n, npoly = 1 << 14, 6
    qs = O.primes_ntt(1 << 50, 2 * n, 2)
    tabs = [V.NttTables(n, q, O.root_2n(q, n)) for q in qs]
    xs = [O.fill_uniform(n, qs[i % 2], seed=300 + i) for i in range(npoly)]
    up, down = V.HostBuffer(npoly * n), V.HostBuffer(npoly * n)
    for i, x in enumerate(xs):
        up.array[i * n:(i + 1) * n] = x
    down.array[:] = 0
    vecs = [ctx.vector(n, False) for _ in range(npoly)]
    batch = V.Batch(ctx)
    for i, v in enumerate(vecs):
        batch.add(V.JOB_NTT_FWD, v, None, v, [tabs[i % 2]])
    batch.run()  # plan and capture before timing-free use below
    copy = H.Stream()
    uploaded, computed = H.Event(), H.Event()
    try:
        for i, v in enumerate(vecs):
            v.write_async(up.array[i * n:(i + 1) * n], 0, copy)
        uploaded.record(copy)
        H.wait(ctx.stream, uploaded)
        batch.launch()
        computed.record(ctx.stream)
        copy.wait(computed)
        for i, v in enumerate(vecs):
            v.read_async(down.array[i * n:(i + 1) * n], 0, copy)
        copy.synchronize()
        for i, x in enumerate(xs):
            want = O.ntt_forward(x, qs[i % 2], tabs[i % 2].w)
            assert np.array_equal(down.array[i * n:(i + 1) * n], want), i
    finally:
        copy.close()
        uploaded.close()
        computed.close()
        batch.close()
        for v in vecs:
            v.close()
        up.close()
        down.close()
        for t in tabs:
            t.close()


# ---------------- RNS elementwise batch ops ----------------

@pytest.mark.parametrize("p", [2, 3, 65537, (1 << 40) + 15])
def test_batch_rns_ops_match_single_calls(ctx, p):
    """CENTER / GTADD / GTSUB over every limb in one batch launch equal the
    oracle's elemmod / elemgtadd / elemgtsub per limb and the single calls."""
    n, L = 1 << 10, 3
    qs = [O.primes_ntt(1 << b, 2 * n, 1)[0] for b in (40, 50, 60)]
    tabs = [V.NttTables(n, q, O.root_2n(q, n)) for q in qs]
    rng = np.random.default_rng(p % 1000)
    # raw 64-bit words: every branch of the reference arithmetic runs
    x = rng.integers(0, U64MAX, L * n, dtype=np.uint64, endpoint=True)
    x[:n] = np.array([O.fill_uniform(n, qs[0], 1)]).ravel()  # limb 0 canonical
    a = vec(ctx, x)
    r_center, r_gtadd, r_gtsub = (ctx.vector(L * n) for _ in range(3))
    gtadd_prm = []
    gtsub_prm = []
    for q in qs:
        gtadd_prm += [q // 2, (1 << 64) - q]
        gtsub_prm += [q // 2, q, p if p > 2 else 7]
    batch = V.Batch(ctx)
    batch.add_rns(V.JOB_CENTER, a, r_center, tabs, [p])
    batch.add_rns(V.JOB_GTADD, a, r_gtadd, tabs, gtadd_prm)
    batch.add_rns(V.JOB_GTSUB, a, r_gtsub, tabs, gtsub_prm)
    batch.run()
    gc, ga, gs = r_center.read(), r_gtadd.read(), r_gtsub.read()
    for i, q in enumerate(qs):
        xi = x[i * n:(i + 1) * n]
        sl = slice(i * n, (i + 1) * n)
        assert np.array_equal(gc[sl], O.elemmod(xi, p, q)), ("center", i)
        assert np.array_equal(ga[sl], O.elemgtadd(xi, gtadd_prm[2 * i],
                                                  gtadd_prm[2 * i + 1])), ("gtadd", i)
        assert np.array_equal(gs[sl], O.elemgtsub(xi, *gtsub_prm[3 * i:3 * i + 3])), \
            ("gtsub", i)
        # the single-call API on a limb view gives the same words
        single = ctx.vector(n)
        a.view(i * n, n).elemmod(single, p, q)
        assert np.array_equal(single.read(), gc[sl]), ("single center", i)
        single.close()
    batch.close()
    for v in (a, r_center, r_gtadd, r_gtsub):
        v.close()
    for t in tabs:
        t.close()


def test_batch_rns_ops_reject_bad_arguments(ctx):
    n = 1 << 8
    q = O.primes_ntt(1 << 50, 2 * n, 1)[0]
    t = V.NttTables(n, q, O.root_2n(q, n))
    a, r = ctx.vector(n), ctx.vector(n)
    batch = V.Batch(ctx)
    with pytest.raises(V.VkhelError):
        batch.add_rns(V.JOB_CENTER, a, r, [t], [1])       # p < 2
    with pytest.raises(V.VkhelError):
        batch.add_rns(V.JOB_GTSUB, a, r, [t], [1, 2, 0])  # mod < 2
    with pytest.raises(V.VkhelError):
        batch.add_rns(V.JOB_NTT_FWD, a, r, [t], [0])      # not an RNS op
    short = ctx.vector(n // 2)
    with pytest.raises(V.VkhelError):
        batch.add_rns(V.JOB_GTADD, a, short, [t], [0, 0])
    batch.close()
    for v in (a, r, short):
        v.close()
    t.close()


# ---------------- on-device table generation ----------------

def _packed_shoup63(r, q):
    wp = (int(r) << 63) // q
    return wp + (wp & 0xFFFFFFFF00000000)


@pytest.mark.parametrize("n,q,w", [
    (2, 5, 2), (8, 17, 3), (8, 17, 1),           # tiny moduli, w = 1 included
    (1 << 10, None, None), (1 << 16, None, None), (1 << 17, None, None),
])
def test_device_tables_match_host_tables(ctx, n, q, w):
    """Tables built on the GPU equal the reference's host arrays (any
    invertible w, primitive or not) and carry exact Shoup companions."""
    if q is None:
        q = O.primes_ntt(1 << 60 if n > 1 << 16 else 1 << 50, 2 * n, 1)[0]
        w = O.root_2n(q, n)
    tab = V.NttTables(n, q, w)
    roots, roots_p, inv, inv_p = tab.read_device(ctx)
    want = O.tables(n, q, w)
    assert np.array_equal(roots, want[0])
    assert np.array_equal(inv, want[1])
    idx = range(n) if n <= 1024 else list(range(64)) + list(range(n - 64, n)) + \
        list(np.random.default_rng(n).integers(0, n, 2000))
    for k in idx:
        assert int(roots_p[k]) == _packed_shoup63(roots[k], q), k
        assert int(inv_p[k]) == _packed_shoup63(inv[k], q), k
    tab.close()


def test_device_created_tables_validate_w_and_transform_alike(ctx):
    n = 1 << 12
    q = O.primes_ntt(1 << 55, 2 * n, 1)[0]
    w = O.root_2n(q, n)
    # w^2 has order n, and 1 has order 1: neither is a primitive 2n-th root.
    for bad in (1, w * w % q):
        with pytest.raises(V.VkhelError) as e:
            V.NttTables.on_device(ctx, n, q, bad)
        assert e.value.code == 1
    host = V.NttTables(n, q, w)
    dev = V.NttTables.on_device(ctx, n, q, w)
    assert dev.root_ok and host.root_ok
    assert not V.NttTables(n, q, w * w % q).root_ok
    # host arrays of a device-made table are computed on request
    for g, e in zip(dev.arrays(), O.tables(n, q, w)):
        assert np.array_equal(g, e)
    rng = np.random.default_rng(7)
    a = rng.integers(0, q, n, dtype=np.uint64)
    outs = []
    for tab in (host, dev):
        v = ctx.from_array(a)
        f = V.Vector(ctx, n)
        b = V.Vector(ctx, n)
        v.forward_transform(f, tab)
        f.inverse_transform(b, tab)
        outs.append((f.read(), b.read()))
        for x in (v, f, b):
            x.close()
    assert np.array_equal(outs[0][0], outs[1][0])
    assert np.array_equal(outs[1][0], O.ntt_forward(a, q, w))
    assert np.array_equal(outs[1][1], a)
    host.close()
    dev.close()


def test_batches_from_several_threads_with_single_calls_between(ctx):
    """Batches built and launched from different threads, while a third
    thread issues single calls on the same context: every result is the
    one each computes alone, and the completion handles report it."""
    import threading
    n = 1 << 12
    q = _prime(12, 50)
    w = O.root_2n(q, n)
    t = V.NttTables(n, q, w)
    rng = np.random.default_rng(5)
    xs = [[rng.integers(0, q, n, dtype=np.uint64) for _ in range(24)]
          for _ in range(2)]
    vs = [[vec(ctx, x) for x in row] for row in xs]
    outs = [[ctx.vector(n) for _ in row] for row in xs]
    ys = [rng.integers(0, q, n, dtype=np.uint64) for _ in range(16)]
    errors = []

    def batch_worker(k):
        try:
            for rep in range(4):
                b = V.Batch(ctx)
                for v, o in zip(vs[k], outs[k]):
                    b.add(V.JOB_NTT_FWD, v, None, o, [t])
                b.launch()
                b.wait()
                assert b.done()
                b.close()
        except Exception as e:  # reported by the main thread
            errors.append(e)

    def single_worker():
        try:
            for y in ys:
                a = vec(ctx, y)
                f = ctx.vector(n)
                a.forward_transform(f, t)
                assert np.array_equal(f.read(), O.ntt_forward(y, q, w))
                a.close()
                f.close()
        except Exception as e:
            errors.append(e)

    th = [threading.Thread(target=batch_worker, args=(k,)) for k in range(2)]
    th.append(threading.Thread(target=single_worker))
    for x in th:
        x.start()
    for x in th:
        x.join(timeout=100)
    assert not any(x.is_alive() for x in th), "a worker thread did not finish"
    assert not errors, errors
    for k in range(2):
        for i in (0, 11, 23):
            assert np.array_equal(outs[k][i].read(), O.ntt_forward(xs[k][i], q, w))
    # query before and after an explicit wait; a rebuilt batch waits for the
    # previous launch before touching its buffers
    b = V.Batch(ctx)
    for v, o in zip(vs[0], outs[0]):
        b.add(V.JOB_NTT_INV, o, None, v, [t])
    b.launch()
    b.add(V.JOB_NTT_FWD, vs[1][0], None, outs[1][0], [t])
    b.launch()
    b.wait()
    assert b.done()
    for i in (0, 5, 23):
        assert np.array_equal(vs[0][i].read(), xs[0][i])
    b.close()
    for row in vs + outs:
        for v in row:
            v.close()
    t.close()


@pytest.mark.parametrize("p", [2, 3, 65537, (1 << 40) + 15])
@pytest.mark.parametrize("logn,tb,mode", [(3, 0, 0), (10, 0, 0), (12, 0, 0),
                                          (16, 12, 2), (16, 11, 1), (17, 13, 2)])
def test_fused_intt_center_matches_intt_then_elemmod(ctx, p, logn, tb, mode):
    """INTT_CENTER (the centered lift applied by the inverse kernel as it
    stores) equals the oracle's INTT followed by elemmod, per limb, for
    moduli of both kernel classes, and the two single calls in sequence."""
    n, L = 1 << logn, 3
    qs = [O.primes_ntt(1 << b, 2 * n, 1)[0] for b in (50, 55, 61)]
    ws = [O.root_2n(q, n) for q in qs]
    tabs = [V.NttTables(n, q, w) for q, w in zip(qs, ws)]
    if tb:
        ctx.set_tile_bits(logn, tb)
        ctx.set_pass_mode(logn, mode)
    try:
        x = np.concatenate([O.fill_uniform(n, q, seed=logn + i)
                            for i, q in enumerate(qs)])
        a = vec(ctx, x)
        r = ctx.vector(L * n)
        b = V.Batch(ctx)
        b.add_rns(V.JOB_INTT_CENTER, a, r, tabs, [p])
        b.run()
        got = r.read()
        for i, (q, w) in enumerate(zip(qs, ws)):
            sl = slice(i * n, (i + 1) * n)
            want = O.elemmod(O.ntt_inverse(x[sl], q, w), p, q)
            assert np.array_equal(got[sl], want), (i, q)
            if i == 0:
                s1, s2 = ctx.vector(n), ctx.vector(n)
                a.view(0, n).inverse_transform(s1, tabs[0])
                s1.elemmod(s2, p, q)
                assert np.array_equal(s2.read(), got[sl])
                s1.close()
                s2.close()
        b.close()
        a.close()
        r.close()
    finally:
        ctx.set_tile_bits(logn, 0)
        ctx.set_pass_mode(logn, 0)
        for t in tabs:
            t.close()


@pytest.mark.parametrize("p", [2, 3, 65537])
def test_intt_center_outside_fused_sizes(ctx, p):
    """n = 2, 4 and 2^21, outside the fused kernels: the inverse transform,
    then the centered lift as a pointwise pass, in the same job."""
    for logn in (1, 2, 21):
        n = 1 << logn
        qs = [O.primes_ntt(1 << b, 2 * n, 1)[0] for b in (50, 61)]
        ws = [O.root_2n(q, n) for q in qs]
        tabs = [V.NttTables(n, q, w) for q, w in zip(qs, ws)]
        x = np.concatenate([O.fill_uniform(n, q, 40 + i) for i, q in enumerate(qs)])
        a, r = vec(ctx, x), ctx.vector(2 * n)
        b = V.Batch(ctx)
        b.add_rns(V.JOB_INTT_CENTER, a, r, tabs, [p])
        b.run()
        got = r.read()
        for i, (q, w) in enumerate(zip(qs, ws)):
            sl = slice(i * n, (i + 1) * n)
            assert np.array_equal(got[sl], O.elemmod(O.ntt_inverse(x[sl], q, w), p, q)), (n, i)
        b.close()
        for v in (a, r):
            v.close()
        for t in tabs:
            t.close()


def test_intt_center_rejects_bad_modulus(ctx):
    n = 1 << 8
    q = O.primes_ntt(1 << 50, 2 * n, 1)[0]
    t = V.NttTables(n, q, O.root_2n(q, n))
    a, r = ctx.vector(n), ctx.vector(n)
    b = V.Batch(ctx)
    with pytest.raises(V.VkhelError):
        b.add_rns(V.JOB_INTT_CENTER, a, r, [t], [1])  # p < 2
    b.close()
    a.close()
    r.close()
    t.close()


@pytest.mark.parametrize("bits", [50, 61])
def test_ntt_above_fused_sizes(ctx, bits):
    """n = 2^21, beyond the fused kernels' 2^20: the per-stage path, per call
    and in a batch (including a polymul, whose pointwise product then runs
    as its own kernel), against the oracle."""
    logn = 21
    n = 1 << logn
    q = _prime(logn, bits)
    w = O.root_2n(q, n)
    t = V.NttTables.on_device(ctx, n, q, w)
    x = O.fill_uniform(n, q, seed=bits)
    y = O.fill_uniform(n, q, seed=bits + 1)
    want = O.ntt_forward(x, q, w)
    a = vec(ctx, x)
    f = ctx.vector(n)
    a.forward_transform(f, t)
    assert np.array_equal(f.read(), want)
    f.inverse_transform(f, t)
    assert np.array_equal(f.read(), x)
    # batch: forward job and a polymul whose product is checked through the
    # convolution theorem on the oracle's transforms
    bv, r, p = vec(ctx, y), ctx.vector(n), ctx.vector(n)
    batch = V.Batch(ctx)
    batch.add(V.JOB_NTT_FWD, a, None, r, [t])
    batch.add(V.JOB_POLYMUL, a, bv, p, [t])
    batch.run()
    assert np.array_equal(r.read(), want)
    prod = O.ntt_inverse(O.elemmul(want, O.ntt_forward(y, q, w), q), q, w)
    assert np.array_equal(p.read(), prod)
    batch.close()
    for v in (a, f, bv, r, p):
        v.close()
    t.close()


def _rescale_oracle(limbs, qs):
    """Exact big-integer statement of RNS rescaling by the last modulus:
    A from the residues (CRT), t the centered residue of A mod q_last,
    A' = (A - t) / q_last exactly; the residues of A' mod the other q."""
    Q = 1
    for q in qs:
        Q *= q
    ql = qs[-1]
    n = len(limbs[0])
    out = [np.zeros(n, dtype=np.uint64) for _ in qs[:-1]]
    coef = [(Q // q) * pow(Q // q, -1, q) for q in qs]
    for j in range(n):
        A = sum(int(limbs[i][j]) * coef[i] for i in range(len(qs))) % Q
        r = A % ql
        t = r - ql if r > ql // 2 else r
        Ap = (A - t) // ql
        assert (A - t) % ql == 0
        assert Ap == (2 * A + ql) // (2 * ql)  # the rounded quotient A / q_last
        for i, q in enumerate(qs[:-1]):
            out[i][j] = Ap % q
    return out


@pytest.mark.parametrize("L", [2, 3, 4])
def test_rns_rescale_matches_big_integer_division(ctx, L):
    n = 1 << 10
    qs = [O.primes_ntt(1 << b, 2 * n, 1)[0] for b in (50, 61, 55, 40)[:L]]
    ws = [O.root_2n(q, n) for q in qs]
    tabs = [V.NttTables(n, q, w) for q, w in zip(qs, ws)]
    rng = np.random.default_rng(L)
    limbs = [rng.integers(0, q, n, dtype=np.uint64) for q in qs]
    # the dropped limb's centered-lift boundary values
    limbs[-1][:4] = [0, qs[-1] // 2, qs[-1] // 2 + 1, qs[-1] - 1]
    a = vec(ctx, np.concatenate(limbs))
    r = ctx.vector((L - 1) * n)
    b = V.Batch(ctx)
    b.add_rns(V.JOB_RESCALE, a, r, tabs)
    b.run()
    want = _rescale_oracle(limbs, qs)
    got = r.read()
    for i in range(L - 1):
        assert np.array_equal(got[i * n:(i + 1) * n], want[i]), i
    # INTT then rescale in one batch: the rescale runs after the transforms
    x = ctx.vector(L * n)
    spec = vec(ctx, np.concatenate([O.ntt_forward(l, q, w)
                                    for l, q, w in zip(limbs, qs, ws)]))
    r2 = ctx.vector((L - 1) * n)
    b2 = V.Batch(ctx)
    b2.add(V.JOB_NTT_INV, spec, None, x, tabs)
    b2.add_rns(V.JOB_RESCALE, x, r2, tabs)
    b2.run()
    assert np.array_equal(r2.read(), got)
    # rejected: one limb, result aliasing the input
    with pytest.raises(V.VkhelError):
        b2.add_rns(V.JOB_RESCALE, a, r, tabs[:1])
    with pytest.raises(V.VkhelError):
        b2.add_rns(V.JOB_RESCALE, a, a, tabs)
    for o in (b, b2):
        o.close()
    for v in (a, r, x, spec, r2):
        v.close()
    for t in tabs:
        t.close()


@pytest.mark.parametrize("logn,bits", [(2, (50, 61)), (4, (50, 61)),
                                       (10, (50, 61, 55, 40)), (14, (59, 50, 61)),
                                       (16, (50, 55, 60, 59)), (21, (50, 61))])
def test_rns_rescale_ntt_form(ctx, logn, bits):
    """Rescaling in NTT form: the INTT of each limb, the big-integer rounded
    division by the last modulus, the NTT of each remaining limb, against one
    RESCALE_NTT job; one- and two-pass sizes, every modulus class, and sizes
    outside the fused kernels (n = 4, 2^21: lift and scaling as pointwise
    passes around the radix-2 transforms)."""
    n = 1 << logn
    qs = [O.primes_ntt(1 << b, 2 * n, 1)[0] for b in bits]
    ws = [O.root_2n(q, n) for q in qs]
    tabs = [V.NttTables(n, q, w) for q, w in zip(qs, ws)]
    L = len(qs)
    spec = [O.fill_uniform(n, q, 300 + i) for i, q in enumerate(qs)]
    coeffs = [O.ntt_inverse(s, q, w) for s, q, w in zip(spec, qs, ws)]
    # the dropped limb's centered-lift boundary values, set in coefficient form
    coeffs[-1][:4] = [0, qs[-1] // 2, qs[-1] // 2 + 1, qs[-1] - 1]
    spec[-1] = O.ntt_forward(coeffs[-1], qs[-1], ws[-1])
    if n > 1 << 12:  # the big-integer oracle on a sample of coefficients
        keep = np.r_[0:64, n - 64:n]
    else:
        keep = np.arange(n)
    want_c = _rescale_oracle([c[keep] for c in coeffs], qs)
    a = vec(ctx, np.concatenate(spec))
    r = ctx.vector((L - 1) * n)
    b = V.Batch(ctx)
    b.add_rns(V.JOB_RESCALE_NTT, a, r, tabs)
    b.run()
    got = r.read()
    for i in range(L - 1):
        got_c = O.ntt_inverse(got[i * n:(i + 1) * n], qs[i], ws[i])
        assert np.array_equal(got_c[keep], want_c[i]), i
    # the whole result against the coefficient-form job on the same limbs
    c = vec(ctx, np.concatenate(coeffs))
    r2 = ctx.vector((L - 1) * n)
    b2 = V.Batch(ctx)
    b2.add_rns(V.JOB_RESCALE, c, r2, tabs)
    b2.run()
    flat = r2.read()
    for i in range(L - 1):
        assert np.array_equal(got[i * n:(i + 1) * n],
                              O.ntt_forward(flat[i * n:(i + 1) * n], qs[i], ws[i])), i
    with pytest.raises(V.VkhelError):
        b2.add_rns(V.JOB_RESCALE_NTT, a, a, tabs)
    with pytest.raises(V.VkhelError):
        b2.add_rns(V.JOB_RESCALE_NTT, a, r, tabs[:1])
    for o in (b, b2):
        o.close()
    for v in (a, r, c, r2):
        v.close()
    for t in tabs:
        t.close()


def test_batch_mixes_every_job_kind(ctx):
    """Every job kind in one batch (forward, inverse, polymul, mul, add, the
    RNS elementwise ops, INTT_CENTER and both rescalings) over sizes 2^5,
    2^12 and 2^15, so every phase runs with several segments and the
    polymul and rescale scratch is carved for many jobs; each result
    against the oracle, then again after a replay."""
    rng = np.random.default_rng(5)
    batch = V.Batch(ctx)
    keep, checks = [], []
    for k in range(30):
        logn = (5, 12, 15)[k % 3]
        n = 1 << logn
        L = 2 + k % 3
        qs = [O.primes_ntt(1 << b, 2 * n, 1)[0] for b in (50, 59, 61, 55)[:L]]
        ws = [O.root_2n(q, n) for q in qs]
        tabs = [V.NttTables(n, q, w) for q, w in zip(qs, ws)]
        xa = [O.fill_uniform(n, q, 900 + 10 * k + i) for i, q in enumerate(qs)]
        xb = [O.fill_uniform(n, q, 950 + 10 * k + i) for i, q in enumerate(qs)]
        a, b = vec(ctx, np.concatenate(xa)), vec(ctx, np.concatenate(xb))
        kind = k % 10
        if kind <= 4:
            r = ctx.vector(L * n)
            batch.add(kind, a, b if kind >= 2 else None, r, tabs)
            fn = [lambda x, y, q, w: O.ntt_forward(x, q, w),
                  lambda x, y, q, w: O.ntt_inverse(x, q, w),
                  lambda x, y, q, w: O.polymul_ntt(x, y, q, w),
                  lambda x, y, q, w: O.elemmul(x, y, q),
                  lambda x, y, q, w: O.addmul_mod(x, y, 1, q)][kind]
            want = [fn(x, y, q, w) for x, y, q, w in zip(xa, xb, qs, ws)]
        elif kind == 5:
            r = ctx.vector(L * n)
            batch.add_rns(V.JOB_CENTER, a, r, tabs, [65537])
            want = [O.elemmod(x, 65537, q) for x, q in zip(xa, qs)]
        elif kind == 6:
            r = ctx.vector(L * n)
            prm = [int(v) for q in qs for v in (q // 3, 12345)]
            batch.add_rns(V.JOB_GTADD, a, r, tabs, prm)
            want = [O.elemgtadd(x, q // 3, 12345) for x, q in zip(xa, qs)]
        elif kind == 7:
            r = ctx.vector(L * n)
            batch.add_rns(V.JOB_INTT_CENTER, a, r, tabs, [3])
            want = [O.elemmod(O.ntt_inverse(x, q, w), 3, q)
                    for x, q, w in zip(xa, qs, ws)]
        else:
            r = ctx.vector((L - 1) * n)
            coeffs = xa if kind == 8 else [O.ntt_inverse(x, q, w)
                                           for x, q, w in zip(xa, qs, ws)]
            keep_i = np.arange(min(n, 64))
            rc = _rescale_oracle([c[keep_i] for c in coeffs], qs)
            batch.add_rns(V.JOB_RESCALE if kind == 8 else V.JOB_RESCALE_NTT,
                          a, r, tabs)
            want = (kind, keep_i, rc, qs, ws)
        keep += [a, b, r] + tabs
        checks.append((k, n, r, want))

    def verify():
        for k, n, r, want in checks:
            got = r.read()
            if isinstance(want, tuple):
                kind, keep_i, rc, qs, ws = want
                for i in range(len(qs) - 1):
                    g = got[i * n:(i + 1) * n]
                    if kind == 9:
                        g = O.ntt_inverse(g, qs[i], ws[i])
                    assert np.array_equal(g[keep_i], rc[i]), (k, i)
            else:
                for i, wv in enumerate(want):
                    assert np.array_equal(got[i * n:(i + 1) * n], wv), (k, i)

    batch.run()
    verify()
    for _, _, r, _ in checks:
        r.write(np.zeros(r.length, dtype=np.uint64))
    batch.run()
    verify()
    batch.close()
    for x in keep:
        x.close()


@pytest.mark.parametrize("mode", [1, 2])
def test_rns_rescale_ntt_pass_modes(ctx, mode):
    """The rescaling prologue and epilogue in both two-pass kernel forms
    (one workgroup per polynomial running both passes, or split passes)
    give the coefficient-form job's words, transformed."""
    logn = 14
    n = 1 << logn
    qs = [O.primes_ntt(1 << b, 2 * n, 1)[0] for b in (61, 50, 59)]
    ws = [O.root_2n(q, n) for q in qs]
    tabs = [V.NttTables(n, q, w) for q, w in zip(qs, ws)]
    spec = np.concatenate([O.fill_uniform(n, q, 77 + i) for i, q in enumerate(qs)])
    coeffs = np.concatenate([O.ntt_inverse(spec[i * n:(i + 1) * n], q, w)
                             for i, (q, w) in enumerate(zip(qs, ws))])
    a, c = vec(ctx, spec), vec(ctx, coeffs)
    r, rc = ctx.vector(2 * n), ctx.vector(2 * n)
    ctx.set_pass_mode(logn, mode)
    try:
        b = V.Batch(ctx)
        b.add_rns(V.JOB_RESCALE_NTT, a, r, tabs)
        b.add_rns(V.JOB_RESCALE, c, rc, tabs)
        b.run()
        got, flat = r.read(), rc.read()
        for i in range(2):
            assert np.array_equal(got[i * n:(i + 1) * n],
                                  O.ntt_forward(flat[i * n:(i + 1) * n], qs[i], ws[i])), i
        b.close()
    finally:
        ctx.set_pass_mode(logn, 0)
        for v in (a, c, r, rc):
            v.close()
        for t in tabs:
            t.close()


def test_batch_with_70000_limbs(ctx):
    """More limbs than a grid dimension holds (65535): a MUL job and an
    NTT_FWD job of 70000 limbs at n = 8 (one pointwise row and one
    transform workgroup per limb), and an NTT_FWD job of 70000 limbs at
    n = 2^12 on 2^11 tiles (two passes, run in chunks), sampled against the
    oracle."""
    L = 70000
    q = O.primes_ntt(1 << 50, 1 << 13, 1)[0]
    rng = np.random.default_rng(70000)
    # 0, the grid's y limit, and chunk edges (2^25 words: 8192 limbs at 2^12)
    sample = rng.choice(L, 24, replace=False).tolist() + [
        0, 8191, 8192, 16383, 16384, 65534, 65535, L - 1]
    for n in (8, 1 << 12):
        t = V.NttTables(n, q, O.root_2n(q, n))
        x = rng.integers(0, q, L * n, dtype=np.uint64)
        y = rng.integers(0, q, L * n, dtype=np.uint64) if n == 8 else None
        a, r = vec(ctx, x), ctx.vector(L * n)
        ctx.set_tile_bits(12, 11)
        try:
            b = V.Batch(ctx)
            b.add(V.JOB_NTT_FWD, a, None, r, [t] * L)
            if n == 8:
                bv, m = vec(ctx, y), ctx.vector(L * n)
                b.add(V.JOB_MUL, a, bv, m, [t] * L)
            b.run()
            got = r.read()
            for i in sample:
                sl = slice(i * n, (i + 1) * n)
                assert np.array_equal(got[sl], O.ntt_forward(x[sl], q, t.w)), (n, i)
            if n == 8:
                assert np.array_equal(m.read(), O.elemmul(x, y, q))
                bv.close()
                m.close()
            b.close()
        finally:
            ctx.set_tile_bits(12, 0)
        a.close()
        r.close()
        t.close()


@pytest.mark.parametrize("logn,L,tb", [(16, 513, 11), (16, 513, 12),
                                       (17, 257, 12)])
def test_two_pass_batch_across_chunk_edge(ctx, logn, L, tb):
    """One chunk of 2^25 words and one of a single polynomial: L limbs
    over three moduli interleaved limb by limb, in one batch job.  The second
    pass walks each chunk's polynomials from the last, and from 2^17 up the
    tasks are grouped by table first; every limb's forward must equal the
    oracle's, and the inverse batch must give the input back."""
    n = 1 << logn
    qs = O.primes_ntt(1 << 50, 2 * n, 3)
    ws = [O.root_2n(q, n) for q in qs]
    tabs = [V.NttTables(n, q, w) for q, w in zip(qs, ws)]
    per = [tabs[i % 3] for i in range(L)]
    rng = np.random.default_rng(L + tb)
    x = np.concatenate([rng.integers(0, qs[i % 3], n, dtype=np.uint64)
                        for i in range(L)])
    a, r, back = vec(ctx, x), ctx.vector(L * n), ctx.vector(L * n)
    ctx.set_tile_bits(logn, tb)
    ctx.set_pass_mode(logn, 2)
    try:
        fwd, inv = V.Batch(ctx), V.Batch(ctx)
        fwd.add(V.JOB_NTT_FWD, a, None, r, per)
        inv.add(V.JOB_NTT_INV, r, None, back, per)
        fwd.run()
        got = r.read()
        for i in range(L):
            sl = slice(i * n, (i + 1) * n)
            assert np.array_equal(got[sl], O.ntt_forward(x[sl], qs[i % 3],
                                                         ws[i % 3])), i
        inv.run()
        assert np.array_equal(back.read(), x)
        fwd.close()
        inv.close()
    finally:
        ctx.set_tile_bits(logn, 0)
        ctx.set_pass_mode(logn, 0)
    for v in (a, r, back):
        v.close()
    for t in tabs:
        t.close()


@pytest.mark.parametrize("tb", [11, 12, 13, 14])
@pytest.mark.parametrize("logn", [14, 16, 17])
def test_one_launch_two_pass_matches_split(ctx, logn, tb):
    """Pass mode 3 (both passes in one launch, k_ntt_handoff: per-XCD ticket
    queues, producer tiles handing each polynomial to its consumer tiles
    through the XCD's L2) gives the split kernels' words, forward and
    inverse, over several primes and polynomial counts (including fewer
    polynomials than XCDs); the batch's completion check passes."""
    n = 1 << logn
    qs = O.primes_ntt(1 << 50, 2 * n, 3) + O.primes_ntt(1 << 60, 2 * n, 1)
    tabs = [V.NttTables(n, q, O.root_2n(q, n)) for q in qs]
    try:
        for npoly in (3, 37):
            print(f"handoff n=2^{logn} tile=2^{tb} polys={npoly}", flush=True)
            polys = []
            for i in range(npoly):
                v = ctx.vector(n, False)
                v.fill_uniform(qs[i % len(qs)], 77 * logn + i)
                polys.append(v)
            want_f, want_i = [], []
            ctx.set_tile_bits(logn, tb)
            for mode in (2, 3):
                ctx.set_pass_mode(logn, mode)
                outs_f = [ctx.vector(n) for _ in polys]
                outs_i = [ctx.vector(n) for _ in polys]
                bt = V.Batch(ctx)
                for i, (v, o) in enumerate(zip(polys, outs_f)):
                    bt.add(V.JOB_NTT_FWD, v, None, o, [tabs[i % len(tabs)]])
                bt.run()
                bi = V.Batch(ctx)
                for i, (o, r) in enumerate(zip(outs_f, outs_i)):
                    bi.add(V.JOB_NTT_INV, o, None, r, [tabs[i % len(tabs)]])
                bi.run()
                got_f = [o.read() for o in outs_f]
                got_i = [r.read() for r in outs_i]
                if mode == 2:
                    want_f, want_i = got_f, got_i
                else:
                    for i in range(npoly):
                        assert np.array_equal(got_f[i], want_f[i]), ("fwd", i)
                        assert np.array_equal(got_i[i], want_i[i]), ("inv", i)
                        assert np.array_equal(got_i[i], polys[i].read()), ("roundtrip", i)
                bt.close()
                bi.close()
                for v in outs_f + outs_i:
                    v.close()
            for v in polys:
                v.close()
    finally:
        ctx.set_pass_mode(logn, 0)
        ctx.set_tile_bits(logn, 0)
        for t in tabs:
            t.close()


def test_map_staging_pool_reuses_buffers(ctx):
    """vkhel_vector_map stages through the context's pool of page-locked
    buffers: after the first cycle, map/unmap cycles allocate nothing (the
    pool's size stays put), concurrent maps get distinct buffers, writes
    through a mapping reach the device at unmap (src/vector.c:270-296), and
    with the pool off idle buffers are freed."""
    import ctypes
    L = V.load()
    n = 1 << 16
    a, b = ctx.vector(n, False), ctx.vector(n, False)
    a.fill_uniform(1 << 60, 5)
    b.fill_uniform(1 << 60, 6)
    xa, xb = a.read(), b.read()
    ctx.set_staging_pool(True)
    for _ in range(3):
        assert np.array_equal(a.map_read(), xa)
    held = ctx.staging_bytes
    assert held >= 8 * n
    for _ in range(10):
        assert np.array_equal(a.map_read(), xa)
        assert np.array_equal(b.map_read(), xb)
    assert ctx.staging_bytes == held  # reused, nothing new allocated
    # two mappings at once: distinct buffers; a write through one reaches
    # the device at unmap
    pa, pb = ctypes.c_void_p(), ctypes.c_void_p()
    L.vkhel_vector_map(a.handle, ctypes.byref(pa), 8 * n)
    L.vkhel_vector_map(b.handle, ctypes.byref(pb), 8 * n)
    assert pa.value != pb.value
    ha = np.ctypeslib.as_array(ctypes.cast(pa, ctypes.POINTER(ctypes.c_uint64)), shape=(n,))
    assert np.array_equal(ha, xa)
    ha[:] = np.arange(n, dtype=np.uint64)
    L.vkhel_vector_unmap(a.handle)
    L.vkhel_vector_unmap(b.handle)
    assert np.array_equal(a.read(), np.arange(n, dtype=np.uint64))
    assert ctx.staging_bytes >= 16 * n
    ctx.set_staging_pool(False)
    assert ctx.staging_bytes == 0
    assert np.array_equal(b.map_read(), xb)
    assert ctx.staging_bytes == 0
    ctx.set_staging_pool(True)
    a.close()
    b.close()
