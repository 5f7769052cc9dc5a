"""Pins the CPU oracle (oracle/) before anything is checked against it.

Sources of truth, in order: the reference's own known answers
(test/numbers.c, test/ntt.c, test/vector.c), fixtures written by the
reference's CPU sources compiled in oracle/_ref (tests/golden/*.json), and
closed forms in Python integers (tests/mathref.py).
"""
import math
import os
import random
import subprocess

import numpy as np
import pytest

from oracle import oracle as O
from tests import mathref as R

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


# ---------------- numbers (src/numbers.c) ----------------

def test_bitlen_known_answers(golden):
    for i in range(1, 65):
        assert O.bitlen(i) == math.floor(math.log2(i)) + 1
    assert O.bitlen(2**63) == 64 and O.bitlen(2**64 - 1) == 64


def test_numbers_known_answers(golden):
    nb = golden["vectors"]["numbers"]
    for a, b, q, want in nb["multiply_mod"]["cases"]:
        assert O.mulmod(a, b, q) == want
    for b, e, q, want in nb["power_mod"]["cases"]:
        assert O.powmod(b, e, q) == want
    mod = nb["is_primitive_root"]["mod"]
    for root, deg, want in nb["is_primitive_root"]["cases"]:
        assert O.is_primitive_root(root, deg, mod) == want
    for a, q, want in nb["inverse_mod"]["cases"]:
        assert O.invmod(a, q) == want


def test_numbers_match_compiled_reference(golden):
    g = golden["numbers"]
    for a, q, want in g["barrett_factor"]:
        assert O.barrett_factor(a, q) == want
    for a, b, q, want in g["multiply_mod"]:
        assert O.mulmod(a, b, q) == want
    for a, e, q, want in g["power_mod"]:
        assert O.powmod(a, e, q) == want
    for a, q, want in g["inverse_mod"]:
        assert O.invmod(a, q) == want


def test_numbers_closed_forms():
    rng = random.Random(7)
    for bits in range(3, 62):
        for _ in range(20):
            q = rng.getrandbits(bits) | (1 << (bits - 1)) | 1
            a, b = rng.randrange(q), rng.randrange(q)
            assert O.mulmod(a, b, q) == a * b % q
            x = rng.getrandbits(64)
            assert O.reduce64(x, q) == x % q
            e = rng.getrandbits(40)
            assert O.powmod(a, e, q) == pow(a, e, q)
            if math.gcd(a, q) == 1 and a > 0:
                assert O.invmod(a, q) == pow(a, -1, q)
            assert O.barrett_factor(a, q) == (a << 64) // q


def test_compiled_reference_numbers_test_passes():
    exe = os.path.join(REPO, "oracle", "_ref", "test_numbers")
    if not os.path.exists(exe):
        pytest.skip("oracle/_ref not built (reference tree absent)")
    out = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stderr
    assert out.stdout.count("passed") == 5


# ---------------- tables (src/ntt_tables.c) ----------------

def test_tables_known_answer(golden):
    t = golden["vectors"]["tables"][0]
    roots, inv, rb, irb = O.tables(t["n"], t["q"], t["w"])
    assert roots.tolist() == t["roots"]


def test_tables_match_compiled_reference(golden):
    for t in golden["tables"]:
        roots, inv, rb, irb = O.tables(t["n"], t["q"], t["w"])
        assert roots.tolist() == t["roots"]
        assert inv.tolist() == t["inv_roots"]
        assert rb.tolist() == t["roots_bf"]
        assert irb.tolist() == t["inv_roots_bf"]


@pytest.mark.parametrize("n,q", [(8, 113), (64, 1125899903827969),
                                 (1024, 1152921504606830593)])
def test_tables_closed_form(n, q):
    w = O.root_2n(q, n)
    roots, inv, rb, irb = O.tables(n, q, w)
    logn = n.bit_length() - 1
    for i in range(n):
        r = pow(w, i, q)
        assert int(roots[R.bitrev(i, logn)]) == r
        assert int(inv[R.bitrev(i, logn)]) == pow(r, -1, q)
        assert int(rb[R.bitrev(i, logn)]) == (r << 64) // q


# ---------------- primes and roots (App. B convention) ----------------

def test_config_primes_and_roots():
    q = 1152921504606830593
    assert O.primes_ntt(1 << 60, 8192, 1) == [q]
    assert O.smallest_generator(q) == R.smallest_generator(q) == 10
    psi = O.root_2n(q, 4096)
    assert psi == 429945184819996456
    assert pow(psi, 4096, q) == q - 1
    assert O.primes_ntt(1 << 50, 1 << 17, 8) == [
        1125899903827969, 1125899902124033, 1125899887312897, 1125899886395393,
        1125899885740033, 1125899884167169, 1125899884036097, 1125899883642881]
    assert O.primes_ntt(1 << 50, 1 << 15, 3) == [
        1125899904679937, 1125899903991809, 1125899903827969]
    p55 = O.primes_ntt(1 << 55, 1 << 16, 16)
    assert p55[0] == 36028797017456641 and p55[-1] == 36028796998844417
    p60 = O.primes_ntt(1 << 60, 1 << 18, 24)
    assert p60[0] == 1152921504606584833 and p60[-1] == 1152921504495697921
    for p in p55 + p60:
        assert all(p % d for d in range(3, 2000, 2))


def test_is_prime_against_trial_division():
    rng = random.Random(3)
    for _ in range(300):
        n = rng.randrange(2, 1 << 24)
        want = n > 1 and all(n % d for d in range(2, int(n**0.5) + 1))
        assert O.is_prime(n) == want


# ---------------- transforms (src/vector.c:513-657) ----------------

def test_transform_known_answers(golden):
    for c in golden["vectors"]["vector"]["transform"]["cases"]:
        fwd = O.ntt_forward(c["coeffs"], c["q"], c["w"])
        assert fwd.tolist() == c["ntt"]
        assert O.ntt_inverse(c["ntt"], c["q"], c["w"]).tolist() == c["coeffs"]


@pytest.mark.parametrize("logn", [1, 2, 3, 4, 5, 6, 7])
@pytest.mark.parametrize("q", [7681, 1125899903827969, 1152921504606584833])
def test_forward_is_bitreversed_negacyclic_dft(logn, q):
    n = 1 << logn
    if (q - 1) % (2 * n):
        pytest.skip("q not 1 mod 2n")
    psi = O.root_2n(q, n)
    a = O.fill_uniform(n, q, seed=logn * 1000 + q % 997)
    want = R.dft_negacyclic_bitrev(a.tolist(), q, psi)
    assert O.ntt_forward(a, q, psi).tolist() == want
    assert O.dft_negacyclic_bitrev(a, q, psi).tolist() == want


@pytest.mark.parametrize("n", [16, 256, 4096])
def test_roundtrip_and_convolution(n):
    q = O.primes_ntt(1 << 60, 2 * n, 1)[0]
    psi = O.root_2n(q, n)
    a = O.fill_uniform(n, q, seed=11)
    b = O.fill_uniform(n, q, seed=12)
    assert np.array_equal(O.ntt_inverse(O.ntt_forward(a, q, psi), q, psi), a)
    c = O.polymul_ntt(a, b, q, psi)
    if n <= 256:
        assert c.tolist() == R.negacyclic_schoolbook(a.tolist(), b.tolist(), q)
    assert np.array_equal(c, O.negacyclic_schoolbook(a, b, q))


def test_forward_noncanonical_inputs_first_stage_reduces():
    # nttfwdbutterfly.comp:56-57: (X + WY) % q and a subtract loop, so
    # inputs >= q act as x mod q unless X + WY wraps past 2^64.
    n, q = 64, 1125899903827969
    psi = O.root_2n(q, n)
    a = O.fill_uniform(n, q, seed=5)
    big = a + np.uint64(q) * np.uint64(1000)
    assert np.array_equal(O.ntt_forward(big, q, psi), O.ntt_forward(a, q, psi))


def test_inverse_rejects_noncanonical():
    with pytest.raises(ValueError):
        O.ntt_inverse([200, 0, 0, 0], 113, 18)


def test_planted_mistakes_are_caught():
    """The closed-form checks fail on plausible oracle mistakes."""
    n, q = 16, 1125899903827969
    psi = O.root_2n(q, n)
    a = O.fill_uniform(n, q, seed=9)
    want = R.dft_negacyclic_bitrev(a.tolist(), q, psi)
    # wrong root (psi^3 is also a primitive 2n-th root)
    assert O.ntt_forward(a, q, pow(psi, 3, q)).tolist() != want
    # natural instead of bit-reversed output order
    nat = [sum(int(x) * pow(psi, (2 * i + 1) * j, q) for j, x in enumerate(a)) % q
           for i in range(n)]
    assert nat != want
    # cyclic instead of negacyclic convolution
    b = O.fill_uniform(n, q, seed=10)
    cyc = [sum(int(a[i]) * int(b[(k - i) % n]) for i in range(n)) % q for k in range(n)]
    assert cyc != O.polymul_ntt(a, b, q, psi).tolist()


# ---------------- elementwise (src/vector.c:298-511) ----------------

def test_elementwise_known_answers(golden):
    v = golden["vectors"]["vector"]
    f = v["elemfma"]
    for c in f["cases"]:
        assert O.elemfma(f["a"], f["b"], c["multiplier"], f["mod"]).tolist() == c["expected"]
    for c in v["elemmul"]["cases"]:
        assert O.elemmul(c["a"], c["b"], c["mod"]).tolist() == c["expected"]
    g = v["elemgtadd"]
    for c in g["cases"]:
        assert O.elemgtadd(g["a"], c["bound"], c["diff"]).tolist() == c["expected"]
    g = v["elemgtsub"]
    for c in g["cases"]:
        assert O.elemgtsub(g["a"], c["bound"], c["diff"], c["mod"]).tolist() == c["expected"]
    for c in v["elemmod"]["cases"]:
        assert O.elemmod(c["a"], c["mod"], c["q"]).tolist() == c["expected"]


def test_elementwise_closed_forms():
    rng = np.random.default_rng(1)
    for q in [769, 1125891450734593, 1152921504606830593, (1 << 61) - 1]:
        a = rng.integers(0, q, 500, dtype=np.uint64)
        b = rng.integers(0, q, 500, dtype=np.uint64)
        raw = rng.integers(0, 2**64 - 1, 500, dtype=np.uint64, endpoint=True)
        assert O.elemmul(a, b, q).tolist() == [int(x) * int(y) % q for x, y in zip(a, b)]
        # elemmul reduces arbitrary 64-bit operands first
        assert O.elemmul(raw, b, q).tolist() == [int(x) * int(y) % q for x, y in zip(raw, b)]
        bound, diff = q // 3, rng.integers(0, 2**63)
        want = [((int(x) - int(diff)) if int(x) > bound else int(x)) % q for x in raw]
        assert O.elemgtsub(raw, bound, int(diff), q).tolist() == want
        assert O.elemgtadd(raw, bound, int(diff)).tolist() == [
            (int(x) + int(diff)) & R.M64 if int(x) > bound else int(x) for x in raw]
        # centred lift x - q [x > q/2], reduced mod p
        for p in [2, 3, 5, 65537, 1000003]:
            want = [((int(x) - q) if int(x) > q // 2 else int(x)) % p for x in a]
            assert O.elemmod(a, p, q).tolist() == want
        c = int(rng.integers(0, q))
        assert O.elemmulconst(raw, c, q).tolist() == [int(x) * c % q for x in raw]


def elemfma_inputs(rng, q, count=3000):
    """Raw 64-bit a (half of them near 2^64) and canonical b."""
    a = np.concatenate([rng.integers(0, 1 << 64, count, dtype=np.uint64),
                        np.uint64((1 << 64) - 1) - rng.integers(0, q, count,
                                                                 dtype=np.uint64),
                        rng.integers(0, q, count, dtype=np.uint64)])
    b = rng.integers(0, q, a.size, dtype=np.uint64)
    return a, b


def test_elemfma_wrap_branches_pinned_by_case_analysis():
    """elemfma's two `q - x` corrections (elemfma.comp:47-48, 51-52) against
    the case analysis of tests/mathref.py (derived from the Shoup quotient's
    error, not from the shader's statements); every case is reached."""
    rng = np.random.default_rng(2)
    seen = [0, 0, 0, 0]
    for q in [3, 97, 769, 1125891450734593, 1152921504606830593,
              2305843009213693951 - 30]:
        if not O.is_prime(q):
            q = O.primes_ntt(q + 1, 2, 1)[0]
        a, b = elemfma_inputs(rng, q)
        for m in [1, 2, 100 % q or 1, int(rng.integers(1, q)), q - 1]:
            want, cases = R.elemfma_closed_form(a.tolist(), b.tolist(), m, q)
            assert O.elemfma(a, b, m, q).tolist() == want, (q, m)
            seen = [s + c for s, c in zip(seen, cases)]
            # the true (a m + b) mod q exactly in the first case
            if cases[0]:
                ok = [(int(x) * m % q) + int(y) < q and
                      ((int(x) * m // q) == ((int(x) * ((m << 64) // q)) >> 64))
                      for x, y in zip(a, b)]
                got = O.elemfma(a, b, m, q)
                for g, x, y, k in zip(got.tolist(), a.tolist(), b.tolist(), ok):
                    if k:
                        assert g == (x * m + y) % q
        assert O.addmul_mod(a, b, 5, q).tolist() == [
            (int(x) * 5 + int(y)) % q for x, y in zip(a, b)]
    assert all(seen), seen


def test_elemfma_closed_form_rejects_planted_mistakes():
    """A restatement that kept `p - q` (the intended correction) or dropped
    the sum's correction would differ from the reference's words."""
    rng = np.random.default_rng(3)
    q = 1152921504606830593
    a, b = elemfma_inputs(rng, q, 500)
    want, _ = R.elemfma_closed_form(a.tolist(), b.tolist(), 1, q)
    intended = [(int(x) + int(y)) % q for x, y in zip(a, b)]
    assert want != intended
    unwrapped = [((int(x) % q) + int(y)) for x, y in zip(a, b)]
    assert want != unwrapped


@pytest.mark.parametrize("bits", [50, 60, 61])
@pytest.mark.parametrize("logn", [1, 2, 3, 4, 5, 6])
def test_forward_first_stage_wrap_pinned_by_closed_form(logn, bits):
    """The forward transform on words near 2^64, where the first stage's
    64-bit X + WY wraps (nttfwdbutterfly.comp:56), against the closed-form
    DFT of the corrected canonical vector (tests/mathref.py); wrapped and
    unwrapped sums both occur."""
    n = 1 << logn
    q = O.primes_ntt(1 << bits, 2 * n, 1)[0]
    psi = O.root_2n(q, n)
    rng = np.random.default_rng(logn * 100 + bits)
    wraps = 0
    for trial in range(12):
        near = np.uint64((1 << 64) - 1) - rng.integers(0, q, n, dtype=np.uint64)
        raw = rng.integers(0, 1 << 64, n, dtype=np.uint64)
        x = np.where(rng.random(n) < 0.6, near, raw)
        want, w = R.ntt_forward_noncanonical(x.tolist(), q, psi)
        wraps += w
        assert O.ntt_forward(x, q, psi).tolist() == want
        # a restatement without the wrap correction would differ
        if w:
            assert R.dft_negacyclic_bitrev([int(v) % q for v in x], q, psi) != want
    assert 0 < wraps < 12 * n // 2


def test_fill_uniform_is_seeded_and_in_range():
    q = 1125899903827969
    x = O.fill_uniform(10000, q, 42)
    assert np.array_equal(x, O.fill_uniform(10000, q, 42))
    assert not np.array_equal(x, O.fill_uniform(10000, q, 43))
    assert int(x.max()) < q and int(x.min()) >= 0
    # mean of uniform [0,q) within 3 sigma
    assert abs(float(x.mean()) / q - 0.5) < 3 * (1 / math.sqrt(12 * 10000))


def test_cpu_bench_runner_matches_oracle():
    n = 1024
    qs = O.primes_ntt(1 << 50, 2 * n, 2)
    ws = [O.root_2n(q, n) for q in qs]
    polys = np.stack([O.fill_uniform(n, qs[i % 2], i) for i in range(6)])
    before = polys.copy()
    secs = O.bench_ntt_batch(polys, n, qs, ws, threads=2)
    assert secs > 0
    # forward+inverse is the identity
    assert np.array_equal(polys, before)


def test_cpu_polymul_runner_matches_schoolbook():
    """The threaded polymul timing runner's products against the schoolbook
    negacyclic product (pair p modulo qs[p % 3])."""
    n = 64
    qs = O.primes_ntt(1 << 55, 2 * n, 3)
    ws = [O.root_2n(q, n) for q in qs]
    a = np.stack([O.fill_uniform(n, qs[i % 3], 10 + i) for i in range(7)])
    b = np.stack([O.fill_uniform(n, qs[i % 3], 50 + i) for i in range(7)])
    secs, out = O.bench_polymul_batch(a, b, n, qs, ws, threads=3)
    assert secs > 0
    for i in range(7):
        assert np.array_equal(out[i], O.negacyclic_schoolbook(a[i], b[i], qs[i % 3])), i
