"""CPU checks of the drop-in boundary: libvkhel.so loads, exports exactly the
vkhel_* functions its headers declare, and its host-only logic (table
construction, argument validation) matches the reference.  No call here
touches a GPU.
"""
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as O
from vkhel_amd import lib as L

REF_API = [
    "vkhel_ctx_create", "vkhel_ctx_destroy", "vkhel_ntt_tables_create",
    "vkhel_ntt_tables_destroy", "vkhel_vector_create", "vkhel_vector_create2",
    "vkhel_vector_destroy", "vkhel_vector_dup", "vkhel_vector_copy_from_host",
    "vkhel_vector_map", "vkhel_vector_unmap", "vkhel_vector_elemfma",
    "vkhel_vector_elemmod", "vkhel_vector_elemmul", "vkhel_vector_elemgtadd",
    "vkhel_vector_elemgtsub", "vkhel_vector_forward_transform",
    "vkhel_vector_inverse_transform",
]


def _exports():
    out = subprocess.run(["nm", "-D", "--defined-only", L.LIB_PATH],
                         capture_output=True, text=True, check=True).stdout
    return sorted({line.split()[-1] for line in out.splitlines() if line.strip()})


def test_library_loads():
    lib = L.load()
    assert lib is not None


def test_reference_api_declared_and_exported():
    declared = set(L.header_symbols())
    exported = set(_exports())
    for name in REF_API:
        assert name in declared, name
        assert name in exported, name


def test_every_declared_symbol_is_exported_and_bound():
    declared = set(L.header_symbols())
    exported = set(_exports())
    assert declared <= exported, declared - exported
    assert declared == set(L.SIGNATURES), declared ^ set(L.SIGNATURES)


def test_version_script_hides_internals():
    # vkhel.syms: global vkhel_*; local *
    others = [s for s in _exports() if not s.startswith("vkhel_")]
    assert others == [], others


def test_public_header_compiles_as_c_and_cpp(tmp_path):
    inc = os.path.join(L.REPO, "include")
    src = tmp_path / "t.c"
    src.write_text('#include <vkhel/vkhel.h>\n#include <vkhel/vkhel_ext.h>\n'
                   'int main(void) { return vkhel_ext_last_error(); }\n')
    for cc in (["gcc", "-std=c11"], ["g++", "-x", "c++"]):
        subprocess.run(cc + ["-I", inc, "-fsyntax-only", str(src)], check=True)


# ---- host-only product logic: table builder and validation ----

def _tables(n, q, w):
    from vkhel_amd import NttTables
    return NttTables(n, q, w)


def test_tables_match_reference_fixtures(golden):
    for t in golden["tables"]:
        tab = _tables(t["n"], t["q"], t["w"])
        r, ir, rb, irb = tab.arrays()
        assert r.tolist() == t["roots"]
        assert ir.tolist() == t["inv_roots"]
        assert rb.tolist() == t["roots_bf"]
        assert irb.tolist() == t["inv_roots_bf"]
        tab.close()


@pytest.mark.parametrize("n,bound", [(1 << 10, 1 << 60), (1 << 16, 1 << 50),
                                     (1 << 17, 1 << 60)])
def test_tables_match_oracle_at_config_sizes(n, bound):
    q = O.primes_ntt(bound, 2 * n, 1)[0]
    w = O.root_2n(q, n)
    tab = _tables(n, q, w)
    got = tab.arrays()
    want = O.tables(n, q, w)
    for g, e in zip(got, want):
        assert np.array_equal(g, e)
    tab.close()


@pytest.mark.parametrize("bound", [1 << 20, 1 << 50, 1 << 60])
def test_table_chain_is_exact_powers(bound):
    """roots[bitrev(k)] = w^k exactly: the reference's repeated Barrett
    products are canonical below q < 2^61, which is what lets the device
    build every entry independently from binary powers of w."""
    n = 1 << 9
    for q in O.primes_ntt(bound, 2 * n, 3):
        w = O.root_2n(q, n)
        r, ir, _, _ = _tables(n, q, w).arrays()
        rev = [int(format(k, "09b")[::-1], 2) for k in range(n)]
        assert [int(r[rev[k]]) for k in range(n)] == [pow(w, k, q) for k in range(n)]
        assert [int(ir[rev[k]]) for k in range(n)] == [pow(w, -k, q) for k in range(n)]


def _small_reduce_consts(q):
    # restates numbers.h small_reduce_consts
    sh = max(q.bit_length() - 7, 0)
    return sh, (1 << (32 + sh)) // q


@pytest.mark.parametrize("q", [5, 7, 17, 97, 127, 7681, 12289, 65537,
                               (1 << 50) - 27, 1125899906826241,
                               (1 << 56) - 5])
def test_final_reduction_quotient_is_exact_or_one_less(q):
    """The forward's last-stage reduction of v < 81q: the estimate
    floor((v >> sh) * m / 2^32) is floor(v/q) or floor(v/q) - 1, and fits
    32 bits."""
    sh, m = _small_reduce_consts(q)
    assert m < 1 << 32
    rng = np.random.default_rng(q % 1000)
    vs = range(81 * q) if q < 200 else \
        [int(x) for x in rng.integers(0, 81 * q, 20000, dtype=np.uint64)] + \
        [k * q + d for k in range(81) for d in (0, 1, q - 1) if k * q + d < 81 * q]
    for v in vs:
        t = v >> sh
        assert t < 1 << 14
        qe = (t * m) >> 32
        assert v // q - 1 <= qe <= v // q, (v, qe)


def _mad_u64_u32(a, b, c):
    # v_mad_u64_u32: a, b 32-bit, c 64-bit, result mod 2^64
    assert a < 1 << 32 and b < 1 << 32 and c < 1 << 64
    return (a * b + c) % (1 << 64)


def _mad_i32_i24(a, b, c):
    # v_mad_i32_i24: the low 24 bits of a and b as signed numbers, their
    # product plus c, mod 2^32
    def s24(x):
        x &= (1 << 24) - 1
        return x - (1 << 24) if x >> 23 else x
    return (s24(a) * s24(b) + c) % (1 << 32)


def _sub_small_multiple(x, k, q):
    # the two device instructions of modarith.h sub_small_multiple
    nq = (1 << 64) - q
    lo = _mad_u64_u32(k, nq & 0xFFFFFFFF, x)
    hi = _mad_i32_i24(k, nq >> 32, lo >> 32)
    return hi << 32 | (lo & 0xFFFFFFFF)


@pytest.mark.parametrize("q", [5, 97, 65537, (1 << 50) - 27, 1125899906826241,
                               (1 << 55) - 55, (1 << 55) - 1,
                               (1 << 54) + 1, (1 << 32) + 15])
def test_sub_small_multiple_is_exact(q):
    """The forward's final reduction takes k*q off v (k <= 80, v < 81q) with
    one 64-bit and one 24-bit multiply-add, which is exact for odd q below
    2^55 (the small class): checked against x - k*q in exact integers,
    including q whose high word is 2^23 - 1, the 24-bit limit."""
    assert q % 2 == 1 and q < 1 << 55
    rng = np.random.default_rng(q % 997)
    cases = [(int(x), int(k)) for x, k in zip(
        rng.integers(0, 81 * q, 4000, dtype=np.uint64) if 81 * q < 1 << 64
        else [], rng.integers(0, 81, 4000))]
    cases += [(k * q + d, k) for k in range(81) for d in (0, 1, q - 1)]
    cases += [(81 * q - 1, 80), (0, 0), (q, 1)]
    for x, k in cases:
        if k * q > x:
            continue  # the estimate never exceeds floor(v / q)
        assert _sub_small_multiple(x, k, q) == x - k * q, (x, k)
    # one step past the limit: a high word of 2^23 no longer fits 24 bits
    q = (1 << 55) + 1
    x, k = 80 * q + 5, 80
    assert _sub_small_multiple(x, k, q) != x - k * q


def test_tables_root_check():
    n = 1 << 8
    q = O.primes_ntt(1 << 40, 2 * n, 1)[0]
    w = O.root_2n(q, n)
    assert _tables(n, q, w).root_ok
    assert pow(w, n, q) == q - 1
    for bad in (1, w * w % q, q - 1):
        assert not _tables(n, q, bad).root_ok


@pytest.mark.parametrize("n,q,w", [
    (3, 113, 18),                 # n not a power of two
    (1, 113, 18),                 # n < 2
    (4, 112, 18),                 # even q
    (4, (1 << 61) + 1, 3),        # q beyond the Barrett range
    (4, 113, 113),                # w >= q
    (4, 7 * 13, 13),              # w not invertible mod q
    (128, 113, 18),               # n >= q
])
def test_invalid_tables_rejected(n, q, w):
    from vkhel_amd import NttTables, VkhelError
    with pytest.raises(VkhelError) as e:
        NttTables(n, q, w)
    assert e.value.code == 1


def test_abort_mode_default_matches_reference_asserts(tmp_path):
    # With the default error mode an invalid call aborts the process, as the
    # reference's assert() does.
    code = ("import ctypes; L = ctypes.CDLL(%r); "
            "L.vkhel_ntt_tables_create.restype = ctypes.c_void_p; "
            "L.vkhel_ntt_tables_create(3, 113, 18)" % L.LIB_PATH)
    r = subprocess.run(["python3", "-c", code], capture_output=True, text=True)
    assert r.returncode != 0
    assert "power of two" in r.stderr
