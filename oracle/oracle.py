"""ctypes wrapper of the CPU oracle -- TEST INFRASTRUCTURE ONLY.

Loads ``oracle/liboracle.so`` (built by ``make -C oracle``).  Only tests/,
``__graft_entry__.smoke()`` and bench.py's CPU-baseline leg import this
module; the product package ``vkhel_amd`` never does.

Every function mirrors one in ``vkhel_oracle.h`` and takes/returns numpy
``uint64`` arrays.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_u64 = ctypes.c_uint64
_p64 = ctypes.POINTER(ctypes.c_uint64)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so missing: run make -C oracle")
        L = ctypes.CDLL(path)
        sig = {
            "or_bitlen": (_u64, [_u64]),
            "or_barrett_factor": (_u64, [_u64, _u64]),
            "or_mulmod": (_u64, [_u64, _u64, _u64]),
            "or_reduce64": (_u64, [_u64, _u64]),
            "or_powmod": (_u64, [_u64, _u64, _u64]),
            "or_invmod": (_u64, [_u64, _u64]),
            "or_is_primitive_root": (ctypes.c_int, [_u64, _u64, _u64]),
            "or_tables_create": (None, [_u64, _u64, _u64, _p64, _p64, _p64, _p64]),
            "or_ntt_forward": (ctypes.c_int, [_p64, _p64, _u64, _u64, _u64]),
            "or_ntt_inverse": (ctypes.c_int, [_p64, _p64, _u64, _u64, _u64]),
            "or_elemfma": (None, [_p64, _p64, _p64, ctypes.c_size_t, _u64, _u64]),
            "or_elemmul": (None, [_p64, _p64, _p64, ctypes.c_size_t, _u64]),
            "or_elemgtadd": (None, [_p64, _p64, ctypes.c_size_t, _u64, _u64]),
            "or_elemgtsub": (None, [_p64, _p64, ctypes.c_size_t, _u64, _u64, _u64]),
            "or_elemmod": (None, [_p64, _p64, ctypes.c_size_t, _u64, _u64]),
            "or_elemmulconst": (None, [_p64, _p64, ctypes.c_size_t, _u64, _u64]),
            "or_addmul_mod": (None, [_p64, _p64, _p64, ctypes.c_size_t, _u64, _u64]),
            "or_dft_negacyclic_bitrev": (None, [_p64, _p64, _u64, _u64, _u64]),
            "or_negacyclic_schoolbook": (None, [_p64, _p64, _p64, _u64, _u64]),
            "or_polymul_ntt": (ctypes.c_int, [_p64, _p64, _p64, _u64, _u64, _u64]),
            "or_is_prime": (ctypes.c_int, [_u64]),
            "or_smallest_generator": (_u64, [_u64]),
            "or_primes_ntt": (ctypes.c_size_t, [_u64, _u64, ctypes.c_size_t, _p64]),
            "or_root_2n": (_u64, [_u64, _u64]),
            "or_fill_uniform": (None, [_p64, ctypes.c_size_t, _u64, _u64]),
            "or_bench_ntt_batch": (ctypes.c_double, [_p64, ctypes.c_size_t, _u64,
                                                     _p64, _p64, ctypes.c_size_t,
                                                     ctypes.c_int]),
            "or_bench_polymul_batch": (ctypes.c_double, [_p64, _p64, _p64,
                                                         ctypes.c_size_t, _u64,
                                                         _p64, _p64,
                                                         ctypes.c_size_t,
                                                         ctypes.c_int]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def _arr(x):
    a = np.ascontiguousarray(np.asarray(x, dtype=np.uint64))
    return a


def _ptr(a):
    return a.ctypes.data_as(_p64)


def bitlen(v):
    return int(lib().or_bitlen(v))


def barrett_factor(factor, mod):
    return int(lib().or_barrett_factor(factor, mod))


def mulmod(a, b, mod):
    return int(lib().or_mulmod(a, b, mod))


def reduce64(x, mod):
    return int(lib().or_reduce64(x, mod))


def powmod(b, e, mod):
    return int(lib().or_powmod(b, e, mod))


def invmod(a, mod):
    return int(lib().or_invmod(a, mod))


def is_primitive_root(root, degree, mod):
    return bool(lib().or_is_primitive_root(root, degree, mod))


def tables(n, q, w):
    """(roots, inv_roots, roots_bf, inv_roots_bf) as in src/ntt_tables.c."""
    out = [np.zeros(n, dtype=np.uint64) for _ in range(4)]
    lib().or_tables_create(n, q, w, *[_ptr(a) for a in out])
    return tuple(out)


def ntt_forward(a, q, w):
    a = _arr(a)
    out = np.empty_like(a)
    if lib().or_ntt_forward(_ptr(a), _ptr(out), a.size, q, w):
        raise ValueError("oracle forward NTT precondition failed")
    return out


def ntt_inverse(a, q, w):
    a = _arr(a)
    out = np.empty_like(a)
    if lib().or_ntt_inverse(_ptr(a), _ptr(out), a.size, q, w):
        raise ValueError("oracle inverse NTT precondition failed")
    return out


def _binary(fn, a, b, *args):
    a, b = _arr(a), _arr(b)
    out = np.empty_like(a)
    fn(_ptr(a), _ptr(b), _ptr(out), a.size, *args)
    return out


def _unary(fn, a, *args):
    a = _arr(a)
    out = np.empty_like(a)
    fn(_ptr(a), _ptr(out), a.size, *args)
    return out


def elemfma(a, b, multiplier, mod):
    return _binary(lib().or_elemfma, a, b, multiplier, mod)


def elemmul(a, b, mod):
    return _binary(lib().or_elemmul, a, b, mod)


def addmul_mod(a, b, multiplier, mod):
    return _binary(lib().or_addmul_mod, a, b, multiplier, mod)


def elemgtadd(a, bound, diff):
    return _unary(lib().or_elemgtadd, a, bound, diff)


def elemgtsub(a, bound, diff, mod):
    return _unary(lib().or_elemgtsub, a, bound, diff, mod)


def elemmod(a, mod, q):
    return _unary(lib().or_elemmod, a, mod, q)


def elemmulconst(a, c, mod):
    return _unary(lib().or_elemmulconst, a, c, mod)


def dft_negacyclic_bitrev(a, q, psi):
    a = _arr(a)
    out = np.empty_like(a)
    lib().or_dft_negacyclic_bitrev(_ptr(a), _ptr(out), a.size, q, psi)
    return out


def negacyclic_schoolbook(a, b, q):
    a, b = _arr(a), _arr(b)
    out = np.empty_like(a)
    lib().or_negacyclic_schoolbook(_ptr(a), _ptr(b), _ptr(out), a.size, q)
    return out


def polymul_ntt(a, b, q, w):
    a, b = _arr(a), _arr(b)
    out = np.empty_like(a)
    if lib().or_polymul_ntt(_ptr(a), _ptr(b), _ptr(out), a.size, q, w):
        raise ValueError("oracle polymul precondition failed")
    return out


def is_prime(n):
    return bool(lib().or_is_prime(n))


def smallest_generator(q):
    return int(lib().or_smallest_generator(q))


def primes_ntt(bound, m, count):
    out = np.zeros(count, dtype=np.uint64)
    k = lib().or_primes_ntt(bound, m, count, _ptr(out))
    return [int(x) for x in out[:k]]


def root_2n(q, n):
    return int(lib().or_root_2n(q, n))


def fill_uniform(length, q, seed):
    out = np.empty(length, dtype=np.uint64)
    lib().or_fill_uniform(_ptr(out), length, q, seed)
    return out


def bench_ntt_batch(polys, n, qs, ws, threads):
    """Forward+inverse NTT of every row of ``polys`` in place; seconds."""
    assert polys.dtype == np.uint64 and polys.flags.c_contiguous
    qs, ws = _arr(qs), _arr(ws)
    return float(lib().or_bench_ntt_batch(_ptr(polys), polys.size // n, n,
                                          _ptr(qs), _ptr(ws), qs.size, threads))


def bench_polymul_batch(a, b, n, qs, ws, threads):
    """Products of the rows of ``a`` and ``b`` (pair p modulo qs[p % L]);
    (seconds, products)."""
    assert a.dtype == np.uint64 and a.flags.c_contiguous and a.shape == b.shape
    out = np.empty_like(a)
    qs, ws = _arr(qs), _arr(ws)
    secs = float(lib().or_bench_polymul_batch(_ptr(a), _ptr(np.ascontiguousarray(b)),
                                              _ptr(out), a.size // n, n, _ptr(qs),
                                              _ptr(ws), qs.size, threads))
    return secs, out
