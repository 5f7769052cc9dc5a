"""Write tests/golden/ref_numbers.json and ref_tables.json.

Runs in the build container only: it loads oracle/_ref/libvkhel_ref_cpu.so,
which ``make -C oracle ref`` compiles from the reference's own
src/numbers.c and src/ntt_tables.c where they lie under /root/reference.
The outputs are data (inputs and the reference's outputs) and are committed;
the tests compare the oracle and the HIP library against them without
needing the reference.

    python tests/golden/gen_ref_fixtures.py
"""
import ctypes
import json
import math
import os
import random

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
SO = os.path.join(REPO, "oracle", "_ref", "libvkhel_ref_cpu.so")

u64 = ctypes.c_uint64


class RefTables(ctypes.Structure):
    # include/priv/ntt_tables.h:6-15
    _fields_ = [("n", u64), ("q", u64), ("w", u64),
                ("roots", ctypes.POINTER(u64)),
                ("inv_roots", ctypes.POINTER(u64)),
                ("roots_bf", ctypes.POINTER(u64)),
                ("inv_roots_bf", ctypes.POINTER(u64))]


def load():
    L = ctypes.CDLL(SO)
    L.nt_compute_barrett_factor.restype = u64
    L.nt_compute_barrett_factor.argtypes = [u64, u64, u64]
    L.nt_multiply_mod.restype = u64
    L.nt_multiply_mod.argtypes = [u64, u64, u64, u64]
    L.nt_power_mod.restype = u64
    L.nt_power_mod.argtypes = [u64, u64, u64]
    L.nt_inverse_mod.restype = u64
    L.nt_inverse_mod.argtypes = [u64, u64]
    L.nt_is_primitive_root.restype = ctypes.c_bool
    L.nt_is_primitive_root.argtypes = [u64, u64, u64]
    L.vkhel_ntt_tables_create.restype = ctypes.POINTER(RefTables)
    L.vkhel_ntt_tables_create.argtypes = [u64, u64, u64]
    return L


def main():
    L = load()
    rng = random.Random(20261202)
    numbers = {"multiply_mod": [], "power_mod": [], "inverse_mod": [],
               "barrett_factor": []}
    # Moduli from 3 to 61 bits (the Barrett validity range, src/numbers.c:8).
    for bits in range(3, 62):
        for _ in range(6):
            q = rng.getrandbits(bits) | (1 << (bits - 1)) | 1
            a, b = rng.randrange(q), rng.randrange(q)
            n = q.bit_length()
            mu = L.nt_compute_barrett_factor(1 << (n - 2), q, n)
            numbers["barrett_factor"].append([a, q, L.nt_compute_barrett_factor(a, q, n)])
            numbers["multiply_mod"].append([a, b, q, L.nt_multiply_mod(a, b, q, mu)])
            e = rng.getrandbits(64)
            numbers["power_mod"].append([a, e, q, L.nt_power_mod(a, e, q)])
            # The reference's ext-Euclid divides by zero unless gcd(a, q) == 1.
            if a > 0 and math.gcd(a, q) == 1:
                numbers["inverse_mod"].append([a, q, L.nt_inverse_mod(a, q)])
    with open(os.path.join(HERE, "ref_numbers.json"), "w") as f:
        json.dump({"source": "oracle/_ref built from reference src/numbers.c",
                   **numbers}, f)

    tables = []
    for n, q, w in [(4, 113, 18),
                    (16, 2251799813685313, 110968848420801),
                    # 50-bit prime == 1 mod 2^17, psi from its smallest generator
                    (64, 1125899903827969, None),
                    (256, 1152921504606830593, None)]:
        if w is None:
            g = next(g for g in range(2, q) if all(
                pow(g, (q - 1) // p, q) != 1 for p in _factors(q - 1)))
            w = pow(g, (q - 1) // (2 * n), q)
        t = L.vkhel_ntt_tables_create(n, q, w).contents
        tables.append({"n": n, "q": q, "w": w,
                       "roots": [t.roots[i] for i in range(n)],
                       "inv_roots": [t.inv_roots[i] for i in range(n)],
                       "roots_bf": [t.roots_bf[i] for i in range(n)],
                       "inv_roots_bf": [t.inv_roots_bf[i] for i in range(n)]})
    with open(os.path.join(HERE, "ref_tables.json"), "w") as f:
        json.dump({"source": "oracle/_ref built from reference src/ntt_tables.c",
                   "tables": tables}, f)


def _factors(m):
    out, p = [], 2
    while p * p <= m:
        if m % p == 0:
            out.append(p)
            while m % p == 0:
                m //= p
        p += 1
    if m > 1:
        out.append(m)
    return out


if __name__ == "__main__":
    main()
