"""Closed forms in plain Python integers, used to pin the oracle.

Nothing here follows the reference's code path: each function states what
the mathematics forces for canonical inputs, so a dropped term, a wrong
sign, a wrong twiddle index or a transposed operand in the oracle (or in the
HIP library) fails against it.
"""
M64 = (1 << 64) - 1


def bitrev(i, bits):
    r = 0
    for _ in range(bits):
        r = (r << 1) | (i & 1)
        i >>= 1
    return r


def dft_negacyclic_bitrev(a, q, psi):
    """A[i] = sum_j a_j psi^((2 bitrev(i) + 1) j) mod q."""
    n = len(a)
    logn = n.bit_length() - 1
    out = []
    for i in range(n):
        e = 2 * bitrev(i, logn) + 1
        out.append(sum(int(x) * pow(psi, e * j, q) for j, x in enumerate(a)) % q)
    return out


def negacyclic_schoolbook(a, b, q):
    n = len(a)
    c = [0] * n
    for i in range(n):
        for j in range(n):
            k = i + j
            if k < n:
                c[k] += int(a[i]) * int(b[j])
            else:
                c[k - n] -= int(a[i]) * int(b[j])
    return [x % q for x in c]


def elemfma_reference_wrap(a, b, m, q):
    """elemfma.comp:36-55 evaluated in Python ints with explicit 2^64
    wrapping: Shoup product with m' = floor(m 2^64 / q), then `q - x` in
    both correction branches (the reference's quirk)."""
    mp = ((m << 64) // q) & M64
    out = []
    for x, y in zip(a, b):
        x, y = int(x), int(y)
        prod = (x * m - ((x * mp) >> 64) * q) & M64
        if prod >= q:
            prod = (q - prod) & M64
        s = (prod + y) & M64
        out.append((q - s) & M64 if s >= q else s)
    return out


def smallest_generator(q):
    m, fs, p = q - 1, [], 2
    while p * p <= m:
        if m % p == 0:
            fs.append(p)
            while m % p == 0:
                m //= p
        p += 1 if p == 2 else 2
    if m > 1:
        fs.append(m)
    g = 2
    while any(pow(g, (q - 1) // f, q) == 1 for f in fs):
        g += 1
    return g
