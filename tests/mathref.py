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


def elemfma_closed_form(a, b, m, q):
    """What the reference's elemfma(a, b, m, q) returns (elemfma.comp:36-55:
    a Shoup product by m' = floor(m 2^64 / q), then each correction written
    `q - x` where `x - q` was meant), for any 64-bit a, b < q and 1 <= m < q,
    derived by cases from the mathematics instead of by evaluating the
    shader's statements:

    With r = a m mod q, Q = floor(a m / q) and the Shoup quotient
    Qs = floor(a m' / 2^64), a m / q - a m' / 2^64 < a / 2^64 < 1 gives
    Qs in {Q, Q - 1}, so the product before its correction is
    P = r + (Q - Qs) q.
      Qs = Q:      P = r < q; the sum r + b < 2q.  Below q it is the
                   result; otherwise the correction gives
                   q - (r + b) mod 2^64 = 2^64 - ((a m + b) mod q) (0 when
                   r + b = q).
      Qs = Q - 1:  P = r + q >= q is corrected to q - P = -r mod 2^64, so
                   the sum is b - r mod 2^64: b - r when b >= r (the
                   result, (b - a m) mod q); otherwise 2^64 - (r - b) >= q
                   and the result is q - (2^64 - (r - b)) = q + r - b.
    Returns the words and the count of each case, so a test can show every
    branch was reached."""
    mp = (m << 64) // q
    out, cases = [], [0, 0, 0, 0]
    for x, y in zip(a, b):
        x, y = int(x), int(y)
        r, Q, Qs = x * m % q, x * m // q, (x * mp) >> 64
        assert Qs in (Q, Q - 1)
        if Qs == Q:
            if r + y < q:
                out.append(r + y)
                cases[0] += 1
            else:
                out.append((q - (r + y)) & M64)
                cases[1] += 1
        elif y >= r:
            out.append(y - r)
            cases[2] += 1
        else:
            out.append(q + r - y)
            cases[3] += 1
    return out, cases


def ntt_forward_noncanonical(x, q, psi):
    """The reference's forward transform (src/vector.c:513-574) of arbitrary
    64-bit words x, from the closed form: every butterfly is exact mod q
    except the first stage's X + WY, whose sum is taken modulo 2^64 before
    its reduction (nttfwdbutterfly.comp:56): where X + WY >= 2^64 (X = x_j,
    WY = W x_{j+n/2} mod q, W = roots[1] = psi^(n/2)) that output is the
    exact residue minus r = 2^64 mod q.  The stages after the first are
    linear, and the first stage maps z with z_j = r/2, z_{j+n/2} = r/(2W)
    (mod q) to (r, 0) at (j, j + n/2).  So the output is the negacyclic DFT
    (closed form) of x mod q minus such a z for every wrapped j.  Returns
    the words and the number of wrapped sums."""
    n = len(x)
    h = n // 2
    W = pow(psi, h, q)
    r = (1 << 64) % q
    half, half_w = r * pow(2, -1, q) % q, r * pow(2 * W, -1, q) % q
    xt = [int(v) % q for v in x]
    wraps = 0
    for j in range(h):
        if int(x[j]) + W * int(x[j + h]) % q >= 1 << 64:
            wraps += 1
            xt[j] = (xt[j] - half) % q
            xt[j + h] = (xt[j + h] - half_w) % q
    return dft_negacyclic_bitrev(xt, q, psi), wraps


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
