// Host-side number theory of the library (table construction and per-call
// constants).  Semantics follow the reference's src/numbers.c so that tables
// and launch constants are bit-identical to what it would produce.
#pragma once

#include <cstdint>

namespace vk {

typedef unsigned __int128 u128;

// include/priv/numbers.h:11-13: bit length (named ceil_log2 upstream).
inline uint64_t bitlen(uint64_t v) {
	return v ? 64 - (uint64_t) __builtin_clzll(v) : 0;
}

// src/numbers.c:30-34: floor(f * 2^64 / q) truncated to 64 bits -- the
// Shoup companion of f when f < q.
inline uint64_t barrett_factor(uint64_t f, uint64_t q) {
	return (uint64_t) (((u128) f << 64) / q);
}

// The 63-bit Shoup companion wp = floor(f * 2^63 / q) as the device
// kernels take it: high word doubled, (2 * (wp >> 32)) << 32 | (wp mod 2^32).
// wp < 2^63 for f < q, so the doubled word fits.
inline uint64_t shoup63_factor(uint64_t f, uint64_t q) {
	const uint64_t wp = barrett_factor(f, q) >> 1;
	return wp + (wp & 0xffffffff00000000ull);
}

// Barrett constant floor(2^(bits+62) / q) of elemmul.c:159-162 and
// elemgtsub.c:154-157 (the factor argument 2^(bits-2) shifted by 64).
inline uint64_t barrett_mu(uint64_t q) {
	return barrett_factor((uint64_t) 1 << (bitlen(q) - 2), q);
}

// src/numbers.c:5-40: 128-bit Barrett, alpha = 62, beta = -2, one
// conditional subtraction.  (66 - bits) & 63 mirrors x86's shift-count
// masking at bits == 2.
inline uint64_t mulmod_barrett(uint64_t a, uint64_t b, uint64_t q) {
	const u128 p = (u128) a * b;
	const uint64_t hi = (uint64_t) (p >> 64), lo = (uint64_t) p;
	const uint64_t n = bitlen(q);
	const uint64_t c = (hi << ((66 - n) & 63)) + (lo >> (n - 2));
	const uint64_t qh = (uint64_t) (((u128) c * barrett_mu(q)) >> 64);
	uint64_t z = lo - qh * q;
	return z >= q ? z - q : z;
}

inline uint64_t mulmod_exact(uint64_t a, uint64_t b, uint64_t q) {
	return (uint64_t) (((u128) a * b) % q);
}

inline uint64_t gcd_u64(uint64_t a, uint64_t b) {
	while (b) {
		const uint64_t t = a % b;
		a = b;
		b = t;
	}
	return a;
}

// src/numbers.c:71-99: extended Euclid on int64.  Only called with
// gcd(a, q) == 1 and q < 2^62 (checked by the callers), where the
// reference's loop is well defined.
inline uint64_t invmod_euclid(uint64_t input, uint64_t q) {
	if (q == 1)
		return 0;
	int64_t a = (int64_t) input, b = (int64_t) q, y = 0, x = 1;
	while (a > 1) {
		const int64_t quot = a / b, rem = a % b;
		a = b;
		b = rem;
		const int64_t t = y;
		y = x - quot * y;
		x = t;
	}
	if (x < 0)
		x += (int64_t) q;
	return (uint64_t) x;
}

// 2^64 mod q.
// Constants of a one-multiply reduction of v < 81q to [0, q): with
// L = bitlen(q), shift = max(L - 7, 0) and m = floor(2^(32 + shift) / q),
// T = v >> shift is below 81 * 2^7 < 2^14 and Q = floor(T * m / 2^32)
// satisfies floor(v/q) - 1 <= Q <= floor(v/q): Q <= T 2^shift / q <= v/q,
// and T m / 2^32 > v/q - 2^shift/q - T/2^32 > v/q - 1 (2^shift/q <= 1/5).
// So v - Q q < 2q and one conditional subtraction finishes; m < 2^32.
inline void small_reduce_consts(uint64_t q, uint32_t &shift, uint32_t &m) {
	const uint64_t l = bitlen(q);
	shift = (uint32_t) (l > 7 ? l - 7 : 0);
	m = (uint32_t) ((((u128) 1) << (32 + shift)) / q);
}

inline uint64_t pow2_64_mod(uint64_t q) {
	return (uint64_t) ((((u128) 1) << 64) % q);
}

}  // namespace vk
