/*
 * vkhel CPU oracle -- TEST INFRASTRUCTURE ONLY (see vkhel_oracle.h).
 *
 * Scalar restatement of the reference's arithmetic.  Nothing in the product
 * library (vkhel_amd/) links this file.  Built by oracle/Makefile into
 * oracle/liboracle.so.
 */
#define _GNU_SOURCE
#include "vkhel_oracle.h"

#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

typedef unsigned __int128 u128;

static uint64_t mulhi64(uint64_t a, uint64_t b) {
	return (uint64_t) (((u128) a * b) >> 64);
}

/* include/priv/numbers.h:11-13 -- named "ceil_log2" upstream but it is the
 * bit length 64 - clz(v) (test/numbers.c:12-18 checks floor(log2 v) + 1). */
uint64_t or_bitlen(uint64_t v) {
	return v ? 64 - (uint64_t) __builtin_clzll(v) : 0;
}

/* src/numbers.c:30-34: floor(factor * 2^64 / mod), truncated to 64 bits.
 * With factor = w < q this is the Shoup companion w'; with factor =
 * 2^(n-2) it is the Barrett constant mu = floor(2^(n+62) / q). */
uint64_t or_barrett_factor(uint64_t factor, uint64_t mod) {
	return (uint64_t) (((u128) factor << 64) / mod);
}

static uint64_t barrett_mu(uint64_t mod) {
	const uint64_t n = or_bitlen(mod);
	return or_barrett_factor((uint64_t) 1 << (n - 2), mod);
}

/* src/numbers.c:5-28 (compute_mod) and elemmul.comp:49-60 (reduce128):
 * Barrett with alpha = 62, beta = -2, n = bitlen(mod), exactly one
 * conditional subtraction.  The shift count 64 - (n - 2) equals 64 when
 * n == 2; x86 masks shift counts to 6 bits, so the compiled reference
 * computes hi << 0 there, and this restatement does the same. */
static uint64_t barrett128(uint64_t hi, uint64_t lo, uint64_t mod) {
	const uint64_t n = or_bitlen(mod);
	const uint64_t mu = barrett_mu(mod);
	const uint64_t num_c = (hi << ((66 - n) & 63)) + (lo >> (n - 2));
	const uint64_t q_hat = mulhi64(num_c, mu);
	uint64_t z = lo - q_hat * mod;
	if (z >= mod)
		z -= mod;
	return z;
}

/* elemmul.comp:36-47, elemgtsub.comp:36-47 (reduce64): the 128-bit
 * Barrett above with hi = 0. */
uint64_t or_reduce64(uint64_t x, uint64_t mod) {
	const uint64_t n = or_bitlen(mod);
	const uint64_t mu = barrett_mu(mod);
	const uint64_t q_hat = mulhi64(x >> (n - 2), mu);
	uint64_t z = x - q_hat * mod;
	if (z >= mod)
		z -= mod;
	return z;
}

/* src/numbers.c:36-40 */
uint64_t or_mulmod(uint64_t a, uint64_t b, uint64_t mod) {
	const u128 p = (u128) a * b;
	return barrett128((uint64_t) (p >> 64), (uint64_t) p, mod);
}

/* src/numbers.c:42-59: base %= mod, then square-and-multiply. */
uint64_t or_powmod(uint64_t base, uint64_t exp, uint64_t mod) {
	uint64_t result = 1;
	base %= mod;
	while (exp > 0) {
		if (exp & 1)
			result = or_mulmod(result, base, mod);
		base = or_mulmod(base, base, mod);
		exp >>= 1;
	}
	return result;
}

/* src/numbers.c:61-69: root^(degree/2) == mod - 1; degree a power of two. */
int or_is_primitive_root(uint64_t root, uint64_t degree, uint64_t mod) {
	if (root == 0)
		return 0;
	if (__builtin_popcountll(degree) != 1)
		return 0;
	return or_powmod(root, degree / 2, mod) == mod - 1;
}

/* src/numbers.c:71-99: extended Euclid on int64 with C truncating
 * division; a negative result is lifted by +mod once. */
uint64_t or_invmod(uint64_t input, uint64_t mod) {
	if (mod == 1)
		return 0;
	int64_t a = (int64_t) input, b = (int64_t) mod, y = 0, x = 1;
	while (a > 1) {
		const int64_t quot = a / b, rem = a % b;
		a = b;
		b = rem;
		const int64_t t = y;
		y = x - quot * y;
		x = t;
	}
	if (x < 0)
		x += (int64_t) mod;
	return (uint64_t) x;
}

/* src/ntt_tables.c:8-15 */
static uint64_t reverse_bits(uint64_t a, uint64_t width) {
	uint64_t r = 0;
	for (uint64_t b = 0; b < width; b++) {
		r = (r << 1) | (a & 1);
		a >>= 1;
	}
	return r;
}

/* src/ntt_tables.c:17-44: roots[bitrev(i)] = w^i by repeated Barrett
 * multiplication, inv_roots[same] = ext-Euclid inverse, then Shoup
 * companions floor(x * 2^64 / q) of both arrays. */
void or_tables_create(uint64_t n, uint64_t q, uint64_t w, uint64_t *roots,
		uint64_t *inv_roots, uint64_t *roots_bf, uint64_t *inv_roots_bf) {
	const uint64_t logn = or_bitlen(n) - 1;
	roots[0] = 1;
	inv_roots[0] = or_invmod(1, q);
	uint64_t prev = 1;
	for (uint64_t i = 1; i < n; i++) {
		const uint64_t idx = reverse_bits(i, logn);
		roots[idx] = or_mulmod(prev, w, q);
		inv_roots[idx] = or_invmod(roots[idx], q);
		prev = roots[idx];
	}
	for (uint64_t i = 0; i < n; i++) {
		roots_bf[i] = or_barrett_factor(roots[i], q);
		inv_roots_bf[i] = or_barrett_factor(inv_roots[i], q);
	}
}

static int pow2_ok(uint64_t n) {
	return n >= 2 && (n & (n - 1)) == 0;
}

/* src/vector.c:533-566 stage loop + nttfwdbutterfly.comp:31-58 butterfly.
 * Stage m reads `in` (first stage) or `out`, block i uses roots[m + i]. */
int or_ntt_forward_t(const uint64_t *in, uint64_t *out, uint64_t n,
		uint64_t q, const uint64_t *roots, const uint64_t *roots_bf) {
	if (!pow2_ok(n))
		return -1;
	const uint64_t *src = in;
	uint64_t t = n / 2;
	for (uint64_t m = 1; m < n; m *= 2) {
		for (uint64_t i = 0; i < m; i++) {
			const uint64_t w = roots[m + i], wp = roots_bf[m + i];
			const uint64_t base = 2 * i * t;
			for (uint64_t pos = 0; pos < t; pos++) {
				const uint64_t xi = base + pos, yi = xi + t;
				const uint64_t X = src[xi], Y = src[yi];
				uint64_t wy = Y * w - mulhi64(Y, wp) * q;
				if (wy >= q)
					wy -= q;
				uint64_t ty = X;
				while (ty < wy)
					ty += q;
				ty -= wy;
				out[xi] = (X + wy) % q;
				out[yi] = ty % q;
			}
		}
		src = out;
		t /= 2;
	}
	return 0;
}

/* src/vector.c:596-649 stage loop + nttrevbutterfly.comp:31-58, then the
 * elemmulconst pass by inv(n) (elemmulconst.comp:35-49).  Precondition:
 * every input < q (the subtract loop at nttrevbutterfly.comp:44-48 is only
 * meaningful there). */
int or_ntt_inverse_t(const uint64_t *in, uint64_t *out, uint64_t n,
		uint64_t q, const uint64_t *inv_roots, const uint64_t *inv_roots_bf) {
	if (!pow2_ok(n))
		return -1;
	for (uint64_t i = 0; i < n; i++)
		if (in[i] >= q)
			return -1;
	const uint64_t *src = in;
	uint64_t t = 1;
	for (uint64_t m = n / 2; m >= 1; m /= 2) {
		for (uint64_t i = 0; i < m; i++) {
			const uint64_t w = inv_roots[m + i], wp = inv_roots_bf[m + i];
			const uint64_t base = 2 * i * t;
			for (uint64_t pos = 0; pos < t; pos++) {
				const uint64_t xi = base + pos, yi = xi + t;
				const uint64_t X = src[xi], Y = src[yi];
				uint64_t ty = X;
				while (ty < Y)
					ty += q;
				ty -= Y;
				uint64_t wy = ty * w - mulhi64(ty, wp) * q;
				if (wy >= q)
					wy -= q;
				out[xi] = (X + Y) % q;
				out[yi] = wy;
			}
		}
		src = out;
		t *= 2;
	}
	or_elemmulconst(out, out, n, or_invmod(n, q), q);
	return 0;
}

struct tables {
	uint64_t *r, *ir, *rb, *irb;
};

static int tables_alloc(struct tables *t, uint64_t n, uint64_t q,
		uint64_t w) {
	t->r = malloc(4 * n * sizeof(uint64_t));
	if (!t->r)
		return -1;
	t->ir = t->r + n;
	t->rb = t->r + 2 * n;
	t->irb = t->r + 3 * n;
	or_tables_create(n, q, w, t->r, t->ir, t->rb, t->irb);
	return 0;
}

int or_ntt_forward(const uint64_t *in, uint64_t *out, uint64_t n, uint64_t q,
		uint64_t w) {
	struct tables t;
	if (!pow2_ok(n) || tables_alloc(&t, n, q, w))
		return -1;
	const int rc = or_ntt_forward_t(in, out, n, q, t.r, t.rb);
	free(t.r);
	return rc;
}

int or_ntt_inverse(const uint64_t *in, uint64_t *out, uint64_t n, uint64_t q,
		uint64_t w) {
	struct tables t;
	if (!pow2_ok(n) || tables_alloc(&t, n, q, w))
		return -1;
	const int rc = or_ntt_inverse_t(in, out, n, q, t.ir, t.irb);
	free(t.r);
	return rc;
}

/* elemfma.comp:36-55 with the push constants of src/kernels/elemfma.c:
 * 159-169 (m' = floor(m * 2^64 / q)).  Restated as written, including the
 * `mod - x` in both correction branches (:47-48 and :51-52), which wraps
 * modulo 2^64 whenever the intermediate reaches q. */
void or_elemfma(const uint64_t *a, const uint64_t *b, uint64_t *out,
		size_t len, uint64_t multiplier, uint64_t mod) {
	const uint64_t mp = or_barrett_factor(multiplier, mod);
	for (size_t i = 0; i < len; i++) {
		uint64_t prod = a[i] * multiplier - mulhi64(a[i], mp) * mod;
		if (prod >= mod)
			prod = mod - prod;
		const uint64_t sum = prod + b[i];
		out[i] = sum >= mod ? mod - sum : sum;
	}
}

/* elemmul.comp:62-73 with src/kernels/elemmul.c:159-168. */
void or_elemmul(const uint64_t *a, const uint64_t *b, uint64_t *out,
		size_t len, uint64_t mod) {
	for (size_t i = 0; i < len; i++) {
		const u128 p = (u128) or_reduce64(a[i], mod) * or_reduce64(b[i], mod);
		out[i] = barrett128((uint64_t) (p >> 64), (uint64_t) p, mod);
	}
}

/* elemgtadd.comp:20-30 */
void or_elemgtadd(const uint64_t *a, uint64_t *out, size_t len,
		uint64_t bound, uint64_t diff) {
	for (size_t i = 0; i < len; i++)
		out[i] = a[i] > bound ? a[i] + diff : a[i];
}

/* elemgtsub.comp:49-67 with src/kernels/elemgtsub.c:154-165. */
void or_elemgtsub(const uint64_t *a, uint64_t *out, size_t len,
		uint64_t bound, uint64_t diff, uint64_t mod) {
	const uint64_t dr = or_reduce64(diff, mod);
	for (size_t i = 0; i < len; i++) {
		const uint64_t r = or_reduce64(a[i], mod);
		if (a[i] > bound) {
			const uint64_t z = r + mod - dr;
			out[i] = z >= mod ? z - mod : z;
		} else {
			out[i] = r;
		}
	}
}

/* src/vector.c:360-368: mod 2 uses elemmodbytwo.comp:19-27 with bound q/2;
 * any other mod uses elemgtsub(bound = q/2, diff = q, mod). */
void or_elemmod(const uint64_t *a, uint64_t *out, size_t len, uint64_t mod,
		uint64_t q) {
	if (mod == 2) {
		const uint64_t sb = q / 2;
		for (size_t i = 0; i < len; i++) {
			const uint64_t u = a[i] & 1;
			out[i] = a[i] > sb ? 1 - u : u;
		}
	} else {
		or_elemgtsub(a, out, len, q / 2, q, mod);
	}
}

/* elemmulconst.comp:35-49 with src/kernels/elemmulconst.c:154-163. */
void or_elemmulconst(const uint64_t *a, uint64_t *out, size_t len,
		uint64_t c, uint64_t mod) {
	const uint64_t cp = or_barrett_factor(c, mod);
	for (size_t i = 0; i < len; i++) {
		uint64_t p = a[i] * c - mulhi64(a[i], cp) * mod;
		if (p >= mod)
			p -= mod;
		out[i] = p;
	}
}

void or_addmul_mod(const uint64_t *a, const uint64_t *b, uint64_t *out,
		size_t len, uint64_t multiplier, uint64_t mod) {
	for (size_t i = 0; i < len; i++)
		out[i] = (uint64_t) (((u128) a[i] * multiplier + b[i]) % mod);
}

/* ---------------- independent mathematics ---------------- */

static uint64_t mulmod_plain(uint64_t a, uint64_t b, uint64_t q) {
	return (uint64_t) (((u128) a * b) % q);
}

static uint64_t powmod_plain(uint64_t b, uint64_t e, uint64_t q) {
	uint64_t r = 1 % q;
	b %= q;
	while (e) {
		if (e & 1)
			r = mulmod_plain(r, b, q);
		b = mulmod_plain(b, b, q);
		e >>= 1;
	}
	return r;
}

void or_dft_negacyclic_bitrev(const uint64_t *a, uint64_t *out, uint64_t n,
		uint64_t q, uint64_t psi) {
	const uint64_t logn = or_bitlen(n) - 1;
	for (uint64_t i = 0; i < n; i++) {
		const uint64_t e = 2 * reverse_bits(i, logn) + 1;
		const uint64_t step = powmod_plain(psi, e, q);
		uint64_t acc = 0, p = 1;
		for (uint64_t j = 0; j < n; j++) {
			acc = (uint64_t) (((u128) acc + mulmod_plain(a[j] % q, p, q)) % q);
			p = mulmod_plain(p, step, q);
		}
		out[i] = acc;
	}
}

void or_negacyclic_schoolbook(const uint64_t *a, const uint64_t *b,
		uint64_t *out, uint64_t n, uint64_t q) {
	uint64_t *acc = calloc(n, sizeof(uint64_t));
	for (uint64_t i = 0; i < n; i++) {
		const uint64_t ai = a[i] % q;
		for (uint64_t j = 0; j < n; j++) {
			const uint64_t p = mulmod_plain(ai, b[j] % q, q);
			const uint64_t k = i + j;
			/* X^n = -1: terms past degree n-1 wrap with a sign flip. */
			if (k < n)
				acc[k] = (uint64_t) (((u128) acc[k] + p) % q);
			else
				acc[k - n] = (uint64_t) (((u128) acc[k - n] + q - p) % q);
		}
	}
	memcpy(out, acc, n * sizeof(uint64_t));
	free(acc);
}

int or_polymul_ntt(const uint64_t *a, const uint64_t *b, uint64_t *out,
		uint64_t n, uint64_t q, uint64_t w) {
	struct tables t;
	if (!pow2_ok(n) || tables_alloc(&t, n, q, w))
		return -1;
	uint64_t *fa = malloc(2 * n * sizeof(uint64_t));
	uint64_t *fb = fa + n;
	int rc = or_ntt_forward_t(a, fa, n, q, t.r, t.rb);
	rc |= or_ntt_forward_t(b, fb, n, q, t.r, t.rb);
	or_elemmul(fa, fb, fa, n, q);
	rc |= or_ntt_inverse_t(fa, out, n, q, t.ir, t.irb);
	free(fa);
	free(t.r);
	return rc;
}

/* Deterministic Miller-Rabin; these 12 bases are exact below 3.3e24. */
int or_is_prime(uint64_t n) {
	static const uint64_t bases[] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29,
		31, 37};
	if (n < 2)
		return 0;
	for (size_t i = 0; i < 12; i++) {
		if (n == bases[i])
			return 1;
		if (n % bases[i] == 0)
			return 0;
	}
	uint64_t d = n - 1;
	int s = 0;
	while (!(d & 1)) {
		d >>= 1;
		s++;
	}
	for (size_t i = 0; i < 12; i++) {
		uint64_t x = powmod_plain(bases[i], d, n);
		if (x == 1 || x == n - 1)
			continue;
		int comp = 1;
		for (int r = 1; r < s; r++) {
			x = mulmod_plain(x, x, n);
			if (x == n - 1) {
				comp = 0;
				break;
			}
		}
		if (comp)
			return 0;
	}
	return 1;
}

static uint64_t gcd64(uint64_t a, uint64_t b) {
	while (b) {
		const uint64_t t = a % b;
		a = b;
		b = t;
	}
	return a;
}

/* Pollard-Brent rho; n is odd and composite. */
static uint64_t rho(uint64_t n) {
	for (uint64_t c = 1;; c++) {
		uint64_t y = 2, x = 2, g = 1, qprod = 1, ys = 2;
		uint64_t r = 1;
		while (g == 1) {
			x = y;
			for (uint64_t i = 0; i < r; i++)
				y = (mulmod_plain(y, y, n) + c) % n;
			for (uint64_t k = 0; k < r && g == 1; k += 64) {
				ys = y;
				for (uint64_t i = 0; i < 64 && i < r - k; i++) {
					y = (mulmod_plain(y, y, n) + c) % n;
					qprod = mulmod_plain(qprod, x > y ? x - y : y - x, n);
				}
				g = gcd64(qprod, n);
			}
			r *= 2;
		}
		if (g == n) {
			do {
				ys = (mulmod_plain(ys, ys, n) + c) % n;
				g = gcd64(x > ys ? x - ys : ys - x, n);
			} while (g == 1);
		}
		if (g != n)
			return g;
	}
}

static size_t factor_into(uint64_t n, uint64_t *f, size_t nf) {
	if (n == 1)
		return nf;
	if (or_is_prime(n)) {
		for (size_t i = 0; i < nf; i++)
			if (f[i] == n)
				return nf;
		f[nf] = n;
		return nf + 1;
	}
	const uint64_t d = rho(n);
	nf = factor_into(d, f, nf);
	return factor_into(n / d, f, nf);
}

/* Distinct prime factors of n (at most 15 for 64-bit n). */
static size_t distinct_factors(uint64_t n, uint64_t *f) {
	size_t nf = 0;
	for (uint64_t p = 2; p < 1000 && p * p <= n; p++) {
		if (n % p == 0) {
			f[nf++] = p;
			while (n % p == 0)
				n /= p;
		}
	}
	return factor_into(n, f, nf);
}

uint64_t or_smallest_generator(uint64_t q) {
	uint64_t f[64];
	const size_t nf = distinct_factors(q - 1, f);
	for (uint64_t g = 2; g < q; g++) {
		int ok = 1;
		for (size_t i = 0; i < nf && ok; i++)
			ok = powmod_plain(g, (q - 1) / f[i], q) != 1;
		if (ok)
			return g;
	}
	return 0;
}

size_t or_primes_ntt(uint64_t bound, uint64_t m, size_t count,
		uint64_t *out) {
	size_t k = 0;
	if (bound < m + 2)
		return 0;
	for (uint64_t c = ((bound - 2) / m) * m + 1; c > 1 && k < count; c -= m) {
		if (or_is_prime(c))
			out[k++] = c;
		if (c <= m)
			break;
	}
	return k;
}

uint64_t or_root_2n(uint64_t q, uint64_t n) {
	const uint64_t g = or_smallest_generator(q);
	if (!g || (q - 1) % (2 * n))
		return 0;
	return powmod_plain(g, (q - 1) / (2 * n), q);
}

uint64_t or_splitmix64(uint64_t *state) {
	uint64_t z = (*state += 0x9e3779b97f4a7c15ULL);
	z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
	z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
	return z ^ (z >> 31);
}

void or_fill_uniform(uint64_t *out, size_t len, uint64_t q, uint64_t seed) {
	uint64_t st = seed;
	const uint64_t bits = or_bitlen(q);
	for (size_t i = 0; i < len;) {
		const uint64_t x = bits >= 64 ? or_splitmix64(&st)
			: or_splitmix64(&st) >> (64 - bits);
		if (x < q)
			out[i++] = x;
	}
}

/* ---------------- CPU baseline runner ---------------- */

/* A fast scalar NTT with the same canonical outputs as or_ntt_forward_t /
 * or_ntt_inverse_t on canonical inputs: Shoup products, conditional
 * subtractions instead of `%`.  Used only to time the CPU baseline; the
 * tests check it against the restatement above. */
static void fwd_fast(uint64_t *a, uint64_t n, uint64_t q, const uint64_t *r,
		const uint64_t *rb) {
	uint64_t t = n / 2;
	for (uint64_t m = 1; m < n; m *= 2, t /= 2) {
		for (uint64_t i = 0; i < m; i++) {
			const uint64_t w = r[m + i], wp = rb[m + i];
			uint64_t *x = a + 2 * i * t, *y = x + t;
			for (uint64_t j = 0; j < t; j++) {
				uint64_t wy = y[j] * w - mulhi64(y[j], wp) * q;
				wy -= wy >= q ? q : 0;
				const uint64_t s = x[j] + wy, d = x[j] + q - wy;
				x[j] = s >= q ? s - q : s;
				y[j] = d >= q ? d - q : d;
			}
		}
	}
}

static void inv_fast(uint64_t *a, uint64_t n, uint64_t q, const uint64_t *r,
		const uint64_t *rb, uint64_t ninv, uint64_t ninvp) {
	uint64_t t = 1;
	for (uint64_t m = n / 2; m >= 1; m /= 2, t *= 2) {
		for (uint64_t i = 0; i < m; i++) {
			const uint64_t w = r[m + i], wp = rb[m + i];
			uint64_t *x = a + 2 * i * t, *y = x + t;
			for (uint64_t j = 0; j < t; j++) {
				const uint64_t s = x[j] + y[j], d = x[j] + q - y[j];
				uint64_t wy = d * w - mulhi64(d, wp) * q;
				wy -= wy >= q ? q : 0;
				x[j] = s >= q ? s - q : s;
				y[j] = wy;
			}
		}
	}
	for (uint64_t i = 0; i < n; i++) {
		uint64_t p = a[i] * ninv - mulhi64(a[i], ninvp) * q;
		a[i] = p >= q ? p - q : p;
	}
}

struct bench_arg {
	uint64_t *polys;
	size_t first, last;
	uint64_t n;
	const uint64_t *qs;
	size_t nprimes;
	struct tables *tabs;
};

static void *bench_worker(void *p) {
	struct bench_arg *a = p;
	for (size_t k = a->first; k < a->last; k++) {
		const size_t pi = k % a->nprimes;
		const uint64_t q = a->qs[pi];
		const struct tables *t = &a->tabs[pi];
		uint64_t *poly = a->polys + k * a->n;
		fwd_fast(poly, a->n, q, t->r, t->rb);
		const uint64_t ninv = or_invmod(a->n, q);
		inv_fast(poly, a->n, q, t->ir, t->irb, ninv,
				or_barrett_factor(ninv, q));
	}
	return NULL;
}

double or_bench_ntt_batch(uint64_t *polys, size_t npolys, uint64_t n,
		const uint64_t *qs, const uint64_t *ws, size_t nprimes,
		int threads) {
	if (!pow2_ok(n) || nprimes == 0 || threads < 1)
		return -1.0;
	struct tables *tabs = calloc(nprimes, sizeof(*tabs));
	for (size_t i = 0; i < nprimes; i++)
		tables_alloc(&tabs[i], n, qs[i], ws[i]);
	pthread_t *th = calloc((size_t) threads, sizeof(pthread_t));
	struct bench_arg *args = calloc((size_t) threads, sizeof(*args));
	struct timespec t0, t1;
	clock_gettime(CLOCK_MONOTONIC, &t0);
	for (int i = 0; i < threads; i++) {
		args[i] = (struct bench_arg) {polys, npolys * (size_t) i / threads,
			npolys * (size_t) (i + 1) / threads, n, qs, nprimes, tabs};
		pthread_create(&th[i], NULL, bench_worker, &args[i]);
	}
	for (int i = 0; i < threads; i++)
		pthread_join(th[i], NULL);
	clock_gettime(CLOCK_MONOTONIC, &t1);
	for (size_t i = 0; i < nprimes; i++)
		free(tabs[i].r);
	free(tabs);
	free(th);
	free(args);
	return (double) (t1.tv_sec - t0.tv_sec) + 1e-9 * (t1.tv_nsec - t0.tv_nsec);
}

struct polymul_arg {
	const uint64_t *a, *b;
	uint64_t *out;
	size_t first, last;
	uint64_t n;
	const uint64_t *qs;
	size_t nprimes;
	struct tables *tabs;
};

static void *polymul_worker(void *p) {
	struct polymul_arg *g = p;
	uint64_t *fb = malloc(g->n * sizeof(uint64_t));
	for (size_t k = g->first; k < g->last; k++) {
		const size_t pi = k % g->nprimes;
		const uint64_t q = g->qs[pi];
		const struct tables *t = &g->tabs[pi];
		uint64_t *fa = g->out + k * g->n;
		memcpy(fa, g->a + k * g->n, g->n * sizeof(uint64_t));
		memcpy(fb, g->b + k * g->n, g->n * sizeof(uint64_t));
		fwd_fast(fa, g->n, q, t->r, t->rb);
		fwd_fast(fb, g->n, q, t->r, t->rb);
		or_elemmul(fa, fb, fa, g->n, q);
		const uint64_t ninv = or_invmod(g->n, q);
		inv_fast(fa, g->n, q, t->ir, t->irb, ninv, or_barrett_factor(ninv, q));
	}
	free(fb);
	return NULL;
}

double or_bench_polymul_batch(const uint64_t *a, const uint64_t *b,
		uint64_t *out, size_t npairs, uint64_t n, const uint64_t *qs,
		const uint64_t *ws, size_t nprimes, int threads) {
	if (!pow2_ok(n) || nprimes == 0 || threads < 1)
		return -1.0;
	struct tables *tabs = calloc(nprimes, sizeof(*tabs));
	for (size_t i = 0; i < nprimes; i++)
		tables_alloc(&tabs[i], n, qs[i], ws[i]);
	pthread_t *th = calloc((size_t) threads, sizeof(pthread_t));
	struct polymul_arg *args = calloc((size_t) threads, sizeof(*args));
	struct timespec t0, t1;
	clock_gettime(CLOCK_MONOTONIC, &t0);
	for (int i = 0; i < threads; i++) {
		args[i] = (struct polymul_arg) {a, b, out, npairs * (size_t) i / threads,
			npairs * (size_t) (i + 1) / threads, n, qs, nprimes, tabs};
		pthread_create(&th[i], NULL, polymul_worker, &args[i]);
	}
	for (int i = 0; i < threads; i++)
		pthread_join(th[i], NULL);
	clock_gettime(CLOCK_MONOTONIC, &t1);
	for (size_t i = 0; i < nprimes; i++)
		free(tabs[i].r);
	free(tabs);
	free(th);
	free(args);
	return (double) (t1.tv_sec - t0.tv_sec) + 1e-9 * (t1.tv_nsec - t0.tv_nsec);
}
