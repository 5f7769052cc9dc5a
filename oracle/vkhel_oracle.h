/*
 * vkhel CPU oracle -- TEST INFRASTRUCTURE ONLY.
 *
 * A scalar C restatement of lolzballs/vkhel's arithmetic, used as the
 * checker for the HIP library in vkhel_amd/.  Only tests/, the smoke entry
 * point and bench.py's CPU-baseline leg may load it; the product library
 * never links or calls it.
 *
 * Every function cites the reference file:line it restates.  Two groups:
 *
 *  (1) Restatements of reference semantics (bit-exact by construction):
 *      or_bitlen, or_barrett_factor, or_mulmod, or_powmod, or_invmod,
 *      or_is_primitive_root, or_tables_create, or_ntt_forward,
 *      or_ntt_inverse, or_elem*.
 *  (2) Independent mathematics used to pin (1) without trusting it:
 *      or_dft_negacyclic_bitrev (closed form), or_negacyclic_schoolbook,
 *      or_is_prime / or_smallest_generator / or_primes_ntt (number theory
 *      from first principles), splitmix64 input generation.
 *
 * How each function of group (1) is pinned (nothing here is "parity
 * unpinned" except where noted):
 *   or_bitlen, or_mulmod, or_powmod, or_invmod, or_is_primitive_root:
 *       test/numbers.c:12-94 known answers, brute force with Python ints,
 *       and the reference numbers.c compiled into oracle/_ref.
 *   or_tables_create: test/ntt.c:19-21, the compiled reference
 *       ntt_tables.c in oracle/_ref, and w^i recomputed with Python pow().
 *   or_ntt_forward / or_ntt_inverse: test/vector.c:243-329 vectors, the
 *       closed-form bit-reversed negacyclic DFT, and the convolution
 *       theorem against the schoolbook product.
 *   or_elemmul, or_elemgtadd, or_elemgtsub, or_elemmod: test/vector.c
 *       :85-241 vectors plus closed forms with Python ints.
 *   or_elemfma: test/vector.c:39-83 vectors; its wrap branch (reference
 *       quirk, elemfma.comp:47-48,51-52) is pinned by a Python restatement
 *       with explicit 2^64 wrapping, not by a reference run.
 *   or_ntt_inverse on inputs >= q: parity unpinned (the reference loop at
 *       nttrevbutterfly.comp:44-48 is undefined-ish there); the oracle
 *       rejects such inputs.
 *   or_bench_ntt_batch, or_bench_polymul_batch (timing runners for the
 *       CPU baseline): checked by the forward+inverse identity and against
 *       schoolbook products.
 */
#ifndef VKHEL_ORACLE_H
#define VKHEL_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- number theory (src/numbers.c) ---- */
uint64_t or_bitlen(uint64_t v);
uint64_t or_barrett_factor(uint64_t factor, uint64_t mod);
uint64_t or_mulmod(uint64_t a, uint64_t b, uint64_t mod);
uint64_t or_reduce64(uint64_t x, uint64_t mod);
uint64_t or_powmod(uint64_t base, uint64_t exp, uint64_t mod);
uint64_t or_invmod(uint64_t a, uint64_t mod);
int or_is_primitive_root(uint64_t root, uint64_t degree, uint64_t mod);

/* ---- tables (src/ntt_tables.c) ---- */
/* Fills four caller-provided arrays of n words each. */
void or_tables_create(uint64_t n, uint64_t q, uint64_t w, uint64_t *roots,
		uint64_t *inv_roots, uint64_t *roots_bf, uint64_t *inv_roots_bf);

/* ---- transforms (src/vector.c:513-657) ---- */
/* Tables are looked up internally (built per call); `in` and `out` may
 * alias.  Return 0 on success, -1 on a precondition failure. */
int or_ntt_forward(const uint64_t *in, uint64_t *out, uint64_t n, uint64_t q,
		uint64_t w);
int or_ntt_inverse(const uint64_t *in, uint64_t *out, uint64_t n, uint64_t q,
		uint64_t w);
/* Same, with caller-built tables (so batches do not rebuild them). */
int or_ntt_forward_t(const uint64_t *in, uint64_t *out, uint64_t n,
		uint64_t q, const uint64_t *roots, const uint64_t *roots_bf);
int or_ntt_inverse_t(const uint64_t *in, uint64_t *out, uint64_t n,
		uint64_t q, const uint64_t *inv_roots, const uint64_t *inv_roots_bf);

/* ---- elementwise (src/vector.c:298-511 + shaders) ---- */
void or_elemfma(const uint64_t *a, const uint64_t *b, uint64_t *out,
		size_t len, uint64_t multiplier, uint64_t mod);
void or_elemmul(const uint64_t *a, const uint64_t *b, uint64_t *out,
		size_t len, uint64_t mod);
void or_elemgtadd(const uint64_t *a, uint64_t *out, size_t len,
		uint64_t bound, uint64_t diff);
void or_elemgtsub(const uint64_t *a, uint64_t *out, size_t len,
		uint64_t bound, uint64_t diff, uint64_t mod);
void or_elemmod(const uint64_t *a, uint64_t *out, size_t len, uint64_t mod,
		uint64_t q);
void or_elemmulconst(const uint64_t *a, uint64_t *out, size_t len,
		uint64_t c, uint64_t mod);
/* Correct (a*m + b) mod q for canonical a, b (the batch API's "add"/"fma"
 * op; not a reference entry point). */
void or_addmul_mod(const uint64_t *a, const uint64_t *b, uint64_t *out,
		size_t len, uint64_t multiplier, uint64_t mod);

/* ---- independent mathematics ---- */
/* A[i] = sum_j a_j psi^((2 bitrev(i) + 1) j) mod q, O(n^2). */
void or_dft_negacyclic_bitrev(const uint64_t *a, uint64_t *out, uint64_t n,
		uint64_t q, uint64_t psi);
/* c = a * b mod (X^n + 1, q), O(n^2), plain 128-bit % arithmetic. */
void or_negacyclic_schoolbook(const uint64_t *a, const uint64_t *b,
		uint64_t *out, uint64_t n, uint64_t q);
/* NTT(a), NTT(b), elemmul, INTT -- the reference's own pipeline. */
int or_polymul_ntt(const uint64_t *a, const uint64_t *b, uint64_t *out,
		uint64_t n, uint64_t q, uint64_t w);

int or_is_prime(uint64_t n);
uint64_t or_smallest_generator(uint64_t q);
/* Up to `count` primes q < bound with q == 1 mod m, descending. */
size_t or_primes_ntt(uint64_t bound, uint64_t m, size_t count, uint64_t *out);
/* psi = g^((q-1)/2n) for the smallest generator g of (Z/qZ)^*. */
uint64_t or_root_2n(uint64_t q, uint64_t n);

/* ---- synthetic inputs ---- */
uint64_t or_splitmix64(uint64_t *state);
/* `len` words uniform in [0,q) by rejection on the top bitlen(q) bits of
 * each splitmix64 draw, stream seeded with `seed`. */
void or_fill_uniform(uint64_t *out, size_t len, uint64_t q, uint64_t seed);

/* ---- CPU baseline runner ---- */
/* Forward then inverse NTT of `npolys` polynomials of n words in place,
 * polynomial p using prime index p % nprimes; `threads` POSIX threads.
 * Returns wall seconds. */
double or_bench_ntt_batch(uint64_t *polys, size_t npolys, uint64_t n,
		const uint64_t *qs, const uint64_t *ws, size_t nprimes,
		int threads);
/* Negacyclic products out_p = a_p * b_p mod (X^n + 1, q) of `npairs` pairs
 * of n words (forward transforms of both, pointwise product, inverse),
 * pair p using prime index p % nprimes; `threads` POSIX threads.  Returns
 * wall seconds. */
double or_bench_polymul_batch(const uint64_t *a, const uint64_t *b,
		uint64_t *out, size_t npairs, uint64_t n, const uint64_t *qs,
		const uint64_t *ws, size_t nprimes, int threads);

#ifdef __cplusplus
}
#endif

#endif
