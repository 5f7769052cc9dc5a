// Device-side 64-bit modular arithmetic for gfx950.  gfx950 has no 64x64
// multiplier: every 64-bit product below lowers to v_mad_u64_u32 /
// v_mul_hi_u32 / v_mul_lo_u32 chains, all full rate on this chip.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace vk {

__device__ __forceinline__ uint64_t mulhi64(uint64_t a, uint64_t b) {
	return __umul64hi(a, b);
}

// Makes a value opaque to the optimiser (emits no instruction), so the
// arithmetic around it is kept as written instead of being reassociated.
__device__ __forceinline__ uint64_t opaque(uint64_t v) {
	asm("" : "+v"(v));
	return v;
}

__device__ __forceinline__ uint32_t opaque32(uint32_t v) {
	asm("" : "+v"(v));
	return v;
}

// A value the compiler must treat as produced here: quantities derived from
// it are computed after this point instead of being hoisted and kept live.
__device__ __forceinline__ uint32_t opaque32_here(uint32_t v) {
	asm volatile("" : "+v"(v));
	return v;
}

// Same as opaque() for a wave-uniform value kept in scalar registers.
__device__ __forceinline__ uint64_t opaque_uniform(uint64_t v) {
	asm("" : "+s"(v));
	return v;
}

// x - m if x >= m else x (a compare, a 64-bit subtraction, two selects).
__device__ __forceinline__ uint64_t csub(uint64_t x, uint64_t m) {
	uint64_t d;
	return __builtin_sub_overflow(x, m, &d) ? x : d;
}

// csub with nm = 2^64 - m given: the difference is one 64-bit add, and the
// select a compare (four instructions instead of five).  nm must come from
// opaque_uniform(), or the add is folded back into a subtraction.
__device__ __forceinline__ uint64_t csub_n(uint64_t x, uint64_t m,
		uint64_t nm) {
	const uint64_t d = x + nm;
	return x < m ? x : d;
}

// (a - b) mod q for a, b in [0, q).
__device__ __forceinline__ uint64_t sub_mod(uint64_t a, uint64_t b,
		uint64_t q) {
	return a >= b ? a - b : a + q - b;
}

// Shoup product with a 63-bit companion wp = floor(w * 2^63 / q) = p1*2^32
// + p0, on 32-bit halves.  The companion is passed packed as
// wpk = (2*p1) << 32 | p0 (shoup63_factor), since 2*p1 is what the
// quotient needs and p1 = (wpk >> 33) is one shift.
// For y < 2^63 and w < q < 2^61 it returns R = y*w - Q*q,
// Q = floor(y * wp / 2^63), and R lies in [0, 2q):
// y*w/q - Q < 1 + y*(w*2^63/q - wp)/2^63 < 2.
// Because y1 = y >> 32 and p1 are both below 2^31, the middle
// partial-product sum S = y0*p1 + y1*p0 + floor(y0*p0 / 2^32) stays below
// 2^64, so Q = 2*y1*p1 + floor(S / 2^31) exactly, from one high multiply,
// three multiply-adds and one shift, with no carry chain.  R is formed modulo 2^64 as
// y*w + Q*nq with nq = 2^64 - q.
// LOW = false drops the low partial product floor(y0*p0 / 2^32) from S:
// that moves S / 2^31 down by less than 2, so the quotient is Q, Q - 1 or
// Q - 2 and R lies in [0, 4q), for one multiply fewer (shoup63_lazy).
template <bool LOW>
__device__ __forceinline__ uint64_t shoup63_impl(uint64_t y, uint64_t w,
		uint64_t wpk, uint64_t nq) {
	const uint32_t y0 = (uint32_t) y, y1 = (uint32_t) (y >> 32);
	const uint32_t p0 = (uint32_t) wpk, p1x2 = (uint32_t) (wpk >> 32);
	const uint32_t p1 = p1x2 >> 1;
	// (the opaque() points keep each sum a single multiply-add)
	const uint64_t s0 = opaque((uint64_t) y0 * p1 +
			(LOW ? __umulhi(y0, p0) : 0u));
	const uint64_t s = (uint64_t) y1 * p0 + s0;
	const uint64_t qt = (uint64_t) y1 * p1x2 + (s >> 31);
	const uint32_t w0 = (uint32_t) w, w1 = (uint32_t) (w >> 32);
	const uint32_t n0 = (uint32_t) nq, n1 = (uint32_t) (nq >> 32);
	const uint32_t q0 = (uint32_t) qt, q1 = (uint32_t) (qt >> 32);
	const uint64_t r = (uint64_t) y0 * w0 + (uint64_t) q0 * n0;
	// Cross terms of the high word, summed in the low halves of a chain of
	// 64-bit multiply-adds (their high halves are never used).
	uint64_t c = opaque((uint64_t) y0 * w1);
	c = opaque((uint64_t) y1 * w0 + c);
	c = opaque((uint64_t) q0 * n1 + c);
	c = (uint64_t) q1 * n0 + c;
	const uint32_t rh = opaque32((uint32_t) (r >> 32) + (uint32_t) c);
	return ((uint64_t) rh << 32) | (uint32_t) r;
}

__device__ __forceinline__ uint64_t shoup63(uint64_t y, uint64_t w,
		uint64_t wpk, uint64_t nq) {
	return shoup63_impl<true>(y, w, wpk, nq);
}

// As shoup63 with R in [0, 4q) instead of [0, 2q).
__device__ __forceinline__ uint64_t shoup63_lazy(uint64_t y, uint64_t w,
		uint64_t wpk, uint64_t nq) {
	return shoup63_impl<false>(y, w, wpk, nq);
}

// elemfma.comp:36-55 on one word: x*m mod q by the Shoup factor
// mp = floor(m * 2^64 / q) with its correction written `q - p` (which wraps
// modulo 2^64 where p - q was meant), then the sum with y corrected the same
// way (src/kernels/elemfma.c:159-169 builds m and mp).  Kept exactly, so the
// result equals the reference's on every input.
__device__ __forceinline__ uint64_t elemfma_word(uint64_t x, uint64_t y,
		uint64_t m, uint64_t mp, uint64_t q) {
	uint64_t p = x * m - __umul64hi(x, mp) * q;
	if (p >= q)
		p = q - p;
	const uint64_t s = p + y;
	return s >= q ? q - s : s;
}

// x - k*q modulo 2^64 for k < 2^23 and q < 2^55, given nq = 2^64 - q.
// With q = q1*2^32 + q0 (q odd, so q0 != 0), nq's low word is 2^32 - q0
// and its high word 2^32 - 1 - q1, which as a signed 24-bit number is
// -(q1 + 1) (q1 < 2^23): one 64-bit multiply-add forms x + k*(2^32 - q0),
// one 24-bit multiply-add the high word's - k*(q1 + 1).  (The general
// product k*nq takes two 32x32 multiplies and register moves.)
__device__ __forceinline__ uint64_t sub_small_multiple(uint64_t x, uint32_t k,
		uint64_t nq) {
	const uint64_t lo = (uint64_t) k * (uint32_t) nq + x;
	const uint32_t hi = (uint32_t) (lo >> 32) +
		(uint32_t) __mul24((int) k, (int) (uint32_t) (nq >> 32));
	return ((uint64_t) hi << 32) | (uint32_t) lo;
}

// Reference 64-bit Barrett (elemmul.comp:36-47, elemgtsub.comp:36-47):
// mu = floor(2^(bits+62) / q), shift = bits - 2, one subtraction.  Exact
// (canonical) for every 64-bit x once bits >= 3.
__device__ __forceinline__ uint64_t barrett64(uint64_t x, uint64_t q,
		uint64_t mu, uint32_t shift) {
	return csub(x - mulhi64(x >> shift, mu) * q, q);
}

// Reference 128-bit Barrett (elemmul.comp:49-60, src/numbers.c:5-28).
// (64 - shift) & 63 reproduces x86 shift-count masking at bits == 2.
__device__ __forceinline__ uint64_t barrett128(uint64_t hi, uint64_t lo,
		uint64_t q, uint64_t mu, uint32_t shift) {
	const uint64_t c = (hi << ((64 - shift) & 63)) + (lo >> shift);
	return csub(lo - mulhi64(c, mu) * q, q);
}

}  // namespace vk
