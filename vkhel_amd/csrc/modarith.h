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

// x - m if x >= m else x.
__device__ __forceinline__ uint64_t csub(uint64_t x, uint64_t m) {
	return x >= m ? x - m : x;
}

// Makes a value opaque to the optimiser (emits no instruction), so the
// arithmetic around it is kept as written instead of being re-matched into
// compare/select or borrow chains.
__device__ __forceinline__ uint64_t opaque(uint64_t v) {
	asm("" : "+v"(v));
	return v;
}

__device__ __forceinline__ uint32_t opaque32(uint32_t v) {
	asm("" : "+v"(v));
	return v;
}

// Same for a wave-uniform value kept in scalar registers.
__device__ __forceinline__ uint64_t opaque_uniform(uint64_t v) {
	asm("" : "+s"(v));
	return v;
}

// csub for x, m < 2^63 without a compare, given nm = 2^64 - m: d = x - m
// wraps to >= 2^63 exactly when x < m, and the sign word then adds m back.
// Two 64-bit adds, a shift and two ANDs; nothing writes a scalar register
// (a compare or a borrow would, and costs wait states before its reader).
__device__ __forceinline__ uint64_t csub_lt63(uint64_t x, uint64_t m,
		uint64_t nm) {
	const uint64_t d = x + nm;
	const uint32_t mask = opaque32((uint32_t) ((int32_t) (d >> 32) >> 31));
	const uint64_t back = ((uint64_t) ((uint32_t) (m >> 32) & mask) << 32) |
		((uint32_t) m & mask);
	return d + back;
}

// (a - b) mod q for a, b in [0, q).
__device__ __forceinline__ uint64_t sub_mod(uint64_t a, uint64_t b,
		uint64_t q) {
	return a >= b ? a - b : a + q - b;
}

// Shoup product: y * w - floor(y * wp / 2^64) * q lies in [0, 2q) for any
// 64-bit y when w < q and wp = floor(w * 2^64 / q) (the reference's
// nttfwdbutterfly.comp:44-48 arithmetic before its final subtraction).
__device__ __forceinline__ uint64_t shoup_lazy(uint64_t y, uint64_t w,
		uint64_t wp, uint64_t q) {
	return y * w - mulhi64(y, wp) * q;
}

__device__ __forceinline__ uint64_t shoup(uint64_t y, uint64_t w,
		uint64_t wp, uint64_t q) {
	return csub(shoup_lazy(y, w, wp, q), q);
}

// Shoup product on 32-bit halves, result in [0, 2q) like shoup_lazy:
//   Q = floor(y * wp / 2^64) exactly, R = y*w + Q*(2^64 - q) mod 2^64.
// nq = 2^64 - q.  Written out so every partial product is one
// v_mad_u64_u32 / v_mul_hi_u32 / v_mul_lo_u32 and the final subtraction
// folds into multiply-adds (no borrow chain).
__device__ __forceinline__ uint64_t shoup32(uint64_t y, uint64_t w,
		uint64_t wp, uint64_t nq) {
	const uint32_t y0 = (uint32_t) y, y1 = (uint32_t) (y >> 32);
	const uint32_t p0 = (uint32_t) wp, p1 = (uint32_t) (wp >> 32);
	const uint64_t a = __umulhi(y0, p0);
	const uint64_t b = (uint64_t) y0 * p1 + a;             // < 2^64
	const uint64_t m = (uint64_t) y1 * p0 + (uint32_t) b;  // < 2^64
	const uint64_t qh = (uint64_t) y1 * p1 + (b >> 32) + (m >> 32);
	const uint32_t w0 = (uint32_t) w, w1 = (uint32_t) (w >> 32);
	const uint32_t n0 = (uint32_t) nq, n1 = (uint32_t) (nq >> 32);
	const uint32_t q0 = (uint32_t) qh, q1 = (uint32_t) (qh >> 32);
	const uint64_t r = (uint64_t) y0 * w0 + (uint64_t) q0 * n0;
	const uint32_t rh = (uint32_t) (r >> 32) + y0 * w1 + y1 * w0 + q0 * n1 +
		q1 * n0;
	return ((uint64_t) rh << 32) | (uint32_t) r;
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
