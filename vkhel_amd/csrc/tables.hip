// On-device construction of the twiddle tables the transforms read.
//
// The reference builds roots[bitrev(i)] = w^i and the inverse table on the
// host, entry by entry (src/ntt_tables.c:17-44), and the device copy would
// then cross PCIe.  Here every entry is computed independently on the GPU:
// entry k of the reference order holds w^e with e = bitrev(k) over log2 n
// bits, which the kernel forms from the binary powers w^(2^j) passed in the
// launch arguments (at most log2 n modular products per entry).  The
// reference's repeated Barrett products (src/numbers.c:5-40, alpha = 62,
// beta = -2) are exact for operands below q < 2^61, so its chain yields
// exactly w^i mod q and both constructions give the same words.
#include "internal.h"
#include "modarith.h"
#include "numbers.h"

namespace vk {

// q^-1 mod 2^64 for odd q (Newton's iteration doubles the correct low bits
// from the 3 bits of q * q = 1 mod 8).
uint64_t inv_mod_2_64(uint64_t q) {
	uint64_t x = q;
	for (int i = 0; i < 5; i++)
		x *= 2 - q * x;
	return x;
}

// Exact a * b mod q for a, b < q: the reference's 128-bit Barrett with one
// conditional subtraction, which is canonical on that domain.
__device__ __forceinline__ uint64_t gen_mulmod(uint64_t a, uint64_t b,
		const TableGenArgs &g) {
	return barrett128(__umul64hi(a, b), a * b, g.q, g.mu, g.shift);
}

// Packed 63-bit Shoup companion of r < q (the host's shoup63_factor).
// B = floor(r * 2^64 / q) satisfies B * q = r * 2^64 - rem with
// rem = r * (2^64 mod q) mod q, so modulo 2^64 B = -rem * q^-1: the
// division is exact and needs no 128-bit divide.
__device__ __forceinline__ uint64_t gen_shoup63(uint64_t r,
		const TableGenArgs &g) {
	const uint64_t b = (0 - gen_mulmod(r, g.r64, g)) * g.qinv;
	const uint64_t wp = b >> 1;
	return wp + (wp & 0xffffffff00000000ull);
}

// One thread per reference index k in [0, n): writes {w^e, companion} and
// {w^-e, companion} at the interleaved position of k (tw_pos).  Position 0,
// which no transform reads, receives w^0.
__global__ __launch_bounds__(256) void k_gen_tables(TableGenArgs g,
		ulonglong2 *fwd, ulonglong2 *inv) {
	const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
	if (k >= (1u << g.logn))
		return;
	const uint32_t e = __brev(k) >> (32 - g.logn);
	uint64_t r = 1, ri = 1;
#pragma unroll
	for (int j = 0; j < kMaxTableLog; j++) {
		if (j < (int) g.logn && ((e >> j) & 1)) {
			r = gen_mulmod(r, g.wpow[j], g);
			ri = gen_mulmod(ri, g.wipow[j], g);
		}
	}
	const uint32_t m = k ? 1u << (31 - __clz(k)) : 0;
	const uint32_t p = k ? tw_pos(m, k - m) : 0;
	fwd[p] = make_ulonglong2(r, gen_shoup63(r, g));
	inv[p] = make_ulonglong2(ri, gen_shoup63(ri, g));
}

void table_gen_args(uint64_t n, uint64_t q, uint64_t w, TableGenArgs &g) {
	g = TableGenArgs();
	g.q = q;
	g.mu = barrett_mu(q);
	g.shift = (uint32_t) (bitlen(q) - 2);
	g.r64 = pow2_64_mod(q);
	g.qinv = inv_mod_2_64(q);
	g.logn = (uint32_t) (bitlen(n) - 1);
	uint64_t p = w, pi = invmod_euclid(w, q);
	for (uint32_t j = 0; j < g.logn; j++) {
		g.wpow[j] = p;
		g.wipow[j] = pi;
		p = mulmod_exact(p, p, q);
		pi = mulmod_exact(pi, pi, q);
	}
}

hipError_t launch_gen_tables(const TableGenArgs &g, ulonglong2 *fwd,
		ulonglong2 *inv, hipStream_t st) {
	const uint32_t n = 1u << g.logn;
	hipLaunchKernelGGL(k_gen_tables, dim3((n + 255) / 256), dim3(256), 0, st,
			g, fwd, inv);
	return hipGetLastError();
}

}  // namespace vk
