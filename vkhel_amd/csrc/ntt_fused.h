// Stage-fused NTT engine for gfx950 (device templates).  Included by the
// per-tile-size instantiation units ntt_fused_kt*.hip.
#pragma once

#include <array>
#include <utility>

#include "handoff.h"
#include "internal.h"
#include "modarith.h"

namespace vk {

typedef void (*FusedKernel)(NttTask, const NttTask *);

typedef void (*PassKernel)(NttTask, const NttTask *, uint32_t);

// Flags, the third argument of the pass kernels.  kSecondPass: the
// second pass of a direction.  kReversed: workgroup row y takes polynomial
// gridDim.y - 1 - y, so that a pass starts on the polynomials the previous
// launch wrote last, which are the ones still held in the last-level cache.
constexpr uint32_t kSecondPass = 1, kReversed = 2;

// Both passes of a two-pass size in one persistent grid (k_ntt_handoff):
// (uniform, tasks, polynomials, state words, spin limit).
typedef void (*HandoffKernel)(NttTask, const NttTask *, uint32_t, uint32_t *,
		uint32_t, uint32_t);

// One compiled (log2 n, tile size) variant, per modulus class (index
// kClassGeneral / kClassMid / kClassSmall, modulus_class()): whole-transform
// kernels for both directions (one workgroup per polynomial), the two
// passes as separate kernels (one workgroup per tile: the default for
// two-pass sizes), and launch shape.
struct FusedEntry {
	FusedKernel fwd[3], inv[3];
	// [split][class]: split 0 takes the top ntt_k1 stages in pass 1, split 1
	// the top alt_k1 (the same kernels as split 0 where there is none)
	PassKernel fwd_a[2][3], fwd_b[2][3], inv_b[2][3], inv_a[2][3];
	// the middle of a polynomial product (k_ntt_pm; pm_ok: this size and
	// tile allow it)
	PassKernel pm[3];
	bool pm_ok;
	// one-launch two-pass transforms (two-pass sizes only)
	HandoffKernel fwd_h[3], inv_h[3];
	int threads;
	int tiles;  // tiles per pass (1: one-tile size, passes unused)
	int k1[2];  // stages of pass 1 per split choice (0 for one-tile sizes)
};

constexpr int kMaxFusedLogn = 20;

constexpr int kMinFusedLogn = 3;
constexpr int kFusedEntries = kMaxFusedLogn - kMinFusedLogn + 1;


// Lazy butterflies.  Forward (Cooley-Tukey): x, y in [0, 4q) -> [0, 4q).
// Inverse (Gentleman-Sande): x, y in [0, 2q) -> [0, 2q).  Both stay
// congruent mod q to the reference's canonical butterflies
// (nttfwdbutterfly.comp:44-57, nttrevbutterfly.comp:44-57).  Values stay
// below 4q < 2^63 (q < 2^61), as shoup63 requires.  Twiddle pairs hold w
// and its packed 63-bit Shoup companion (shoup63_factor).
struct LazyK {
	uint64_t two_q, n_two_q, nq;  // 2q, 2^64 - 2q, 2^64 - q
	uint64_t four_q, n_four_q;
};

__device__ __forceinline__ LazyK lazy_consts(const NttConsts &c) {
	return LazyK{c.two_q, opaque_uniform(0 - c.two_q), opaque_uniform(0 - c.q),
		c.two_q << 1, opaque_uniform(0 - (c.two_q << 1))};
}

// 2^64 - 2^h q, h = 3..5, for the inverse's end-of-round reductions.
struct LazyInvK {
	uint64_t n[6];
};

__device__ __forceinline__ LazyInvK lazy_inv_consts(const NttConsts &c) {
	LazyInvK k;
	k.n[0] = k.n[1] = k.n[2] = 0;
#pragma unroll
	for (int h = 3; h <= 5; h++)
		k.n[h] = opaque_uniform(0 - (c.q << h));
	return k;
}

__device__ __forceinline__ void bf_fwd(uint64_t &x, uint64_t &y, ulonglong2 w,
		const LazyK &k) {
	const uint64_t xr = csub_n(x, k.two_q, k.n_two_q);
	const uint64_t t = shoup63(y, w.x, w.y, k.nq);
	x = xr + t;
	y = xr + k.two_q - t;
}

// Forward butterfly without the input reduction (q < kNoReduceQ), its
// product from shoup63_lazy (below 4q): each stage grows the bound by less
// than 4q.
__device__ __forceinline__ void bf_fwd_nr(uint64_t &x, uint64_t &y,
		ulonglong2 w, const LazyK &k) {
	const uint64_t xr = x;
	const uint64_t t = shoup63_lazy(y, w.x, w.y, k.nq);
	x = xr + t;
	y = xr + k.four_q - t;
}

__device__ __forceinline__ void bf_inv(uint64_t &x, uint64_t &y, ulonglong2 w,
		const LazyK &k) {
	const uint64_t s = csub_n(x + y, k.two_q, k.n_two_q);
	const uint64_t d = x + k.two_q - y;
	x = s;
	y = shoup63(d, w.x, w.y, k.nq);
}

// First inverse stage: the inverse's inputs are below q (above q the
// reference's butterfly, nttrevbutterfly.comp:44-48, is not meaningful), so
// the sum is already below 2q and needs no reduction.  LAZY: the mid
// class's product, below 4q.
template <bool LAZY>
__device__ __forceinline__ void bf_inv_first(uint64_t &x, uint64_t &y,
		ulonglong2 w, const LazyK &k) {
	const uint64_t s = x + y;
	const uint64_t d = x + k.two_q - y;
	x = s;
	y = LAZY ? shoup63_lazy(d, w.x, w.y, k.nq) : shoup63(d, w.x, w.y, k.nq);
}

// Mid-class forward butterfly (q < kMidQ): x, y in [0, 8q) -> [0, 8q),
// with the input reduced below 4q and a shoup63_lazy product (below 4q):
// 8q < 2^63 keeps the product's input in range.
__device__ __forceinline__ void bf_fwd_mid(uint64_t &x, uint64_t &y,
		ulonglong2 w, const LazyK &k) {
	const uint64_t xr = csub_n(x, k.four_q, k.n_four_q);
	const uint64_t t = shoup63_lazy(y, w.x, w.y, k.nq);
	x = xr + t;
	y = xr + k.four_q - t;
}

// Mid-class inverse butterfly: x, y in [0, 4q) -> [0, 4q).
__device__ __forceinline__ void bf_inv_mid(uint64_t &x, uint64_t &y,
		ulonglong2 w, const LazyK &k) {
	const uint64_t s = csub_n(x + y, k.four_q, k.n_four_q);
	const uint64_t d = x + k.four_q - y;
	x = s;
	y = shoup63_lazy(d, w.x, w.y, k.nq);
}

// First forward stage when both inputs are already below q: the lazy
// butterfly without its input reduction (for the small and mid classes the
// product of shoup63_lazy, below 4q).  The reference's first stage gives
// the canonical forms of these same residues.
template <int MC>
__device__ __forceinline__ void bf_fwd_first_canon(uint64_t &x, uint64_t &y,
		ulonglong2 w, const LazyK &k) {
	if constexpr (MC != kClassGeneral) {
		bf_fwd_nr(x, y, w, k);
	} else {
		const uint64_t t = shoup63(y, w.x, w.y, k.nq);
		const uint64_t xr = x;
		x = xr + t;
		y = xr + k.two_q - t;
	}
}

// First forward stage on arbitrary 64-bit words, exactly as the reference:
// (X + WY) mod 2^64 mod q and (X - WY) mod q with WY = W*Y mod q canonical
// (nttfwdbutterfly.comp:44-57 on unreduced input words).
__device__ __forceinline__ void bf_fwd_first(uint64_t &x, uint64_t &y,
		ulonglong2 w, const NttConsts &c) {
	const uint64_t X = x;
	const uint64_t yr = barrett64(y, c.q, c.mu, c.bshift);
	const uint64_t wy = csub(shoup63(yr, w.x, w.y, 0 - c.q), c.q);
	const uint64_t xr = barrett64(X, c.q, c.mu, c.bshift);
	uint64_t ox = csub(xr + wy, c.q);
	if (X + wy < X)
		ox = sub_mod(ox, c.r64, c.q);
	x = ox;
	y = sub_mod(xr, wy, c.q);
}

// Last inverse stage (m = 1) with the reference's final multiplication by
// n^-1 (src/vector.c:633-638) folded in; canonical outputs.  ninv_p and
// w1n_p are 63-bit Shoup companions.
__device__ __forceinline__ void bf_inv_last(uint64_t &x, uint64_t &y,
		const NttConsts &c, const LazyK &k) {
	const uint64_t s = csub_n(x + y, k.two_q, k.n_two_q);
	const uint64_t d = x + k.two_q - y;
	x = csub_n(shoup63(s, c.ninv, c.ninv_p, k.nq), c.q, k.nq);
	y = csub_n(shoup63(d, c.w1n, c.w1n_p, k.nq), c.q, k.nq);
}

// Inverse epilogue: elemmod(x, p, q) of a canonical result x
// (src/vector.c:342-386): for p = 2 the reference's elemmodbytwo with bound
// q/2 (elemmodbytwo.comp:19-27), else elemgtsub(x, q/2, q, p)
// (elemgtsub.comp:49-67), the centered lift of x reduced mod p.
__device__ __forceinline__ uint64_t post_center(uint64_t x,
		const NttConsts &c) {
	const uint64_t half = c.q >> 1;
	if (c.post_p == 2) {
		const uint64_t u = x & 1;
		return x > half ? 1 - u : u;
	}
	const uint64_t r = barrett64(x, c.post_p, c.post_mu, c.post_shift);
	if (x > half)
		return csub(r + c.post_p - barrett64(c.q, c.post_p, c.post_mu,
				c.post_shift), c.post_p);
	return r;
}

// The forward transform's optional prologue and epilogue (NttConsts
// lift_* and scale_*), used by NTT-form RNS rescaling: the centered lift
// of the dropped limb's words into Z_q as they load (non-canonical words
// take the first stage's exact reduction), and (a - y) * w mod q as the
// canonical results store.
__device__ __forceinline__ uint64_t lift_word(uint64_t x, const NttConsts &c) {
	return x > c.lift_bound ? x + c.lift_add : x;
}

__device__ __forceinline__ uint64_t scale_word(uint64_t a, uint64_t y,
		const NttConsts &c) {
	const uint64_t d = a >= y ? a - y : a + c.q - y;
	return csub(shoup63(d, c.scale_w, c.scale_wp, 0 - c.q), c.q);
}

// Inverse butterflies for q < kNoReduceQ with the sums left unreduced
// inside a round.  x and y lie below bx*q and by*q; C = by*q keeps the
// difference non-negative, and the product input x + C - y stays below
// (bx + by) q <= 64q < 2^63.  The sum's bound is bx + by, the product's
// (shoup63_lazy) 4q.
__device__ __forceinline__ void bf_inv_nr(uint64_t &x, uint64_t &y,
		ulonglong2 w, uint64_t C, const LazyK &k) {
	const uint64_t s = x + y;
	const uint64_t d = x + C - y;
	x = s;
	y = shoup63_lazy(d, w.x, w.y, k.nq);
}

__device__ __forceinline__ void bf_inv_last_nr(uint64_t &x, uint64_t &y,
		uint64_t C, const NttConsts &c, const LazyK &k) {
	const uint64_t s = x + y;
	const uint64_t d = x + C - y;
	x = csub_n(shoup63(s, c.ninv, c.ninv_p, k.nq), c.q, k.nq);
	y = csub_n(shoup63(d, c.w1n, c.w1n_p, k.nq), c.q, k.nq);
}

// Bound, in units of q, of register e of an inverse round after its stages
// d <= upto (stages whose bit is clear in `active` are absent), when every
// register entered the round below bin*q: a product output is below 4q, a
// sum below the sum of its operands' bounds.  Both operands of a sum have
// the same lower register bits, hence the same history and bound, so a sum
// doubles it.  All bounds are powers of two.  (Plain loops, not recursion,
// so that calls with constant arguments fold.)
__host__ __device__ constexpr int inv_bound(int bin, int active, int upto,
		int e) {
	int b = bin;
	for (int s = 0; s <= upto; s++)
		if ((active >> s) & 1)
			b = ((e >> s) & 1) ? 4 : 2 * b;
	return b;
}

__host__ __device__ constexpr int log2_exact(int x) {
	int l = 0;
	while ((1 << l) < x)
		l++;
	return l;
}

// Tiles hold 2^KT words of one polynomial, 8 per thread.  A pass runs a
// contiguous range of butterfly stages on every tile of the polynomial:
//   PASS 0: one tile = the whole polynomial (n <= 2^TB), all stages;
//   PASS 1: 2^k1 rows x 2^cb columns (row stride 2^k2), the top k1 stages;
//   PASS 2: 2^rb rows x 2^k2 contiguous words, the low k2 stages.
// Everything below is fixed at compile time by (LOGN, PASS, TB), TB being
// the tile size (log2 words) of the two-pass path.
// Pass 1 takes the top log2 n - 9 stages, so that pass 2 works on rows of
// 512 words (one wave per row, three full rounds), when pass 1's tiles still
// have column runs of at least 16 words (128 bytes); otherwise the top
// floor(log2 n / 2) stages.  None for one-tile sizes.
constexpr int ntt_k1(int logn, int tile_bits) {
	return logn <= tile_bits ? 0
		: (logn - 9 >= 1 && tile_bits - (logn - 9) >= 4 && logn - 9 < logn / 2)
			? logn - 9 : logn / 2;
}

// A stage split is usable on tiles of 2^tb words when pass 1's tiles keep
// column runs of at least 16 words and a pass-2 row fits in one tile.
constexpr bool split_ok(int logn, int tb, int k1) {
	return k1 >= 1 && tb - k1 >= 4 && logn - k1 <= tb;
}

// The alternative split the tuner may pick (the `split` knob, 1): one
// column stage more than ntt_k1, else one fewer; -1 where neither fits.
constexpr int alt_k1(int logn, int tb) {
	return logn <= tb ? -1
		: split_ok(logn, tb, ntt_k1(logn, tb) + 1) ? ntt_k1(logn, tb) + 1
		: split_ok(logn, tb, ntt_k1(logn, tb) - 1) ? ntt_k1(logn, tb) - 1 : -1;
}

// Stage split of split choice s (0: ntt_k1, 1: alt_k1, or ntt_k1 where no
// alternative exists).
constexpr int split_k1(int logn, int tb, int s) {
	return s == 1 && alt_k1(logn, tb) > 0 ? alt_k1(logn, tb) : ntt_k1(logn, tb);
}

template <int LOGN, int PASS, int TB, int SPLIT = 0>
struct Geo {
	static constexpr int P = PASS;
	static constexpr int K1 = split_k1(LOGN, TB, SPLIT);
	static constexpr int K2 = LOGN - K1;
	static constexpr int KT = PASS == 0 ? LOGN : TB;
	static constexpr int CB = TB - K1;  // pass 1 columns per tile
	static constexpr int LO = PASS == 1 ? CB : 0;
	static constexpr int HI = PASS == 0 ? LOGN : PASS == 1 ? TB : K2;
	// low local bits contiguous in memory, and the stride of the rest
	static constexpr int CBITS = PASS == 1 ? CB : KT;
	static constexpr int JSHIFT = PASS == 1 ? K2 : 0;
	// global stage bit = local bit + SHIFT
	static constexpr int SHIFT = PASS == 1 ? K2 - CB : 0;
	static constexpr int ROUNDS = (HI - LO + 2) / 3;
	// threads' worth of work (8 words each), and the physical threads of
	// a workgroup that run it (a workgroup holds at most 1024)
	static constexpr int THREADS = 1 << (KT - 3);
	static constexpr int PHYS = THREADS > 1024 ? 1024 : THREADS;
	static constexpr int VT = THREADS / PHYS;
	static constexpr int NTILES = PASS == 0 ? 1 : 1 << (LOGN - TB);
};

// Rounds of a pass: 3-stage groups over local bits [LO, HI), forward
// top-down, inverse bottom-up.  Bit 9 is where rounds stop being
// wave-local (the exchange rule below), so the groups are cut at bit 9 when
// that needs no extra round: then at most the rounds touching bits >= 9
// exchange across waves.
constexpr int ceil3(int x) {
	return (x + 2) / 3;
}

template <class G>
struct RoundPlan {
	static constexpr int M = G::LO > 9 ? G::LO : G::HI < 9 ? G::HI : 9;
	static constexpr int NU = ceil3(G::HI - M);  // groups above the cut
	static constexpr int NL = ceil3(M - G::LO);  // groups below it
	static constexpr bool CUT = NU + NL == ceil3(G::HI - G::LO);
	static constexpr int ROUNDS = ceil3(G::HI - G::LO);
};

// Field position and stage range of round R.
template <class G, bool INV, int R>
struct Round {
	using P = RoundPlan<G>;
	// plain partition
	static constexpr int BHI_F = G::HI - 3 * R;
	static constexpr int BLO_F = BHI_F - 3 >= G::LO ? BHI_F - 3 : G::LO;
	static constexpr int BLO_I = G::LO + 3 * R;
	static constexpr int BHI_I = BLO_I + 3 <= G::HI ? BLO_I + 3 : G::HI;
	// partition cut at bit P::M
	static constexpr int CF_HI = R < P::NU ? G::HI - 3 * R : P::M - 3 * (R - P::NU);
	static constexpr int CF_FLOOR = R < P::NU ? P::M : G::LO;
	static constexpr int CF_LO = CF_HI - 3 >= CF_FLOOR ? CF_HI - 3 : CF_FLOOR;
	static constexpr int CI_LO = R < P::NL ? G::LO + 3 * R : P::M + 3 * (R - P::NL);
	static constexpr int CI_CEIL = R < P::NL ? P::M : G::HI;
	static constexpr int CI_HI = CI_LO + 3 <= CI_CEIL ? CI_LO + 3 : CI_CEIL;
	static constexpr int BLO = P::CUT ? (INV ? CI_LO : CF_LO) : (INV ? BLO_I : BLO_F);
	static constexpr int BHI = P::CUT ? (INV ? CI_HI : CF_HI) : (INV ? BHI_I : BHI_F);
	// The 3-bit field holding the round's stages: at BLO (or as high as the
	// tile allows), or at bit 6 when a short round can sit there entirely
	// below bit 9 and so stay inside each wave.
	static constexpr int F0 = BLO + 3 <= G::KT ? BLO : G::KT - 3;
	static constexpr int F = F0 > 6 && BHI - 3 <= 6 ? 6 : F0;
	// Lanes of a wave are contiguous in memory for 2^RUN words; direct
	// global access is used when that covers a 128-byte line.
	static constexpr int RUN = F < G::CBITS ? F : G::CBITS;
	static constexpr bool DIRECT = RUN >= 4;
};

// Global-memory accesses through address-space-1 pointers, so they compile
// to global_load/store (task pointers are generic; flat accesses would also
// hold the LDS wait counter).
typedef __attribute__((address_space(1))) uint64_t gu64;
typedef __attribute__((address_space(1))) const uint64_t gcu64;
typedef unsigned long long u64x2 __attribute__((ext_vector_type(2)));
typedef __attribute__((address_space(1))) const u64x2 gcu128;

// Word indices within one polynomial stay below 2^28, so the byte offset
// is formed in 32 bits: the compiler then addresses with the wave-uniform
// base and a 32-bit per-lane offset instead of a 64-bit address per lane.
typedef __attribute__((address_space(1))) const char gcchar;
typedef __attribute__((address_space(1))) char gchar;

__device__ __forceinline__ uint64_t gload(const uint64_t *p, uint32_t i) {
	return *(gcu64 *) ((gcchar *) p + (i << 3));
}

// A transform's input words, read once.  One-tile transforms load them
// with the nontemporal hint (measured on MI355X, 2^25-word batches: 3-6%
// faster at n = 2^10, 0-2% at 2^12 and 2^13); two-pass transforms keep
// ordinary loads (with the hint there, 2^16 measured up to 4% slower: the
// second pass reads the first pass's output from the last-level cache,
// and nontemporal stores were slower still).
template <int PASS>
__device__ __forceinline__ uint64_t gload_in(const uint64_t *p, uint32_t i) {
	if constexpr (PASS == 0)
		return __builtin_nontemporal_load((gcu64 *) ((gcchar *) p + (i << 3)));
	else
		return gload(p, i);
}

__device__ __forceinline__ void gstore(uint64_t *p, uint32_t i, uint64_t v) {
	*(gu64 *) ((gchar *) p + (i << 3)) = v;
}

// Twiddle pair i, addressed like gload: uniform base plus a 32-bit byte
// offset (the fused kernels serve n <= 2^20, so i << 4 < 2^24).
__device__ __forceinline__ ulonglong2 gload_tw(const ulonglong2 *p,
		uint32_t i) {
	const u64x2 v = *(gcu128 *) ((gcchar *) p + (i << 4));
	return make_ulonglong2(v.x, v.y);
}

// One padding word per 8 keeps the radix-8 access patterns conflict free.
__host__ __device__ __forceinline__ uint32_t pad(uint32_t l) {
	return l + (l >> 3);
}

template <class G>
__host__ __device__ __forceinline__ uint32_t tile_addr(uint32_t l,
		uint32_t jbase) {
	if (G::CBITS >= G::KT)
		return jbase + l;
	return ((l >> G::CBITS) << G::JSHIFT) + (l & ((1u << G::CBITS) - 1)) +
		jbase;
}

// The tile-local index of a round's element e is l0 | (e << F), and
// round_base leaves bits F..F+2 of l0 clear, so the two bit fields are
// disjoint.  Both pad() and tile_addr() act on disjoint fields additively
// (shifts and masks with no carries between the fields), so an element's
// LDS index and global offset are the base's plus a compile-time constant;
// spelled out, they compile to immediate instruction offsets instead of
// per-element address arithmetic.
template <class G>
__host__ __device__ constexpr uint32_t tile_off(uint32_t l) {
	return G::CBITS >= G::KT ? l :
		((l >> G::CBITS) << G::JSHIFT) + (l & ((1u << G::CBITS) - 1));
}

// Local index of element 0 of thread t in a round whose 3-bit field sits at
// local bit f; element e is at l0 + (e << f).
template <int F>
__device__ __forceinline__ uint32_t round_base(uint32_t t) {
	return ((t >> F) << (F + 3)) | (t & ((1u << F) - 1));
}

// Twiddle slots of one round: bit 2 of the field uses slot 0, bit 1 slots
// 1-2, bit 0 slots 3-6.
__host__ __device__ constexpr int tw_slot(int d, int gi) {
	return (d == 2 ? 0 : d == 1 ? 1 : 3) + gi;
}

// Loads every twiddle pair round R uses into w (slots as tw_slot).
template <class G, bool INV, int R, int LOGN>
__device__ __forceinline__ void load_round_tw(ulonglong2 w[7], uint32_t t,
		uint32_t gbase, const ulonglong2 *tw) {
	using RD = Round<G, INV, R>;
	constexpr int F = RD::F;
	const uint32_t l0 = round_base<F>(t);
#pragma unroll
	for (int d = 2; d >= 0; d--) {
		const int lb = F + d;
		if (lb < RD::BLO || lb >= RD::BHI)
			continue;
		const int b = lb + G::SHIFT;
		// Group index of the thread's first butterfly at this stage, in a
		// stage block of M entries.  It is a multiple of the stage's
		// butterfly-group count 4 >> d (the round field's bits of l0 are
		// zero, gbase is a multiple of the tile), so the thread's entries sit
		// at one interleaved position plus multiples of M/4.
		const uint32_t M = 1u << (LOGN - 1 - b);
		const uint32_t STRIDE = M < 4 ? 1 : M >> 2;
		uint32_t p0 = tw_pos(M, (gbase + l0) >> (lb + 1));
		// With the field at bit 6 or above, the lanes of a wave differ only
		// in index bits 0-5 (round_base keeps thread bits 0-5 there), which
		// the shift drops: the entries are the same for the whole wave, and a
		// wave-uniform index lets them load through the scalar cache.
		if constexpr (F >= 6 && G::THREADS >= 64)
			p0 = __builtin_amdgcn_readfirstlane(p0);
#pragma unroll
		for (int gi = 0; gi < (4 >> d); gi++)
			w[tw_slot(d, gi)] = gload_tw(tw, p0 + gi * STRIDE);
	}
}

// LDS exchanges.  In a round whose 3-bit field sits at local bit F <= 6, the
// 64 lanes of wave w hold exactly the 512 indices [512w, 512w + 512) (lane
// bits fill index bits below 9 around the field, wave bits are index bits 9
// and up).  An exchange between two such rounds therefore stays inside each
// wave and needs no workgroup barrier: LDS operations of one wave complete in
// program order, so a compiler-level wave fence suffices.  Each thread only
// rewrites the LDS words it read in the same round, so no exchange can
// overwrite words another wave has still to read.
constexpr int kWaveLocalMaxField = 6;

__device__ __forceinline__ void wave_sync() {
	__builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
	__builtin_amdgcn_wave_barrier();
	__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <bool CROSS>
__device__ __forceinline__ void exchange_sync() {
	if constexpr (CROSS)
		__syncthreads();
	else
		wave_sync();
}

// Lane-contiguous (staged) global access of a tile: thread t moves words
// base(t) + 64k, k = 0..7, with base(t) = 512*(t/64) + t%64, so each wave
// touches only its own 512-word block and each instruction 64 consecutive
// words.  base(t) and 64k occupy disjoint bits (base's bits 6-8 are zero),
// and both the padding and the tile's column split act on disjoint bit
// fields additively, so the padded LDS index and the global index of word k
// are the base's plus a constant per k; the offsets are spelled out because
// the compiler cannot see that the adds do not carry.
template <class G>
struct Staged {
	static constexpr bool WAVE = G::THREADS >= 64;
	static constexpr uint32_t PAD_STEP = 72;  // pad(x + 64) - pad(x)
	// tile_addr of 64k without the tile base
	static constexpr uint32_t g_off(int k) {
		const uint32_t x = 64u * (uint32_t) k;
		return G::CBITS >= G::KT ? x :
			((x >> G::CBITS) << G::JSHIFT) + (x & ((1u << G::CBITS) - 1));
	}

	__host__ __device__ static uint32_t base(uint32_t t) {
		return ((t >> 6) << 9) | (t & 63);
	}
	// padded LDS index and global index of word k of thread t
	__host__ __device__ static uint32_t lds_at(uint32_t t, int k) {
		if constexpr (WAVE)
			return pad(base(t)) + k * PAD_STEP;
		else
			return pad(t + k * G::THREADS);
	}
	__host__ __device__ static uint32_t global_at(uint32_t t, int k,
			uint32_t jbase) {
		if constexpr (WAVE)
			return tile_addr<G>(base(t), jbase) + g_off(k);
		else
			return tile_addr<G>(t + k * G::THREADS, jbase);
	}
};

// The pointwise product of a polynomial product, fused into the inverse
// transform's first loads: (x * y) mod q for canonical spectra x, y, by the
// reference's 128-bit Barrett reduction (elemmul.comp:49-60), canonical.
__device__ __forceinline__ uint64_t spectra_product(uint64_t x, uint64_t y,
		const NttConsts &c) {
	return barrett128(mulhi64(x, y), x * y, c.q, c.mu, c.bshift);
}

template <class G, bool INV, int R>
constexpr bool round_wave_local() {
	return Round<G, INV, R>::F <= kWaveLocalMaxField;
}

// TRAIL: end with a workgroup barrier, for callers that reuse the LDS tile.
// MC: modulus class (modulus_class(q)); the small class's forward
// butterflies run without input reduction.
// IO flags of a sequence of rounds: the first round's operands are
// already in v in its own layout (no loads), and the last round leaves its
// results in v (no stores).
// kIoStreamIn: the first round's global loads carry the nontemporal hint
// whatever the pass.
enum { kIoPreloaded = 1, kIoKeep = 2, kIoStreamIn = 4 };

// Tiles of 2^14 words have 2048 threads' worth of work (8 words each) but
// a workgroup holds at most 1024: physical thread tp then runs the round's
// work of virtual threads tp and tp + 1024 one after the other (G::VT).
// A physical wave always runs the same two virtual waves, each of which
// touches only its own 512-word block in wave-local rounds, so the
// exchange rules above hold unchanged; the synchronisations of a round
// come after all its virtual threads.
template <class G, bool INV, int R, int LOGN, bool TRAIL, int MC, int IO = 0>
__device__ __forceinline__ void run_round(uint64_t v[8], ulonglong2 w[7],
		uint32_t tp, const uint64_t *src, const uint64_t *src2, uint64_t *dst,
		uint64_t *lds, uint32_t jbase, uint32_t gbase, const ulonglong2 *tw,
		const NttConsts &c, bool final) {
	using RD = Round<G, INV, R>;
	constexpr int F = RD::F;
	constexpr bool FIRST_ROUND = R == 0;
	constexpr bool LAST_ROUND = R == G::ROUNDS - 1;
	constexpr bool LOCAL = round_wave_local<G, INV, R>();
	constexpr bool PRE = FIRST_ROUND && (IO & kIoPreloaded);
	constexpr bool KEEP = LAST_ROUND && (IO & kIoKeep);
	// pass whose load flavour the first round's global loads take
	constexpr int LP = (IO & kIoStreamIn) ? 0 : G::P;
	constexpr uint32_t PHYS = G::PHYS;
	constexpr int VT = G::VT;
	static_assert(VT == 1 || !(IO & (kIoPreloaded | kIoKeep)),
			"values kept in registers across rounds need one virtual thread");
	// workgroups hold exactly G::PHYS threads, or 64 when a one-tile size
	// needs fewer
	const bool active = G::THREADS >= 64 || tp < (uint32_t) G::THREADS;

	// ---- operands ----
	// the first loads of an inverse transform may form a pointwise product
	constexpr bool PRODUCT = INV && FIRST_ROUND && G::P != 1;
	// the forward's optional prologue, on the loads of its input
	constexpr bool LIFT = !INV && FIRST_ROUND && G::P != 2 && !PRE;
	if (FIRST_ROUND && !RD::DIRECT && !PRE) {
		// lane-contiguous global loads staged through LDS
		if (active) {
#pragma unroll
			for (int i = 0; i < VT; i++) {
				const uint32_t t = tp + i * PHYS;
				if (PRODUCT && src2) {
					// four products in flight: all eight would not fit in
					// eight waves' registers
#pragma unroll 4
					for (int k = 0; k < 8; k++) {
						const uint32_t gi = Staged<G>::global_at(t, k, jbase);
						lds[Staged<G>::lds_at(t, k)] = spectra_product(
								gload_in<LP>(src, gi), gload_in<LP>(src2, gi), c);
					}
				} else if (LIFT && c.lift_add) {
#pragma unroll
					for (int k = 0; k < 8; k++)
						lds[Staged<G>::lds_at(t, k)] = lift_word(gload_in<LP>(src,
								Staged<G>::global_at(t, k, jbase)), c);
				} else {
#pragma unroll
					for (int k = 0; k < 8; k++)
						lds[Staged<G>::lds_at(t, k)] = gload_in<LP>(src,
								Staged<G>::global_at(t, k, jbase));
				}
			}
		}
		exchange_sync<!LOCAL>();
	}
#pragma unroll
	for (int i = 0; i < VT; i++) {
		const uint32_t t = tp + i * PHYS;
		const uint32_t l0 = round_base<F>(t);
		if (active) {
			if constexpr (PRE) {
				// operands already in v
			} else if (FIRST_ROUND && RD::DIRECT && PRODUCT && src2) {
				const uint32_t g0 = tile_addr<G>(l0, jbase);
#pragma unroll
				for (int e = 0; e < 8; e++) {
					const uint32_t gi = g0 + tile_off<G>((uint32_t) e << F);
					v[e] = spectra_product(gload_in<LP>(src, gi),
							gload_in<LP>(src2, gi), c);
				}
			} else if (FIRST_ROUND && RD::DIRECT && LIFT && c.lift_add) {
				const uint32_t g0 = tile_addr<G>(l0, jbase);
#pragma unroll
				for (int e = 0; e < 8; e++)
					v[e] = lift_word(gload_in<LP>(src, g0 +
							tile_off<G>((uint32_t) e << F)), c);
			} else if (FIRST_ROUND && RD::DIRECT) {
				const uint32_t g0 = tile_addr<G>(l0, jbase);
#pragma unroll
				for (int e = 0; e < 8; e++)
					v[e] = gload_in<LP>(src, g0 + tile_off<G>((uint32_t) e << F));
			} else {
				const uint32_t p0 = pad(l0);
#pragma unroll
				for (int e = 0; e < 8; e++)
					v[e] = lds[p0 + pad((uint32_t) e << F)];
			}
			// every twiddle of the round, issued before any arithmetic
			// (loading the next round's before the barrier instead costs
			// registers beyond eight waves per SIMD and measured no faster)
			load_round_tw<G, INV, R, LOGN>(w, t, gbase, tw);
		}

		// ---- butterflies ----
		if (active) {
			const LazyK lk = lazy_consts(c);
			// inverse with deferred sum reductions (q < kNoReduceQ): stages
			// present in this round, and the bound its inputs enter with
			// (below q at the transform's first round, below 8q after any
			// other)
			constexpr bool NR = MC == kClassSmall;
			constexpr bool MID = MC == kClassMid;
			constexpr bool LAZY_INV = INV && NR;
			constexpr int ACTIVE = (F >= RD::BLO && F < RD::BHI ? 1 : 0) |
				(F + 1 >= RD::BLO && F + 1 < RD::BHI ? 2 : 0) |
				(F + 2 >= RD::BLO && F + 2 < RD::BHI ? 4 : 0);
			constexpr int BIN = FIRST_ROUND && G::P != 1 ? 1 : 8;
			constexpr int TOP = LOGN - 1 - G::SHIFT;  // local bit of the last stage
			constexpr bool HAS_LAST = INV && TOP >= F && TOP <= F + 2 &&
				TOP >= RD::BLO && TOP < RD::BHI;
#pragma unroll
			for (int k = 0; k < 3; k++) {
				const int d = INV ? k : 2 - k;
				const int lb = F + d;
				if (lb < RD::BLO || lb >= RD::BHI)
					continue;
				const int b = lb + G::SHIFT;
				if (!INV && b == LOGN - 1) {
					// Inputs below q (the usual case) take the lazy butterfly;
					// a wave holding any larger word follows the reference's
					// exact reduction.
					bool canon = true;
#pragma unroll
					for (int e = 0; e < 8; e++)
						canon &= v[e] < c.q;
					if (__all(canon)) {
#pragma unroll
						for (int e = 0; e < 4; e++)
							bf_fwd_first_canon<MC>(v[e], v[e + 4], w[0], lk);
					} else {
#pragma unroll
						for (int e = 0; e < 4; e++)
							bf_fwd_first(v[e], v[e + 4], w[0], c);
					}
				} else if (INV && b == LOGN - 1) {
#pragma unroll
					for (int e = 0; e < 4; e++) {
						if constexpr (LAZY_INV)
							bf_inv_last_nr(v[e], v[e + 4], c.q << log2_exact(
									inv_bound(BIN, ACTIVE, d - 1, e + 4)), c, lk);
						else if constexpr (MID)
							bf_inv_last_nr(v[e], v[e + 4], lk.four_q, c, lk);
						else
							bf_inv_last(v[e], v[e + 4], c, lk);
					}
				} else {
#pragma unroll
					for (int gi = 0; gi < (4 >> d); gi++) {
#pragma unroll
						for (int li = 0; li < (1 << d); li++) {
							const int e = (gi << (d + 1)) | li;
							if (LAZY_INV)
								bf_inv_nr(v[e], v[e | (1 << d)], w[tw_slot(d, gi)],
										c.q << log2_exact(inv_bound(BIN, ACTIVE, d - 1,
												e | (1 << d))), lk);
							else if (INV && b == 0)
								bf_inv_first<MID>(v[e], v[e | (1 << d)], w[tw_slot(d, gi)],
										lk);
							else if (INV && MID)
								bf_inv_mid(v[e], v[e | (1 << d)], w[tw_slot(d, gi)], lk);
							else if (INV)
								bf_inv(v[e], v[e | (1 << d)], w[tw_slot(d, gi)], lk);
							else if (NR)
								bf_fwd_nr(v[e], v[e | (1 << d)], w[tw_slot(d, gi)], lk);
							else if (MID)
								bf_fwd_mid(v[e], v[e | (1 << d)], w[tw_slot(d, gi)], lk);
							else
								bf_fwd(v[e], v[e | (1 << d)], w[tw_slot(d, gi)], lk);
						}
					}
				}
			}
			if constexpr (LAZY_INV && !HAS_LAST) {
				// back below 8q for the next round (whose registers each mix
				// values of every register here): halve each bound by
				// conditional subtractions of (bound / 2) q, ..., 8q.  Within
				// a round the bounds then reach 64q (sums) and the products'
				// inputs stay below 64q < 2^63.
				const LazyInvK ik = lazy_inv_consts(c);
#pragma unroll
				for (int e = 0; e < 8; e++) {
#pragma unroll
					for (int h = 5; h >= 3; h--) {
						if ((1 << h) < inv_bound(BIN, ACTIVE, 2, e))
							v[e] = csub_n(v[e], c.q << h, ik.n[h]);
					}
				}
			}
			if constexpr (HAS_LAST) {
				// the transform's results, canonical: the optional epilogue
				if (c.post_p) {
#pragma unroll
					for (int e = 0; e < 8; e++)
						v[e] = post_center(v[e], c);
				}
			}
			if (!INV && LAST_ROUND && final) {
				if (NR) {
					// from below 81q: a quotient estimate qe <= 80, off by at
					// most one (small_reduce_consts), its multiple of q taken
					// off with a 24-bit high word (q < kNoReduceQ = 2^55), then
					// one subtraction
#pragma unroll
					for (int e = 0; e < 8; e++) {
						const uint32_t qe = __umulhi((uint32_t) (v[e] >> c.red_shift),
								c.red_m);
						v[e] = csub_n(sub_small_multiple(v[e], qe, lk.nq), c.q, lk.nq);
					}
				} else if (MID) {
					// from below 8q: the same estimate, a full product by q
#pragma unroll
					for (int e = 0; e < 8; e++) {
						const uint32_t qe = __umulhi((uint32_t) (v[e] >> c.red_shift),
								c.red_m);
						v[e] = csub_n(v[e] + (uint64_t) qe * lk.nq, c.q, lk.nq);
					}
				} else {
#pragma unroll
					for (int e = 0; e < 8; e++)
						v[e] = csub_n(csub_n(v[e], lk.two_q, lk.n_two_q), c.q, lk.nq);
				}
			}
		}

		// ---- results of this virtual thread ----
		if constexpr (KEEP) {
			// results stay in v
		} else if constexpr (LAST_ROUND && RD::DIRECT) {
			// (a forward transform's final round holds its field at bit 0
			// and always stores through LDS, where the epilogue is applied)
			static_assert(INV || G::P == 1, "forward final round stores directly");
			if (active) {
				// store addresses formed here, not held from the kernel's start
				const uint32_t g0 = tile_addr<G>(round_base<F>(opaque32_here(t)),
						jbase);
#pragma unroll
				for (int e = 0; e < 8; e++)
					gstore(dst, g0 + tile_off<G>((uint32_t) e << F), v[e]);
			}
		} else {
			// to LDS: the next round's exchange, or the staged store below
			if (active) {
				const uint32_t p0 = pad(l0);
#pragma unroll
				for (int e = 0; e < 8; e++)
					lds[p0 + pad((uint32_t) e << F)] = v[e];
			}
		}
	}

	// ---- results: synchronisation and staged stores ----
	if constexpr (KEEP) {
		// results stay in v
	} else if constexpr (LAST_ROUND && RD::DIRECT) {
		if constexpr (TRAIL)
			__syncthreads();
	} else if constexpr (LAST_ROUND) {
		// lane-contiguous global stores staged through LDS
		exchange_sync<!LOCAL>();
		if (active) {
#pragma unroll
			for (int i = 0; i < VT; i++) {
				const uint32_t ts = opaque32_here(tp + i * PHYS);
				if (!INV && final && c.scale_w) {
#pragma unroll
					for (int k = 0; k < 8; k++) {
						const uint32_t gi = Staged<G>::global_at(ts, k, jbase);
						gstore(dst, gi, scale_word(gload(src2, gi),
								lds[Staged<G>::lds_at(ts, k)], c));
					}
				} else {
#pragma unroll
					for (int k = 0; k < 8; k++)
						gstore(dst, Staged<G>::global_at(ts, k, jbase),
								lds[Staged<G>::lds_at(ts, k)]);
				}
			}
		}
		if constexpr (TRAIL)
			__syncthreads();
	} else {
		exchange_sync<!(LOCAL && round_wave_local<G, INV, R + 1>())>();
	}
}

template <class G, bool INV, int LOGN, bool TRAIL, int MC, int IO, int... R>
__device__ __forceinline__ void run_rounds_v(std::integer_sequence<int, R...>,
		uint64_t v[8], uint32_t t, const uint64_t *src, const uint64_t *src2,
		uint64_t *dst, uint64_t *lds, uint32_t jbase, uint32_t gbase,
		const ulonglong2 *tw, const NttConsts &c, bool final) {
	ulonglong2 w[7];
	(run_round<G, INV, R, LOGN, TRAIL, MC, IO>(v, w, t, src, src2, dst, lds,
			jbase, gbase, tw, c, final), ...);
}

template <class G, bool INV, int LOGN, bool TRAIL, int MC, int... R>
__device__ __forceinline__ void run_rounds(std::integer_sequence<int, R...> seq,
		uint32_t t, const uint64_t *src, const uint64_t *src2, uint64_t *dst,
		uint64_t *lds, uint32_t jbase, uint32_t gbase, const ulonglong2 *tw,
		const NttConsts &c, bool final) {
	uint64_t v[8];
	run_rounds_v<G, INV, LOGN, TRAIL, MC, 0>(seq, v, t, src, src2, dst, lds,
			jbase, gbase, tw, c, final);
}

template <int LOGN, int PASS, int TB, bool INV, int MC>
__device__ __forceinline__ void run_pass(const uint64_t *src,
		const uint64_t *src2, uint64_t *dst, uint64_t *lds, const ulonglong2 *tw,
		const NttConsts &c, bool final) {
	using G = Geo<LOGN, PASS, TB>;
	const uint32_t t = threadIdx.x;
	for (uint32_t i = 0; i < (uint32_t) G::NTILES; i++) {
		// pass 1 tiles start at column i * 2^cb; pass 2 tiles at row block
		// i (2^TB contiguous words), which is also their twiddle offset.
		const uint32_t jbase = PASS == 1 ? i << G::CB : i << TB;
		const uint32_t gbase = PASS == 2 ? jbase : 0;
		run_rounds<G, INV, LOGN, true, MC>(
				std::make_integer_sequence<int, G::ROUNDS>(), t, src, src2, dst, lds,
				jbase, gbase, tw, c, final);
	}
}

template <int LOGN, int TB>
struct FusedShape {
	static constexpr int KT = LOGN <= TB ? LOGN : TB;
	static constexpr int THREADS = (1 << (KT - 3)) < 64 ? 64
		: (1 << (KT - 3)) > 1024 ? 1024 : 1 << (KT - 3);
	static constexpr int LDS_WORDS = (1 << KT) + (1 << (KT - 3));
};

// Waves per SIMD a whole-transform kernel is compiled for.  One-tile sizes
// up to 2^13 words fit eight waves per SIMD in LDS (2^13: two 1024-thread
// workgroups of 72 KiB per CU; 2^12: four of 36 KiB), which holds when the
// kernel keeps to 64 VGPRs; left to itself the compiler gives the inverse
// 72, and one workgroup fewer fits per CU.  A 2^14-word tile (144 KiB)
// holds one workgroup per CU, four waves per SIMD, and two-pass sizes (the
// whole polynomial per workgroup, pass mode 1) are left to the compiler.
constexpr int fused_waves(int logn, int tb) {
	return logn <= 13 && logn <= tb ? 8 : 1;
}

// One workgroup per task.  With `tasks` null, workgroup i transforms the
// i-th polynomial after `uniform` (same tables, contiguous polynomials).
template <int LOGN, bool INV, int TB, int MC>
__global__ __launch_bounds__((FusedShape<LOGN, TB>::THREADS))
__attribute__((amdgpu_waves_per_eu(fused_waves(LOGN, TB)))) void k_ntt_fused(
		NttTask uniform, const NttTask *tasks) {
	__shared__ uint64_t lds[FusedShape<LOGN, TB>::LDS_WORDS];
	NttTask k;
	if (tasks) {
		k = tasks[blockIdx.x];
	} else {
		k = uniform;
		const uint64_t off = (uint64_t) blockIdx.x << LOGN;
		k.in += off;
		k.out += off;
		if (k.in2)
			k.in2 += off;
	}
	if constexpr (LOGN <= TB) {
		run_pass<LOGN, 0, TB, INV, MC>(k.in, k.in2, k.out, lds, k.tw, k.c, true);
	} else if constexpr (!INV) {
		run_pass<LOGN, 1, TB, INV, MC>(k.in, nullptr, k.out, lds, k.tw, k.c,
				false);
		run_pass<LOGN, 2, TB, INV, MC>(k.out, k.in2, k.out, lds, k.tw, k.c,
				true);
	} else {
		run_pass<LOGN, 2, TB, INV, MC>(k.in, k.in2, k.out, lds, k.tw, k.c, false);
		run_pass<LOGN, 1, TB, INV, MC>(k.out, nullptr, k.out, lds, k.tw, k.c,
				true);
	}
}

// One pass of a two-pass size, one workgroup per (tile, polynomial):
// blockIdx.x = tile, blockIdx.y = polynomial (counted from the last with
// kReversed).  The first pass of a direction reads `in`, the second
// (kSecondPass) reads and writes `out`.
// Waves per SIMD the pass kernels are compiled for: eight (at most 64
// VGPRs), which every shape fits without spilling except pass 1 of 2^20 on
// 2^13-word tiles (ten stages over eight columns), which needs about 96.
// (2^14-word tiles take 144 KiB of LDS, one workgroup of 16 waves per CU:
// four waves per SIMD, so registers are not the limit there.)
constexpr int pass_waves(int logn, int pass, int tb) {
	return tb == 14 ? 4 : logn == 20 && pass == 1 && tb == 13 ? 5 : 8;
}

template <int LOGN, int PASS, bool INV, int TB, int MC, int SPLIT = 0>
__global__ __launch_bounds__((FusedShape<LOGN, TB>::THREADS))
__attribute__((amdgpu_waves_per_eu(pass_waves(LOGN, PASS, TB)))) void k_ntt_pass(
		NttTask uniform, const NttTask *tasks, uint32_t flags) {
	__shared__ uint64_t lds[FusedShape<LOGN, TB>::LDS_WORDS];
	if constexpr (LOGN > TB) {
		const bool second = flags & kSecondPass;
		const uint32_t y = flags & kReversed ? gridDim.y - 1 - blockIdx.y
				: blockIdx.y;
		NttTask k;
		if (tasks) {
			k = tasks[y];
		} else {
			k = uniform;
			const uint64_t off = (uint64_t) y << LOGN;
			k.in += off;
			k.out += off;
			if (k.in2)
				k.in2 += off;
		}
		using G = Geo<LOGN, PASS, TB, SPLIT>;
		const uint32_t i = blockIdx.x;
		const uint32_t jbase = PASS == 1 ? i << G::CB : i << TB;
		const uint32_t gbase = PASS == 2 ? jbase : 0;
		// in2: the product operand of an inverse's first pass, the
		// epilogue operand of a forward's second
		const bool with_in2 = INV ? !second : second;
		run_rounds<G, INV, LOGN, false, MC>(
				std::make_integer_sequence<int, G::ROUNDS>(), threadIdx.x,
				second ? k.out : k.in, with_in2 ? k.in2 : nullptr, k.out, lds, jbase,
				gbase, k.tw, k.c, second);
	}
}

// ---- both passes in one launch (handoff.h) ----
//
// The XCD (its own L2) this wave runs on.
__device__ __forceinline__ uint32_t xcc_id() {
	uint32_t v;
	asm volatile("s_getreg_b32 %0, hwreg(HW_REG_XCC_ID)" : "=s"(v));
	return v & (kXcds - 1);
}

// An atomic read of *p performed where the atomics are (a compare-and-swap
// that never changes the word), not a load a cache above may satisfy.
__device__ __forceinline__ uint32_t atomic_peek(uint32_t *p) {
	uint32_t e = 0;
	__hip_atomic_compare_exchange_strong(p, &e, 0u, __ATOMIC_RELAXED,
			__ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
	return e;
}

// One workgroup per ticket: each workgroup claims the next ticket of its
// XCD's queue (handoff_decode) and runs that tile of the polynomial's first
// or second pass; the grid gives every XCD as many workgroups as the
// longest queue has tickets, and the ones left without a ticket exit.  (A
// persistent loop over tickets needs about twice the registers.)  A
// producer tile publishes by waiting for its stores to complete (they are
// then in the XCD's L2: the vector L1 writes through) and incrementing the
// polynomial's count.  A consumer tile first waits for the count to reach
// the tile count, polling at most spin_limit times (then it sets the error
// flag and goes on, and no later wait starts, so the launch always ends),
// and invalidates its CU's vector L1 so that its loads see the L2.  Tasks
// come from the device array (read-only: restrict lets them load through
// the scalar cache).
template <int LOGN, bool INV, int TB, int MC>
__global__ __launch_bounds__((FusedShape<LOGN, TB>::THREADS))
__attribute__((amdgpu_waves_per_eu(pass_waves(LOGN, 1, TB)))) void k_ntt_handoff(
		NttTask, const NttTask *__restrict__ tasks, uint32_t npoly,
		uint32_t *state, uint32_t spin_limit, uint32_t ahead) {
	__shared__ uint64_t lds[FusedShape<LOGN, TB>::LDS_WORDS];
	__shared__ uint32_t s_ticket;
	if constexpr (LOGN > TB) {
		using G1 = Geo<LOGN, 1, TB>;
		using G2 = Geo<LOGN, 2, TB>;
		constexpr uint32_t T = G1::NTILES;
		const uint32_t t = threadIdx.x;
		const uint32_t x = xcc_id();
		if (t == 0)
			s_ticket = __hip_atomic_fetch_add(state + x * kHandoffStride, 1u,
					__ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
		__syncthreads();
		// (uniform, so the task loads through the scalar cache)
		const uint32_t k = __builtin_amdgcn_readfirstlane(s_ticket);
		if (k >= handoff_queue_tickets(npoly, x, T))
			return;
		const HandoffWork w = handoff_decode(k, x, npoly, T, ahead);
		uint32_t *count = state + kHandoffCounts + x * handoff_count_stride(npoly) +
			w.poly / kXcds;
		const NttTask tk = tasks[w.poly];
		if (!w.producer) {
			// (once any wait has given up, none waits again)
			if (t == 0 && !atomic_peek(state + kHandoffErr)) {
				uint32_t spins = 0;
				while (atomic_peek(count) < T) {
					if (++spins >= spin_limit) {
						__hip_atomic_fetch_or(state + kHandoffErr, 1u, __ATOMIC_RELAXED,
								__HIP_MEMORY_SCOPE_AGENT);
						break;
					}
					__builtin_amdgcn_s_sleep(2);
				}
			}
			__syncthreads();
			asm volatile("buffer_inv sc0\n\ts_waitcnt vmcnt(0)" ::: "memory");
		}
		// forward: producer = column pass (1), consumer = row pass (2);
		// inverse: producer = row pass (2), consumer = column pass (1)
		if (w.producer == INV) {
			const uint32_t jb = w.tile << TB;
			run_rounds<G2, INV, LOGN, false, MC>(
					std::make_integer_sequence<int, G2::ROUNDS>(), t,
					w.producer ? tk.in : tk.out, tk.in2, tk.out, lds, jb, jb, tk.tw,
					tk.c, !w.producer);
		} else {
			run_rounds<G1, INV, LOGN, false, MC>(
					std::make_integer_sequence<int, G1::ROUNDS>(), t,
					w.producer ? tk.in : tk.out, nullptr, tk.out, lds,
					w.tile << G1::CB, 0, tk.tw, tk.c, !w.producer);
		}
		if (w.producer) {
			asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
			__syncthreads();
			if (t == 0)
				__hip_atomic_fetch_add(count, 1u, __ATOMIC_RELAXED,
						__HIP_MEMORY_SCOPE_AGENT);
		}
		if (t == 0)
			__hip_atomic_fetch_add(state + handoff_done_word(x), 1u,
					__ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
	}
}

// The middle of a polynomial product a*b fused into one kernel: b's and
// a's forward rounds (twiddles tw2; b from in2, a from in) with canonical
// results kept in registers, their pointwise product, then the inverse
// rounds (twiddles tw) from those registers.  Neither spectrum is stored
// and loaded back.  One-tile sizes run the whole product, one workgroup per
// task (blockIdx.x), reading a and b and writing the result: 24 bytes per
// output word.  Two-pass sizes run the row pass of all three transforms,
// one workgroup per (tile, task) = (blockIdx.x, blockIdx.y), on a and b
// after their forward column passes, and the inverse column pass follows.  Possible when the forward's last round and the inverse's first
// both hold their field at bit 0: thread t then holds words 8t..8t+7 in
// both, and both are wave-local, so a wave fence orders the forward's LDS
// reads before the inverse's first exchange writes.
// Two-pass sizes only with rows of 512 words (three rounds): longer rows
// (n = 2^19, 2^20) spill the product's registers at eight waves per SIMD
// and measured no faster than the separate kernels.
template <int LOGN, int TB>
constexpr bool pm_fusable() {
	using G = Geo<LOGN, LOGN <= TB ? 0 : 2, TB>;
	return G::VT == 1 && (LOGN <= TB || G::K2 == 9) &&
		Round<G, false, G::ROUNDS - 1>::F == 0 && Round<G, true, 0>::F == 0;
}

// Words i, i + 1 of p (i even).
__device__ __forceinline__ u64x2 gload2(const uint64_t *p, uint32_t i) {
	return *(gcu128 *) ((gcchar *) p + (i << 3));
}

// One-tile sizes leave the register budget to the compiler (it takes 123-128
// VGPRs; any lower cap spills); the row kernel holds both forward results
// (16 VGPRs more than a pass kernel) and runs at five waves per SIMD, where
// it needs no scratch (87 VGPRs).  At six it spilled 52 bytes per thread
// and measured 5-6% slower on 2^27-word batches of products at n = 2^15 to
// 2^17 on MI355X; at four (on 2^12 tiles the same two 512-thread
// workgroups per CU, with 128 VGPRs) 1-3% slower at 2^15 and 2^16.
constexpr int pm_waves(int logn, int tb) {
	return logn <= tb ? 1 : 5;
}

template <int LOGN, int TB, int MC>
__global__ __launch_bounds__((FusedShape<LOGN, TB>::THREADS))
__attribute__((amdgpu_waves_per_eu(pm_waves(LOGN, TB)))) void k_ntt_pm(
		NttTask uniform, const NttTask *tasks, uint32_t) {
	__shared__ uint64_t lds[FusedShape<LOGN, TB>::LDS_WORDS];
	if constexpr (pm_fusable<LOGN, TB>()) {
		constexpr int PASS = LOGN <= TB ? 0 : 2;
		using G = Geo<LOGN, PASS, TB>;
		const uint32_t task = PASS == 0 ? blockIdx.x : blockIdx.y;
		NttTask k;
		if (tasks) {
			k = tasks[task];
		} else {
			k = uniform;
			const uint64_t off = (uint64_t) task << LOGN;
			k.in += off;
			k.out += off;
			k.in2 += off;
		}
		const uint32_t t = threadIdx.x;
		const uint32_t jbase = PASS == 0 ? 0 : blockIdx.x << TB;
		const auto seq = std::make_integer_sequence<int, G::ROUNDS>();
		// b's forward rounds, then a's, both kept in registers in the same
		// layout (thread t holds words 8t..8t+7 after the last round).  Both
		// operands are read once, with the nontemporal hint, so that they
		// leave the L2 to the two twiddle tables the kernel rereads (2^15 to
		// 2^17 measured 1.5-3% faster than with ordinary loads).
		uint64_t vb[8], v[8];
		run_rounds_v<G, false, LOGN, false, MC, kIoKeep | kIoStreamIn>(seq, vb, t, k.in2,
				nullptr, k.out, lds, jbase, jbase, k.tw2, k.c, true);
		// (a's first round may write LDS words other waves read in b's last)
		__syncthreads();
		run_rounds_v<G, false, LOGN, false, MC, kIoKeep | kIoStreamIn>(seq, v, t, k.in,
				nullptr, k.out, lds, jbase, jbase, k.tw2, k.c, true);
		if (G::THREADS >= 64 || t < (uint32_t) G::THREADS) {
#pragma unroll
			for (int e = 0; e < 8; e++)
				v[e] = spectra_product(v[e], vb[e], k.c);
		}
		wave_sync();
		run_rounds_v<G, true, LOGN, false, MC, kIoPreloaded>(seq, v, t, nullptr,
				nullptr, k.out, lds, jbase, jbase, k.tw, k.c, PASS == 0);
	}
}

// The pass kernels of one (size, pass, direction, tile, split) for the
// three modulus classes; the alternative split is instantiated only where
// it exists (ALT_SPLIT: 1 there, else 0, naming the default kernels).
#define PASS_ROW(LG, PS, INV, TB, SP) \
	{k_ntt_pass<LG, PS, INV, TB, 0, SP>, k_ntt_pass<LG, PS, INV, TB, 1, SP>, \
		k_ntt_pass<LG, PS, INV, TB, 2, SP>}
#define ALT_SPLIT(LG, TB) (alt_k1(LG, TB) > 0 ? 1 : 0)

// Kernel table for tile size TB, entry i serving log2 n = i + 3.
template <int TB, int... L>
constexpr std::array<FusedEntry, sizeof...(L)> make_fused_table(
		std::integer_sequence<int, L...>) {
	return {FusedEntry{
		{k_ntt_fused<L + kMinFusedLogn, false, TB, 0>,
			k_ntt_fused<L + kMinFusedLogn, false, TB, 1>,
			k_ntt_fused<L + kMinFusedLogn, false, TB, 2>},
		{k_ntt_fused<L + kMinFusedLogn, true, TB, 0>,
			k_ntt_fused<L + kMinFusedLogn, true, TB, 1>,
			k_ntt_fused<L + kMinFusedLogn, true, TB, 2>},
		{PASS_ROW(L + kMinFusedLogn, 1, false, TB, 0),
			PASS_ROW(L + kMinFusedLogn, 1, false, TB, ALT_SPLIT(L + kMinFusedLogn, TB))},
		{PASS_ROW(L + kMinFusedLogn, 2, false, TB, 0),
			PASS_ROW(L + kMinFusedLogn, 2, false, TB, ALT_SPLIT(L + kMinFusedLogn, TB))},
		{PASS_ROW(L + kMinFusedLogn, 2, true, TB, 0),
			PASS_ROW(L + kMinFusedLogn, 2, true, TB, ALT_SPLIT(L + kMinFusedLogn, TB))},
		{PASS_ROW(L + kMinFusedLogn, 1, true, TB, 0),
			PASS_ROW(L + kMinFusedLogn, 1, true, TB, ALT_SPLIT(L + kMinFusedLogn, TB))},
		{k_ntt_pm<L + kMinFusedLogn, TB, 0>,
			k_ntt_pm<L + kMinFusedLogn, TB, 1>,
			k_ntt_pm<L + kMinFusedLogn, TB, 2>},
		pm_fusable<L + kMinFusedLogn, TB>(),
		{k_ntt_handoff<L + kMinFusedLogn, false, TB, 0>,
			k_ntt_handoff<L + kMinFusedLogn, false, TB, 1>,
			k_ntt_handoff<L + kMinFusedLogn, false, TB, 2>},
		{k_ntt_handoff<L + kMinFusedLogn, true, TB, 0>,
			k_ntt_handoff<L + kMinFusedLogn, true, TB, 1>,
			k_ntt_handoff<L + kMinFusedLogn, true, TB, 2>},
		FusedShape<L + kMinFusedLogn, TB>::THREADS,
		1 << (L + kMinFusedLogn > TB ? L + kMinFusedLogn - TB : 0),
		{ntt_k1(L + kMinFusedLogn, TB),
			split_k1(L + kMinFusedLogn, TB, 1)}}...};
}

// Per-tile-size tables (ntt_fused_kt*.hip).  Sizes n <= 2^TB use the
// one-tile kernel and are identical across tables.
const FusedEntry *fused_table_kt11();
const FusedEntry *fused_table_kt12();
const FusedEntry *fused_table_kt13();
const FusedEntry *fused_table_kt14();

}  // namespace vk
