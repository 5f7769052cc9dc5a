// Negacyclic NTT / INTT of vkhel (src/vector.c:513-657) for gfx950.
//
// The reference issues one dispatch per butterfly group (n - 1 dispatches
// and log2(n) fenced submits per transform).  Here (ntt_fused.h):
//
//   * n <= 2^tile: one workgroup per polynomial stages it once in LDS and
//     runs every stage there, three stages per register round (radix-8 per
//     thread).
//   * larger n: two tiled passes, each its own grid of one workgroup per
//     (tile, polynomial): pass 1 runs the top floor(log2 n / 2) stages on
//     column tiles, pass 2 the low stages on row tiles.  Tiles of 2^11,
//     2^12 or 2^13 words are chosen per size (autotuned, or a measured
//     default).  Whole-polynomial workgroups running both passes remain as
//     a tunable mode.
//
// Twiddles come from device tables in the reference's stage blocks
// (roots[m + i], src/vector.c:536-547; interleaved within each block, see
// tw_pos) as {w, 63-bit Shoup companion} pairs.  Results are canonical and
// equal the reference's stage-by-stage output bit for bit: every butterfly
// is exact mod q, so only the order of independent butterflies changes.
// Sizes above 2^20 and the RADIX2 algorithm use the radix-2 kernels below.
#include <algorithm>
#include <array>
#include <mutex>
#include <unordered_map>
#include <utility>
#include <vector>

#include "internal.h"
#include "modarith.h"
#include "ntt_fused.h"

namespace vk {

// ---------------------------------------------------------------- radix-2

// One forward stage over `total` = batch * n/2 butterflies.  Stage s has
// m = 2^s groups of t = n >> (s + 1) butterflies (src/vector.c:536-547,
// nttfwdbutterfly.comp:31-58).  FIRST reduces arbitrary 64-bit inputs the
// way the reference's first stage does: (X + WY) mod 2^64 mod q and
// (X - WY) mod q.
template <bool FIRST>
__global__ __launch_bounds__(256) void k_fwd_stage(const uint64_t *in,
		uint64_t *out, const ulonglong2 *tw, NttConsts c, uint32_t s,
		uint64_t total) {
	const uint32_t hb = c.logn - 1 - s;  // log2 t
	const uint64_t half = (uint64_t) 1 << (c.logn - 1);
	for (uint64_t idx = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x;
			idx < total; idx += (uint64_t) gridDim.x * blockDim.x) {
		const uint64_t poly = idx >> (c.logn - 1), b = idx & (half - 1);
		const uint64_t grp = b >> hb, pos = b & (((uint64_t) 1 << hb) - 1);
		const uint64_t x = (poly << c.logn) + (grp << (hb + 1)) + pos;
		const uint64_t y = x + ((uint64_t) 1 << hb);
		const ulonglong2 w = tw[tw_pos(1u << s, (uint32_t) grp)];
		const uint64_t X = in[x], Y = in[y];
		// twiddle pairs hold 63-bit Shoup companions (inputs below 2^63)
		const uint64_t wy = csub(shoup63(FIRST ? barrett64(Y, c.q, c.mu,
						c.bshift) : Y, w.x, w.y, 0 - c.q), c.q);
		uint64_t ox, oy;
		if (FIRST) {
			const uint64_t xr = barrett64(X, c.q, c.mu, c.bshift);
			ox = csub(xr + wy, c.q);
			if (X + wy < X)  // the reference's 64-bit sum wrapped
				ox = sub_mod(ox, c.r64, c.q);
			oy = sub_mod(xr, wy, c.q);
		} else {
			ox = csub(X + wy, c.q);
			oy = sub_mod(X, wy, c.q);
		}
		out[x] = ox;
		out[y] = oy;
	}
}

// One inverse stage (src/vector.c:599-630, nttrevbutterfly.comp:31-58).
// LAST (s == 0, m == 1) folds the final multiplication by n^-1
// (src/vector.c:633-638) into both outputs.
template <bool LAST>
__global__ __launch_bounds__(256) void k_inv_stage(const uint64_t *in,
		uint64_t *out, const ulonglong2 *tw, NttConsts c, uint32_t s,
		uint64_t total) {
	const uint32_t hb = c.logn - 1 - s;
	const uint64_t half = (uint64_t) 1 << (c.logn - 1);
	for (uint64_t idx = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x;
			idx < total; idx += (uint64_t) gridDim.x * blockDim.x) {
		const uint64_t poly = idx >> (c.logn - 1), b = idx & (half - 1);
		const uint64_t grp = b >> hb, pos = b & (((uint64_t) 1 << hb) - 1);
		const uint64_t x = (poly << c.logn) + (grp << (hb + 1)) + pos;
		const uint64_t y = x + ((uint64_t) 1 << hb);
		const uint64_t X = in[x], Y = in[y];
		const uint64_t sum = csub(X + Y, c.q), d = sub_mod(X, Y, c.q);
		if (LAST) {
			out[x] = csub(shoup63(sum, c.ninv, c.ninv_p, 0 - c.q), c.q);
			out[y] = csub(shoup63(d, c.w1n, c.w1n_p, 0 - c.q), c.q);
		} else {
			const ulonglong2 w = tw[tw_pos(1u << s, (uint32_t) grp)];
			out[x] = sum;
			out[y] = csub(shoup63(d, w.x, w.y, 0 - c.q), c.q);
		}
	}
}

static int launch_radix2_tw(const ulonglong2 *tw, const NttConsts &c,
		const uint64_t *in, uint64_t *out, uint32_t batch, bool inverse,
		hipStream_t st);

static int launch_radix2(const DevTables &t, const uint64_t *in, uint64_t *out,
		uint32_t batch, bool inverse, hipStream_t st) {
	return launch_radix2_tw(inverse ? t.inv : t.fwd, t.c, in, out, batch,
			inverse, st);
}

static int launch_radix2_tw(const ulonglong2 *tw, const NttConsts &c,
		const uint64_t *in, uint64_t *out, uint32_t batch, bool inverse,
		hipStream_t st) {
	const uint32_t logn = c.logn;
	const uint64_t total = (uint64_t) batch << (logn - 1);
	const unsigned grid = (unsigned) std::min<uint64_t>((total + 255) / 256,
			16384);
	if (!inverse) {
		for (uint32_t s = 0; s < logn; s++) {
			if (s == 0)
				hipLaunchKernelGGL(k_fwd_stage<true>, dim3(grid), dim3(256), 0, st,
						in, out, tw, c, s, total);
			else
				hipLaunchKernelGGL(k_fwd_stage<false>, dim3(grid), dim3(256), 0, st,
						out, out, tw, c, s, total);
		}
	} else {
		for (int s = (int) logn - 1; s >= 0; s--) {
			const uint64_t *src = (s == (int) logn - 1) ? in : out;
			if (s == 0)
				hipLaunchKernelGGL(k_inv_stage<true>, dim3(grid), dim3(256), 0, st,
						src, out, tw, c, 0u, total);
			else
				hipLaunchKernelGGL(k_inv_stage<false>, dim3(grid), dim3(256), 0, st,
						src, out, tw, c, (uint32_t) s, total);
		}
	}
	return hip_ok(hipGetLastError(), "ntt radix2 launch") ? VKHEL_OK
		: VKHEL_ERR_DEVICE;
}

static bool fused_supported(uint32_t logn) {
	return logn >= (uint32_t) kMinFusedLogn && logn <= (uint32_t) kMaxFusedLogn;
}

// Tile sizes (log2 words) the fused kernels are compiled for.
static const FusedEntry &fused_entry(uint32_t logn, int tile_bits) {
	const FusedEntry *t = tile_bits == 11 ? fused_table_kt11()
		: tile_bits == 12 ? fused_table_kt12() : tile_bits == 13 ? fused_table_kt13()
		: fused_table_kt14();
	return t[logn - kMinFusedLogn];
}

NttTask make_task(const DevTables &t, const uint64_t *in, uint64_t *out,
		bool inverse) {
	NttTask k;
	k.in = in;
	k.out = out;
	k.tw = inverse ? t.inv : t.fwd;
	k.c = t.c;
	k.in2 = nullptr;
	k.tw2 = nullptr;
	return k;
}

int launch_task_ntt(const NttTask &k, bool inverse, hipStream_t st) {
	return launch_radix2_tw(k.tw, k.c, k.in, k.out, 1, inverse, st);
}

bool ntt_fused_supported(uint32_t logn) {
	return fused_supported(logn);
}

// Untuned default tile per size, from batched timings on MI355X (2^25 words
// per batch, split passes): 2^12 tiles from n = 2^14 to 2^17, 2^13 above;
// sizes up to 2^13 run in one pass on a 2^13 tile.
static int default_tile_bits(uint32_t logn) {
	if (logn <= 13)
		return 13;
	return logn <= 17 ? 12 : 13;
}

int tile_bits_for(const vkhel_ctx *ctx, uint32_t logn) {
	const int tb = logn < ctx->tile_bits.size() ? ctx->tile_bits[logn] : 0;
	return tb ? tb : default_tile_bits(logn);
}

// Split two-pass transforms run in chunks of at most 2^25 words (256 MiB)
// of polynomials, both passes of one chunk before the next, and the second
// pass walks the chunk's polynomials from the last (kReversed): it starts
// on what the first pass wrote last, still held in the 256 MB last-level
// cache, and the next launch's first pass, walking forward, starts where
// this one ended.  Measured on MI355X (forward, 2^16 words, 2^11 tiles,
// variants run back to back on one box): 512 polynomials 0.212 ms in two
// 2^24-word chunks in order, 0.214 with the second pass reversed, 0.200 in
// one reversed chunk; 2048 polynomials 0.920 -> 0.876 ms (2^25-word
// reversed chunks against 2^24 in order); 256 x 2^17 0.238 -> 0.224,
// 32 x 2^20 0.348 -> 0.332, 1024 x 2^15 0.199 -> 0.189.  Smaller chunks add
// launch tails (2^23-word reversed chunks: 0.236 at 512 x 2^16).  Chunks
// also keep the grid's y dimension within its limit of 65535.
constexpr uint64_t kChunkWords = 1ull << 25;

// The kernels of a two-pass size on a context: pass 1 (columns) from the
// tile_bits table, pass 2 (rows) from the tile_bits2 one, both with the
// context's stage split.  Both passes must split the stages alike; where
// the second tile's table would not, pass 2 takes the first tile.
struct PassPlan {
	const FusedEntry *e1, *e2;
	int split;
};

static PassPlan pass_plan(const vkhel_ctx *ctx, uint32_t logn, int tile_bits) {
	const int s = logn < ctx->split.size() ? ctx->split[logn] : 0;
	const FusedEntry &e1 = fused_entry(logn, tile_bits);
	const int tb2 = logn < ctx->tile_bits2.size() ? ctx->tile_bits2[logn] : 0;
	const FusedEntry *e2 = tb2 ? &fused_entry(logn, tb2) : &e1;
	if (e2->tiles <= 1 || e2->k1[s] != e1.k1[s])
		e2 = &e1;
	return PassPlan{&e1, e2, s};
}

// Launches `ntasks` transforms of one size and direction: either through
// the device task array `dev_tasks` or, when it is null, `uniform` and the
// ntasks - 1 polynomials contiguous after it.  Split mode runs the two
// passes as separate grids of one workgroup per tile (pass_plan's
// kernels).
static int launch_fused_impl(const vkhel_ctx *ctx, const NttTask &uniform,
		const NttTask *dev_tasks, uint32_t ntasks, uint32_t logn, int tile_bits,
		bool split, bool inverse, int mc, hipStream_t st) {
	const FusedEntry &e = fused_entry(logn, tile_bits);
	if (split && e.tiles > 1) {
		const PassPlan pp = pass_plan(ctx, logn, tile_bits);
		const FusedEntry &e1 = *pp.e1, &e2 = *pp.e2;
		const int s = pp.split;
		const uint32_t chunk = (uint32_t) std::max<uint64_t>(1,
				std::min<uint64_t>(kChunkWords >> logn, 65535));
		for (uint32_t first = 0; first < ntasks; first += chunk) {
			const uint32_t count = std::min(chunk, ntasks - first);
			NttTask u = uniform;
			const NttTask *d = dev_tasks;
			if (d) {
				d += first;
			} else {
				const uint64_t off = (uint64_t) first << logn;
				u.in += off;
				u.out += off;
				if (u.in2)
					u.in2 += off;
			}
			// forward: columns, then rows; inverse: rows, then columns
			if (!inverse) {
				hipLaunchKernelGGL(e1.fwd_a[s][mc], dim3(e1.tiles, count),
						dim3(e1.threads), 0, st, u, d, 0u);
				hipLaunchKernelGGL(e2.fwd_b[s][mc], dim3(e2.tiles, count),
						dim3(e2.threads), 0, st, u, d, kSecondPass | kReversed);
			} else {
				hipLaunchKernelGGL(e2.inv_b[s][mc], dim3(e2.tiles, count),
						dim3(e2.threads), 0, st, u, d, 0u);
				hipLaunchKernelGGL(e1.inv_a[s][mc], dim3(e1.tiles, count),
						dim3(e1.threads), 0, st, u, d, kSecondPass | kReversed);
			}
		}
	} else {
		hipLaunchKernelGGL(inverse ? e.inv[mc] : e.fwd[mc],
				dim3(ntasks), dim3(e.threads), 0, st, uniform, dev_tasks);
	}
	return hip_ok(hipGetLastError(), "ntt fused launch") ? VKHEL_OK
		: VKHEL_ERR_DEVICE;
}

// Two-pass sizes run their passes as separate grids (one workgroup per tile)
// unless tuned otherwise: at every size and batch measured, the split grids'
// smaller register footprint and full-chip occupancy beat one workgroup per
// polynomial.
static int pass_mode_of(const vkhel_ctx *ctx, uint32_t logn) {
	return logn < ctx->pass_mode.size() ? ctx->pass_mode[logn] : 0;
}

static bool use_split(const vkhel_ctx *ctx, uint32_t logn) {
	return pass_mode_of(ctx, logn) != 1;
}

// Zeroes the state words of a one-launch two-pass transform: a kernel, not
// hipMemsetAsync, so that it is safe inside captured graphs.
__global__ __launch_bounds__(256) void k_handoff_reset(uint32_t *p, uint32_t n) {
	for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n;
			i += gridDim.x * blockDim.x)
		p[i] = 0;
}

// Polls a consumer tile makes before it gives up and sets the error flag
// (each poll is an atomic round trip and a short sleep, about a microsecond;
// a producer tile takes tens of microseconds at most).
constexpr uint32_t kHandoffSpinLimit = 1u << 20;

uint64_t handoff_state_bytes(uint32_t ntasks) {
	return 4ull * handoff_state_words(ntasks);
}

// (the one-launch kernels exist for the default split and one tile for
// both passes)
int fused_split_k1(uint32_t logn, int tb, int s) {
	if (!fused_supported(logn) || tb < 11 || tb > 14 || s < 0 || s > 1)
		return 0;
	return fused_entry(logn, tb).k1[s];
}

bool handoff_state_ok(const vkhel_ctx *ctx, const uint32_t *state,
		uint32_t npoly, uint32_t logn) {
	std::vector<uint32_t> h(kHandoffCounts);
	if (hipMemcpy(h.data(), state, 4 * h.size(), hipMemcpyDeviceToHost) !=
			hipSuccess)
		return false;
	const uint32_t tiles = handoff_tiles(logn, tile_bits_for(ctx, logn));
	if (h[handoff_err_word()] != 0)
		return false;
	for (uint32_t x = 0; x < kXcds; x++)
		if (h[handoff_done_word(x)] != handoff_queue_tickets(npoly, x, tiles))
			return false;
	return true;
}

bool uses_handoff(const vkhel_ctx *ctx, uint32_t logn) {
	return fused_supported(logn) && pass_mode_of(ctx, logn) == 3 &&
		fused_entry(logn, tile_bits_for(ctx, logn)).tiles > 1 &&
		ctx->split[logn] == 0 && (ctx->tile_bits2[logn] == 0 ||
			ctx->tile_bits2[logn] == tile_bits_for(ctx, logn));
}

// Both passes of `ntasks` two-pass transforms in one persistent launch
// (k_ntt_handoff), its state in `state` (handoff_state_bytes(ntasks)).
static int launch_handoff(const vkhel_ctx *ctx, const FusedEntry &e,
		const NttTask *dev_tasks, uint32_t ntasks, bool inverse, int mc,
		uint32_t *state, hipStream_t st) {
	const uint32_t words = handoff_state_words(ntasks);
	hipLaunchKernelGGL(k_handoff_reset, dim3(std::min<uint32_t>((words + 255) / 256,
					1024)), dim3(256), 0, st, state, words);
	const HandoffKernel k = inverse ? e.inv_h[mc] : e.fwd_h[mc];
	// one workgroup per ticket: placed round-robin over the XCDs, this grid
	// gives every XCD as many workgroups as the longest queue (XCD 0's) has
	// tickets
	const uint64_t grid = (uint64_t) kXcds *
		handoff_queue_tickets(ntasks, 0, (uint32_t) e.tiles);
	if (grid > 0x7fffffffull)
		return fail(VKHEL_ERR_INVALID, "ntt handoff: %u polynomials is too many "
				"for one launch", ntasks);
	// look-ahead: about the polynomials the resident workgroups of one XCD
	// hold tickets of (four 512-thread workgroups per CU, 32 CUs per XCD)
	// unless set on the context
	const uint32_t ahead = ctx->handoff_ahead > 0 ? (uint32_t) ctx->handoff_ahead
		: std::max<uint32_t>(1, (uint32_t) (ctx->cu_count / kXcds) *
				std::max(1, 2048 / e.threads) / (2 * e.tiles) + 1);
	const NttTask none{};
	hipLaunchKernelGGL(k, dim3((uint32_t) grid), dim3(e.threads), 0, st, none,
			dev_tasks, ntasks, state, kHandoffSpinLimit, ahead);
	return hip_ok(hipGetLastError(), "ntt handoff launch") ? VKHEL_OK
		: VKHEL_ERR_DEVICE;
}

int launch_fused_tasks(const vkhel_ctx *ctx, const NttTask *dev_tasks,
		uint32_t ntasks, uint32_t logn, bool inverse, int mc, hipStream_t st,
		uint32_t *handoff_state, uint64_t handoff_bytes) {
	if (ntasks == 0)
		return VKHEL_OK;
	if (!fused_supported(logn))
		return fail(VKHEL_ERR_INVALID, "fused NTT: unsupported log2 n = %u", logn);
	const int tb = tile_bits_for(ctx, logn);
	const FusedEntry &e = fused_entry(logn, tb);
	if (pass_mode_of(ctx, logn) == 3 && e.tiles > 1 && dev_tasks &&
			handoff_state && handoff_bytes >= handoff_state_bytes(ntasks))
		return launch_handoff(ctx, e, dev_tasks, ntasks, inverse, mc,
				handoff_state, st);
	const NttTask none{};
	return launch_fused_impl(ctx, none, dev_tasks, ntasks, logn, tb,
			use_split(ctx, logn), inverse, mc, st);
}

// The context's tile when the fused product kernel exists for it (not for
// 2^14-word tiles, whose threads each run two groups of words), else the
// largest tile where it does.
int polymul_fused_tile(const vkhel_ctx *ctx, uint32_t logn) {
	if (!fused_supported(logn))
		return 0;
	const int tb = tile_bits_for(ctx, logn);
	if (fused_entry(logn, tb).pm_ok)
		return tb;
	for (int alt = 13; alt >= 11; alt--)
		if (fused_entry(logn, alt).pm_ok)
			return alt;
	return 0;
}

// Fused polynomial products from device task arrays (col_tasks null for
// one-tile sizes; otherwise two per product, a's column pass then b's) or,
// with both arrays null, from `ucol_a`, `ucol_b` and `upm` for one product
// (the kernels' uniform mode, nothing read from device memory).
static int polymul_impl(const NttTask &ucol_a, const NttTask &ucol_b,
		const NttTask &upm, const NttTask *col_tasks, const NttTask *pm_tasks,
		bool cols, uint32_t ntasks, uint32_t logn, int tile_bits, int mc,
		hipStream_t st) {
	if (ntasks == 0)
		return VKHEL_OK;
	if (!fused_supported(logn))
		return fail(VKHEL_ERR_INVALID, "fused polymul: unsupported log2 n = %u",
				logn);
	const FusedEntry &e = fused_entry(logn, tile_bits);
	if (!e.pm_ok || (e.tiles > 1) != cols)
		return fail(VKHEL_ERR_INVALID,
				"fused polymul: unsupported log2 n = %u on 2^%d tiles", logn,
				tile_bits);
	if (e.tiles == 1) {
		hipLaunchKernelGGL(e.pm[mc], dim3(ntasks), dim3(e.threads), 0, st, upm,
				pm_tasks, 0u);
	} else {
		// Chunks of 2^24 words per operand, so that the row kernel reads
		// the column passes' output while it is still cached.  (Walking the
		// row kernel's products from the last, or 2^25 words per operand,
		// measured the same on MI355X: 2^27-word batches at n = 2^15 and
		// 2^17, within 2% either way.)
		const uint32_t chunk = (uint32_t) std::max<uint64_t>(1,
				std::min<uint64_t>(kChunkWords >> (logn + 1), 65535));
		for (uint32_t first = 0; first < ntasks; first += chunk) {
			const uint32_t count = std::min(chunk, ntasks - first);
			const dim3 grid(e.tiles, count);
			const NttTask *pt = pm_tasks ? pm_tasks + first : nullptr;
			// the forward column passes of a and b
			if (col_tasks) {
				hipLaunchKernelGGL(e.fwd_a[0][mc], dim3(e.tiles, 2 * count),
						dim3(e.threads), 0, st, ucol_a, col_tasks + 2 * first, 0u);
			} else {
				hipLaunchKernelGGL(e.fwd_a[0][mc], grid, dim3(e.threads), 0, st, ucol_a,
						(const NttTask *) nullptr, 0u);
				hipLaunchKernelGGL(e.fwd_a[0][mc], grid, dim3(e.threads), 0, st, ucol_b,
						(const NttTask *) nullptr, 0u);
			}
			hipLaunchKernelGGL(e.pm[mc], grid, dim3(e.threads), 0, st, upm, pt, 0u);
			hipLaunchKernelGGL(e.inv_a[0][mc], grid, dim3(e.threads), 0, st, upm, pt,
					kSecondPass);
		}
	}
	return hip_ok(hipGetLastError(), "fused polymul launch") ? VKHEL_OK
		: VKHEL_ERR_DEVICE;
}

int launch_polymul_tasks(const NttTask *col_tasks, const NttTask *pm_tasks,
		uint32_t ntasks, uint32_t logn, int tile_bits, int mc, hipStream_t st) {
	const NttTask none{};
	return polymul_impl(none, none, none, col_tasks, pm_tasks,
			col_tasks != nullptr, ntasks, logn, tile_bits, mc, st);
}

int launch_polymul_one(const NttTask &col_a, const NttTask &col_b,
		const NttTask &pm, bool cols, uint32_t logn, int tile_bits, int mc,
		hipStream_t st) {
	return polymul_impl(col_a, col_b, pm, nullptr, nullptr, cols, 1, logn,
			tile_bits, mc, st);
}

int launch_task_fused(const vkhel_ctx *ctx, const NttTask &k, bool inverse,
		hipStream_t st) {
	const uint32_t logn = k.c.logn;
	if (!fused_supported(logn))
		return fail(VKHEL_ERR_INVALID, "fused NTT: unsupported log2 n = %u", logn);
	return launch_fused_impl(ctx, k, nullptr, 1, logn, tile_bits_for(ctx, logn),
			use_split(ctx, logn), inverse, modulus_class(k.c.q), st);
}

int launch_ntt(vkhel_ctx *ctx, const DevTables &t, const uint64_t *in,
		uint64_t *out, uint32_t batch, bool inverse, hipStream_t s) {
	if (batch == 0)
		return VKHEL_OK;
	const uint32_t logn = t.c.logn;
	if (ctx->ntt_algo == VKHEL_NTT_RADIX2 || !fused_supported(logn))
		return launch_radix2(t, in, out, batch, inverse, s);
	return launch_fused_impl(ctx, make_task(t, in, out, inverse), nullptr, batch,
			logn,
			tile_bits_for(ctx, logn), use_split(ctx, logn), inverse,
			modulus_class(t.c.q), s);
}

static void transform(const vkhel_vector *op, vkhel_vector *res,
		vkhel_ntt_tables *ntt, bool inverse) {
	const char *fn = inverse ? "inverse_transform" : "forward_transform";
	if (!op || !res || !ntt)
		return (void) fail(VKHEL_ERR_INVALID, "%s: null argument", fn);
	if (op->ctx != res->ctx)
		return (void) fail(VKHEL_ERR_INVALID, "%s: vectors from different contexts",
				fn);
	if (op->length < ntt->n || res->length < ntt->n)
		return (void) fail(VKHEL_ERR_INVALID,
				"%s: vectors hold %llu/%llu words, tables need n=%llu", fn,
				(unsigned long long) op->length, (unsigned long long) res->length,
				(unsigned long long) ntt->n);
	vkhel_ctx *ctx = res->ctx;
	const DevTables *t = device_tables(ctx, ntt);
	if (!t)
		return;
	if (launch_ntt(ctx, *t, op->dev, res->dev, 1, inverse, ctx->stream) ==
			VKHEL_OK)
		hip_ok(hipStreamSynchronize(ctx->stream), "sync");
}

}  // namespace vk

extern "C" {

// src/vector.c:513-574
void vkhel_vector_forward_transform(const struct vkhel_vector *operand,
		struct vkhel_vector *result, struct vkhel_ntt_tables *ntt) {
	vk::transform(operand, result, ntt, false);
}

// src/vector.c:576-657
void vkhel_vector_inverse_transform(const struct vkhel_vector *operand,
		struct vkhel_vector *result, struct vkhel_ntt_tables *ntt) {
	vk::transform(operand, result, ntt, true);
}

}  // extern "C"
