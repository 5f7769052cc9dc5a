// Autotuner for the fused NTT kernels.  For each transform size it times
// every compiled tile size (2^11, 2^12, 2^13 words, i.e. the stage-fusion
// split and workgroup shape of the two-pass path) in both pass modes on a
// synthetic batch held in device memory, in interleaved rounds, and keeps
// the fastest forward+inverse pair for the context.  The reference has no counterpart: its kernels have one fixed
// shape (64-thread workgroups, one dispatch per butterfly group).
#include <algorithm>
#include <cstdio>
#include <vector>

#include "internal.h"
#include "numbers.h"

namespace vk {

// Fills `n` words with splitmix64-derived values below q (synthetic data;
// values only need to be valid residues).
__global__ void k_fill_residues(uint64_t *p, uint64_t n, uint64_t q,
		uint64_t seed) {
	for (uint64_t i = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x; i < n;
			i += (uint64_t) gridDim.x * blockDim.x) {
		uint64_t z = seed + (i + 1) * 0x9e3779b97f4a7c15ull;
		z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
		z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
		z ^= z >> 31;
		p[i] = z % q;
	}
}

static const int kTileChoices[] = {11, 12, 13, 14};
constexpr int kNumTiles = 4;
static const int kTuneRounds = 4;

}  // namespace vk

using namespace vk;

extern "C" {

int vkhel_ext_tile_bits(struct vkhel_ctx *ctx, uint32_t logn) {
	if (!ctx || logn >= 64)
		return fail(VKHEL_ERR_INVALID, "tile_bits: bad argument"), -1;
	return tile_bits_for(ctx, logn);
}

int vkhel_ext_set_tile_bits(struct vkhel_ctx *ctx, uint32_t logn,
		int tile_bits) {
	if (!ctx || logn >= 64 || (tile_bits != 0 && (tile_bits < 11 ||
					tile_bits > 14)))
		return fail(VKHEL_ERR_INVALID, "set_tile_bits: tile_bits must be 0 or "
				"11..14");
	ctx->tile_bits[logn] = tile_bits;
	return VKHEL_OK;
}

int vkhel_ext_pass_mode(struct vkhel_ctx *ctx, uint32_t logn) {
	if (!ctx || logn >= 64)
		return fail(VKHEL_ERR_INVALID, "pass_mode: bad argument"), -1;
	return ctx->pass_mode[logn];
}

int vkhel_ext_set_pass_mode(struct vkhel_ctx *ctx, uint32_t logn, int mode) {
	if (!ctx || logn >= 64 || mode < 0 || mode > 3)
		return fail(VKHEL_ERR_INVALID, "set_pass_mode: mode must be 0..3");
	ctx->pass_mode[logn] = mode;
	return VKHEL_OK;
}

int vkhel_ext_set_handoff_ahead(struct vkhel_ctx *ctx, int polys) {
	if (!ctx || polys < 0 || polys > 4096)
		return fail(VKHEL_ERR_INVALID, "set_handoff_ahead: 0..4096 polynomials");
	ctx->handoff_ahead = polys;
	return VKHEL_OK;
}

int vkhel_ext_autotune(struct vkhel_ctx *ctx, uint32_t logn_lo,
		uint32_t logn_hi, uint32_t batch) {
	if (!ctx || logn_lo < 3 || logn_hi > 20 || logn_lo > logn_hi)
		return fail(VKHEL_ERR_INVALID, "autotune: sizes must lie in [2^3, 2^20]");
	// A 50-bit NTT prime (1 mod 2^21) and synthetic twiddles: timing does
	// not depend on the values, only on the residues staying below q.
	const uint64_t q = 1125899906826241ull;
	hipStream_t s = ctx->stream;
	for (uint32_t logn = logn_lo; logn <= logn_hi; logn++) {
		const uint64_t n = 1ull << logn;
		// only sizes above the smallest tile have more than one variant
		if (logn <= 11)
			continue;
		const uint32_t polys = batch ? batch : (uint32_t) std::max<uint64_t>(
				(uint64_t) ctx->cu_count * 2, (1ull << 25) >> logn);
		uint64_t *data = nullptr;
		ulonglong2 *tw = nullptr;
		if (!hip_ok(hipMalloc(&data, polys * n * 8), "autotune alloc") ||
				!hip_ok(hipMalloc(&tw, n * sizeof(ulonglong2)), "autotune alloc")) {
			(void) hipFree(data);
			return vkhel_ext_last_error();
		}
		k_fill_residues<<<1024, 256, 0, s>>>(data, polys * n, q, logn);
		std::vector<ulonglong2> htw(n);
		for (uint64_t i = 0; i < n; i++) {
			const uint64_t w = (0x9e3779b97f4a7c15ull * (i + 1)) % q;
			htw[i] = make_ulonglong2(w, shoup63_factor(w, q));
		}
		NttConsts c{};
		c.q = q;
		c.two_q = 2 * q;
		c.r64 = pow2_64_mod(q);
		c.mu = barrett_mu(q);
		c.bshift = (uint32_t) (bitlen(q) - 2);
		c.logn = logn;
		c.ninv = invmod_euclid(n, q);
		c.ninv_p = shoup63_factor(c.ninv, q);
		c.w1n = c.ninv;
		c.w1n_p = c.ninv_p;
		small_reduce_consts(q, c.red_shift, c.red_m);
		hipEvent_t e0, e1;
		if (!hip_ok(hipMemcpyAsync(tw, htw.data(), n * sizeof(ulonglong2),
						hipMemcpyHostToDevice, s), "autotune upload") ||
				!hip_ok(hipEventCreate(&e0), "event") ||
				!hip_ok(hipEventCreate(&e1), "event")) {
			(void) hipFree(data);
			(void) hipFree(tw);
			return vkhel_ext_last_error();
		}
		DevTables t;
		t.n = n;
		t.fwd = tw;
		t.inv = tw;
		t.c = c;
		const int saved = ctx->tile_bits[logn];
		const int saved_mode = ctx->pass_mode[logn];
		// Candidates: tile 2^11 / 2^12 / 2^13 x whole-polynomial / split
		// passes (one-tile sizes have a single kernel).  Each is timed in
		// several interleaved rounds and keeps its fastest, so a transient
		// slowdown of the device does not decide the choice.
		int cand_tb[2 * kNumTiles], cand_mode[2 * kNumTiles], ncand = 0;
		for (int choice = 0; choice < 2 * kNumTiles; choice++) {
			const int tb = kTileChoices[choice % kNumTiles];
			const int mode = 1 + choice / kNumTiles;  // whole polynomial, then split
			// one-tile sizes have one kernel (the same in every table that
			// holds the size in one tile): time it once
			const int one_tile_tb = logn <= 13 ? 13 : 14;
			if (logn <= (uint32_t) tb && (mode == 1 || tb != one_tile_tb))
				continue;
			cand_tb[ncand] = tb;
			cand_mode[ncand++] = mode;
		}
		float cand_ms[2 * kNumTiles];
		for (int i = 0; i < ncand; i++)
			cand_ms[i] = 1e30f;
		int rc = VKHEL_OK;
		for (int round = 0; round < kTuneRounds && rc == VKHEL_OK; round++) {
			for (int i = 0; i < ncand && rc == VKHEL_OK; i++) {
				ctx->tile_bits[logn] = cand_tb[i];
				ctx->pass_mode[logn] = cand_mode[i];
				// warm-up
				rc = launch_ntt(ctx, t, data, data, polys, false, s);
				if (rc == VKHEL_OK)
					rc = launch_ntt(ctx, t, data, data, polys, true, s);
				(void) hipEventRecord(e0, s);
				const int reps = 3;
				for (int rep = 0; rep < reps && rc == VKHEL_OK; rep++) {
					rc = launch_ntt(ctx, t, data, data, polys, false, s);
					if (rc == VKHEL_OK)
						rc = launch_ntt(ctx, t, data, data, polys, true, s);
				}
				(void) hipEventRecord(e1, s);
				float ms = 0;
				if (rc != VKHEL_OK || !hip_ok(hipEventSynchronize(e1), "sync") ||
						!hip_ok(hipEventElapsedTime(&ms, e0, e1), "elapsed")) {
					rc = rc != VKHEL_OK ? rc : vkhel_ext_last_error();
					break;
				}
				cand_ms[i] = std::min(cand_ms[i], ms);
			}
		}
		if (rc != VKHEL_OK) {
			ctx->tile_bits[logn] = saved;
			ctx->pass_mode[logn] = saved_mode;
			(void) hipEventDestroy(e0);
			(void) hipEventDestroy(e1);
			(void) hipFree(data);
			(void) hipFree(tw);
			return rc;
		}
		int best = 0;
		for (int i = 1; i < ncand; i++)
			if (cand_ms[i] < cand_ms[best])
				best = i;
		const int best_tb = cand_tb[best], best_mode = cand_mode[best];
		ctx->tile_bits[logn] = best_tb;
		ctx->pass_mode[logn] = best_mode;
		(void) hipEventDestroy(e0);
		(void) hipEventDestroy(e1);
		(void) hipFree(data);
		(void) hipFree(tw);
	}
	return VKHEL_OK;
}

// Text format: one "log2n tile_bits pass_mode" triple per line.
int vkhel_ext_autotune_save(struct vkhel_ctx *ctx, const char *path) {
	if (!ctx || !path)
		return fail(VKHEL_ERR_INVALID, "autotune_save: null argument");
	FILE *f = fopen(path, "w");
	if (!f)
		return fail(VKHEL_ERR_INVALID, "autotune_save: cannot open %s", path);
	for (uint32_t l = 0; l < ctx->tile_bits.size(); l++)
		if (ctx->tile_bits[l] || ctx->pass_mode[l])
			fprintf(f, "%u %d %d\n", l, ctx->tile_bits[l], ctx->pass_mode[l]);
	fclose(f);
	return VKHEL_OK;
}

int vkhel_ext_autotune_load(struct vkhel_ctx *ctx, const char *path) {
	if (!ctx || !path)
		return fail(VKHEL_ERR_INVALID, "autotune_load: null argument");
	FILE *f = fopen(path, "r");
	if (!f)
		return fail(VKHEL_ERR_INVALID, "autotune_load: cannot open %s", path);
	unsigned l;
	int tb, mode;
	int rc = VKHEL_OK;
	while (fscanf(f, "%u %d %d", &l, &tb, &mode) == 3) {
		if (l >= ctx->tile_bits.size() || (tb != 0 && (tb < 11 || tb > 14)) ||
				mode < 0 || mode > 3) {
			rc = fail(VKHEL_ERR_INVALID, "autotune_load: bad entry %u %d %d", l, tb,
					mode);
			break;
		}
		ctx->tile_bits[l] = tb;
		ctx->pass_mode[l] = mode;
	}
	fclose(f);
	return rc;
}

}  // extern "C"
