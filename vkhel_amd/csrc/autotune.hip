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

static const int kTuneRounds = 4;

static const int kTileChoices[] = {11, 12, 13, 14};
constexpr int kNumTiles = 4;

// One tuning choice for a size: pass-1 tile, pass mode, stage split and
// row-pass tile (0: the pass-1 tile).
struct Candidate {
	int tb, mode, split, tb2;
	void apply(vkhel_ctx *ctx, uint32_t logn) const {
		ctx->tile_bits[logn] = tb;
		ctx->pass_mode[logn] = mode;
		ctx->split[logn] = split;
		ctx->tile_bits2[logn] = tb2;
	}
};

// A synthetic batch of one size on the device, timed through the batch
// path (device task array, one-launch state) that batches and serving use.
struct Tuning {
	uint64_t *data = nullptr;
	ulonglong2 *tw = nullptr;
	NttTask *tasks = nullptr;
	uint32_t *state = nullptr;
	uint64_t state_bytes = 0;
	uint32_t polys = 0, logn = 0;
	int mc = 0;
	hipEvent_t e0 = nullptr, e1 = nullptr;

	int setup(vkhel_ctx *ctx, uint32_t lg, uint32_t np, uint64_t q) {
		logn = lg;
		polys = np;
		const uint64_t n = 1ull << logn;
		hipStream_t s = ctx->stream;
		state_bytes = handoff_state_bytes(polys);
		if (!hip_ok(hipMalloc(&data, polys * n * 8), "autotune alloc") ||
				!hip_ok(hipMalloc(&tw, n * sizeof(ulonglong2)), "autotune alloc") ||
				!hip_ok(hipMalloc(&tasks, polys * sizeof(NttTask)), "autotune alloc") ||
				!hip_ok(hipMalloc(&state, state_bytes), "autotune alloc") ||
				!hip_ok(hipEventCreate(&e0), "event") ||
				!hip_ok(hipEventCreate(&e1), "event"))
			return release(), vkhel_ext_last_error();
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
		mc = modulus_class(q);
		DevTables t;
		t.n = n;
		t.fwd = tw;
		t.inv = tw;
		t.c = c;
		std::vector<NttTask> h(polys);
		for (uint32_t i = 0; i < polys; i++)
			h[i] = make_task(t, data + i * n, data + i * n, false);
		if (!hip_ok(hipMemcpyAsync(tw, htw.data(), n * sizeof(ulonglong2),
						hipMemcpyHostToDevice, s), "autotune upload") ||
				!hip_ok(hipMemcpyAsync(tasks, h.data(), polys * sizeof(NttTask),
						hipMemcpyHostToDevice, s), "autotune upload") ||
				!hip_ok(hipStreamSynchronize(s), "sync"))
			return release(), vkhel_ext_last_error();
		return VKHEL_OK;
	}

	void release() {
		(void) hipFree(data);
		(void) hipFree(tw);
		(void) hipFree(tasks);
		(void) hipFree(state);
		if (e0)
			(void) hipEventDestroy(e0);
		if (e1)
			(void) hipEventDestroy(e1);
		data = nullptr;
		tw = nullptr;
		tasks = nullptr;
		state = nullptr;
		e0 = e1 = nullptr;
	}

	int run(vkhel_ctx *ctx, bool inverse) {
		return launch_fused_tasks(ctx, tasks, polys, logn, inverse, mc,
				ctx->stream, state, state_bytes);
	}

	static bool has_alt_split(uint32_t lg, int tb) {
		return fused_split_k1(lg, tb, 1) != fused_split_k1(lg, tb, 0);
	}

	// A row-pass tile tb2 can follow pass-1 tile tb: both tables split the
	// size in two passes with the same stage split.
	static bool rows_ok(uint32_t lg, int tb, int tb2, int sp) {
		return lg > (uint32_t) tb2 && fused_split_k1(lg, tb2, sp) ==
			fused_split_k1(lg, tb, sp) && (sp == 0 || has_alt_split(lg, tb));
	}

	// Times every candidate (forward + inverse, three repetitions after a
	// warm-up) in kTuneRounds interleaved rounds, keeping each one's fastest
	// round so that a transient slowdown does not decide; a one-launch
	// candidate whose schedule check fails is dropped.  Leaves the fastest
	// in *best.
	int race(vkhel_ctx *ctx, const std::vector<Candidate> &cands, int *best) {
		std::vector<float> ms(cands.size(), 1e30f);
		std::vector<bool> bad(cands.size(), false);
		hipStream_t s = ctx->stream;
		for (int round = 0; round < kTuneRounds; round++) {
			for (size_t i = 0; i < cands.size(); i++) {
				if (bad[i])
					continue;
				cands[i].apply(ctx, logn);
				int rc = run(ctx, false);
				if (rc == VKHEL_OK)
					rc = run(ctx, true);
				(void) hipEventRecord(e0, s);
				for (int rep = 0; rep < 3 && rc == VKHEL_OK; rep++) {
					rc = run(ctx, false);
					if (rc == VKHEL_OK)
						rc = run(ctx, true);
				}
				(void) hipEventRecord(e1, s);
				float t = 0;
				if (rc != VKHEL_OK || !hip_ok(hipEventSynchronize(e1), "sync") ||
						!hip_ok(hipEventElapsedTime(&t, e0, e1), "elapsed"))
					return rc != VKHEL_OK ? rc : vkhel_ext_last_error();
				if (cands[i].mode == 3 && uses_handoff(ctx, logn) &&
						!handoff_state_ok(ctx, state, polys, logn)) {
					bad[i] = true;
					continue;
				}
				ms[i] = std::min(ms[i], t);
			}
		}
		*best = -1;
		for (size_t i = 0; i < cands.size(); i++)
			if (!bad[i] && (*best < 0 || ms[i] < ms[*best]))
				*best = (int) i;
		if (*best < 0)
			return fail(VKHEL_ERR_DEVICE, "autotune: no candidate ran correctly");
		return VKHEL_OK;
	}
};

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
		Tuning tu;
		int rc = tu.setup(ctx, logn, polys, q);
		if (rc != VKHEL_OK)
			return rc;
		const int saved_tb = ctx->tile_bits[logn], saved_mode = ctx->pass_mode[logn];
		const int saved_split = ctx->split[logn], saved_tb2 = ctx->tile_bits2[logn];
		ctx->split[logn] = 0;
		ctx->tile_bits2[logn] = 0;
		// 1. tile x pass mode: whole polynomial per workgroup (1), a grid
		// per pass (2), both passes in one launch (3, two-pass sizes)
		std::vector<Candidate> cands;
		for (int i = 0; i < kNumTiles; i++) {
			const int tb = kTileChoices[i];
			for (int mode = 1; mode <= 3; mode++) {
				// one-tile sizes have one kernel (the same in every table
				// that holds the size in one tile): time it once
				const int one_tile_tb = logn <= 13 ? 13 : 14;
				if (logn <= (uint32_t) tb && (mode != 2 || tb != one_tile_tb))
					continue;
				cands.push_back(Candidate{tb, mode, 0, 0});
			}
		}
		int best = -1;
		rc = tu.race(ctx, cands, &best);
		// 2. for split passes: the alternative stage split, and a row-pass
		// tile of its own
		if (rc == VKHEL_OK && cands[best].mode == 2 &&
				logn > (uint32_t) cands[best].tb) {
			const Candidate b = cands[best];
			std::vector<Candidate> more{b};
			if (tu.has_alt_split(logn, b.tb))
				more.push_back(Candidate{b.tb, 2, 1, 0});
			for (int i = 0; i < kNumTiles; i++) {
				const int tb2 = kTileChoices[i];
				for (int sp = 0; sp <= 1; sp++)
					if (tb2 != b.tb && tu.rows_ok(logn, b.tb, tb2, sp))
						more.push_back(Candidate{b.tb, 2, sp, tb2});
			}
			int best2 = 0;
			if (more.size() > 1)
				rc = tu.race(ctx, more, &best2);
			if (rc == VKHEL_OK)
				cands[best] = more[best2];
		}
		tu.release();
		if (rc != VKHEL_OK) {
			ctx->tile_bits[logn] = saved_tb;
			ctx->pass_mode[logn] = saved_mode;
			ctx->split[logn] = saved_split;
			ctx->tile_bits2[logn] = saved_tb2;
			return rc;
		}
		cands[best].apply(ctx, logn);
	}
	return VKHEL_OK;
}

// Text format: one "log2n tile_bits pass_mode split tile_bits2" line per
// tuned size (files with three fields per line, from before the split and
// row-pass tile were tuned, load too).
int vkhel_ext_autotune_save(struct vkhel_ctx *ctx, const char *path) {
	if (!ctx || !path)
		return fail(VKHEL_ERR_INVALID, "autotune_save: null argument");
	FILE *f = fopen(path, "w");
	if (!f)
		return fail(VKHEL_ERR_INVALID, "autotune_save: cannot open %s", path);
	for (uint32_t l = 0; l < ctx->tile_bits.size(); l++)
		if (ctx->tile_bits[l] || ctx->pass_mode[l] || ctx->split[l] ||
				ctx->tile_bits2[l])
			fprintf(f, "%u %d %d %d %d\n", l, ctx->tile_bits[l], ctx->pass_mode[l],
					ctx->split[l], ctx->tile_bits2[l]);
	fclose(f);
	return VKHEL_OK;
}

int vkhel_ext_autotune_load(struct vkhel_ctx *ctx, const char *path) {
	if (!ctx || !path)
		return fail(VKHEL_ERR_INVALID, "autotune_load: null argument");
	FILE *f = fopen(path, "r");
	if (!f)
		return fail(VKHEL_ERR_INVALID, "autotune_load: cannot open %s", path);
	int rc = VKHEL_OK;
	char line[256];
	while (fgets(line, sizeof line, f)) {
		unsigned l;
		int tb, mode, sp = 0, tb2 = 0;
		const int got = sscanf(line, "%u %d %d %d %d", &l, &tb, &mode, &sp, &tb2);
		if (got <= 0)
			continue;
		if ((got != 3 && got != 5) || l >= ctx->tile_bits.size() ||
				(tb != 0 && (tb < 11 || tb > 14)) || mode < 0 || mode > 3 ||
				sp < 0 || sp > 1 || (tb2 != 0 && (tb2 < 11 || tb2 > 14))) {
			rc = fail(VKHEL_ERR_INVALID, "autotune_load: bad entry \"%s\"", line);
			break;
		}
		ctx->tile_bits[l] = tb;
		ctx->pass_mode[l] = mode;
		ctx->split[l] = sp;
		ctx->tile_bits2[l] = tb2;
	}
	fclose(f);
	return rc;
}

int vkhel_ext_split(struct vkhel_ctx *ctx, uint32_t logn) {
	if (!ctx || logn >= 64)
		return fail(VKHEL_ERR_INVALID, "split: bad argument"), -1;
	return ctx->split[logn];
}

int vkhel_ext_set_split(struct vkhel_ctx *ctx, uint32_t logn, int split) {
	if (!ctx || logn >= 64 || split < 0 || split > 1)
		return fail(VKHEL_ERR_INVALID, "set_split: split must be 0 or 1");
	ctx->split[logn] = split;
	return VKHEL_OK;
}

int vkhel_ext_tile_bits2(struct vkhel_ctx *ctx, uint32_t logn) {
	if (!ctx || logn >= 64)
		return fail(VKHEL_ERR_INVALID, "tile_bits2: bad argument"), -1;
	return ctx->tile_bits2[logn];
}

int vkhel_ext_set_tile_bits2(struct vkhel_ctx *ctx, uint32_t logn,
		int tile_bits) {
	if (!ctx || logn >= 64 || (tile_bits != 0 && (tile_bits < 11 ||
					tile_bits > 14)))
		return fail(VKHEL_ERR_INVALID, "set_tile_bits2: tile_bits must be 0 or "
				"11..14");
	ctx->tile_bits2[logn] = tile_bits;
	return VKHEL_OK;
}

}  // extern "C"
