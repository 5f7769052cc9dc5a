// Negacyclic NTT / INTT of vkhel (src/vector.c:513-657) for gfx950.
//
// The reference issues one dispatch per butterfly group (n - 1 dispatches
// and log2(n) fenced submits per transform).  Here the whole butterfly
// network of a polynomial runs inside one workgroup:
//
//   * n <= 2^13: the polynomial is staged once in LDS and every stage runs
//     there, three stages per register round (radix-8 per thread).
//   * n >  2^13: two tiled passes.  Pass A runs the top k1 stages on
//     2^k1-row x 2^cb-column tiles, pass B the low k2 stages on 2^rb-row x
//     2^k2 tiles (k1 + k2 = log2 n, every tile 2^13 words).  Pass B reads
//     what pass A wrote through the same workgroup, so __syncthreads() is
//     the only synchronisation and the intermediate stays in L2/MALL.
//
// Twiddles come from the HBM table in the reference's order (roots[m + i],
// src/ntt_tables.c), as {w, w'} Shoup pairs.  Results are canonical and
// equal the reference's stage-by-stage output bit for bit: every butterfly
// is exact mod q, so only the order of independent butterflies changes.
#include "internal.h"
#include "modarith.h"

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
		const ulonglong2 w = tw[((uint64_t) 1 << s) + grp];
		const uint64_t X = in[x], Y = in[y];
		const uint64_t wy = shoup(Y, w.x, w.y, c.q);
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
			out[x] = shoup(sum, c.ninv, c.ninv_p, c.q);
			out[y] = shoup(d, c.w1n, c.w1n_p, c.q);
		} else {
			const ulonglong2 w = tw[((uint64_t) 1 << s) + grp];
			out[x] = sum;
			out[y] = shoup(d, w.x, w.y, c.q);
		}
	}
}

static int launch_radix2(const DevTables &t, const uint64_t *in, uint64_t *out,
		uint32_t batch, bool inverse, hipStream_t st) {
	const uint32_t logn = t.c.logn;
	const uint64_t total = (uint64_t) batch << (logn - 1);
	const unsigned grid = (unsigned) std::min<uint64_t>((total + 255) / 256,
			16384);
	if (!inverse) {
		for (uint32_t s = 0; s < logn; s++) {
			if (s == 0)
				hipLaunchKernelGGL(k_fwd_stage<true>, dim3(grid), dim3(256), 0, st,
						in, out, t.fwd, t.c, s, total);
			else
				hipLaunchKernelGGL(k_fwd_stage<false>, dim3(grid), dim3(256), 0, st,
						out, out, t.fwd, t.c, s, total);
		}
	} else {
		for (int s = (int) logn - 1; s >= 0; s--) {
			const uint64_t *src = (s == (int) logn - 1) ? in : out;
			if (s == 0)
				hipLaunchKernelGGL(k_inv_stage<true>, dim3(grid), dim3(256), 0, st,
						src, out, t.inv, t.c, 0u, total);
			else
				hipLaunchKernelGGL(k_inv_stage<false>, dim3(grid), dim3(256), 0, st,
						src, out, t.inv, t.c, (uint32_t) s, total);
		}
	}
	return hip_ok(hipGetLastError(), "ntt radix2 launch") ? VKHEL_OK
		: VKHEL_ERR_DEVICE;
}

// ------------------------------------------------------------------ fused

constexpr int kTileBits = 13;                // 8192 words per tile
constexpr int kTileThreads = 1 << (kTileBits - 3);
constexpr int kTileLds = (1 << kTileBits) + (1 << (kTileBits - 3));

// One padding word per 8 keeps radix-8 strides bank-conflict free.
__device__ __forceinline__ uint32_t pad(uint32_t l) {
	return l + (l >> 3);
}

// Where a tile lives in the polynomial and which stages it runs.
struct Tile {
	uint32_t kt;     // log2 tile size
	uint32_t cbits;  // low local bits that are contiguous in memory
	uint32_t jshift; // memory shift of the bits above cbits
	uint64_t jbase;  // memory offset of local index 0
	uint32_t shift;  // global stage bit = local bit + shift
	uint64_t gbase;  // twiddle offset (pass B: row block start)
	uint32_t lo, hi; // local stage bits [lo, hi) to run
};

__device__ __forceinline__ uint64_t tile_addr(const Tile &g, uint32_t l) {
	if (g.cbits >= g.kt)
		return g.jbase + l;
	return ((uint64_t) (l >> g.cbits) << g.jshift) +
		(l & ((1u << g.cbits) - 1)) + g.jbase;
}

// Local index of element e (3 bits) of thread t in a round whose 3-bit
// field sits at local bit f.
__device__ __forceinline__ uint32_t insert3(uint32_t t, uint32_t e,
		uint32_t f) {
	return ((t >> f) << (f + 3)) | (e << f) | (t & ((1u << f) - 1));
}

template <bool INV>
__device__ __forceinline__ void butterfly(uint64_t &x, uint64_t &y,
		ulonglong2 w, const NttConsts &c, bool edge) {
	if (!INV) {
		const uint64_t X = x;
		const uint64_t wy = shoup(y, w.x, w.y, c.q);
		if (edge) {  // first forward stage: arbitrary 64-bit inputs
			const uint64_t xr = barrett64(X, c.q, c.mu, c.bshift);
			uint64_t ox = csub(xr + wy, c.q);
			if (X + wy < X)
				ox = sub_mod(ox, c.r64, c.q);
			x = ox;
			y = sub_mod(xr, wy, c.q);
		} else {
			x = csub(X + wy, c.q);
			y = sub_mod(X, wy, c.q);
		}
	} else {
		const uint64_t sum = csub(x + y, c.q), d = sub_mod(x, y, c.q);
		if (edge) {  // last inverse stage, times n^-1
			x = shoup(sum, c.ninv, c.ninv_p, c.q);
			y = shoup(d, c.w1n, c.w1n_p, c.q);
		} else {
			x = sum;
			y = shoup(d, w.x, w.y, c.q);
		}
	}
}

// Runs the stages of local bits [blo, bhi) (a subset of the 3-bit field at
// f) on this thread's 8 elements.
template <bool INV>
__device__ __forceinline__ void round3(uint64_t v[8], uint32_t t, uint32_t f,
		uint32_t blo, uint32_t bhi, const Tile &g, const ulonglong2 *tw,
		const NttConsts &c) {
#pragma unroll
	for (int k = 0; k < 3; k++) {
		const int d = INV ? k : 2 - k;  // field bit of this stage
		const uint32_t lb = f + d;
		if (lb < blo || lb >= bhi)
			continue;
		const uint32_t b = lb + g.shift;  // global stage bit
		const bool edge = b == c.logn - 1;
		const uint64_t m = (uint64_t) 1 << (c.logn - 1 - b);
#pragma unroll
		for (int gi = 0; gi < (4 >> d); gi++) {
			const uint32_t l0 = insert3(t, (uint32_t) (gi << (d + 1)), f);
			const ulonglong2 w = tw[m + ((g.gbase + l0) >> (lb + 1))];
#pragma unroll
			for (int li = 0; li < (1 << d); li++) {
				const int e = (gi << (d + 1)) | li;
				butterfly<INV>(v[e], v[e | (1 << d)], w, c, edge);
			}
		}
	}
}

template <bool INV>
__device__ void run_tile(const uint64_t *src, uint64_t *dst, uint64_t *lds,
		const Tile &g, const ulonglong2 *tw, const NttConsts &c) {
	const uint32_t t = threadIdx.x, nt = blockDim.x;
	const uint32_t size = 1u << g.kt;
	for (uint32_t p = t; p < size / 2; p += nt) {
		const uint32_t l = 2 * p;
		const ulonglong2 x = *(const ulonglong2 *) (src + tile_addr(g, l));
		lds[pad(l)] = x.x;
		lds[pad(l + 1)] = x.y;
	}
	__syncthreads();
	const int span = (int) (g.hi - g.lo);
	const int rounds = (span + 2) / 3;
	for (int r = 0; r < rounds; r++) {
		uint32_t blo, bhi;
		if (!INV) {  // top-down
			bhi = g.hi - 3 * r;
			blo = bhi >= g.lo + 3 ? bhi - 3 : g.lo;
		} else {  // bottom-up
			blo = g.lo + 3 * r;
			bhi = blo + 3 <= g.hi ? blo + 3 : g.hi;
		}
		const uint32_t f = blo + 3 <= g.kt ? blo : g.kt - 3;
		if (t < (size >> 3)) {
			uint64_t v[8];
#pragma unroll
			for (int e = 0; e < 8; e++)
				v[e] = lds[pad(insert3(t, e, f))];
			round3<INV>(v, t, f, blo, bhi, g, tw, c);
#pragma unroll
			for (int e = 0; e < 8; e++)
				lds[pad(insert3(t, e, f))] = v[e];
		}
		__syncthreads();
	}
	for (uint32_t p = t; p < size / 2; p += nt) {
		const uint32_t l = 2 * p;
		*(ulonglong2 *) (dst + tile_addr(g, l)) =
			make_ulonglong2(lds[pad(l)], lds[pad(l + 1)]);
	}
	__syncthreads();
}

// Pass A geometry: stages of the top k1 global bits on column chunk ci.
__device__ __forceinline__ Tile tile_cols(uint32_t logn, uint32_t k1,
		uint32_t ci) {
	const uint32_t k2 = logn - k1, cb = kTileBits - k1;
	Tile g;
	g.kt = kTileBits;
	g.cbits = cb;
	g.jshift = k2;
	g.jbase = (uint64_t) ci << cb;
	g.shift = k2 - cb;
	g.gbase = 0;
	g.lo = cb;
	g.hi = kTileBits;
	return g;
}

// Pass B geometry: stages of the low k2 global bits on row block hi.
__device__ __forceinline__ Tile tile_rows(uint32_t logn, uint32_t k1,
		uint32_t hi) {
	const uint32_t k2 = logn - k1, rb = kTileBits - k2;
	Tile g;
	g.kt = kTileBits;
	g.cbits = kTileBits;
	g.jshift = 0;
	g.jbase = (uint64_t) hi << (rb + k2);
	g.shift = 0;
	g.gbase = g.jbase;
	g.lo = 0;
	g.hi = k2;
	return g;
}

// Fused path limits: 2^3 <= n <= 2^24 (two passes of at most 12 + 12 stages
// with at least one contiguous column per pass-A tile row).
static bool fused_supported(uint32_t logn) {
	return logn >= 3 && logn <= 24;
}

NttTask make_task(const DevTables &t, const uint64_t *in, uint64_t *out,
		bool inverse) {
	NttTask k;
	k.in = in;
	k.out = out;
	k.tw = inverse ? t.inv : t.fwd;
	k.c = t.c;
	k.k1 = t.c.logn <= (uint32_t) kTileBits ? 0 : t.c.logn / 2;
	return k;
}

// One workgroup per task.  With `tasks` null, workgroup i transforms the
// i-th polynomial after `uniform` (same tables, contiguous polynomials).
template <bool INV>
__global__ __launch_bounds__(kTileThreads) void k_ntt_fused(NttTask uniform,
		const NttTask *tasks) {
	__shared__ uint64_t lds[kTileLds];
	NttTask k;
	if (tasks) {
		k = tasks[blockIdx.x];
	} else {
		k = uniform;
		const uint64_t off = (uint64_t) blockIdx.x << k.c.logn;
		k.in += off;
		k.out += off;
	}
	const NttConsts &c = k.c;
	const uint32_t logn = c.logn;
	if (logn <= (uint32_t) kTileBits) {
		Tile g;
		g.kt = logn;
		g.cbits = logn;
		g.jshift = 0;
		g.jbase = 0;
		g.shift = 0;
		g.gbase = 0;
		g.lo = 0;
		g.hi = logn;
		run_tile<INV>(k.in, k.out, lds, g, k.tw, c);
		return;
	}
	const uint32_t k1 = k.k1, k2 = logn - k1;
	const uint32_t ncol = 1u << (k2 - (kTileBits - k1));
	const uint32_t nrow = 1u << (k1 - (kTileBits - k2));
	if (!INV) {
		for (uint32_t i = 0; i < ncol; i++)
			run_tile<INV>(k.in, k.out, lds, tile_cols(logn, k1, i), k.tw, c);
		for (uint32_t i = 0; i < nrow; i++)
			run_tile<INV>(k.out, k.out, lds, tile_rows(logn, k1, i), k.tw, c);
	} else {
		for (uint32_t i = 0; i < nrow; i++)
			run_tile<INV>(k.in, k.out, lds, tile_rows(logn, k1, i), k.tw, c);
		for (uint32_t i = 0; i < ncol; i++)
			run_tile<INV>(k.out, k.out, lds, tile_cols(logn, k1, i), k.tw, c);
	}
}

static uint32_t fused_threads(uint32_t logn) {
	return logn >= (uint32_t) kTileBits ? kTileThreads
		: std::max<uint32_t>(1u << (logn - 3), 64);
}

int launch_fused_tasks(const NttTask *dev_tasks, uint32_t ntasks,
		uint32_t max_logn, bool inverse, hipStream_t st) {
	if (ntasks == 0)
		return VKHEL_OK;
	const NttTask none{};
	const dim3 block(fused_threads(max_logn));
	if (inverse)
		hipLaunchKernelGGL(k_ntt_fused<true>, dim3(ntasks), block, 0, st, none,
				dev_tasks);
	else
		hipLaunchKernelGGL(k_ntt_fused<false>, dim3(ntasks), block, 0, st, none,
				dev_tasks);
	return hip_ok(hipGetLastError(), "ntt fused launch") ? VKHEL_OK
		: VKHEL_ERR_DEVICE;
}

static int launch_fused(const DevTables &t, const uint64_t *in, uint64_t *out,
		uint32_t batch, bool inverse, hipStream_t st) {
	const NttTask k = make_task(t, in, out, inverse);
	const dim3 block(fused_threads(t.c.logn));
	if (inverse)
		hipLaunchKernelGGL(k_ntt_fused<true>, dim3(batch), block, 0, st, k,
				(const NttTask *) nullptr);
	else
		hipLaunchKernelGGL(k_ntt_fused<false>, dim3(batch), block, 0, st, k,
				(const NttTask *) nullptr);
	return hip_ok(hipGetLastError(), "ntt fused launch") ? VKHEL_OK
		: VKHEL_ERR_DEVICE;
}

int launch_ntt(vkhel_ctx *ctx, const DevTables &t, const uint64_t *in,
		uint64_t *out, uint32_t batch, bool inverse, hipStream_t s) {
	if (batch == 0)
		return VKHEL_OK;
	const uint32_t logn = t.c.logn;
	if (ctx->ntt_algo == VKHEL_NTT_RADIX2 || !fused_supported(logn))
		return launch_radix2(t, in, out, batch, inverse, s);
	return launch_fused(t, in, out, batch, inverse, s);
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
