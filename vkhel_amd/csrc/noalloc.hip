// Polynomial products that allocate no device memory, for callers that
// cannot call hipMalloc: code under HIP graph capture, tight serving loops,
// harnesses that forbid allocation.  The batch path's POLYMUL job keeps b's
// spectrum in scratch it allocates (batch.hip); here that scratch is a
// static device array, one per GPU, of a fixed maximum size, and every
// launch takes its task by value, so no task array is uploaded either.
//
// The array lives on each device for the life of the process.  Host code
// never passes its name to a kernel (in host code the name is a host-side
// stand-in, not a device address): each device's address is looked up once
// with hipGetSymbolAddress, with that device current, when a context is
// created on it, and only that address is used afterwards.
#include <mutex>

#include "internal.h"
#include "numbers.h"

namespace vk {

// b's spectra: nlimbs * n words per call, at most 2^22 words (32 MiB).
constexpr uint64_t kNoAllocScratchWords = 1ull << 22;
static_assert(kNoAllocScratchWords == VKHEL_POLYMUL_NOALLOC_MAX_WORDS,
		"the header states the scratch maximum");

static __device__ __align__(256) uint64_t scratch[kNoAllocScratchWords];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static uint64_t *g_scratch_ptr[kMaxDevices];

int noalloc_init(int device) {
	if (device < 0 || device >= kMaxDevices)
		return fail(VKHEL_ERR_INVALID, "scratch: device %d out of range", device);
	std::lock_guard<std::mutex> g(g_scratch_mu);
	if (g_scratch_ptr[device])
		return VKHEL_OK;
	void *p = nullptr;
	if (!hip_ok(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)),
				"hipGetSymbolAddress(scratch)"))
		return vkhel_ext_last_error();
	if (!p)
		return fail(VKHEL_ERR_DEVICE, "scratch: null device address");
	g_scratch_ptr[device] = (uint64_t *) p;
	return VKHEL_OK;
}

static uint64_t *scratch_ptr(int device) {
	std::lock_guard<std::mutex> g(g_scratch_mu);
	return device >= 0 && device < kMaxDevices ? g_scratch_ptr[device] : nullptr;
}

// The device tables of `t` on ctx if they are already resident (generating
// them would allocate), else nullptr with the error recorded.
static const DevTables *resident_tables(vkhel_ctx *ctx,
		const vkhel_ntt_tables *t) {
	std::lock_guard<std::mutex> g(ctx->mu);
	auto it = ctx->tables.find(t->id);
	if (it == ctx->tables.end())
		return fail(VKHEL_ERR_INVALID, "polymul_noalloc: tables (n=%llu, q=%llu) "
				"are not resident on this context; upload them first with "
				"vkhel_ext_tables_upload", (unsigned long long) t->n,
				(unsigned long long) t->q), nullptr;
	return it->second;
}

// One limb of a*b mod (X^n + 1, q), the same kernels and order as a batch's
// POLYMUL job (batch.hip plan_tasks), with b's spectrum in `tmp`.
static int polymul_limb(vkhel_ctx *ctx, const DevTables &t, const uint64_t *a,
		const uint64_t *b, uint64_t *res, uint64_t *tmp, hipStream_t st) {
	const uint32_t logn = t.c.logn;
	const int tb = polymul_fused_tile(ctx, logn);
	if (tb) {
		// both forwards (after their column passes, a's in place in res and
		// b's into tmp, for two-pass sizes), the product and the inverse in
		// the fused kernels
		const bool cols = logn > (uint32_t) tb;
		const NttTask col_a = make_task(t, a, res, false);
		const NttTask col_b = make_task(t, b, tmp, false);
		NttTask pm = make_task(t, cols ? res : a, res, true);
		pm.in2 = cols ? tmp : b;
		pm.tw2 = t.fwd;
		return launch_polymul_one(col_a, col_b, pm, cols, logn, tb,
				modulus_class(t.c.q), st);
	}
	int rc = launch_ntt(ctx, t, b, tmp, 1, false, st);
	if (rc != VKHEL_OK)
		return rc;
	rc = launch_ntt(ctx, t, a, res, 1, false, st);
	if (rc != VKHEL_OK)
		return rc;
	if (ntt_fused_supported(logn)) {
		// the fused inverse multiplies the two spectra as it loads them
		NttTask inv = make_task(t, res, res, true);
		inv.in2 = tmp;
		return launch_task_fused(ctx, inv, true, st);
	}
	rc = launch_elemmul(res, tmp, res, t.n, t.c.q, st);
	if (rc != VKHEL_OK)
		return rc;
	return launch_ntt(ctx, t, res, res, 1, true, st);
}

}  // namespace vk

using namespace vk;

extern "C" int vkhel_ext_polymul_noalloc(struct vkhel_ctx *ctx,
		const struct vkhel_vector *a, const struct vkhel_vector *b,
		struct vkhel_vector *res, struct vkhel_ntt_tables *const *tables,
		uint32_t nlimbs, void *stream) {
	if (!ctx || !a || !b || !res || !tables || nlimbs == 0)
		return fail(VKHEL_ERR_INVALID, "polymul_noalloc: null argument");
	if (a->ctx != ctx || b->ctx != ctx || res->ctx != ctx)
		return fail(VKHEL_ERR_INVALID,
				"polymul_noalloc: vector from another context");
	if (b == res)
		return fail(VKHEL_ERR_INVALID, "polymul_noalloc: result aliases b");
	for (uint32_t i = 0; i < nlimbs; i++)
		if (!tables[i] || tables[i]->n != tables[0]->n)
			return fail(VKHEL_ERR_INVALID,
					"polymul_noalloc: limb tables must all exist and share n");
	const uint64_t n = tables[0]->n;
	const uint64_t words = (uint64_t) nlimbs * n;
	if (words > kNoAllocScratchWords)
		return fail(VKHEL_ERR_INVALID,
				"polymul_noalloc: %u limbs x n=%llu = %llu words exceed the "
				"scratch maximum of %llu words", nlimbs, (unsigned long long) n,
				(unsigned long long) words,
				(unsigned long long) kNoAllocScratchWords);
	if (a->length < words || b->length < words || res->length < words)
		return fail(VKHEL_ERR_INVALID,
				"polymul_noalloc: vectors need %llu words (%u limbs x n)",
				(unsigned long long) words, nlimbs);
	uint64_t *sp = scratch_ptr(ctx->device);
	if (!sp)
		return fail(VKHEL_ERR_DEVICE,
				"polymul_noalloc: scratch of device %d not set up", ctx->device);
	// every limb's tables resident before anything is enqueued
	for (uint32_t l = 0; l < nlimbs; l++)
		if (!resident_tables(ctx, tables[l]))
			return vkhel_ext_last_error();
	hipStream_t st = stream ? (hipStream_t) stream : ctx->stream;
	for (uint32_t l = 0; l < nlimbs; l++) {
		const DevTables *t = resident_tables(ctx, tables[l]);
		if (!t)
			return vkhel_ext_last_error();
		const uint64_t off = l * n;
		const int rc = polymul_limb(ctx, *t, a->dev + off, b->dev + off,
				res->dev + off, sp + off, st);
		if (rc != VKHEL_OK)
			return rc;
	}
	return VKHEL_OK;
}
