// Pointwise modular ops of vkhel (src/vector.c:298-511) as streaming gfx950
// kernels.  Each is HBM-bound (16 or 24 bytes per element), so the kernels
// move two words per lane per access (dwordx4) over a grid-stride loop and
// keep the reference's per-element arithmetic exactly.
#include "internal.h"
#include "modarith.h"
#include "numbers.h"

namespace vk {

// elemmul.comp:62-73: Barrett64 on both inputs, Barrett128 on the product.
struct OpMul {
	uint64_t q, mu;
	uint32_t shift;
	__device__ uint64_t operator()(uint64_t x, uint64_t y) const {
		const uint64_t a = barrett64(x, q, mu, shift);
		const uint64_t b = barrett64(y, q, mu, shift);
		return barrett128(mulhi64(a, b), a * b, q, mu, shift);
	}
};

// elemfma.comp:36-55 including its `q - x` corrections (wrapping), so the
// result equals the reference's bit for bit on every input.
struct OpFma {
	uint64_t q, m, mp;
	__device__ uint64_t operator()(uint64_t x, uint64_t y) const {
		return elemfma_word(x, y, m, mp, q);
	}
};

// (x + y) mod q for x, y < q (the batch API's ADD op).
struct OpAddMod {
	uint64_t q;
	__device__ uint64_t operator()(uint64_t x, uint64_t y) const {
		return csub(x + y, q);
	}
};

// elemgtadd.comp:20-30
struct OpGtAdd {
	uint64_t bound, diff;
	__device__ uint64_t operator()(uint64_t x) const {
		return x > bound ? x + diff : x;
	}
};

// elemgtsub.comp:49-67
struct OpGtSub {
	uint64_t bound, diff, q, mu;
	uint32_t shift;
	__device__ uint64_t operator()(uint64_t x) const {
		const uint64_t r = barrett64(x, q, mu, shift);
		if (x > bound)
			return csub(r + q - barrett64(diff, q, mu, shift), q);
		return r;
	}
};

// elemmodbytwo.comp:19-27
struct OpModTwo {
	uint64_t sb;
	__device__ uint64_t operator()(uint64_t x) const {
		const uint64_t u = x & 1;
		return x > sb ? 1 - u : u;
	}
};

// VEC: every pointer is 16-byte aligned and two words move per access;
// otherwise (a vector view at an odd word offset) one word per access.
template <class Op, bool VEC>
__global__ __launch_bounds__(256) void k_binary(const uint64_t *a,
		const uint64_t *b, uint64_t *out, uint64_t len, Op op) {
	const uint64_t stride = (uint64_t) gridDim.x * blockDim.x;
	uint64_t i = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x;
	if constexpr (VEC) {
		for (const uint64_t pairs = len >> 1; i < pairs; i += stride) {
			const ulonglong2 x = ((const ulonglong2 *) a)[i];
			const ulonglong2 y = ((const ulonglong2 *) b)[i];
			((ulonglong2 *) out)[i] = make_ulonglong2(op(x.x, y.x), op(x.y, y.y));
		}
		if ((len & 1) && blockIdx.x == 0 && threadIdx.x == 0)
			out[len - 1] = op(a[len - 1], b[len - 1]);
	} else {
		for (; i < len; i += stride)
			out[i] = op(a[i], b[i]);
	}
}

template <class Op, bool VEC>
__global__ __launch_bounds__(256) void k_unary(const uint64_t *a,
		uint64_t *out, uint64_t len, Op op) {
	const uint64_t stride = (uint64_t) gridDim.x * blockDim.x;
	uint64_t i = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x;
	if constexpr (VEC) {
		for (const uint64_t pairs = len >> 1; i < pairs; i += stride) {
			const ulonglong2 x = ((const ulonglong2 *) a)[i];
			((ulonglong2 *) out)[i] = make_ulonglong2(op(x.x), op(x.y));
		}
		if ((len & 1) && blockIdx.x == 0 && threadIdx.x == 0)
			out[len - 1] = op(a[len - 1]);
	} else {
		for (; i < len; i += stride)
			out[i] = op(a[i]);
	}
}

static bool aligned16(const void *p) {
	return ((uintptr_t) p & 15) == 0;
}

// Synthetic residues: high half of splitmix64(seed, i) * q.
__global__ __launch_bounds__(256) void k_fill_uniform(uint64_t *p, uint64_t n,
		uint64_t q, uint64_t seed) {
	for (uint64_t i = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x; i < n;
			i += (uint64_t) gridDim.x * blockDim.x) {
		uint64_t z = seed + (i + 1) * 0x9e3779b97f4a7c15ull;
		z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
		z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
		z ^= z >> 31;
		p[i] = mulhi64(z, q);
	}
}

// One access per operand per thread over the whole vector where the grid
// allows (up to 2^20 workgroups): on MI355X, elemmul over 2^27 words ran
// at 5.2-5.3 TB/s with 65536 workgroups (four grid-stride iterations)
// against 4.5 TB/s with 8192, and unrolling several accesses per thread
// did not help.
static unsigned grid_for(uint64_t len) {
	const uint64_t blocks = ((len >> 1) + 255) / 256;
	return (unsigned) std::max<uint64_t>(1, std::min<uint64_t>(blocks, 1u << 20));
}

template <class Op>
int launch_binary(const uint64_t *a, const uint64_t *b, uint64_t *out,
		uint64_t len, Op op, hipStream_t s) {
	if (len == 0)
		return VKHEL_OK;
	if (aligned16(a) && aligned16(b) && aligned16(out))
		hipLaunchKernelGGL((k_binary<Op, true>), dim3(grid_for(len)), dim3(256), 0,
				s, a, b, out, len, op);
	else
		hipLaunchKernelGGL((k_binary<Op, false>), dim3(grid_for(2 * len)),
				dim3(256), 0, s, a, b, out, len, op);
	return hip_ok(hipGetLastError(), "launch") ? VKHEL_OK : VKHEL_ERR_DEVICE;
}

template <class Op>
int launch_unary(const uint64_t *a, uint64_t *out, uint64_t len, Op op,
		hipStream_t s) {
	if (len == 0)
		return VKHEL_OK;
	if (aligned16(a) && aligned16(out))
		hipLaunchKernelGGL((k_unary<Op, true>), dim3(grid_for(len)), dim3(256), 0,
				s, a, out, len, op);
	else
		hipLaunchKernelGGL((k_unary<Op, false>), dim3(grid_for(2 * len)),
				dim3(256), 0, s, a, out, len, op);
	return hip_ok(hipGetLastError(), "launch") ? VKHEL_OK : VKHEL_ERR_DEVICE;
}

OpMul make_mul(uint64_t q) {
	return OpMul{q, barrett_mu(q), (uint32_t) (bitlen(q) - 2)};
}

int launch_elemmul(const uint64_t *a, const uint64_t *b, uint64_t *out,
		uint64_t len, uint64_t q, hipStream_t s) {
	return launch_binary(a, b, out, len, make_mul(q), s);
}

int launch_addmod(const uint64_t *a, const uint64_t *b, uint64_t *out,
		uint64_t len, uint64_t q, hipStream_t s) {
	return launch_binary(a, b, out, len, OpAddMod{q}, s);
}

int launch_elemfma(const uint64_t *a, const uint64_t *b, uint64_t *out,
		uint64_t len, uint64_t m, uint64_t q, hipStream_t s) {
	return launch_binary(a, b, out, len, OpFma{q, m, barrett_factor(m, q)}, s);
}

// Shared argument checks: same context (src/vector.c:302, 345, 392, 432,
// 474) and inputs at least as long as the result, whose length sets the
// dispatch size (src/kernels/elemmul.c:164).
static bool check_vectors(const char *fn, bool binary, const vkhel_vector *a,
		const vkhel_vector *b, const vkhel_vector *res) {
	if (!a || !res || (binary && !b))
		return fail(VKHEL_ERR_INVALID, "%s: null vector", fn), false;
	if (a->ctx != res->ctx || (b && b->ctx != res->ctx))
		return fail(VKHEL_ERR_INVALID, "%s: vectors from different contexts",
				fn), false;
	if (a->length < res->length || (b && b->length < res->length))
		return fail(VKHEL_ERR_INVALID,
				"%s: input shorter than result (%llu < %llu)", fn,
				(unsigned long long) std::min(a->length,
					b ? b->length : a->length),
				(unsigned long long) res->length), false;
	return true;
}

static void finish(vkhel_ctx *ctx, int rc) {
	if (rc == VKHEL_OK)
		hip_ok(hipStreamSynchronize(ctx->stream), "sync");
}

}  // namespace vk

using namespace vk;

extern "C" {

// src/vector.c:298-340, push constants src/kernels/elemfma.c:159-169.
void vkhel_vector_elemfma(const struct vkhel_vector *a,
		const struct vkhel_vector *b, struct vkhel_vector *res,
		uint64_t multiplier, uint64_t mod) {
	if (!check_vectors("elemfma", true, a, b, res))
		return;
	if (mod < 2)
		return (void) fail(VKHEL_ERR_INVALID, "elemfma: mod must be >= 2");
	finish(res->ctx, launch_elemfma(a->dev, b->dev, res->dev, res->length,
				multiplier, mod, res->ctx->stream));
}

// src/vector.c:342-386: mod 2 -> elemmodbytwo(q/2), else elemgtsub(q/2, q).
void vkhel_vector_elemmod(const struct vkhel_vector *a,
		struct vkhel_vector *res, uint64_t mod, uint64_t q) {
	if (!check_vectors("elemmod", false, a, nullptr, res))
		return;
	int rc;
	if (mod == 2) {
		rc = launch_unary(a->dev, res->dev, res->length, OpModTwo{q / 2},
				res->ctx->stream);
	} else {
		if (mod < 2)
			return (void) fail(VKHEL_ERR_INVALID, "elemmod: mod must be >= 2");
		const OpGtSub op{q / 2, q, mod, barrett_mu(mod),
			(uint32_t) (bitlen(mod) - 2)};
		rc = launch_unary(a->dev, res->dev, res->length, op, res->ctx->stream);
	}
	finish(res->ctx, rc);
}

// src/vector.c:388-427, push constants src/kernels/elemmul.c:159-168.
void vkhel_vector_elemmul(const struct vkhel_vector *a,
		const struct vkhel_vector *b, struct vkhel_vector *res, uint64_t mod) {
	if (!check_vectors("elemmul", true, a, b, res))
		return;
	if (mod < 2)
		return (void) fail(VKHEL_ERR_INVALID, "elemmul: mod must be >= 2");
	finish(res->ctx, launch_elemmul(a->dev, b->dev, res->dev, res->length, mod,
				res->ctx->stream));
}

// src/vector.c:429-468
void vkhel_vector_elemgtadd(const struct vkhel_vector *a,
		struct vkhel_vector *res, uint64_t bound, uint64_t diff) {
	if (!check_vectors("elemgtadd", false, a, nullptr, res))
		return;
	finish(res->ctx, launch_unary(a->dev, res->dev, res->length,
				OpGtAdd{bound, diff}, res->ctx->stream));
}

// src/vector.c:470-511, push constants src/kernels/elemgtsub.c:154-165.
void vkhel_vector_elemgtsub(const struct vkhel_vector *a,
		struct vkhel_vector *res, uint64_t bound, uint64_t diff, uint64_t mod) {
	if (!check_vectors("elemgtsub", false, a, nullptr, res))
		return;
	if (mod < 2)
		return (void) fail(VKHEL_ERR_INVALID, "elemgtsub: mod must be >= 2");
	const OpGtSub op{bound, diff, mod, barrett_mu(mod),
		(uint32_t) (bitlen(mod) - 2)};
	finish(res->ctx, launch_unary(a->dev, res->dev, res->length, op,
				res->ctx->stream));
}

int vkhel_ext_vector_fill_uniform(struct vkhel_vector *v, uint64_t offset,
		uint64_t count, uint64_t q, uint64_t seed) {
	if (!v || offset > v->length || count > v->length - offset || q == 0)
		return fail(VKHEL_ERR_INVALID, "fill_uniform: bad range or modulus");
	if (count == 0)
		return VKHEL_OK;
	const unsigned grid = (unsigned) std::min<uint64_t>((count + 255) / 256, 8192);
	hipLaunchKernelGGL(k_fill_uniform, dim3(grid), dim3(256), 0, v->ctx->stream,
			v->dev + offset, count, q, seed);
	if (!hip_ok(hipGetLastError(), "fill launch") ||
			!hip_ok(hipStreamSynchronize(v->ctx->stream), "sync"))
		return VKHEL_ERR_DEVICE;
	return VKHEL_OK;
}

}  // extern "C"
