// HIP runtime layer: replaces the reference's Vulkan context, VMA allocator
// and staging transfers (src/vulkan.c, src/vector.c:1-296) and its host
// table builder (src/ntt_tables.c).
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <set>
#include <vector>

#include "internal.h"
#include "numbers.h"

namespace vk {

static int g_abort_on_error = 1;
static thread_local int t_err = VKHEL_OK;
static thread_local char t_msg[512];

int fail(int code, const char *fmt, ...) {
	va_list ap;
	va_start(ap, fmt);
	vsnprintf(t_msg, sizeof t_msg, fmt, ap);
	va_end(ap);
	t_err = code;
	if (__atomic_load_n(&g_abort_on_error, __ATOMIC_RELAXED)) {
		fprintf(stderr, "vkhel: error %d: %s\n", code, t_msg);
		abort();
	}
	return code;
}

void clear_error() {
	t_err = VKHEL_OK;
	t_msg[0] = 0;
}

bool hip_ok(hipError_t e, const char *what) {
	if (e == hipSuccess)
		return true;
	fail(e == hipErrorOutOfMemory ? VKHEL_ERR_NOMEM : VKHEL_ERR_DEVICE,
			"%s: %s", what, hipGetErrorString(e));
	return false;
}

// Live contexts, so destroying a tables object can drop its device copies.
static std::mutex g_ctx_mu;
static std::set<vkhel_ctx *> g_ctxs;
static uint64_t g_next_tables_id = 1;

static void free_dev_tables(DevTables *d) {
	if (d) {
		(void) hipFree(d->alloc);
		delete d;
	}
}

const DevTables *device_tables(vkhel_ctx *ctx, const vkhel_ntt_tables *t) {
	if (!t)
		return fail(VKHEL_ERR_INVALID, "null tables"), nullptr;
	std::lock_guard<std::mutex> g(ctx->mu);
	auto it = ctx->tables.find(t->id);
	if (it != ctx->tables.end())
		return it->second;

	const uint64_t n = t->n, q = t->q;
	// {w, packed 63-bit Shoup companion} pairs for both directions, stage
	// blocks interleaved as tw_pos() lays them out, computed on the GPU.
	TableGenArgs ga;
	table_gen_args(n, q, t->w, ga);
	DevTables *d = new DevTables;
	d->n = n;
	if (!hip_ok(hipMalloc(&d->alloc, 2 * n * sizeof(ulonglong2)),
				"hipMalloc(tables)")) {
		delete d;
		return nullptr;
	}
	ulonglong2 *dev = (ulonglong2 *) d->alloc;
	if (!hip_ok(launch_gen_tables(ga, dev, dev + n, ctx->table_stream),
				"table generation") ||
			!hip_ok(hipStreamSynchronize(ctx->table_stream),
				"table generation sync")) {
		free_dev_tables(d);
		return nullptr;
	}
	d->fwd = (const ulonglong2 *) d->alloc;
	d->inv = d->fwd + n;
	NttConsts &c = d->c;
	c.q = q;
	c.two_q = 2 * q;
	c.r64 = pow2_64_mod(q);
	c.mu = barrett_mu(q);
	c.bshift = (uint32_t) (bitlen(q) - 2);
	c.logn = (uint32_t) (bitlen(n) - 1);
	// src/vector.c:633: n^-1 by the reference's ext-Euclid.
	c.ninv = invmod_euclid(n % q, q);
	c.ninv_p = shoup63_factor(c.ninv, q);
	// inv_roots[1] = w^-bitrev(1) = w^-(n/2)
	c.w1n = mulmod_exact(ga.wipow[ga.logn - 1], c.ninv, q);
	c.w1n_p = shoup63_factor(c.w1n, q);
	small_reduce_consts(q, c.red_shift, c.red_m);
	ctx->tables.emplace(t->id, d);
	return d;
}

}  // namespace vk

using namespace vk;

extern "C" {

void vkhel_ext_set_abort_on_error(int enabled) {
	__atomic_store_n(&g_abort_on_error, enabled ? 1 : 0, __ATOMIC_RELAXED);
}

int vkhel_ext_last_error(void) {
	return t_err;
}

const char *vkhel_ext_last_error_message(void) {
	return t_msg;
}

void vkhel_ext_clear_error(void) {
	clear_error();
}

// src/vkhel.c:4-8 + src/vulkan.c:153-206: device 0, one queue (here one
// default-priority stream).
struct vkhel_ctx *vkhel_ctx_create() {
	int count = 0;
	if (!hip_ok(hipGetDeviceCount(&count), "hipGetDeviceCount"))
		return nullptr;
	if (count < 1)
		return fail(VKHEL_ERR_DEVICE, "no HIP device"), nullptr;
	vkhel_ctx *ctx = new vkhel_ctx;
	ctx->device = 0;
	hipDeviceProp_t p;
	if (!hip_ok(hipSetDevice(0), "hipSetDevice") ||
			!hip_ok(hipGetDeviceProperties(&p, 0), "hipGetDeviceProperties") ||
			!hip_ok(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking),
				"hipStreamCreate") ||
			!hip_ok(hipStreamCreateWithFlags(&ctx->table_stream,
					hipStreamNonBlocking), "hipStreamCreate")) {
		if (ctx->stream)
			(void) hipStreamDestroy(ctx->stream);
		delete ctx;
		return nullptr;
	}
	ctx->device_name = p.gcnArchName;
	ctx->cu_count = p.multiProcessorCount;
	if (noalloc_init(ctx->device) != VKHEL_OK) {
		(void) hipStreamDestroy(ctx->stream);
		(void) hipStreamDestroy(ctx->table_stream);
		delete ctx;
		return nullptr;
	}
	std::lock_guard<std::mutex> g(g_ctx_mu);
	g_ctxs.insert(ctx);
	return ctx;
}

void vkhel_ctx_destroy(struct vkhel_ctx *ctx) {
	if (!ctx)
		return;
	{
		std::lock_guard<std::mutex> g(g_ctx_mu);
		g_ctxs.erase(ctx);
	}
	(void) hipStreamSynchronize(ctx->stream);
	for (auto &kv : ctx->tables)
		free_dev_tables(kv.second);
	for (auto &b : ctx->staging)
		(void) hipHostFree(b.p);
	(void) hipStreamDestroy(ctx->stream);
	(void) hipStreamDestroy(ctx->table_stream);
	delete ctx;
}

// src/ntt_tables.c:65-80 without the arrays.  Validation the reference
// leaves to asserts or undefined behaviour is done up front: n a power of
// two >= 2 (else bitrev indices overflow the arrays), 5 <= q < 2^61 odd (the
// Barrett validity range, src/numbers.c:8), w < q and gcd(w, q) == 1 (else
// the ext-Euclid loop divides by zero), n < q (inverse of n).
static vkhel_ntt_tables *tables_new(uint64_t n, uint64_t q, uint64_t w) {
	if (n < 2 || (n & (n - 1)) || n > (1ull << kMaxTableLog))
		return fail(VKHEL_ERR_INVALID,
				"ntt tables: n=%llu must be a power of two in [2, 2^28]",
				(unsigned long long) n), nullptr;
	if (q < 5 || !(q & 1) || q >= (1ull << 61) || n >= q)
		return fail(VKHEL_ERR_INVALID,
				"ntt tables: q=%llu must be odd, n < q and 5 <= q < 2^61",
				(unsigned long long) q), nullptr;
	if (w >= q || gcd_u64(w, q) != 1)
		return fail(VKHEL_ERR_INVALID,
				"ntt tables: w=%llu must be < q and invertible mod q",
				(unsigned long long) w), nullptr;
	vkhel_ntt_tables *t = (vkhel_ntt_tables *) calloc(1, sizeof *t);
	if (!t)
		return fail(VKHEL_ERR_NOMEM, "ntt tables: host allocation"), nullptr;
	t->n = n;
	t->q = q;
	t->w = w;
	// w^n by n = 2^k squarings; -1 means w has order exactly 2n.
	uint64_t p = w;
	for (uint64_t m = 1; m < n; m <<= 1)
		p = mulmod_exact(p, p, q);
	t->root_ok = p == q - 1;
	std::lock_guard<std::mutex> g(g_ctx_mu);
	t->id = g_next_tables_id++;
	return t;
}

// src/ntt_tables.c:17-44: the host arrays of the reference's struct.
static bool tables_fill_host(vkhel_ntt_tables *t) {
	const uint64_t n = t->n, q = t->q, w = t->w;
	uint64_t *mem = (uint64_t *) malloc(4 * n * sizeof(uint64_t));
	if (!mem)
		return fail(VKHEL_ERR_NOMEM, "ntt tables: host allocation"), false;
	const uint64_t logn = bitlen(n) - 1;
	uint64_t *roots = mem, *inv_roots = mem + n;
	// roots[bitrev(i)] = w^i by repeated Barrett products, as upstream.
	// The inverse table uses w^-i = (w^-1)^i: both it and the reference's
	// per-entry ext-Euclid yield the unique canonical inverse.
	const uint64_t winv = invmod_euclid(w, q);
	uint64_t r = 1, ri = 1;
	for (uint64_t i = 0; i < n; i++) {
		uint64_t idx = 0;
		for (uint64_t b = 0, v = i; b < logn; b++, v >>= 1)
			idx = (idx << 1) | (v & 1);
		roots[idx] = r;
		inv_roots[idx] = ri;
		r = mulmod_barrett(r, w, q);
		ri = mulmod_exact(ri, winv, q);
	}
	for (uint64_t i = 0; i < n; i++) {
		mem[2 * n + i] = barrett_factor(roots[i], q);
		mem[3 * n + i] = barrett_factor(inv_roots[i], q);
	}
	t->roots_of_unity = roots;
	t->inv_roots_of_unity = inv_roots;
	t->roots_barrett_factors = mem + 2 * n;
	t->inv_roots_barrett_factors = mem + 3 * n;
	return true;
}

// Guards the lazy host arrays of tables made by
// vkhel_ext_tables_create_device.
static std::mutex g_host_tables_mu;

struct vkhel_ntt_tables *vkhel_ntt_tables_create(uint64_t n, uint64_t q,
		uint64_t w) {
	vkhel_ntt_tables *t = tables_new(n, q, w);
	if (t && !tables_fill_host(t)) {
		free(t);
		return nullptr;
	}
	return t;
}

// src/ntt_tables.c:82-87 (which leaks the struct; this frees it) plus the
// device copies in every live context.  A device copy may be in use on any
// stream of the device (batch, serving workers, callers' streams), so the
// whole device is waited for before it is freed; callers must still not
// destroy tables that a queued batch launch or server request names.
void vkhel_ntt_tables_destroy(struct vkhel_ntt_tables *t) {
	if (!t)
		return;
	{
		std::lock_guard<std::mutex> g(g_ctx_mu);
		for (vkhel_ctx *ctx : g_ctxs) {
			std::lock_guard<std::mutex> g2(ctx->mu);
			auto it = ctx->tables.find(t->id);
			if (it != ctx->tables.end()) {
				(void) hipSetDevice(ctx->device);
				(void) hipDeviceSynchronize();
				free_dev_tables(it->second);
				ctx->tables.erase(it);
			}
		}
	}
	free(t->roots_of_unity);
	free(t);
}

static vkhel_vector *vector_new(vkhel_ctx *ctx, uint64_t length, bool zero) {
	if (!ctx)
		return fail(VKHEL_ERR_INVALID, "vector_create: null ctx"), nullptr;
	vkhel_vector *v = (vkhel_vector *) calloc(1, sizeof *v);
	if (!v)
		return fail(VKHEL_ERR_NOMEM, "vector_create: host allocation"), nullptr;
	v->ctx = ctx;
	v->length = length;
	const size_t bytes = (length ? length : 1) * sizeof(uint64_t);
	if (!hip_ok(hipMalloc(&v->dev, bytes), "hipMalloc(vector)")) {
		free(v);
		return nullptr;
	}
	if (zero && !(hip_ok(hipMemsetAsync(v->dev, 0, bytes, ctx->stream),
					"hipMemsetAsync") &&
				hip_ok(hipStreamSynchronize(ctx->stream), "sync"))) {
		(void) hipFree(v->dev);
		free(v);
		return nullptr;
	}
	return v;
}

// src/vector.c:206-229
struct vkhel_vector *vkhel_vector_create(struct vkhel_ctx *ctx,
		uint64_t length) {
	return vector_new(ctx, length, true);
}

struct vkhel_vector *vkhel_vector_create2(struct vkhel_ctx *ctx,
		uint64_t length, bool zero) {
	return vector_new(ctx, length, zero);
}

// A page-locked staging buffer of at least `bytes` from the context's pool
// (smallest free one that fits; allocated when none does), or nullptr with
// the error recorded.
static void *staging_get(vkhel_ctx *ctx, size_t bytes) {
	std::lock_guard<std::mutex> g(ctx->staging_mu);
	vkhel_ctx::Staging *best = nullptr;
	for (auto &b : ctx->staging)
		if (!b.busy && b.bytes >= bytes && (!best || b.bytes < best->bytes))
			best = &b;
	if (best) {
		best->busy = true;
		return best->p;
	}
	void *p = nullptr;
	if (!hip_ok(hipHostMalloc(&p, bytes, 0), "hipHostMalloc(map)"))
		return nullptr;
	ctx->staging.push_back(vkhel_ctx::Staging{p, bytes, true});
	return p;
}

// Returns a staging buffer to the pool (or frees it with the pool off).
static void staging_put(vkhel_ctx *ctx, void *p) {
	std::lock_guard<std::mutex> g(ctx->staging_mu);
	for (size_t i = 0; i < ctx->staging.size(); i++) {
		if (ctx->staging[i].p != p)
			continue;
		if (ctx->staging_pool) {
			ctx->staging[i].busy = false;
		} else {
			(void) hipHostFree(p);
			ctx->staging.erase(ctx->staging.begin() + i);
		}
		return;
	}
}

// src/vector.c:231-235
void vkhel_vector_destroy(struct vkhel_vector *v) {
	if (!v)
		return;
	(void) hipStreamSynchronize(v->ctx->stream);
	if (v->staging)
		staging_put(v->ctx, v->staging);
	if (!v->view)
		(void) hipFree(v->dev);
	free(v);
}

// src/vector.c:249-260: uninitialised allocation + device-to-device copy.
struct vkhel_vector *vkhel_vector_dup(struct vkhel_vector *src) {
	if (!src)
		return fail(VKHEL_ERR_INVALID, "dup: null vector"), nullptr;
	vkhel_vector *v = vector_new(src->ctx, src->length, false);
	if (!v)
		return nullptr;
	if (!hip_ok(hipMemcpyAsync(v->dev, src->dev, src->length * 8,
					hipMemcpyDeviceToDevice, src->ctx->stream), "dup copy") ||
			!hip_ok(hipStreamSynchronize(src->ctx->stream), "sync")) {
		vkhel_vector_destroy(v);
		return nullptr;
	}
	return v;
}

// src/vector.c:262-268 maps, overwrites all length words, and unmaps
// (writing them back); the net effect is this one host-to-device copy.
void vkhel_vector_copy_from_host(struct vkhel_vector *v, const uint64_t *e) {
	if (!v || (!e && v->length))
		return (void) fail(VKHEL_ERR_INVALID, "copy_from_host: null argument");
	hip_ok(hipMemcpyAsync(v->dev, e, v->length * 8, hipMemcpyHostToDevice,
				v->ctx->stream), "copy_from_host") &&
		hip_ok(hipStreamSynchronize(v->ctx->stream), "sync");
}

// src/vector.c:270-283: a host buffer holding the device contents.  The
// staging buffer is max(size, length*8) bytes so callers that pass a short
// size (examples/example.c:8,32,40) still get a valid mapping; the first
// min(size, length*8) bytes are copied from the device.  The reference
// allocates a staging buffer per map (src/vector.c:274-278); here it comes
// from the context's pool of page-locked buffers, so steady-state map /
// unmap cycles allocate nothing and copy by DMA.
void vkhel_vector_map(struct vkhel_vector *v, void **mem, size_t size) {
	if (!v || !mem)
		return (void) fail(VKHEL_ERR_INVALID, "map: null argument");
	if (v->staging)
		return (void) fail(VKHEL_ERR_INVALID, "map: vector already mapped");
	const size_t words = std::max<size_t>((size + 7) / 8, v->length);
	void *p = staging_get(v->ctx, std::max<size_t>(words, 1) * 8);
	if (!p)
		return;
	const size_t copy = std::min<size_t>(size, v->length * 8);
	if (copy && !(hip_ok(hipMemcpyAsync(p, v->dev, copy, hipMemcpyDeviceToHost,
						v->ctx->stream), "map copy") &&
				hip_ok(hipStreamSynchronize(v->ctx->stream), "sync"))) {
		staging_put(v->ctx, p);
		return;
	}
	v->staging = p;
	v->staging_words = words;
	*mem = p;
}

// src/vector.c:285-296: writes length words back, frees the staging.
void vkhel_vector_unmap(struct vkhel_vector *v) {
	if (!v || !v->staging)
		return (void) fail(VKHEL_ERR_INVALID, "unmap: vector not mapped");
	hip_ok(hipMemcpyAsync(v->dev, v->staging, v->length * 8,
				hipMemcpyHostToDevice, v->ctx->stream), "unmap copy") &&
		hip_ok(hipStreamSynchronize(v->ctx->stream), "sync");
	staging_put(v->ctx, v->staging);
	v->staging = nullptr;
	v->staging_words = 0;
}

struct vkhel_vector *vkhel_ext_vector_view(struct vkhel_vector *base,
		uint64_t offset, uint64_t length) {
	if (!base || offset > base->length || length > base->length - offset)
		return fail(VKHEL_ERR_INVALID, "vector_view: range out of bounds"),
			nullptr;
	vkhel_vector *v = (vkhel_vector *) calloc(1, sizeof *v);
	if (!v)
		return fail(VKHEL_ERR_NOMEM, "vector_view: host allocation"), nullptr;
	v->ctx = base->ctx;
	v->length = length;
	v->dev = base->dev + offset;
	v->view = true;
	return v;
}

int vkhel_ext_set_staging_pool(struct vkhel_ctx *ctx, int enabled) {
	if (!ctx)
		return fail(VKHEL_ERR_INVALID, "set_staging_pool: null ctx");
	std::lock_guard<std::mutex> g(ctx->staging_mu);
	ctx->staging_pool = enabled ? 1 : 0;
	if (!enabled) {
		// free what no mapping holds
		for (size_t i = 0; i < ctx->staging.size();) {
			if (!ctx->staging[i].busy) {
				(void) hipHostFree(ctx->staging[i].p);
				ctx->staging.erase(ctx->staging.begin() + i);
			} else {
				i++;
			}
		}
	}
	return VKHEL_OK;
}

uint64_t vkhel_ext_staging_bytes(struct vkhel_ctx *ctx) {
	if (!ctx)
		return 0;
	std::lock_guard<std::mutex> g(ctx->staging_mu);
	uint64_t n = 0;
	for (auto &b : ctx->staging)
		n += b.bytes;
	return n;
}

void *vkhel_ext_ctx_stream(struct vkhel_ctx *ctx) {
	return ctx ? (void *) ctx->stream : nullptr;
}

int vkhel_ext_ctx_synchronize(struct vkhel_ctx *ctx) {
	if (!ctx)
		return fail(VKHEL_ERR_INVALID, "null ctx");
	return hip_ok(hipStreamSynchronize(ctx->stream), "sync") ? VKHEL_OK
		: VKHEL_ERR_DEVICE;
}

const char *vkhel_ext_ctx_device_name(struct vkhel_ctx *ctx) {
	return ctx ? ctx->device_name.c_str() : "";
}

uint64_t *vkhel_ext_vector_device_ptr(struct vkhel_vector *v) {
	return v ? v->dev : nullptr;
}

uint64_t vkhel_ext_vector_length(const struct vkhel_vector *v) {
	return v ? v->length : 0;
}

int vkhel_ext_vector_read(const struct vkhel_vector *v, uint64_t offset,
		uint64_t count, uint64_t *host) {
	if (!v || (!host && count) || offset > v->length ||
			count > v->length - offset)
		return fail(VKHEL_ERR_INVALID, "vector_read: range out of bounds");
	if (count && !(hip_ok(hipMemcpyAsync(host, v->dev + offset, count * 8,
						hipMemcpyDeviceToHost, v->ctx->stream), "read") &&
				hip_ok(hipStreamSynchronize(v->ctx->stream), "sync")))
		return VKHEL_ERR_DEVICE;
	return VKHEL_OK;
}

int vkhel_ext_vector_write(struct vkhel_vector *v, uint64_t offset,
		uint64_t count, const uint64_t *host) {
	if (!v || (!host && count) || offset > v->length ||
			count > v->length - offset)
		return fail(VKHEL_ERR_INVALID, "vector_write: range out of bounds");
	if (count && !(hip_ok(hipMemcpyAsync(v->dev + offset, host, count * 8,
						hipMemcpyHostToDevice, v->ctx->stream), "write") &&
				hip_ok(hipStreamSynchronize(v->ctx->stream), "sync")))
		return VKHEL_ERR_DEVICE;
	return VKHEL_OK;
}

void *vkhel_ext_host_alloc(size_t bytes) {
	void *p = nullptr;
	if (!hip_ok(hipHostMalloc(&p, std::max<size_t>(bytes, 1), 0),
				"hipHostMalloc(host_alloc)"))
		return nullptr;
	return p;
}

void vkhel_ext_host_free(void *p) {
	if (p)
		(void) hip_ok(hipHostFree(p), "hipHostFree");
}

int vkhel_ext_vector_write_async(struct vkhel_vector *v, uint64_t offset,
		uint64_t count, const uint64_t *host, void *stream) {
	if (!v || (!host && count) || offset > v->length ||
			count > v->length - offset)
		return fail(VKHEL_ERR_INVALID, "vector_write_async: range out of bounds");
	hipStream_t s = stream ? (hipStream_t) stream : v->ctx->stream;
	if (count && !hip_ok(hipMemcpyAsync(v->dev + offset, host, count * 8,
					hipMemcpyHostToDevice, s), "write_async"))
		return VKHEL_ERR_DEVICE;
	return VKHEL_OK;
}

int vkhel_ext_vector_read_async(const struct vkhel_vector *v, uint64_t offset,
		uint64_t count, uint64_t *host, void *stream) {
	if (!v || (!host && count) || offset > v->length ||
			count > v->length - offset)
		return fail(VKHEL_ERR_INVALID, "vector_read_async: range out of bounds");
	hipStream_t s = stream ? (hipStream_t) stream : v->ctx->stream;
	if (count && !hip_ok(hipMemcpyAsync(host, v->dev + offset, count * 8,
					hipMemcpyDeviceToHost, s), "read_async"))
		return VKHEL_ERR_DEVICE;
	return VKHEL_OK;
}

int vkhel_ext_tables_info(const struct vkhel_ntt_tables *t, uint64_t *n,
		uint64_t *q, uint64_t *w) {
	if (!t)
		return fail(VKHEL_ERR_INVALID, "null tables");
	if (n)
		*n = t->n;
	if (q)
		*q = t->q;
	if (w)
		*w = t->w;
	return VKHEL_OK;
}

int vkhel_ext_tables_copy(const struct vkhel_ntt_tables *t, uint64_t *roots,
		uint64_t *inv_roots, uint64_t *roots_bf, uint64_t *inv_roots_bf) {
	if (!t)
		return fail(VKHEL_ERR_INVALID, "null tables");
	std::lock_guard<std::mutex> g(g_host_tables_mu);
	if (!t->roots_of_unity &&
			!tables_fill_host(const_cast<vkhel_ntt_tables *>(t)))
		return vkhel_ext_last_error();
	const size_t b = t->n * sizeof(uint64_t);
	if (roots)
		memcpy(roots, t->roots_of_unity, b);
	if (inv_roots)
		memcpy(inv_roots, t->inv_roots_of_unity, b);
	if (roots_bf)
		memcpy(roots_bf, t->roots_barrett_factors, b);
	if (inv_roots_bf)
		memcpy(inv_roots_bf, t->inv_roots_barrett_factors, b);
	return VKHEL_OK;
}

int vkhel_ext_tables_upload(struct vkhel_ctx *ctx, struct vkhel_ntt_tables *t) {
	if (!ctx)
		return fail(VKHEL_ERR_INVALID, "null ctx");
	return device_tables(ctx, t) ? VKHEL_OK : vkhel_ext_last_error();
}

struct vkhel_ntt_tables *vkhel_ext_tables_create_device(struct vkhel_ctx *ctx,
		uint64_t n, uint64_t q, uint64_t w) {
	if (!ctx)
		return fail(VKHEL_ERR_INVALID, "null ctx"), nullptr;
	vkhel_ntt_tables *t = tables_new(n, q, w);
	if (!t)
		return nullptr;
	if (!t->root_ok) {
		free(t);
		return fail(VKHEL_ERR_INVALID,
				"ntt tables: w=%llu is not a primitive 2n-th root of unity "
				"mod q (w^n != -1)", (unsigned long long) w), nullptr;
	}
	if (!device_tables(ctx, t)) {
		vkhel_ntt_tables_destroy(t);
		return nullptr;
	}
	return t;
}

int vkhel_ext_tables_root_ok(const struct vkhel_ntt_tables *t) {
	if (!t)
		return fail(VKHEL_ERR_INVALID, "null tables"), -1;
	return t->root_ok;
}

int vkhel_ext_tables_read_device(struct vkhel_ctx *ctx,
		const struct vkhel_ntt_tables *t, uint64_t *roots,
		uint64_t *roots_shoup, uint64_t *inv_roots, uint64_t *inv_roots_shoup) {
	if (!ctx || !t)
		return fail(VKHEL_ERR_INVALID, "null ctx or tables");
	const DevTables *d = device_tables(ctx, t);
	if (!d)
		return vkhel_ext_last_error();
	const uint64_t n = t->n;
	std::vector<ulonglong2> h(2 * n);
	if (!hip_ok(hipMemcpy(h.data(), d->alloc, 2 * n * sizeof(ulonglong2),
					hipMemcpyDeviceToHost), "read tables"))
		return vkhel_ext_last_error();
	for (uint64_t k = 0; k < n; k++) {
		const uint64_t m = k ? 1ull << (bitlen(k) - 1) : 0;
		const uint64_t p = k ? tw_pos((uint32_t) m, (uint32_t) (k - m)) : 0;
		if (roots)
			roots[k] = h[p].x;
		if (roots_shoup)
			roots_shoup[k] = h[p].y;
		if (inv_roots)
			inv_roots[k] = h[n + p].x;
		if (inv_roots_shoup)
			inv_roots_shoup[k] = h[n + p].y;
	}
	return VKHEL_OK;
}

int vkhel_ext_set_ntt_algo(struct vkhel_ctx *ctx, int algo) {
	if (!ctx || algo < VKHEL_NTT_AUTO || algo > VKHEL_NTT_FUSED)
		return fail(VKHEL_ERR_INVALID, "set_ntt_algo: bad argument");
	ctx->ntt_algo = algo;
	return VKHEL_OK;
}

}  // extern "C"
