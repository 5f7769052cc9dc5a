// Batched execution of many RNS polynomial jobs (vkhel_ext.h).  The
// reference has no batch API: a caller would issue one synchronous
// vkhel_vector_* call per polynomial (src/vector.c:513-657), each paying
// log2(n) fenced submits.  Here a job list turns into three device phases,
// each one launch over every polynomial limb of every job:
//
//   1. forward transforms   (NTT_FWD jobs, both operands of POLYMUL jobs)
//   2. pointwise products   (MUL, ADD jobs, the spectra of POLYMUL jobs)
//   3. inverse transforms   (NTT_INV jobs, POLYMUL results)
//
// The launch sequence is captured once into a hipGraph and replayed while
// the job list is unchanged.
#include <algorithm>
#include <vector>

#include "internal.h"
#include "modarith.h"
#include "numbers.h"

namespace vk {

struct Job {
	int op;
	const vkhel_vector *a, *b;
	vkhel_vector *res;
	std::vector<const vkhel_ntt_tables *> tables;
};

// One pointwise limb: out = a (op) b over len words mod q.
struct PointTask {
	const uint64_t *a, *b;
	uint64_t *out;
	uint64_t len, q, mu;
	uint32_t shift, op;  // op: 0 = mul (elemmul semantics), 1 = add
};

__global__ __launch_bounds__(256) void k_pointwise_tasks(const PointTask *tasks) {
	const PointTask t = tasks[blockIdx.y];
	const uint64_t pairs = t.len >> 1;
	const uint64_t stride = (uint64_t) gridDim.x * blockDim.x;
	for (uint64_t i = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x;
			i < pairs; i += stride) {
		const ulonglong2 x = ((const ulonglong2 *) t.a)[i];
		const ulonglong2 y = ((const ulonglong2 *) t.b)[i];
		ulonglong2 r;
		if (t.op == 0) {
			const uint64_t a0 = barrett64(x.x, t.q, t.mu, t.shift);
			const uint64_t b0 = barrett64(y.x, t.q, t.mu, t.shift);
			const uint64_t a1 = barrett64(x.y, t.q, t.mu, t.shift);
			const uint64_t b1 = barrett64(y.y, t.q, t.mu, t.shift);
			r.x = barrett128(mulhi64(a0, b0), a0 * b0, t.q, t.mu, t.shift);
			r.y = barrett128(mulhi64(a1, b1), a1 * b1, t.q, t.mu, t.shift);
		} else {
			r.x = csub(x.x + y.x, t.q);
			r.y = csub(x.y + y.y, t.q);
		}
		((ulonglong2 *) t.out)[i] = r;
	}
	if ((t.len & 1) && blockIdx.x == 0 && threadIdx.x == 0) {
		const uint64_t k = t.len - 1;
		if (t.op == 0) {
			const uint64_t a0 = barrett64(t.a[k], t.q, t.mu, t.shift);
			const uint64_t b0 = barrett64(t.b[k], t.q, t.mu, t.shift);
			t.out[k] = barrett128(mulhi64(a0, b0), a0 * b0, t.q, t.mu, t.shift);
		} else {
			t.out[k] = csub(t.a[k] + t.b[k], t.q);
		}
	}
}

}  // namespace vk

namespace vk {
// A run of tasks with one log2 n inside a direction's task array.
struct Segment {
	uint32_t off, count, logn;
};
}  // namespace vk

struct vkhel_batch {
	vkhel_ctx *ctx;
	std::vector<vk::Job> jobs;
	bool dirty = true;
	// Device state built at launch.
	vk::NttTask *d_fwd = nullptr, *d_inv = nullptr;
	vk::PointTask *d_pt = nullptr;
	std::vector<vk::NttTask> h_fwd, h_inv;
	std::vector<vk::Segment> seg_fwd, seg_inv;
	uint32_t n_pt = 0;
	uint64_t pt_maxlen = 0;
	uint64_t *scratch = nullptr;
	size_t scratch_words = 0;
	hipGraph_t graph = nullptr;
	hipGraphExec_t exec = nullptr;
};

namespace vk {

static void release_device(vkhel_batch *b) {
	if (b->exec)
		(void) hipGraphExecDestroy(b->exec);
	if (b->graph)
		(void) hipGraphDestroy(b->graph);
	(void) hipFree(b->d_fwd);
	(void) hipFree(b->d_inv);
	(void) hipFree(b->d_pt);
	b->exec = nullptr;
	b->graph = nullptr;
	b->d_fwd = b->d_inv = nullptr;
	b->d_pt = nullptr;
}

template <class T>
static bool upload(const std::vector<T> &h, T **d) {
	if (h.empty())
		return true;
	if (!hip_ok(hipMalloc(d, h.size() * sizeof(T)), "hipMalloc(batch tasks)"))
		return false;
	return hip_ok(hipMemcpy(*d, h.data(), h.size() * sizeof(T),
				hipMemcpyHostToDevice), "upload batch tasks");
}

// Sorts tasks by log2 n and records the runs.
static std::vector<Segment> segment(std::vector<NttTask> &tasks) {
	std::stable_sort(tasks.begin(), tasks.end(),
			[](const NttTask &a, const NttTask &b) { return a.c.logn < b.c.logn; });
	std::vector<Segment> segs;
	for (uint32_t i = 0; i < tasks.size(); i++) {
		if (segs.empty() || segs.back().logn != tasks[i].c.logn)
			segs.push_back(Segment{i, 0, tasks[i].c.logn});
		segs.back().count++;
	}
	return segs;
}

static int launch_segments(const vkhel_ctx *ctx,
		const std::vector<Segment> &segs, const std::vector<NttTask> &host,
		const NttTask *dev, bool inverse, hipStream_t s) {
	for (const Segment &g : segs) {
		int rc = VKHEL_OK;
		if (ntt_fused_supported(g.logn)) {
			rc = launch_fused_tasks(ctx, dev + g.off, g.count, g.logn, inverse, s);
		} else {
			for (uint32_t i = 0; i < g.count && rc == VKHEL_OK; i++)
				rc = launch_task_ntt(host[g.off + i], inverse, s);
		}
		if (rc != VKHEL_OK)
			return rc;
	}
	return VKHEL_OK;
}

static int enqueue(vkhel_batch *b, hipStream_t s) {
	int rc = launch_segments(b->ctx, b->seg_fwd, b->h_fwd, b->d_fwd, false, s);
	if (rc == VKHEL_OK && b->n_pt) {
		const uint64_t blocks = std::min<uint64_t>((b->pt_maxlen / 2 + 255) / 256,
				64);
		hipLaunchKernelGGL(k_pointwise_tasks, dim3((unsigned) std::max<uint64_t>(
						blocks, 1), b->n_pt), dim3(256), 0, s, b->d_pt);
		rc = hip_ok(hipGetLastError(), "pointwise launch") ? VKHEL_OK
			: VKHEL_ERR_DEVICE;
	}
	if (rc == VKHEL_OK)
		rc = launch_segments(b->ctx, b->seg_inv, b->h_inv, b->d_inv, true, s);
	return rc;
}

// Turns the job list into device task arrays and a captured graph.
static int build(vkhel_batch *b) {
	release_device(b);
	vkhel_ctx *ctx = b->ctx;
	std::vector<NttTask> fwd, inv;
	std::vector<PointTask> pt;
	uint64_t pt_maxlen = 0;
	size_t scratch_words = 0;
	for (const Job &j : b->jobs)
		if (j.op == VKHEL_JOB_POLYMUL)
			scratch_words += j.tables.size() * j.tables[0]->n;
	if (scratch_words > b->scratch_words) {
		(void) hipFree(b->scratch);
		b->scratch = nullptr;
		b->scratch_words = 0;
		if (!hip_ok(hipMalloc(&b->scratch, scratch_words * 8),
					"hipMalloc(batch scratch)"))
			return vkhel_ext_last_error();
		b->scratch_words = scratch_words;
	}
	uint64_t *sp = b->scratch;
	for (const Job &j : b->jobs) {
		const uint64_t n = j.tables[0]->n;
		for (size_t l = 0; l < j.tables.size(); l++) {
			const DevTables *t = device_tables(ctx, j.tables[l]);
			if (!t)
				return vkhel_ext_last_error();
			const uint64_t off = l * n;
			const uint64_t q = t->c.q;
			switch (j.op) {
			case VKHEL_JOB_NTT_FWD:
				fwd.push_back(make_task(*t, j.a->dev + off, j.res->dev + off, false));
				break;
			case VKHEL_JOB_NTT_INV:
				inv.push_back(make_task(*t, j.a->dev + off, j.res->dev + off, true));
				break;
			case VKHEL_JOB_POLYMUL: {
				uint64_t *tmp = sp;
				sp += n;
				fwd.push_back(make_task(*t, j.a->dev + off, j.res->dev + off, false));
				fwd.push_back(make_task(*t, j.b->dev + off, tmp, false));
				pt.push_back(PointTask{j.res->dev + off, tmp, j.res->dev + off, n, q,
						barrett_mu(q), (uint32_t) (bitlen(q) - 2), 0});
				inv.push_back(make_task(*t, j.res->dev + off, j.res->dev + off, true));
				break;
			}
			case VKHEL_JOB_MUL:
			case VKHEL_JOB_ADD:
				pt.push_back(PointTask{j.a->dev + off, j.b->dev + off,
						j.res->dev + off, n, q, barrett_mu(q),
						(uint32_t) (bitlen(q) - 2),
						(uint32_t) (j.op == VKHEL_JOB_MUL ? 0 : 1)});
				break;
			}
			pt_maxlen = std::max<uint64_t>(pt_maxlen, n);
		}
	}
	b->seg_fwd = segment(fwd);
	b->seg_inv = segment(inv);
	if (!upload(fwd, &b->d_fwd) || !upload(inv, &b->d_inv) ||
			!upload(pt, &b->d_pt))
		return vkhel_ext_last_error();
	b->h_fwd = std::move(fwd);
	b->h_inv = std::move(inv);
	b->n_pt = (uint32_t) pt.size();
	b->pt_maxlen = pt_maxlen;

	// Capture the launch sequence once; replays skip per-kernel launch cost.
	hipStream_t s = ctx->stream;
	if (!hip_ok(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal),
				"begin capture"))
		return vkhel_ext_last_error();
	const int rc = enqueue(b, s);
	hipGraph_t g = nullptr;
	if (!hip_ok(hipStreamEndCapture(s, &g), "end capture") || rc != VKHEL_OK) {
		if (g)
			(void) hipGraphDestroy(g);
		return rc != VKHEL_OK ? rc : vkhel_ext_last_error();
	}
	b->graph = g;
	if (!hip_ok(hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0),
				"graph instantiate"))
		return vkhel_ext_last_error();
	b->dirty = false;
	return VKHEL_OK;
}

}  // namespace vk

using namespace vk;

extern "C" {

struct vkhel_batch *vkhel_batch_create(struct vkhel_ctx *ctx) {
	if (!ctx)
		return fail(VKHEL_ERR_INVALID, "batch_create: null ctx"), nullptr;
	vkhel_batch *b = new vkhel_batch;
	b->ctx = ctx;
	return b;
}

void vkhel_batch_destroy(struct vkhel_batch *b) {
	if (!b)
		return;
	(void) hipStreamSynchronize(b->ctx->stream);
	release_device(b);
	(void) hipFree(b->scratch);
	delete b;
}

void vkhel_batch_clear(struct vkhel_batch *b) {
	if (!b)
		return;
	b->jobs.clear();
	b->dirty = true;
}

uint64_t vkhel_batch_num_jobs(const struct vkhel_batch *b) {
	return b ? b->jobs.size() : 0;
}

int vkhel_batch_add(struct vkhel_batch *b, int op, const struct vkhel_vector *a,
		const struct vkhel_vector *bv, struct vkhel_vector *res,
		struct vkhel_ntt_tables *const *tables, uint32_t nlimbs) {
	if (!b || !a || !res || !tables || nlimbs == 0)
		return fail(VKHEL_ERR_INVALID, "batch_add: null argument");
	if (op < VKHEL_JOB_NTT_FWD || op > VKHEL_JOB_ADD)
		return fail(VKHEL_ERR_INVALID, "batch_add: unknown op %d", op);
	const bool binary = op >= VKHEL_JOB_POLYMUL;
	if (binary && !bv)
		return fail(VKHEL_ERR_INVALID, "batch_add: op needs a second operand");
	if (a->ctx != b->ctx || res->ctx != b->ctx || (binary && bv->ctx != b->ctx))
		return fail(VKHEL_ERR_INVALID, "batch_add: vector from another context");
	for (uint32_t i = 0; i < nlimbs; i++)
		if (!tables[i] || tables[i]->n != tables[0]->n)
			return fail(VKHEL_ERR_INVALID,
					"batch_add: limb tables must all exist and share n");
	const uint64_t words = (uint64_t) nlimbs * tables[0]->n;
	if (a->length < words || res->length < words ||
			(binary && bv->length < words))
		return fail(VKHEL_ERR_INVALID,
				"batch_add: vectors need %llu words (%u limbs x n)",
				(unsigned long long) words, nlimbs);
	if (op == VKHEL_JOB_POLYMUL && (bv == res))
		return fail(VKHEL_ERR_INVALID, "batch_add: polymul result aliases b");
	Job j;
	j.op = op;
	j.a = a;
	j.b = binary ? bv : nullptr;
	j.res = res;
	j.tables.assign(tables, tables + nlimbs);
	b->jobs.push_back(std::move(j));
	b->dirty = true;
	return VKHEL_OK;
}

int vkhel_batch_launch(struct vkhel_batch *b) {
	if (!b)
		return fail(VKHEL_ERR_INVALID, "batch_launch: null batch");
	if (b->jobs.empty())
		return VKHEL_OK;
	if (b->dirty) {
		const int rc = build(b);
		if (rc != VKHEL_OK)
			return rc;
	}
	return hip_ok(hipGraphLaunch(b->exec, b->ctx->stream), "graph launch")
		? VKHEL_OK : VKHEL_ERR_DEVICE;
}

int vkhel_batch_run(struct vkhel_batch *b) {
	const int rc = vkhel_batch_launch(b);
	if (rc != VKHEL_OK)
		return rc;
	return vkhel_ext_ctx_synchronize(b->ctx);
}

}  // extern "C"
