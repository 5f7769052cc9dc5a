// Batched execution of many RNS polynomial jobs (vkhel_ext.h).  The
// reference has no batch API: a caller would issue one synchronous
// vkhel_vector_* call per polynomial (src/vector.c:513-657), each paying
// log2(n) fenced submits.  Here a job list turns into device phases, each
// a handful of launches over every polynomial limb of every job:
//
//   1. forward transforms   (NTT_FWD jobs, both operands of POLYMUL jobs,
//                            only b where the polymul kernel is fused)
//      fused polymuls       (a's forward, the product and the inverse in
//                            one kernel: k_ntt_pm)
//   2. pointwise products   (MUL, ADD jobs; POLYMUL spectra only for sizes
//                            the fused kernels do not cover)
//   3. inverse transforms   (NTT_INV jobs, POLYMUL results; the fused
//                            inverse multiplies the two spectra as it loads)
//   4. pointwise            (RESCALE; RESCALE_NTT's lift of the dropped limb)
//   5. forward transforms   (RESCALE_NTT's lifted limbs)
//   6. pointwise            (RESCALE_NTT's subtraction and scaling)
//
// Transform tasks are grouped by size: one launch per distinct n, each
// workgroup taking one polynomial limb.  The batch API captures the launch
// sequence once into a hipGraph and replays it while the job list is
// unchanged, or, for a sequence of a few launches, enqueues it again on
// each launch (kDirectLaunchMax).
#include <algorithm>
#include <vector>

#include "handoff.h"
#include "jobs.h"
#include "modarith.h"
#include "numbers.h"

namespace vk {

__device__ __forceinline__ uint64_t point_op(const PointTask &t, uint64_t x,
		uint64_t y) {
	switch (t.op) {
	case kPtMul: {
		const uint64_t a = barrett64(x, t.q, t.mu, t.shift);
		const uint64_t b = barrett64(y, t.q, t.mu, t.shift);
		return barrett128(mulhi64(a, b), a * b, t.q, t.mu, t.shift);
	}
	case kPtAdd:
		return csub(x + y, t.q);
	case kPtFma:
		return elemfma_word(x, y, t.w, t.wp, t.q);
	case kPtGtSub: {  // elemgtsub.comp:49-67
		const uint64_t r = barrett64(x, t.q, t.mu, t.shift);
		if (x > t.bound)
			return csub(r + t.q - barrett64(t.diff, t.q, t.mu, t.shift), t.q);
		return r;
	}
	case kPtGtAdd:  // elemgtadd.comp:20-30
		return x > t.bound ? x + t.diff : x;
	case kPtRescale: {
		// (a - [b]) / q_last mod q: a canonical mod q, b mod q_last
		const uint64_t bm = barrett64(y, t.q, t.mu, t.shift);
		const uint64_t c = y > t.bound ? sub_mod(bm, t.diff, t.q) : bm;
		const uint64_t d = sub_mod(barrett64(x, t.q, t.mu, t.shift), c, t.q);
		return csub(shoup63(d, t.w, t.wp, 0 - t.q), t.q);
	}
	default: {  // kPtModTwo, elemmodbytwo.comp:19-27
		const uint64_t u = x & 1;
		return x > t.bound ? 1 - u : u;
	}
	}
}

// One workgroup row per task (blockIdx.y); unary ops never read b.  Two
// words per access when the task's pointers are 16-byte aligned, one
// otherwise (vector views at odd word offsets).
__global__ __launch_bounds__(256) void k_pointwise_tasks(const PointTask *tasks) {
	const PointTask t = tasks[blockIdx.y];
	const bool binary = t.op == kPtMul || t.op == kPtAdd || t.op == kPtRescale ||
		t.op == kPtFma;
	const uint64_t stride = (uint64_t) gridDim.x * blockDim.x;
	if ((((uintptr_t) t.a | (uintptr_t) t.out |
					(binary ? (uintptr_t) t.b : 0)) & 15) != 0) {
		for (uint64_t i = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x;
				i < t.len; i += stride)
			t.out[i] = point_op(t, t.a[i], binary ? t.b[i] : 0);
		return;
	}
	const uint64_t pairs = t.len >> 1;
	for (uint64_t i = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x;
			i < pairs; i += stride) {
		const ulonglong2 x = ((const ulonglong2 *) t.a)[i];
		const ulonglong2 y = binary ? ((const ulonglong2 *) t.b)[i]
			: make_ulonglong2(0, 0);
		ulonglong2 r;
		r.x = point_op(t, x.x, y.x);
		r.y = point_op(t, x.y, y.y);
		((ulonglong2 *) t.out)[i] = r;
	}
	if ((t.len & 1) && blockIdx.x == 0 && threadIdx.x == 0) {
		const uint64_t k = t.len - 1;
		t.out[k] = point_op(t, t.a[k], binary ? t.b[k] : 0);
	}
}

void DeviceTasks::release() {
	(void) hipFree(fwd);
	(void) hipFree(inv);
	(void) hipFree(pm_col);
	(void) hipFree(pm);
	(void) hipFree(fwd2);
	(void) hipFree(pt);
	(void) hipFree(post);
	(void) hipFree(post2);
	(void) hipFree(handoff);
	handoff = nullptr;
	handoff_bytes = 0;
	fwd = inv = pm_col = pm = fwd2 = nullptr;
	pt = post = post2 = nullptr;
	cap_fwd = cap_inv = cap_pm_col = cap_pm = cap_fwd2 = 0;
	cap_pt = cap_post = cap_post2 = 0;
}

template <class T>
static bool reserve_array(T **d, size_t *cap, size_t n) {
	if (*cap >= n)
		return true;
	(void) hipFree(*d);
	*d = nullptr;
	*cap = 0;
	if (!hip_ok(hipMalloc(d, n * sizeof(T)), "hipMalloc(task array)"))
		return false;
	*cap = n;
	return true;
}

static bool reserve_handoff(DeviceTasks &dt, uint64_t bytes) {
	if (dt.handoff_bytes >= bytes)
		return true;
	(void) hipFree(dt.handoff);
	dt.handoff = nullptr;
	dt.handoff_bytes = 0;
	if (!hip_ok(hipMalloc(&dt.handoff, bytes), "hipMalloc(handoff state)"))
		return false;
	dt.handoff_bytes = bytes;
	return true;
}

bool DeviceTasks::reserve(size_t n) {
	return reserve_array(&fwd, &cap_fwd, n) && reserve_array(&inv, &cap_inv, n) &&
		reserve_array(&pm_col, &cap_pm_col, n) && reserve_array(&pm, &cap_pm, n) &&
		reserve_array(&fwd2, &cap_fwd2, n) && reserve_array(&pt, &cap_pt, n) &&
		reserve_array(&post, &cap_post, n) && reserve_array(&post2, &cap_post2, n) &&
		reserve_handoff(*this, handoff_state_bytes((uint32_t) n));
}

uint64_t DeviceTasks::bytes_for(size_t n) {
	return (uint64_t) n * (5 * sizeof(NttTask) + 3 * sizeof(PointTask)) +
		handoff_state_bytes((uint32_t) n);
}

int make_job(const vkhel_ctx *ctx, int op, const vkhel_vector *a,
		const vkhel_vector *b, vkhel_vector *res,
		vkhel_ntt_tables *const *tables, uint32_t nlimbs, Job *out) {
	if (!a || !res || !tables || nlimbs == 0)
		return fail(VKHEL_ERR_INVALID, "job: null argument");
	if ((op < VKHEL_JOB_NTT_FWD || op > VKHEL_JOB_ADD) && op != VKHEL_JOB_ADD_REF)
		return fail(VKHEL_ERR_INVALID, "job: unknown op %d", op);
	const bool binary = op >= VKHEL_JOB_POLYMUL;
	if (binary && !b)
		return fail(VKHEL_ERR_INVALID, "job: op needs a second operand");
	if (a->ctx != ctx || res->ctx != ctx || (binary && b->ctx != ctx))
		return fail(VKHEL_ERR_INVALID, "job: vector from another context");
	for (uint32_t i = 0; i < nlimbs; i++)
		if (!tables[i] || tables[i]->n != tables[0]->n)
			return fail(VKHEL_ERR_INVALID,
					"job: limb tables must all exist and share n");
	const uint64_t words = (uint64_t) nlimbs * tables[0]->n;
	if (a->length < words || res->length < words ||
			(binary && b->length < words))
		return fail(VKHEL_ERR_INVALID,
				"job: vectors need %llu words (%u limbs x n)",
				(unsigned long long) words, nlimbs);
	if (op == VKHEL_JOB_POLYMUL && b == res)
		return fail(VKHEL_ERR_INVALID, "job: polymul result aliases b");
	out->op = op;
	out->a = a;
	out->b = binary ? b : nullptr;
	out->res = res;
	out->tables.assign(tables, tables + nlimbs);
	return VKHEL_OK;
}

static const uint64_t kNoParams = 0;

int make_rns_job(const vkhel_ctx *ctx, int op, const vkhel_vector *a,
		vkhel_vector *res, vkhel_ntt_tables *const *tables, uint32_t nlimbs,
		const uint64_t *params, Job *out) {
	const bool rescale = op == VKHEL_JOB_RESCALE || op == VKHEL_JOB_RESCALE_NTT;
	if (!a || !res || !tables || nlimbs == 0 || (!params && !rescale))
		return fail(VKHEL_ERR_INVALID, "rns job: null argument");
	if (op < VKHEL_JOB_CENTER || op > VKHEL_JOB_RESCALE_NTT)
		return fail(VKHEL_ERR_INVALID, "rns job: unknown op %d", op);
	if (rescale && !params)
		params = &kNoParams;
	if (a->ctx != ctx || res->ctx != ctx)
		return fail(VKHEL_ERR_INVALID, "rns job: vector from another context");
	for (uint32_t i = 0; i < nlimbs; i++)
		if (!tables[i] || tables[i]->n != tables[0]->n)
			return fail(VKHEL_ERR_INVALID,
					"rns job: limb tables must all exist and share n");
	const uint64_t words = (uint64_t) nlimbs * tables[0]->n;
	// rescaling drops the last limb from the result
	const uint64_t res_words = rescale ? words - tables[0]->n : words;
	if (a->length < words || res->length < res_words)
		return fail(VKHEL_ERR_INVALID,
				"rns job: vectors need %llu and %llu words (%u limbs x n)",
				(unsigned long long) words, (unsigned long long) res_words,
				nlimbs);
	if (rescale) {
		if (nlimbs < 2)
			return fail(VKHEL_ERR_INVALID, "rns job: rescale needs 2+ limbs");
		const uint64_t ql = tables[nlimbs - 1]->q;
		for (uint32_t i = 0; i + 1 < nlimbs; i++)
			if (gcd_u64(ql % tables[i]->q, tables[i]->q) != 1)
				return fail(VKHEL_ERR_INVALID,
						"rns job: rescale needs the last modulus invertible mod "
						"every other (limb %u)", i);
		if (a == res)
			return fail(VKHEL_ERR_INVALID,
					"rns job: rescale reads the dropped limb; res must not be a");
	}
	const bool center = op == VKHEL_JOB_CENTER || op == VKHEL_JOB_INTT_CENTER;
	const size_t np = rescale ? 0 : center ? 1 :
		(op == VKHEL_JOB_GTADD ? 2 : 3) * (size_t) nlimbs;
	if (center && params[0] < 2)
		return fail(VKHEL_ERR_INVALID, "rns job: center needs p >= 2");
	if (op == VKHEL_JOB_GTSUB)
		for (uint32_t i = 0; i < nlimbs; i++)
			if (params[3 * i + 2] < 2)
				return fail(VKHEL_ERR_INVALID, "rns job: gtsub limb %u mod < 2", i);
	out->op = op;
	out->a = a;
	out->b = nullptr;
	out->res = res;
	out->tables.assign(tables, tables + nlimbs);
	out->params.assign(params, params + np);
	return VKHEL_OK;
}

uint64_t job_scratch_words(const Job &j) {
	if (j.op == VKHEL_JOB_RESCALE_NTT)
		return j.tables[0]->n;
	return j.op == VKHEL_JOB_POLYMUL ? j.tables.size() * j.tables[0]->n : 0;
}

uint64_t job_operand_bytes(const Job &j) {
	const uint64_t words = j.tables.size() * j.tables[0]->n;
	const uint64_t operands = j.b ? 3 : 2;
	return 8 * operands * words;
}

uint64_t job_transform_coeffs(const Job &j) {
	const uint64_t words = j.tables.size() * j.tables[0]->n;
	switch (j.op) {
	case VKHEL_JOB_POLYMUL:
		return 3 * words;  // two forward transforms and one inverse
	case VKHEL_JOB_NTT_FWD:
	case VKHEL_JOB_NTT_INV:
	case VKHEL_JOB_INTT_CENTER:
	case VKHEL_JOB_RESCALE_NTT:  // the dropped limb's INTT, the others' NTTs
		return words;
	default:
		return 0;
	}
}

// Sorts tasks by log2 n and modulus class and records the runs.  With
// by_table the tasks of a run of n >= 2^kGroupByTableLogn are also grouped
// by their twiddle table (stable: a table's tasks keep their order), so
// that the workgroups in flight share one table in each XCD's 4 MB L2: a
// row pass reads about 2^logn table entries of 16 bytes per modulus
// (2 MiB at n = 2^17), and eight moduli interleaved task by task would
// want 16 MiB.  Measured on MI355X, 8 moduli interleaved, forward ms:
// 256 x 2^17 0.228 -> 0.221, 128 x 2^18 0.252 -> 0.244, 64 x 2^19
// 0.272 -> 0.264, 32 x 2^20 0.340 -> 0.311.  Smaller sizes
// lose by it (512 x 2^16 0.199 -> 0.214: the polynomials in flight then
// lie far apart in memory instead of side by side) and keep their order;
// so do polynomial products, whose column and row tasks are matched by
// index.
constexpr uint32_t kGroupByTableLogn = 17;

static std::vector<Segment> segment(std::vector<NttTask> &tasks,
		bool by_table) {
	const auto key = [](const NttTask &k) {
		return 3 * k.c.logn + (uint32_t) modulus_class(k.c.q);
	};
	std::stable_sort(tasks.begin(), tasks.end(),
			[&](const NttTask &a, const NttTask &b) {
				const uint32_t ka = key(a), kb = key(b);
				if (ka != kb)
					return ka < kb;
				// same key, so the same n
				return by_table && a.c.logn >= kGroupByTableLogn &&
					(uintptr_t) a.tw < (uintptr_t) b.tw;
			});
	std::vector<Segment> segs;
	for (uint32_t i = 0; i < tasks.size(); i++) {
		const int mc = modulus_class(tasks[i].c.q);
		if (segs.empty() || segs.back().logn != tasks[i].c.logn ||
				segs.back().mc != mc)
			segs.push_back(Segment{i, 0, tasks[i].c.logn, mc, 0, kNoHandoff});
		segs.back().count++;
	}
	return segs;
}

// elemmod(a, pm, q) over n words into out (src/vector.c:342-386: the
// reference's dispatch, modulo-two kernel for pm = 2).
static PointTask center_task(const uint64_t *a, uint64_t *out, uint64_t n,
		uint64_t q, uint64_t pm) {
	PointTask p{a, nullptr, out, n, 0, 0, q / 2, 0, 0, kPtModTwo};
	if (pm != 2) {
		p.op = kPtGtSub;
		p.diff = q;
		p.q = pm;
		p.mu = barrett_mu(pm);
		p.shift = (uint32_t) (bitlen(pm) - 2);
	}
	return p;
}

// Workgroups of 256 threads that give a task of `len` words one two-word
// access per thread (at most 1024: longer tasks loop).
static uint64_t pointwise_blocks(uint64_t len) {
	return std::max<uint64_t>(std::min<uint64_t>((len / 2 + 255) / 256, 1024), 1);
}

// Pointwise tasks in the order their launches take them: longest first
// (they are independent, so any order gives the same words).
static void order_pointwise(std::vector<PointTask> &h) {
	std::stable_sort(h.begin(), h.end(), [](const PointTask &x,
			const PointTask &y) { return x.len > y.len; });
}

int plan_tasks(vkhel_ctx *ctx, const Job *const *jobs, size_t njobs,
		uint64_t *scratch, TaskSet &ts) {
	ts = TaskSet();
	uint64_t *sp = scratch;
	for (size_t ji = 0; ji < njobs; ji++) {
		const Job &j = *jobs[ji];
		const uint64_t n = j.tables[0]->n;
		// device twiddles only for jobs that transform
		const bool transforms = j.op <= VKHEL_JOB_POLYMUL ||
			j.op == VKHEL_JOB_INTT_CENTER || j.op == VKHEL_JOB_RESCALE_NTT;
		// RESCALE_NTT: the dropped limb's inverse transform
		uint64_t *dropped = nullptr;
		if (j.op == VKHEL_JOB_RESCALE_NTT) {
			dropped = sp;
			sp += n;
		}
		for (size_t l = 0; l < j.tables.size(); l++) {
			const DevTables *t = transforms ? device_tables(ctx, j.tables[l])
				: nullptr;
			if (transforms && !t)
				return vkhel_ext_last_error();
			const uint64_t off = l * n;
			const uint64_t q = j.tables[l]->q;
			const PointTask mul{nullptr, nullptr, nullptr, n, q, barrett_mu(q), 0, 0,
				(uint32_t) (bitlen(q) - 2), kPtMul};
			switch (j.op) {
			case VKHEL_JOB_NTT_FWD:
				ts.fwd.push_back(make_task(*t, j.a->dev + off, j.res->dev + off, false));
				break;
			case VKHEL_JOB_NTT_INV:
				ts.inv.push_back(make_task(*t, j.a->dev + off, j.res->dev + off, true));
				break;
			case VKHEL_JOB_RESCALE: {
				// limb l of (a - [a_last]) / q_last, after every transform
				const size_t last = j.tables.size() - 1;
				if (l == last)
					break;
				const uint64_t ql = j.tables[last]->q;
				const uint64_t w = invmod_euclid(ql % q, q);
				PointTask p{j.a->dev + off, j.a->dev + last * n, j.res->dev + off, n,
					q, barrett_mu(q), ql / 2, ql % q, (uint32_t) (bitlen(q) - 2),
					kPtRescale, w, shoup63_factor(w, q)};
				ts.post.push_back(p);
				break;
			}
			case VKHEL_JOB_RESCALE_NTT: {
				// INTT of the last limb; per other limb l: its centered lift
				// into Z_q_l, the forward transform of that into res limb l,
				// then (a_l - res_l) * q_last^-1.  The fused forward kernels
				// apply the lift as they load and the last step as they
				// store; other sizes run them as pointwise passes (elemgtsub
				// with bound q_last / 2, diff q_last; the rescale op with a
				// bound no word exceeds: no lift).
				const size_t last = j.tables.size() - 1;
				if (l == last) {
					ts.inv.push_back(make_task(*t, j.a->dev + off, dropped, true));
					break;
				}
				const uint64_t ql = j.tables[last]->q;
				const uint64_t w = invmod_euclid(ql % q, q);
				if (ntt_fused_supported(t->c.logn)) {
					NttTask f = make_task(*t, dropped, j.res->dev + off, false);
					f.c.lift_bound = ql / 2;
					f.c.lift_add = q - ql % q;
					f.c.scale_w = w;
					f.c.scale_wp = shoup63_factor(w, q);
					f.in2 = j.a->dev + off;
					ts.fwd2.push_back(f);
					break;
				}
				PointTask lift{dropped, nullptr, j.res->dev + off, n, q, barrett_mu(q),
					ql / 2, ql, (uint32_t) (bitlen(q) - 2), kPtGtSub};
				ts.post.push_back(lift);
				ts.fwd2.push_back(make_task(*t, j.res->dev + off, j.res->dev + off,
						false));
				PointTask p{j.a->dev + off, j.res->dev + off, j.res->dev + off, n, q,
					barrett_mu(q), ~0ull, 0, (uint32_t) (bitlen(q) - 2), kPtRescale, w,
					shoup63_factor(w, q)};
				ts.post2.push_back(p);
				break;
			}
			case VKHEL_JOB_INTT_CENTER: {
				// elemmod applied by the inverse kernel as it stores; sizes
				// outside the fused kernels run it as a pointwise pass after
				// the inverse transforms
				NttTask inv = make_task(*t, j.a->dev + off, j.res->dev + off, true);
				const uint64_t pm = j.params[0];
				if (!ntt_fused_supported(t->c.logn)) {
					ts.inv.push_back(inv);
					ts.post.push_back(center_task(j.res->dev + off, j.res->dev + off, n,
							q, pm));
					break;
				}
				inv.c.post_p = pm;
				if (pm != 2) {
					inv.c.post_mu = barrett_mu(pm);
					inv.c.post_shift = (uint32_t) (bitlen(pm) - 2);
				}
				ts.inv.push_back(inv);
				break;
			}
			case VKHEL_JOB_POLYMUL: {
				uint64_t *tmp = sp;
				sp += n;
				if (polymul_fused_tile(ctx, t->c.logn)) {
					// both forwards, the product and the inverse in one
					// kernel; two-pass sizes first run the forward column
					// passes of a (in place in res) and b (into scratch)
					const bool cols = t->c.logn > (uint32_t) polymul_fused_tile(ctx,
							t->c.logn);
					if (cols) {
						ts.pm_col.push_back(make_task(*t, j.a->dev + off,
								j.res->dev + off, false));
						ts.pm_col.push_back(make_task(*t, j.b->dev + off, tmp, false));
					}
					NttTask pm = make_task(*t, cols ? j.res->dev + off
							: j.a->dev + off, j.res->dev + off, true);
					pm.in2 = cols ? tmp : j.b->dev + off;
					pm.tw2 = t->fwd;
					ts.pm.push_back(pm);
					break;
				}
				ts.fwd.push_back(make_task(*t, j.a->dev + off, j.res->dev + off, false));
				ts.fwd.push_back(make_task(*t, j.b->dev + off, tmp, false));
				NttTask inv = make_task(*t, j.res->dev + off, j.res->dev + off, true);
				if (ntt_fused_supported(t->c.logn)) {
					// the fused inverse forms the pointwise product as it loads
					inv.in2 = tmp;
				} else {
					PointTask p = mul;
					p.a = j.res->dev + off;
					p.b = tmp;
					p.out = j.res->dev + off;
					ts.pt.push_back(p);
				}
				ts.inv.push_back(inv);
				break;
			}
			case VKHEL_JOB_MUL:
			case VKHEL_JOB_ADD:
			case VKHEL_JOB_ADD_REF: {
				PointTask p = mul;
				p.a = j.a->dev + off;
				p.b = j.b->dev + off;
				p.out = j.res->dev + off;
				p.op = j.op == VKHEL_JOB_MUL ? kPtMul
					: j.op == VKHEL_JOB_ADD ? kPtAdd : kPtFma;
				if (p.op == kPtFma) {
					// elemfma with multiplier 1 (src/vector.c:298-340)
					p.w = 1;
					p.wp = barrett_factor(1, q);
				}
				ts.pt.push_back(p);
				break;
			}
			case VKHEL_JOB_CENTER:
				ts.pt.push_back(center_task(j.a->dev + off, j.res->dev + off, n, q,
						j.params[0]));
				break;
			case VKHEL_JOB_GTADD:
			case VKHEL_JOB_GTSUB: {
				// src/vector.c:429-511: the same per-limb push constants
				PointTask p{j.a->dev + off, nullptr, j.res->dev + off, n, 0, 0, 0, 0,
					0, kPtGtAdd};
				const uint64_t *pr = j.params.data();
				if (j.op == VKHEL_JOB_GTADD) {
					p.bound = pr[2 * l];
					p.diff = pr[2 * l + 1];
				} else {
					p.op = kPtGtSub;
					p.bound = pr[3 * l];
					p.diff = pr[3 * l + 1];
					p.q = pr[3 * l + 2];
				}
				if (p.op == kPtGtSub) {
					p.mu = barrett_mu(p.q);
					p.shift = (uint32_t) (bitlen(p.q) - 2);
				}
				ts.pt.push_back(p);
				break;
			}
			}
		}
	}
	order_pointwise(ts.pt);
	order_pointwise(ts.post);
	order_pointwise(ts.post2);
	ts.seg_fwd = segment(ts.fwd, true);
	ts.seg_inv = segment(ts.inv, true);
	ts.seg_fwd2 = segment(ts.fwd2, true);
	ts.seg_pm_col = segment(ts.pm_col, false);
	ts.seg_pm = segment(ts.pm, false);
	for (auto *segs : {&ts.seg_pm_col, &ts.seg_pm}) {
		for (Segment &g : *segs)
			g.tb = polymul_fused_tile(ctx, g.logn);
	}
	// segments that run both passes in one launch get state words of
	// their own, so every one can be checked after the launch
	for (auto *segs : {&ts.seg_fwd, &ts.seg_inv, &ts.seg_fwd2}) {
		for (Segment &g : *segs) {
			if (uses_handoff(ctx, g.logn)) {
				g.tb = tile_bits_for(ctx, g.logn);
				g.hoff = (uint32_t) ts.handoff_words;
				ts.handoff_words += handoff_state_bytes(g.count) / 4;
			}
		}
	}
	return VKHEL_OK;
}

int check_handoff(const TaskSet &ts, const DeviceTasks &dt) {
	if (!ts.handoff_words)
		return VKHEL_OK;
	std::vector<uint32_t> h(ts.handoff_words);
	if (!hip_ok(hipMemcpy(h.data(), dt.handoff, 4 * h.size(),
					hipMemcpyDeviceToHost), "read handoff state"))
		return vkhel_ext_last_error();
	for (const auto *segs : {&ts.seg_fwd, &ts.seg_inv, &ts.seg_fwd2}) {
		for (const Segment &g : *segs) {
			if (g.hoff == kNoHandoff)
				continue;
			const uint32_t *st = h.data() + g.hoff;
			const uint32_t tiles = handoff_tiles(g.logn, g.tb);
			bool ok = st[handoff_err_word()] == 0;
			for (uint32_t x = 0; x < kXcds; x++)
				ok = ok && st[handoff_done_word(x)] ==
					handoff_queue_tickets(g.count, x, tiles);
			if (!ok)
				return fail(VKHEL_ERR_DEVICE, "one-launch two-pass transform of %u "
						"polynomials (n = 2^%u) did not complete its schedule "
						"(error flag %u); its results are invalid", g.count, g.logn,
						st[handoff_err_word()]);
		}
	}
	return VKHEL_OK;
}

template <class T>
static bool upload_array(const std::vector<T> &h, T **d, size_t *cap,
		hipStream_t s) {
	if (h.empty())
		return true;
	if (h.size() > *cap) {
		const size_t want = std::max<size_t>(h.size(), 2 * *cap);
		(void) hipFree(*d);
		*d = nullptr;
		*cap = 0;
		if (!hip_ok(hipMalloc(d, want * sizeof(T)), "hipMalloc(task array)"))
			return false;
		*cap = want;
	}
	return hip_ok(hipMemcpyAsync(*d, h.data(), h.size() * sizeof(T),
				hipMemcpyHostToDevice, s), "upload task array");
}

int upload_tasks(const TaskSet &ts, DeviceTasks &dt, hipStream_t s) {
	if (ts.handoff_words && !reserve_handoff(dt, 4 * ts.handoff_words))
		return vkhel_ext_last_error();
	if (!upload_array(ts.fwd, &dt.fwd, &dt.cap_fwd, s) ||
			!upload_array(ts.inv, &dt.inv, &dt.cap_inv, s) ||
			!upload_array(ts.pm_col, &dt.pm_col, &dt.cap_pm_col, s) ||
			!upload_array(ts.pm, &dt.pm, &dt.cap_pm, s) ||
			!upload_array(ts.fwd2, &dt.fwd2, &dt.cap_fwd2, s) ||
			!upload_array(ts.post2, &dt.post2, &dt.cap_post2, s) ||
			!upload_array(ts.pt, &dt.pt, &dt.cap_pt, s) ||
			!upload_array(ts.post, &dt.post, &dt.cap_post, s))
		return vkhel_ext_last_error();
	return VKHEL_OK;
}

static int launch_segments(const vkhel_ctx *ctx,
		const std::vector<Segment> &segs, const std::vector<NttTask> &host,
		const NttTask *dev, bool inverse, const DeviceTasks &dt, hipStream_t s) {
	for (const Segment &g : segs) {
		int rc = VKHEL_OK;
		if (ntt_fused_supported(g.logn)) {
			rc = g.hoff == kNoHandoff
				? launch_fused_tasks(ctx, dev + g.off, g.count, g.logn, inverse, g.mc, s)
				: launch_fused_tasks(ctx, dev + g.off, g.count, g.logn, inverse, g.mc, s,
						dt.handoff + g.hoff, dt.handoff_bytes - 4ull * g.hoff);
		} else {
			for (uint32_t i = 0; i < g.count && rc == VKHEL_OK; i++)
				rc = launch_task_ntt(host[g.off + i], inverse, s);
		}
		if (rc != VKHEL_OK)
			return rc;
	}
	return VKHEL_OK;
}

// One pointwise phase over tasks ordered by order_pointwise: a row of
// workgroups per task, in launches of tasks whose lengths need within 4x
// the same number of workgroups (each launch sized for its longest task),
// so short tasks do not carry the long ones' empty workgroups; rows in
// launches of at most 65535 (the grid's y limit).
static int launch_pointwise(const std::vector<PointTask> &h, const PointTask *d,
		hipStream_t s) {
	for (size_t i = 0; i < h.size();) {
		const uint64_t blocks = pointwise_blocks(h[i].len);
		size_t j = i + 1;
		while (j < h.size() && j - i < 65535 && pointwise_blocks(h[j].len) * 4 >= blocks)
			j++;
		hipLaunchKernelGGL(k_pointwise_tasks, dim3((unsigned) blocks,
				(unsigned) (j - i)), dim3(256), 0, s, d + i);
		if (!hip_ok(hipGetLastError(), "pointwise launch"))
			return VKHEL_ERR_DEVICE;
		i = j;
	}
	return VKHEL_OK;
}

// The fused polymul segments; a two-pass size's column-pass tasks form a
// segment of the same (log2 n, class) key with two tasks per product (a's
// and b's column passes, adjacent), the same jobs in the same order (both
// lists are filled in job order and sorted stably by that key).
static int launch_polymul_segments(const TaskSet &ts, const DeviceTasks &dt,
		hipStream_t s) {
	for (const Segment &g : ts.seg_pm) {
		const NttTask *cols = nullptr;
		for (const Segment &c : ts.seg_pm_col) {
			if (c.logn == g.logn && c.mc == g.mc) {
				if (c.count != 2 * g.count)
					return fail(VKHEL_ERR_INVALID, "polymul plan: column segment "
							"of %u tasks for %u products", c.count, g.count);
				cols = dt.pm_col + c.off;
			}
		}
		const int rc = launch_polymul_tasks(cols, dt.pm + g.off, g.count, g.logn,
				g.tb, g.mc, s);
		if (rc != VKHEL_OK)
			return rc;
	}
	return VKHEL_OK;
}

int enqueue_tasks(const vkhel_ctx *ctx, const TaskSet &ts,
		const DeviceTasks &dt, hipStream_t s) {
	int rc = launch_segments(ctx, ts.seg_fwd, ts.fwd, dt.fwd, false, dt, s);
	if (rc == VKHEL_OK)
		rc = launch_polymul_segments(ts, dt, s);
	if (rc == VKHEL_OK)
		rc = launch_pointwise(ts.pt, dt.pt, s);
	if (rc == VKHEL_OK)
		rc = launch_segments(ctx, ts.seg_inv, ts.inv, dt.inv, true, dt, s);
	if (rc == VKHEL_OK)
		rc = launch_pointwise(ts.post, dt.post, s);
	if (rc == VKHEL_OK)
		rc = launch_segments(ctx, ts.seg_fwd2, ts.fwd2, dt.fwd2, false, dt, s);
	if (rc == VKHEL_OK)
		rc = launch_pointwise(ts.post2, dt.post2, s);
	return rc;
}

}  // namespace vk

struct vkhel_batch {
	vkhel_ctx *ctx;
	std::vector<vk::Job> jobs;
	bool dirty = true;
	vk::TaskSet ts;
	vk::DeviceTasks dt;
	uint64_t *scratch = nullptr;
	size_t scratch_words = 0;
	hipGraph_t graph = nullptr;
	hipGraphExec_t exec = nullptr;
	// Few launches: enqueued directly on every launch instead of replaying
	// the graph (kDirectLaunchMax).
	bool direct = false;
	// Private stream the batch uploads its task arrays on and captures its
	// graph from, so that building a batch never captures work other
	// threads enqueue on the context stream.
	hipStream_t build_stream = nullptr;
	// Recorded on the context stream after every launch: the batch's
	// completion handle, waited on before its buffers are rebuilt.
	hipEvent_t done = nullptr;
	bool launched = false;
};

namespace vk {

static void release_graph(vkhel_batch *b) {
	if (b->exec)
		(void) hipGraphExecDestroy(b->exec);
	if (b->graph)
		(void) hipGraphDestroy(b->graph);
	b->exec = nullptr;
	b->graph = nullptr;
	b->direct = false;
}

// Batches of at most this many kernel launches are enqueued directly on
// each launch: a graph replay costs about 10 us more on the device than
// the same launches from the host stream (measured on MI355X: 512 x 2^16
// forward transforms, two launches, 0.231 ms replayed against 0.214 ms
// enqueued; 8192 x 2^12, one launch, 0.148 against 0.137), while batches
// of many launches (a 1000-job mixed batch, about 50) time the same either
// way and keep the graph for its lower host cost.
constexpr size_t kDirectLaunchMax = 8;

// Waits for the batch's last launch (its buffers and graph are then free
// to change).
static int wait_launched(vkhel_batch *b) {
	if (!b->launched)
		return VKHEL_OK;
	if (!hip_ok(hipEventSynchronize(b->done), "batch wait"))
		return vkhel_ext_last_error();
	b->launched = false;
	return check_handoff(b->ts, b->dt);
}

// Turns the job list into device task arrays and a captured graph.
static int build(vkhel_batch *b) {
	int rc = wait_launched(b);
	if (rc != VKHEL_OK)
		return rc;
	release_graph(b);
	vkhel_ctx *ctx = b->ctx;
	if (!b->build_stream &&
			!hip_ok(hipStreamCreateWithFlags(&b->build_stream,
					hipStreamNonBlocking), "batch stream"))
		return vkhel_ext_last_error();
	uint64_t scratch_words = 0;
	for (const Job &j : b->jobs)
		scratch_words += job_scratch_words(j);
	if (scratch_words > b->scratch_words) {
		(void) hipFree(b->scratch);
		b->scratch = nullptr;
		b->scratch_words = 0;
		if (!hip_ok(hipMalloc(&b->scratch, scratch_words * 8),
					"hipMalloc(batch scratch)"))
			return vkhel_ext_last_error();
		b->scratch_words = scratch_words;
	}
	std::vector<const Job *> jp;
	for (const Job &j : b->jobs)
		jp.push_back(&j);
	rc = plan_tasks(ctx, jp.data(), jp.size(), b->scratch, b->ts);
	hipStream_t s = b->build_stream;
	if (rc == VKHEL_OK)
		rc = upload_tasks(b->ts, b->dt, s);
	if (rc == VKHEL_OK)
		rc = hip_ok(hipStreamSynchronize(s), "sync") ? VKHEL_OK
			: VKHEL_ERR_DEVICE;
	if (rc != VKHEL_OK)
		return rc;

	// Capture the launch sequence once; replays skip per-kernel launch cost.
	if (!hip_ok(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal),
				"begin capture"))
		return vkhel_ext_last_error();
	rc = enqueue_tasks(ctx, b->ts, b->dt, s);
	hipGraph_t g = nullptr;
	if (!hip_ok(hipStreamEndCapture(s, &g), "end capture") || rc != VKHEL_OK) {
		if (g)
			(void) hipGraphDestroy(g);
		return rc != VKHEL_OK ? rc : vkhel_ext_last_error();
	}
	size_t nodes = 0;
	if (!hip_ok(hipGraphGetNodes(g, nullptr, &nodes), "graph nodes")) {
		(void) hipGraphDestroy(g);
		return vkhel_ext_last_error();
	}
	if (nodes <= kDirectLaunchMax) {
		(void) hipGraphDestroy(g);
		b->direct = true;
		b->dirty = false;
		return VKHEL_OK;
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
	(void) wait_launched(b);
	release_graph(b);
	b->dt.release();
	(void) hipFree(b->scratch);
	if (b->done)
		(void) hipEventDestroy(b->done);
	if (b->build_stream)
		(void) hipStreamDestroy(b->build_stream);
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
	if (!b)
		return fail(VKHEL_ERR_INVALID, "batch_add: null batch");
	Job j;
	const int rc = make_job(b->ctx, op, a, bv, res, tables, nlimbs, &j);
	if (rc != VKHEL_OK)
		return rc;
	b->jobs.push_back(std::move(j));
	b->dirty = true;
	return VKHEL_OK;
}

int vkhel_batch_add_rns(struct vkhel_batch *b, int op,
		const struct vkhel_vector *a, struct vkhel_vector *res,
		struct vkhel_ntt_tables *const *tables, uint32_t nlimbs,
		const uint64_t *params) {
	if (!b)
		return fail(VKHEL_ERR_INVALID, "batch_add_rns: null batch");
	Job j;
	const int rc = make_rns_job(b->ctx, op, a, res, tables, nlimbs, params, &j);
	if (rc != VKHEL_OK)
		return rc;
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
	if (!b->done && !hip_ok(hipEventCreateWithFlags(&b->done,
				hipEventDisableTiming), "batch event"))
		return vkhel_ext_last_error();
	if (b->direct) {
		const int rc = enqueue_tasks(b->ctx, b->ts, b->dt, b->ctx->stream);
		if (rc != VKHEL_OK)
			return rc;
	} else if (!hip_ok(hipGraphLaunch(b->exec, b->ctx->stream), "graph launch")) {
		return vkhel_ext_last_error();
	}
	if (!hip_ok(hipEventRecord(b->done, b->ctx->stream), "batch event"))
		return vkhel_ext_last_error();
	b->launched = true;
	return VKHEL_OK;
}

int vkhel_batch_wait(struct vkhel_batch *b) {
	if (!b)
		return fail(VKHEL_ERR_INVALID, "batch_wait: null batch");
	return wait_launched(b);
}

int vkhel_batch_query(struct vkhel_batch *b) {
	if (!b)
		return fail(VKHEL_ERR_INVALID, "batch_query: null batch"), -1;
	if (!b->launched)
		return 1;
	const hipError_t e = hipEventQuery(b->done);
	if (e == hipErrorNotReady)
		return 0;
	if (!hip_ok(e, "batch query"))
		return -1;
	return 1;
}

int vkhel_batch_run(struct vkhel_batch *b) {
	const int rc = vkhel_batch_launch(b);
	if (rc != VKHEL_OK)
		return rc;
	return vkhel_batch_wait(b);
}

}  // extern "C"
