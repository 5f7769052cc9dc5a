// Job model shared by the batch API (batch.hip) and the serving engine
// (serve.hip): a job is one operation on an RNS polynomial (nlimbs limbs of
// n words, limb i modulo tables[i]->q); a set of jobs is planned into three
// device phases of tasks (forward transforms, pointwise, inverse), plus, for
// RNS rescaling, a pointwise phase after the inverse and (NTT-form
// rescaling) a second forward phase and pointwise phase after that.
#pragma once

#include <vector>

#include "internal.h"

namespace vk {

struct Job {
	int op;
	const vkhel_vector *a, *b;
	vkhel_vector *res;
	std::vector<const vkhel_ntt_tables *> tables;
	std::vector<uint64_t> params;  // RNS elementwise ops only
};

// One pointwise limb over len words.  op:
//   kPtMul   out = a.b mod q (elemmul semantics)
//   kPtAdd   out = (a + b) mod q
//   kPtFma   out = elemfma(a, b, w, q), wp = floor(w * 2^64 / q)
//   kPtGtSub out = elemgtsub(a, bound, diff, q)  (also elemmod, p != 2)
//   kPtGtAdd out = elemgtadd(a, bound, diff)
//   kPtModTwo out = elemmod(a, 2, .) with bound = q_limb / 2
//   kPtRescale out = (a - [b]) * w mod q, [b] the centered lift of the
//              dropped limb b (b > bound: b - q_last) reduced mod q,
//              diff = q_last mod q, w = q_last^-1 mod q (wp its packed
//              Shoup companion)
enum { kPtMul = 0, kPtAdd, kPtGtSub, kPtGtAdd, kPtModTwo, kPtRescale, kPtFma };
struct PointTask {
	const uint64_t *a, *b;
	uint64_t *out;
	uint64_t len, q, mu;
	uint64_t bound, diff;
	uint32_t shift, op;
	uint64_t w, wp;
};

// A run of transform tasks with one log2 n and one modulus class
// (modulus_class() of every q); tb: the tile size fused polynomial products
// were planned for (0 for plain transforms, which take the context's).
struct Segment {
	uint32_t off, count, logn;
	int mc;
	int tb;
	// word offset of this segment's state in DeviceTasks::handoff when it
	// runs both passes in one launch (pass mode 3), else kNoHandoff
	uint32_t hoff;
};
constexpr uint32_t kNoHandoff = 0xffffffffu;

struct TaskSet {
	std::vector<NttTask> fwd, inv;
	std::vector<Segment> seg_fwd, seg_inv;
	// fused polynomial products (launch_polymul_tasks): the forward column
	// passes of a (two-pass sizes only, task i for product i of the same
	// key) and the fused products
	std::vector<NttTask> pm_col, pm;
	std::vector<Segment> seg_pm_col, seg_pm;
	std::vector<PointTask> pt, post;  // before / after the inverse phase
	// after `post`: forward transforms, then pointwise tasks
	std::vector<NttTask> fwd2;
	std::vector<Segment> seg_fwd2;
	std::vector<PointTask> post2;
	uint64_t handoff_words = 0;  // state words of all one-launch segments
};

// Device copies of a TaskSet's arrays, grown on demand.
struct DeviceTasks {
	NttTask *fwd = nullptr, *inv = nullptr, *pm_col = nullptr, *pm = nullptr;
	NttTask *fwd2 = nullptr;
	PointTask *pt = nullptr, *post = nullptr, *post2 = nullptr;
	size_t cap_fwd = 0, cap_inv = 0, cap_pm_col = 0, cap_pm = 0, cap_fwd2 = 0;
	size_t cap_pt = 0, cap_post = 0, cap_post2 = 0;
	// state of one-launch two-pass transforms (pass mode 3)
	uint32_t *handoff = nullptr;
	uint64_t handoff_bytes = 0;
	void release();
	// Allocates every array for `tasks` entries up front (false with the
	// error recorded on failure).
	bool reserve(size_t tasks);
	// Device bytes `reserve(tasks)` takes.
	static uint64_t bytes_for(size_t tasks);
};

// Checks a job's arguments and fills `out`.  Returns a vkhel_status.
int make_job(const vkhel_ctx *ctx, int op, const vkhel_vector *a,
		const vkhel_vector *b, vkhel_vector *res,
		vkhel_ntt_tables *const *tables, uint32_t nlimbs, Job *out);
// Checks an RNS job (VKHEL_JOB_CENTER / _GTADD / _GTSUB / _INTT_CENTER /
// _RESCALE / _RESCALE_NTT).
int make_rns_job(const vkhel_ctx *ctx, int op, const vkhel_vector *a,
		vkhel_vector *res, vkhel_ntt_tables *const *tables, uint32_t nlimbs,
		const uint64_t *params, Job *out);
// Device scratch words a job needs (the spectrum of b for POLYMUL, the
// dropped limb's inverse transform for RESCALE_NTT).
uint64_t job_scratch_words(const Job &j);
// Device bytes of a job's operands and result: the serving engine's
// admission cost (its scratch comes from the engine's reserved arenas).
uint64_t job_operand_bytes(const Job &j);
// Transform coefficients a job processes (for throughput accounting).
uint64_t job_transform_coeffs(const Job &j);
// Plans `jobs` into tasks; POLYMUL scratch is carved from `scratch`.
int plan_tasks(vkhel_ctx *ctx, const Job *const *jobs, size_t njobs,
		uint64_t *scratch, TaskSet &ts);
// Copies the task arrays to `dt` (stream-ordered) and enqueues the
// phases on `s`.
int upload_tasks(const TaskSet &ts, DeviceTasks &dt, hipStream_t s);
// After the launch has completed: VKHEL_OK if every one-launch two-pass
// segment ran all its tiles with no wait given up, else the error recorded.
int check_handoff(const TaskSet &ts, const DeviceTasks &dt);
int enqueue_tasks(const vkhel_ctx *ctx, const TaskSet &ts,
		const DeviceTasks &dt, hipStream_t s);

}  // namespace vk
