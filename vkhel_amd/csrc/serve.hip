// Serving engine: continuous batching of NTT / INTT / polymul / pointwise
// and RNS (centered lift, INTT_CENTER, rescaling) requests over
// device-resident operands.  The reference has no serving
// layer (every vkhel_vector_* call blocks on a fence, src/vector.c:326).
//
// Requests are queued on submit and admitted in arrival order while the
// device footprint of admitted-but-unfinished requests stays within the
// HBM budget and the in-flight count within its limit.  `streams` worker
// threads each own a HIP stream, task arrays and a scratch arena, all
// allocated when the server is created (the arenas count against the
// budget; nothing is allocated while serving); a worker that becomes free
// takes every admissible waiting request that fits its arena (up to
// max_batch) as one micro-batch, plans it into the batch executor's
// phases and runs it.  Requests that arrive while other micro-batches are
// on the device are therefore admitted as budget and slots free up, and
// their kernels overlap with the running ones on the other streams.  Each
// request's result depends only on its own operands (transform tasks are
// independent), so it is the same alone and in any mix.
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <deque>
#include <thread>
#include <unordered_map>

#include "jobs.h"

namespace vk {

static uint64_t now_ns() {
	return (uint64_t) std::chrono::duration_cast<std::chrono::nanoseconds>(
			std::chrono::steady_clock::now().time_since_epoch()).count();
}

enum ReqState { kWaiting, kRunning, kDone };

struct Request {
	Job job;
	uint64_t bytes = 0;    // operands and result
	uint64_t scratch = 0;  // words of the worker's arena it needs
	uint64_t limbs = 0;
	int state = kWaiting;
	int status = VKHEL_OK;
	vkhel_request_info info{};
};

struct Worker {
	hipStream_t stream = nullptr;
	DeviceTasks dt;
	TaskSet ts;
	uint64_t *scratch = nullptr;
	uint64_t scratch_words = 0;
};

}  // namespace vk

struct vkhel_server {
	vkhel_ctx *ctx = nullptr;
	vkhel_server_config cfg{};
	std::mutex mu;
	std::condition_variable cv_work, cv_done;
	std::deque<uint64_t> waiting;
	std::unordered_map<uint64_t, vk::Request> reqs;
	uint64_t next_id = 1;
	// device bytes of the workers' scratch arenas (held from creation),
	// and the per-worker arena and task capacity in words / limbs
	uint64_t reserved_bytes = 0;
	uint64_t arena_words = 0;
	uint64_t task_limbs = 0;
	uint64_t inflight_bytes = 0;
	uint32_t inflight = 0;
	bool stop = false;
	vkhel_server_stats stats{};
	std::vector<vk::Worker> workers;
	std::vector<std::thread> threads;
};

namespace vk {

// Bytes of the budget left for requests' operands once the arenas are
// reserved.
static uint64_t operand_budget(const vkhel_server *s) {
	return s->cfg.hbm_budget_bytes - s->reserved_bytes;
}

// True when the oldest waiting request can be admitted now.
static bool head_admissible(const vkhel_server *s) {
	if (s->waiting.empty())
		return false;
	const Request &r = s->reqs.at(s->waiting.front());
	return s->inflight < s->cfg.max_inflight &&
		s->inflight_bytes + r.bytes <= operand_budget(s);
}

static int run_micro_batch(vkhel_server *s, Worker &w,
		const std::vector<const Job *> &jobs) {
	// admission kept the batch within the worker's arena and task arrays
	int rc = plan_tasks(s->ctx, jobs.data(), jobs.size(), w.scratch, w.ts);
	if (rc == VKHEL_OK)
		rc = upload_tasks(w.ts, w.dt, w.stream);
	if (rc == VKHEL_OK)
		rc = enqueue_tasks(s->ctx, w.ts, w.dt, w.stream);
	if (rc == VKHEL_OK && !hip_ok(hipStreamSynchronize(w.stream), "serve sync"))
		rc = VKHEL_ERR_DEVICE;
	if (rc == VKHEL_OK)
		rc = check_handoff(w.ts, w.dt);
	return rc;
}

// Plain per-call service of one request (vkhel_server_config.per_call): the
// library's synchronous single calls, one per limb and step, each waited
// for before the next as a reference caller does (src/vector.c:320-334
// fences every call): NTT / INTT one transform call, POLYMUL two forward
// transforms, elemmul and the inverse, MUL elemmul, ADD the modular add,
// ADD_REF elemfma with multiplier 1.  Other ops run as a one-job batch.
static int run_per_call(vkhel_server *s, Worker &w, const Job &j) {
	if (j.op > VKHEL_JOB_ADD && j.op != VKHEL_JOB_ADD_REF) {
		const Job *one = &j;
		return run_micro_batch(s, w, std::vector<const Job *>(1, one));
	}
	const uint64_t n = j.tables[0]->n;
	for (size_t l = 0; l < j.tables.size(); l++) {
		const uint64_t off = l * n, q = j.tables[l]->q;
		const uint64_t *a = j.a->dev + off;
		const uint64_t *b = j.b ? j.b->dev + off : nullptr;
		uint64_t *r = j.res->dev + off;
		const DevTables *t = nullptr;
		if (j.op <= VKHEL_JOB_POLYMUL && !(t = device_tables(s->ctx, j.tables[l])))
			return vkhel_ext_last_error();
		auto call = [&](int rc) {
			if (rc == VKHEL_OK && !hip_ok(hipStreamSynchronize(w.stream), "sync"))
				rc = VKHEL_ERR_DEVICE;
			return rc;
		};
		int rc = VKHEL_OK;
		switch (j.op) {
		case VKHEL_JOB_NTT_FWD:
		case VKHEL_JOB_NTT_INV:
			rc = call(launch_ntt(s->ctx, *t, a, r, 1, j.op == VKHEL_JOB_NTT_INV,
						w.stream));
			break;
		case VKHEL_JOB_POLYMUL:
			rc = call(launch_ntt(s->ctx, *t, a, r, 1, false, w.stream));
			if (rc == VKHEL_OK)
				rc = call(launch_ntt(s->ctx, *t, b, w.scratch, 1, false, w.stream));
			if (rc == VKHEL_OK)
				rc = call(launch_elemmul(r, w.scratch, r, n, q, w.stream));
			if (rc == VKHEL_OK)
				rc = call(launch_ntt(s->ctx, *t, r, r, 1, true, w.stream));
			break;
		case VKHEL_JOB_MUL:
			rc = call(launch_elemmul(a, b, r, n, q, w.stream));
			break;
		case VKHEL_JOB_ADD:
			rc = call(launch_addmod(a, b, r, n, q, w.stream));
			break;
		default:
			rc = call(launch_elemfma(a, b, r, n, 1, q, w.stream));
			break;
		}
		if (rc != VKHEL_OK)
			return rc;
	}
	return VKHEL_OK;
}

static void worker_loop(vkhel_server *s, int wi) {
	Worker &w = s->workers[wi];
	(void) hipSetDevice(s->ctx->device);
	for (;;) {
		std::vector<uint64_t> ids;
		std::vector<const Job *> jobs;
		{
			std::unique_lock<std::mutex> lk(s->mu);
			s->cv_work.wait(lk, [s] {
				return head_admissible(s) || (s->stop && s->waiting.empty());
			});
			if (s->waiting.empty() && s->stop)
				return;
			const uint64_t t = now_ns();
			uint64_t arena = 0, limbs = 0;
			while (ids.size() < s->cfg.max_batch && head_admissible(s)) {
				const uint64_t id = s->waiting.front();
				Request &r = s->reqs.at(id);
				// the micro-batch must fit this worker's arena and task
				// arrays (a request alone always does: submit checks it)
				if (arena + r.scratch > s->arena_words ||
						limbs + r.limbs > s->task_limbs)
					break;
				arena += r.scratch;
				limbs += r.limbs;
				s->waiting.pop_front();
				r.state = kRunning;
				r.info.admit_ns = t;
				s->inflight++;
				s->inflight_bytes += r.bytes;
				ids.push_back(id);
				jobs.push_back(&r.job);
			}
			s->stats.peak_inflight_bytes = std::max(s->stats.peak_inflight_bytes,
					s->inflight_bytes);
		}
		if (ids.empty())
			continue;
		// Job records stay valid: only the waiter erases a request, and only
		// once it is done.
		const int rc = s->cfg.per_call ? run_per_call(s, w, *jobs[0])
			: run_micro_batch(s, w, jobs);
		const uint64_t t = now_ns();
		{
			std::lock_guard<std::mutex> lk(s->mu);
			uint64_t coeffs = 0;
			for (uint64_t id : ids) {
				Request &r = s->reqs.at(id);
				r.state = kDone;
				r.status = rc;
				r.info.status = rc;
				r.info.complete_ns = t;
				r.info.batch_size = (uint32_t) ids.size();
				s->inflight--;
				s->inflight_bytes -= r.bytes;
				coeffs += job_transform_coeffs(r.job);
			}
			s->stats.completed += ids.size();
			s->stats.batches++;
			s->stats.transform_coeffs += coeffs;
		}
		s->cv_done.notify_all();
		s->cv_work.notify_all();
	}
}

}  // namespace vk

using namespace vk;

extern "C" {

struct vkhel_server *vkhel_server_create(struct vkhel_ctx *ctx,
		const struct vkhel_server_config *cfg) {
	if (!ctx)
		return fail(VKHEL_ERR_INVALID, "server_create: null ctx"), nullptr;
	vkhel_server *s = new vkhel_server;
	s->ctx = ctx;
	if (cfg)
		s->cfg = *cfg;
	if (!s->cfg.hbm_budget_bytes)
		s->cfg.hbm_budget_bytes = 200ull * 1000 * 1000 * 1000;
	if (!s->cfg.max_inflight)
		s->cfg.max_inflight = 4096;
	if (!s->cfg.max_batch)
		s->cfg.max_batch = 256;
	if (!s->cfg.streams)
		s->cfg.streams = 4;
	if (!s->cfg.queue_capacity)
		s->cfg.queue_capacity = 65536;
	if (s->cfg.per_call) {
		// one request at a time, in arrival order: a plain synchronous server
		s->cfg.streams = 1;
		s->cfg.max_batch = 1;
	}
	if (s->cfg.streams > 32) {
		delete s;
		return fail(VKHEL_ERR_INVALID, "server_create: at most 32 streams"),
			nullptr;
	}
	// Arenas: 1 GiB per stream by default, or an eighth of the budget over
	// the streams when that is less; they count against the budget.
	if (!s->cfg.scratch_bytes)
		s->cfg.scratch_bytes = std::min<uint64_t>(1ull << 30,
				s->cfg.hbm_budget_bytes / 8 / s->cfg.streams);
	s->arena_words = s->cfg.scratch_bytes / 8;
	s->reserved_bytes = 8 * s->arena_words * s->cfg.streams;
	if (s->reserved_bytes > s->cfg.hbm_budget_bytes) {
		delete s;
		return fail(VKHEL_ERR_BUDGET, "server_create: %u arenas of %llu bytes "
				"exceed the budget of %llu bytes", s->cfg.streams,
				(unsigned long long) s->cfg.scratch_bytes,
				(unsigned long long) s->cfg.hbm_budget_bytes), nullptr;
	}
	if (!s->cfg.task_limbs)
		s->cfg.task_limbs = 16384;
	s->task_limbs = std::max<uint64_t>(s->cfg.task_limbs, 1);
	s->workers.resize(s->cfg.streams);
	for (Worker &w : s->workers) {
		// every task array holds two tasks per limb (a polymul limb's two
		// forward transforms); RESCALE_NTT needs one per array and limb
		bool ok = hip_ok(hipStreamCreateWithFlags(&w.stream,
					hipStreamNonBlocking), "server stream") &&
			w.dt.reserve(2 * s->task_limbs);
		if (ok && s->arena_words) {
			ok = hip_ok(hipMalloc(&w.scratch, 8 * s->arena_words),
					"hipMalloc(server arena)");
			w.scratch_words = ok ? s->arena_words : 0;
		}
		if (!ok) {
			for (Worker &x : s->workers) {
				if (x.stream)
					(void) hipStreamDestroy(x.stream);
				x.dt.release();
				(void) hipFree(x.scratch);
			}
			delete s;
			return nullptr;
		}
	}
	for (uint32_t i = 0; i < s->cfg.streams; i++)
		s->threads.emplace_back(worker_loop, s, (int) i);
	return s;
}

int vkhel_server_drain(struct vkhel_server *s) {
	if (!s)
		return fail(VKHEL_ERR_INVALID, "server_drain: null server");
	std::unique_lock<std::mutex> lk(s->mu);
	s->cv_done.wait(lk, [s] { return s->waiting.empty() && s->inflight == 0; });
	return VKHEL_OK;
}

void vkhel_server_destroy(struct vkhel_server *s) {
	if (!s)
		return;
	{
		std::lock_guard<std::mutex> lk(s->mu);
		s->stop = true;
	}
	s->cv_work.notify_all();
	for (std::thread &t : s->threads)
		t.join();
	for (Worker &w : s->workers) {
		w.dt.release();
		(void) hipFree(w.scratch);
		(void) hipStreamDestroy(w.stream);
	}
	delete s;
}

}  // extern "C"

// Queues a checked job; returns its id, or 0 with the error recorded.
static uint64_t submit_request(vkhel_server *s, Request &&r) {
	r.bytes = job_operand_bytes(r.job);
	r.scratch = job_scratch_words(r.job);
	r.limbs = r.job.tables.size();
	r.info.submit_ns = now_ns();
	uint64_t id;
	{
		std::lock_guard<std::mutex> lk(s->mu);
		if (s->stop)
			return fail(VKHEL_ERR_INVALID, "server_submit: server stopping"), 0;
		if (r.bytes > operand_budget(s)) {
			s->stats.rejected++;
			return fail(VKHEL_ERR_BUDGET,
					"server_submit: request needs %llu bytes, budget is %llu "
					"(%llu after the %llu reserved for scratch arenas)",
					(unsigned long long) r.bytes,
					(unsigned long long) s->cfg.hbm_budget_bytes,
					(unsigned long long) operand_budget(s),
					(unsigned long long) s->reserved_bytes), 0;
		}
		if (r.scratch > s->arena_words || r.limbs > s->task_limbs) {
			s->stats.rejected++;
			return fail(VKHEL_ERR_BUDGET,
					"server_submit: request needs %llu scratch words and %llu "
					"limbs; a worker holds %llu and %llu",
					(unsigned long long) r.scratch, (unsigned long long) r.limbs,
					(unsigned long long) s->arena_words,
					(unsigned long long) s->task_limbs), 0;
		}
		if (s->waiting.size() >= s->cfg.queue_capacity) {
			s->stats.rejected++;
			return fail(VKHEL_ERR_BUSY, "server_submit: queue full"), 0;
		}
		id = s->next_id++;
		s->reqs.emplace(id, std::move(r));
		s->waiting.push_back(id);
		s->stats.submitted++;
	}
	s->cv_work.notify_one();
	return id;
}

extern "C" {

uint64_t vkhel_server_submit(struct vkhel_server *s, int op,
		const struct vkhel_vector *a, const struct vkhel_vector *b,
		struct vkhel_vector *res, struct vkhel_ntt_tables *const *tables,
		uint32_t nlimbs) {
	if (!s)
		return fail(VKHEL_ERR_INVALID, "server_submit: null server"), 0;
	Request r;
	if (make_job(s->ctx, op, a, b, res, tables, nlimbs, &r.job) != VKHEL_OK)
		return 0;
	return submit_request(s, std::move(r));
}

uint64_t vkhel_server_submit_rns(struct vkhel_server *s, int op,
		const struct vkhel_vector *a, struct vkhel_vector *res,
		struct vkhel_ntt_tables *const *tables, uint32_t nlimbs,
		const uint64_t *params) {
	if (!s)
		return fail(VKHEL_ERR_INVALID, "server_submit_rns: null server"), 0;
	Request r;
	if (make_rns_job(s->ctx, op, a, res, tables, nlimbs, params, &r.job) !=
			VKHEL_OK)
		return 0;
	return submit_request(s, std::move(r));
}

int vkhel_server_wait(struct vkhel_server *s, uint64_t id, uint64_t timeout_ns,
		struct vkhel_request_info *info) {
	if (!s)
		return fail(VKHEL_ERR_INVALID, "server_wait: null server");
	std::unique_lock<std::mutex> lk(s->mu);
	auto it = s->reqs.find(id);
	if (it == s->reqs.end())
		return fail(VKHEL_ERR_INVALID, "server_wait: unknown request %llu",
				(unsigned long long) id);
	auto done = [s, id] { return s->reqs.at(id).state == kDone; };
	if (timeout_ns) {
		if (!s->cv_done.wait_for(lk, std::chrono::nanoseconds(timeout_ns), done))
			return VKHEL_ERR_TIMEOUT;
	} else {
		s->cv_done.wait(lk, done);
	}
	Request &r = s->reqs.at(id);
	const int status = r.status;
	if (info)
		*info = r.info;
	s->reqs.erase(id);
	return status;
}

int vkhel_server_cancel(struct vkhel_server *s, uint64_t id) {
	if (!s)
		return fail(VKHEL_ERR_INVALID, "server_cancel: null server");
	{
		std::lock_guard<std::mutex> lk(s->mu);
		auto it = s->reqs.find(id);
		if (it == s->reqs.end())
			return fail(VKHEL_ERR_INVALID, "server_cancel: unknown request");
		Request &r = it->second;
		if (r.state != kWaiting)
			return fail(VKHEL_ERR_INVALID,
					"server_cancel: request already admitted");
		for (auto w = s->waiting.begin(); w != s->waiting.end(); ++w)
			if (*w == id) {
				s->waiting.erase(w);
				break;
			}
		r.state = kDone;
		r.status = VKHEL_ERR_CANCELLED;
		r.info.status = VKHEL_ERR_CANCELLED;
		r.info.complete_ns = now_ns();
		s->stats.cancelled++;
	}
	s->cv_done.notify_all();
	s->cv_work.notify_all();
	return VKHEL_OK;
}

int vkhel_server_get_stats(struct vkhel_server *s, struct vkhel_server_stats *out) {
	if (!s || !out)
		return fail(VKHEL_ERR_INVALID, "server_stats: null argument");
	std::lock_guard<std::mutex> lk(s->mu);
	*out = s->stats;
	out->reserved_bytes = s->reserved_bytes;
	out->inflight_bytes = s->inflight_bytes;
	out->inflight = s->inflight;
	out->waiting = (uint32_t) s->waiting.size();
	return VKHEL_OK;
}

}  // extern "C"

extern "C" int vkhel_server_replay(struct vkhel_server *s,
		const struct vkhel_job_desc *jobs, uint64_t njobs, double rate_per_s,
		uint64_t seed, struct vkhel_request_info *infos) {
	if (!s || (!jobs && njobs) || (!infos && njobs) || rate_per_s < 0)
		return fail(VKHEL_ERR_INVALID, "server_replay: bad argument");
	std::vector<uint64_t> ids(njobs, 0);
	uint64_t st = seed;
	double t_arrival = 0;
	const uint64_t t0 = now_ns();
	int rc = VKHEL_OK;
	for (uint64_t i = 0; i < njobs; i++) {
		if (rate_per_s > 0) {
			// exponential gap: -ln(U) / rate with U uniform in (0, 1]
			uint64_t z = (st += 0x9e3779b97f4a7c15ull);
			z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
			z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
			z ^= z >> 31;
			const double u = ((double) (z >> 11) + 1.0) * (1.0 / 9007199254740992.0);
			t_arrival += -std::log(u) / rate_per_s;
			const uint64_t due = t0 + (uint64_t) (t_arrival * 1e9);
			while (now_ns() < due)
				std::this_thread::yield();
		}
		const vkhel_job_desc &d = jobs[i];
		ids[i] = vkhel_server_submit(s, d.op, d.a, d.b, d.res, d.tables,
				d.nlimbs);
		if (!ids[i]) {
			rc = vkhel_ext_last_error();
			infos[i] = vkhel_request_info{};
			infos[i].status = rc;
		}
	}
	for (uint64_t i = 0; i < njobs; i++) {
		if (!ids[i])
			continue;
		const int st_i = vkhel_server_wait(s, ids[i], 0, &infos[i]);
		if (st_i != VKHEL_OK && rc == VKHEL_OK)
			rc = st_i;
	}
	return rc;
}
