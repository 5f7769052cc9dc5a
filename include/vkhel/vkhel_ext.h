/*
 * vkhel extensions for MI355X: error control, device-side access, batched
 * transforms and a serving engine.  The reference has no counterpart for
 * any of these (its API is vkhel.h only); they are additions layered on the
 * same handles, exported under the same vkhel_* version script.
 *
 * Functions returning int return 0 (VKHEL_OK) on success and a
 * vkhel_status code otherwise; vkhel_ext_last_error_message() describes
 * the last failure on the calling thread.
 */
#ifndef VKHEL_EXT_H
#define VKHEL_EXT_H

#include <stddef.h>
#include <stdint.h>

#include "vkhel.h"

#ifdef __cplusplus
extern "C" {
#endif

enum vkhel_status {
	VKHEL_OK = 0,
	VKHEL_ERR_INVALID = 1,   /* bad argument (sizes, moduli, handles) */
	VKHEL_ERR_DEVICE = 2,    /* a HIP runtime call failed */
	VKHEL_ERR_NOMEM = 3,     /* device or host allocation failed */
	VKHEL_ERR_BUDGET = 4,    /* serving: request exceeds the HBM budget */
	VKHEL_ERR_CANCELLED = 5, /* serving: request cancelled */
	VKHEL_ERR_TIMEOUT = 6,   /* serving: deadline passed before admission */
	VKHEL_ERR_BUSY = 7,      /* serving: queue full */
};

/* ---- error control ---- */
/* Default 1: print the message and abort(), as the reference's asserts do.
 * With 0, failing calls return early (NULL for constructors) and record
 * the error for the calling thread. */
void vkhel_ext_set_abort_on_error(int enabled);
int vkhel_ext_last_error(void);
const char *vkhel_ext_last_error_message(void);
void vkhel_ext_clear_error(void);

/* ---- device access ---- */
/* The hipStream_t all legacy calls of this context run on. */
void *vkhel_ext_ctx_stream(struct vkhel_ctx *);
int vkhel_ext_ctx_synchronize(struct vkhel_ctx *);
/* Device name, e.g. "gfx950"; valid while the context lives. */
const char *vkhel_ext_ctx_device_name(struct vkhel_ctx *);
uint64_t *vkhel_ext_vector_device_ptr(struct vkhel_vector *);
uint64_t vkhel_ext_vector_length(const struct vkhel_vector *);
/* A vector aliasing words [offset, offset + length) of `base` (e.g. one
 * RNS limb), usable wherever a vector is; destroying it frees only the
 * handle.  It must not outlive `base`. */
struct vkhel_vector *vkhel_ext_vector_view(struct vkhel_vector *base,
		uint64_t offset, uint64_t length);
/* Synchronous device<->host copies of `count` words starting at `offset`. */
int vkhel_ext_vector_read(const struct vkhel_vector *, uint64_t offset,
		uint64_t count, uint64_t *host);
int vkhel_ext_vector_write(struct vkhel_vector *, uint64_t offset,
		uint64_t count, const uint64_t *host);

/* vkhel_vector_map stages through page-locked buffers from a per-context
 * pool (reused by later maps; freed with the context).  enabled 0 frees
 * each buffer at unmap instead (and frees the idle ones now); 1 (default)
 * keeps them.  vkhel_ext_staging_bytes: bytes the pool holds. */
int vkhel_ext_set_staging_pool(struct vkhel_ctx *, int enabled);
uint64_t vkhel_ext_staging_bytes(struct vkhel_ctx *);

/* ---- asynchronous transfers ---- */
/* Page-locked host memory (null on failure); free with vkhel_ext_host_free.
 * Copies from and to it run as DMA without a staging copy and overlap
 * kernels running on other streams. */
void *vkhel_ext_host_alloc(size_t bytes);
void vkhel_ext_host_free(void *);
/* Enqueue a copy of `count` words at `offset` to / from `host` on `stream`
 * (a hipStream_t; null: the context's stream) and return without waiting.
 * The caller orders the copy against kernels on other streams (HIP events)
 * and keeps `host` alive and untouched until the copy completes. */
int vkhel_ext_vector_write_async(struct vkhel_vector *, uint64_t offset,
		uint64_t count, const uint64_t *host, void *stream);
int vkhel_ext_vector_read_async(const struct vkhel_vector *, uint64_t offset,
		uint64_t count, uint64_t *host, void *stream);

/* Fills `count` words from `offset` with synthetic residues uniform in
 * [0, q): word i is the high half of splitmix64(seed, i) * q. */
int vkhel_ext_vector_fill_uniform(struct vkhel_vector *, uint64_t offset,
		uint64_t count, uint64_t q, uint64_t seed);

/* ---- tables ---- */
/* Reads back the reference-layout host tables (n words each), computing
 * them first for tables made by vkhel_ext_tables_create_device. */
int vkhel_ext_tables_info(const struct vkhel_ntt_tables *, uint64_t *n,
		uint64_t *q, uint64_t *w);
int vkhel_ext_tables_copy(const struct vkhel_ntt_tables *, uint64_t *roots,
		uint64_t *inv_roots, uint64_t *roots_bf, uint64_t *inv_roots_bf);
/* Builds the device copy of the tables on `ctx` now instead of at first
 * use.  Device copies are generated on the GPU from (n, q, w); nothing
 * crosses PCIe. */
int vkhel_ext_tables_upload(struct vkhel_ctx *, struct vkhel_ntt_tables *);
/* Tables built on the GPU of `ctx` only (src/ntt_tables.c:65-80 builds
 * them on the host, O(n) per table): the host arrays that
 * vkhel_ext_tables_copy returns are computed on first request.  Unlike
 * vkhel_ntt_tables_create, w must be a primitive 2n-th root of unity mod q
 * (w^n = -1), else VKHEL_ERR_INVALID.  Release with
 * vkhel_ntt_tables_destroy. */
struct vkhel_ntt_tables *vkhel_ext_tables_create_device(struct vkhel_ctx *,
		uint64_t n, uint64_t q, uint64_t w);
/* 1 if the tables' w is a primitive 2n-th root of unity mod q, 0 if not
 * (the reference accepts any invertible w and never checks); -1 with the
 * error set for a null handle. */
int vkhel_ext_tables_root_ok(const struct vkhel_ntt_tables *);
/* Reads the device copy on `ctx` back in the reference's index order
 * (n words each; any pointer may be NULL): roots[k] and inv_roots[k] as in
 * the host arrays, with the packed 63-bit Shoup companions the kernels use,
 * (2 * (wp >> 32)) << 32 | (wp mod 2^32) for wp = floor(root * 2^63 / q). */
int vkhel_ext_tables_read_device(struct vkhel_ctx *,
		const struct vkhel_ntt_tables *, uint64_t *roots,
		uint64_t *roots_shoup, uint64_t *inv_roots, uint64_t *inv_roots_shoup);

/* ---- NTT kernel selection ---- */
enum vkhel_ntt_algo {
	VKHEL_NTT_AUTO = 0,     /* autotuned choice per n (default) */
	VKHEL_NTT_RADIX2 = 1,   /* one launch per stage (validation path) */
	VKHEL_NTT_FUSED = 2,    /* stage-fused LDS/register kernels */
};
int vkhel_ext_set_ntt_algo(struct vkhel_ctx *, int algo);

/* ---- autotuning of the fused NTT kernels ----
 * Sizes above a tile run as two fused passes over tiles of 2^11, 2^12,
 * 2^13 or 2^14 words (the tile size fixes the stage split and workgroup
 * shape; 2^14-word tiles use 144 KiB of LDS and 1024 threads, and run
 * n = 2^14 in one pass).
 * vkhel_ext_autotune times each compiled tile size, with both pass modes
 * (below), for every log2 n in [logn_lo, logn_hi] (3..20) on `batch`
 * synthetic polynomials (0: enough to fill the device) and keeps the
 * fastest per size on this context. */
int vkhel_ext_autotune(struct vkhel_ctx *, uint32_t logn_lo, uint32_t logn_hi,
		uint32_t batch);
/* Tile size (log2 words) in use for log2 n, or -1 on error. */
int vkhel_ext_tile_bits(struct vkhel_ctx *, uint32_t logn);
/* tile_bits 11..14, or 0 for the built-in default. */
int vkhel_ext_set_tile_bits(struct vkhel_ctx *, uint32_t logn, int tile_bits);
/* Pass mode for log2 n: 0 = built-in default (split), 1 = one workgroup
 * per polynomial for both passes, 2 = one grid per pass with one workgroup
 * per tile (split), 3 = both passes in one launch (see below).  The
 * autotuner sets it too. */
int vkhel_ext_pass_mode(struct vkhel_ctx *, uint32_t logn);
int vkhel_ext_set_pass_mode(struct vkhel_ctx *, uint32_t logn, int mode);
/* Pass mode 3 runs both passes of a two-pass size in one launch (batches
 * and serving only; single calls keep mode 2): each XCD serves a queue of
 * tiles in which the first pass of a polynomial runs `polys` polynomials
 * ahead of its second, and hands the polynomial from one pass to the other
 * through its own L2.  0 (default): derived from the occupancy. */
int vkhel_ext_set_handoff_ahead(struct vkhel_ctx *, int polys);
/* Split passes (mode 2) can also be shaped per size: the stage split (0 =
 * the default: pass 1 takes the top log2 n - 9 stages where rows of 512
 * words fit, else floor(log2 n / 2); 1 = one column stage more, or one
 * fewer where that does not fit; sizes with no alternative ignore it) and
 * the row pass's own tile (0 = the tile_bits tile; ignored where that tile
 * would split the stages differently).  vkhel_ext_autotune tunes tile and
 * pass mode, then the split and the row-pass tile, timing each candidate
 * through the batch path. */
int vkhel_ext_split(struct vkhel_ctx *, uint32_t logn);
int vkhel_ext_set_split(struct vkhel_ctx *, uint32_t logn, int split);
int vkhel_ext_tile_bits2(struct vkhel_ctx *, uint32_t logn);
int vkhel_ext_set_tile_bits2(struct vkhel_ctx *, uint32_t logn, int tile_bits);
/* Tuned choices as text lines "log2n tile_bits pass_mode split tile_bits2"
 * (lines of the first three fields load too). */
int vkhel_ext_autotune_save(struct vkhel_ctx *, const char *path);
int vkhel_ext_autotune_load(struct vkhel_ctx *, const char *path);

/* ---- batched transforms ----
 * A batch is a list of jobs executed by one launch sequence on the context
 * stream (planned once while the job list is unchanged; replayed from a
 * captured hipGraph when it has more than a few launches).  A job works on an RNS polynomial: `nlimbs` limbs of n words
 * stored back to back in one vector (limb i at word offset i*n), limb i
 * reduced modulo tables[i]->q.  Every tables[i] of one job has the same n.
 * Vectors must hold at least nlimbs*n words.  The vectors and tables a job
 * names must outlive every launch of the batch that holds it (destroying
 * tables waits for the device, but a launch enqueued later would read
 * freed twiddles).  Inputs of NTT jobs follow the
 * single-call rules: forward accepts any 64-bit words, inverse and polymul
 * operands of the inverse step expect words < q.
 */
enum vkhel_job_op {
	VKHEL_JOB_NTT_FWD = 0,  /* res = NTT(a)                               */
	VKHEL_JOB_NTT_INV = 1,  /* res = INTT(a)                              */
	VKHEL_JOB_POLYMUL = 2,  /* res = INTT(NTT(a) . NTT(b)) = a*b mod X^n+1 */
	VKHEL_JOB_MUL = 3,      /* res = a.b mod q   (elemmul semantics)      */
	VKHEL_JOB_ADD = 4,      /* res = (a + b) mod q for a, b < q           */
	/* RNS elementwise ops (vkhel_batch_add_rns), per limb i, q_i from
	 * tables[i], with the single-call semantics of the named function:   */
	VKHEL_JOB_CENTER = 5,   /* res_i = elemmod(a_i, params[0], q_i)        */
	VKHEL_JOB_GTADD = 6,    /* res_i = elemgtadd(a_i, params[2i], params[2i+1]) */
	VKHEL_JOB_GTSUB = 7,    /* res_i = elemgtsub(a_i, params[3i], params[3i+1],
	                                             params[3i+2])             */
	/* fused: res_i = elemmod(INTT(a_i), params[0], q_i), the centered lift
	 * applied as the inverse transform stores its results (one pass over
	 * memory less than INTT then CENTER) for n from 2^3 to 2^20; other
	 * sizes run the lift as a pointwise pass after the transform */
	VKHEL_JOB_INTT_CENTER = 8,
	/* RNS rescaling by the last modulus (nlimbs >= 2, params unused):
	 * res_i = (a_i - [a_last]) * q_last^-1 mod q_i for i < nlimbs - 1, where
	 * [a_last] is the centered lift of the last limb (a_last - q_last when
	 * a_last > q_last / 2); res holds nlimbs - 1 limbs and is the rounded
	 * quotient A / q_last of the integer A the limbs represent.  Inputs
	 * canonical; res must not alias a.  Runs after every transform of the
	 * batch, so it may consume an INTT of the same batch. */
	VKHEL_JOB_RESCALE = 9,
	/* the same rescaling on a polynomial in NTT form, result in NTT form:
	 * res_i = (a_i - NTT_i([INTT_last(a_last)])) * q_last^-1 mod q_i for
	 * i < nlimbs - 1, [.] the centered lift into Z_q_i; that is NTT_i of
	 * VKHEL_JOB_RESCALE applied to the INTTs of the limbs.  Runs as an
	 * inverse transform of the last limb, the lift into every other limb,
	 * their forward transforms and one pointwise pass, after every other
	 * transform of the batch.  Same checks as VKHEL_JOB_RESCALE. */
	VKHEL_JOB_RESCALE_NTT = 10,
	/* res = elemfma(a, b, 1, q) per limb: the reference's vector add
	 * (src/vector.c:298-340, elemfma.comp:36-55 with multiplier 1) word for
	 * word, including its q - x correction that wraps modulo 2^64 where
	 * (a mod q) + b >= q; any 64-bit inputs.  A batch of these computes
	 * exactly what one vkhel_vector_elemfma(a, b, res, 1, q) call per limb
	 * does; VKHEL_JOB_ADD is the correct (a + b) mod q. */
	VKHEL_JOB_ADD_REF = 11,
};

struct vkhel_batch;
struct vkhel_batch *vkhel_batch_create(struct vkhel_ctx *);
void vkhel_batch_destroy(struct vkhel_batch *);
void vkhel_batch_clear(struct vkhel_batch *);
/* b is ignored (may be NULL) for NTT_FWD / NTT_INV. */
int vkhel_batch_add(struct vkhel_batch *, int op,
		const struct vkhel_vector *a, const struct vkhel_vector *b,
		struct vkhel_vector *res, struct vkhel_ntt_tables *const *tables,
		uint32_t nlimbs);
/* Adds an RNS job (VKHEL_JOB_CENTER / _GTADD / _GTSUB, elementwise: every
 * limb of every such job in the batch runs in the batch's one pointwise
 * launch; _INTT_CENTER, _RESCALE, _RESCALE_NTT as documented above).
 * CENTER is the centered lift of limb i (mod q_i) reduced mod p =
 * params[0]; GTADD / GTSUB take (bound, diff) / (bound, diff, mod) per
 * limb; the rescaling ops take none (params may be NULL).  `params` is
 * copied. */
int vkhel_batch_add_rns(struct vkhel_batch *, int op,
		const struct vkhel_vector *a, struct vkhel_vector *res,
		struct vkhel_ntt_tables *const *tables, uint32_t nlimbs,
		const uint64_t *params);
/* Enqueues the batch on the context stream and returns immediately.  The
 * batch is built (task arrays uploaded, launch sequence captured) on a
 * stream of its own, so batches may be built and launched from several
 * threads while other calls use the same context; one batch object is used
 * by one thread at a time.  Relaunching or changing a batch waits for its
 * previous launch. */
int vkhel_batch_launch(struct vkhel_batch *);
/* Completion handle of the last launch: wait blocks until it finished
 * (VKHEL_OK at once if the batch was not launched), query returns 1 if it
 * finished, 0 if it is still running, -1 on error. */
int vkhel_batch_wait(struct vkhel_batch *);
int vkhel_batch_query(struct vkhel_batch *);
/* vkhel_batch_launch + vkhel_batch_wait. */
int vkhel_batch_run(struct vkhel_batch *);
uint64_t vkhel_batch_num_jobs(const struct vkhel_batch *);

/* ---- serving engine ----
 * Continuous batching over device-resident operands.  Requests use the
 * batch job model.  `streams` workers each own a HIP stream, task arrays
 * and a scratch arena (polymul spectra, rescaling temporaries), all
 * allocated at vkhel_server_create; the arenas count against
 * hbm_budget_bytes and nothing is allocated while serving.  Requests are
 * queued on submit and admitted in arrival order while the operand and
 * result bytes of admitted, unfinished requests stay within the budget
 * left after the arenas and at most max_inflight requests are in flight.
 * A free worker takes every admissible waiting request that fits its arena
 * and task arrays (up to max_batch) as one micro-batch, so requests
 * arriving while others run are admitted as budget and slots free up and
 * overlap with them on the device.  A request's result does not depend on
 * what it was batched with.  Callers keep operands alive and unmodified,
 * and their tables alive, until the request completes.
 */
struct vkhel_server_config {
	uint64_t hbm_budget_bytes; /* 0: 200 GB */
	uint32_t max_inflight;     /* 0: 4096 requests */
	uint32_t max_batch;        /* requests per micro-batch, 0: 256 */
	uint32_t streams;          /* worker streams, 0: 4 (at most 32) */
	uint32_t queue_capacity;   /* waiting requests, 0: 65536 */
	/* scratch arena per stream; 0: 1 GiB, or hbm_budget_bytes / 8 /
	 * streams when that is less */
	uint64_t scratch_bytes;
	/* polynomial limbs per micro-batch (task array capacity), 0: 16384 */
	uint32_t task_limbs;
	/* 1: plain per-call service instead of batching -- one stream, one
	 * request at a time in arrival order, each limb and step a synchronous
	 * single call (the comparison point for latency); 0: batching */
	uint32_t per_call;
};

struct vkhel_request_info {
	int status;            /* VKHEL_OK, VKHEL_ERR_CANCELLED, ... */
	uint64_t submit_ns;    /* steady-clock timestamps */
	uint64_t admit_ns;
	uint64_t complete_ns;
	uint32_t batch_size;   /* requests in the micro-batch that ran it */
};

struct vkhel_server_stats {
	uint64_t submitted, completed, cancelled, rejected;
	uint64_t batches;            /* micro-batches run */
	uint64_t transform_coeffs;   /* NTT coefficients processed */
	uint64_t inflight_bytes, peak_inflight_bytes;
	uint32_t inflight, waiting;
	uint64_t reserved_bytes;     /* scratch arenas held since creation */
};

struct vkhel_server;
struct vkhel_server *vkhel_server_create(struct vkhel_ctx *,
		const struct vkhel_server_config *);
/* Finishes admitted and waiting work, then stops the workers. */
void vkhel_server_destroy(struct vkhel_server *);
/* Returns a request id > 0, or 0 with the error recorded: VKHEL_ERR_BUDGET
 * if the request alone exceeds the operand budget, a worker's arena or its
 * task arrays, VKHEL_ERR_BUSY if the queue is full, VKHEL_ERR_INVALID for
 * bad arguments. */
uint64_t vkhel_server_submit(struct vkhel_server *, int op,
		const struct vkhel_vector *a, const struct vkhel_vector *b,
		struct vkhel_vector *res, struct vkhel_ntt_tables *const *tables,
		uint32_t nlimbs);
/* As vkhel_server_submit for the RNS jobs of vkhel_batch_add_rns
 * (VKHEL_JOB_CENTER ... VKHEL_JOB_RESCALE_NTT, the same arguments). */
uint64_t vkhel_server_submit_rns(struct vkhel_server *, int op,
		const struct vkhel_vector *a, struct vkhel_vector *res,
		struct vkhel_ntt_tables *const *tables, uint32_t nlimbs,
		const uint64_t *params);
/* Waits up to timeout_ns (0: no limit).  Returns the request's status, or
 * VKHEL_ERR_TIMEOUT if it is still pending.  A completed request is
 * released by the wait that returns its status. */
int vkhel_server_wait(struct vkhel_server *, uint64_t id, uint64_t timeout_ns,
		struct vkhel_request_info *info);
/* Cancels a request that has not been admitted yet. */
int vkhel_server_cancel(struct vkhel_server *, uint64_t id);
/* Blocks until nothing is waiting or in flight. */
int vkhel_server_drain(struct vkhel_server *);
int vkhel_server_get_stats(struct vkhel_server *, struct vkhel_server_stats *);

/* Load generator: submits jobs[i] at the i-th arrival of a Poisson process
 * of `rate_per_s` arrivals per second (exponential gaps from a seeded
 * splitmix64 stream; rate 0 submits everything at once), waits for all of
 * them and stores each request's record in infos[i].  Returns VKHEL_OK or
 * the first failing status. */
struct vkhel_job_desc {
	int op;
	const struct vkhel_vector *a;
	const struct vkhel_vector *b;
	struct vkhel_vector *res;
	struct vkhel_ntt_tables *const *tables;
	uint32_t nlimbs;
};
int vkhel_server_replay(struct vkhel_server *, const struct vkhel_job_desc *jobs,
		uint64_t njobs, double rate_per_s, uint64_t seed,
		struct vkhel_request_info *infos);

/* ---- polynomial products without device allocation ----
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory: code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc.  Computes what a one-job batch with VKHEL_JOB_POLYMUL
 * computes (res = a*b mod (X^n + 1, q_i) per limb, the same kernels and
 * bit-identical words), enqueued on `stream` (a hipStream_t; null: the
 * context's stream) without waiting.  It allocates nothing: b's spectra go
 * to a static device array of VKHEL_POLYMUL_NOALLOC_MAX_WORDS words per GPU
 * (32 MiB, held until the process ends) and every kernel takes its task by
 * value.  So nlimbs * n must not exceed that maximum, and every limb's
 * tables must already be resident on the context (vkhel_ext_tables_upload,
 * or any earlier transform with them); otherwise VKHEL_ERR_INVALID and
 * nothing is enqueued.  Every call on one GPU shares its array: calls that
 * overlap on one GPU (other streams, other threads) race, so serialise them
 * on one stream.  res must not alias b. */
#define VKHEL_POLYMUL_NOALLOC_MAX_WORDS (1ull << 22)
int vkhel_ext_polymul_noalloc(struct vkhel_ctx *, const struct vkhel_vector *a,
		const struct vkhel_vector *b, struct vkhel_vector *res,
		struct vkhel_ntt_tables *const *tables, uint32_t nlimbs, void *stream);

#ifdef __cplusplus
}
#endif

#endif
