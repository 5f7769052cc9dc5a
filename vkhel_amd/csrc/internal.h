// Internal types of the MI355X vkhel library.  Host code is C++17 built by
// hipcc for gfx950; nothing here is part of the public ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "vkhel/vkhel.h"
#include "vkhel/vkhel_ext.h"

namespace vk {

// Records an error for the calling thread, prints it, and aborts unless
// vkhel_ext_set_abort_on_error(0) was called.  Returns `code`.
int fail(int code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
void clear_error();

// Returns true on success; otherwise records VKHEL_ERR_DEVICE (or NOMEM).
bool hip_ok(hipError_t e, const char *what);

// Constants one (n, q) transform needs besides its twiddle tables.  Passed
// to kernels by value.
struct NttConsts {
	uint64_t q;
	uint64_t two_q;
	uint64_t r64;     // 2^64 mod q: correction for the wrapped X + WY sum
	uint64_t mu;      // floor(2^(bits+62) / q), bits = bitlen(q)
	uint64_t ninv;    // n^-1 mod q
	uint64_t ninv_p;  // floor(ninv * 2^63 / q), the 63-bit Shoup companion
	uint64_t w1n;     // inv_roots[1] * n^-1 mod q (last inverse stage)
	uint64_t w1n_p;
	// final reduction of values below 81q (forward, q < kNoReduceQ):
	// quotient estimate umulhi(v >> red_shift, red_m), see numbers.h
	uint32_t red_shift, red_m;
	// epilogue of an inverse transform (0: none): the result is replaced
	// by elemmod(result, post_p, q) (src/vector.c:342-386), with the
	// Barrett constants of post_p
	uint64_t post_p, post_mu;
	uint32_t post_shift;
	// prologue of a forward transform (lift_add 0: none): every input word
	// x > lift_bound becomes x + lift_add (with lift_add = q - (q' mod q),
	// the centered lift of a word mod q' into Z_q, left for the first
	// stage to reduce)
	uint64_t lift_bound, lift_add;
	// epilogue of a forward transform (scale_w 0: none): the result y at
	// word i becomes (in2[i] - y) * scale_w mod q, scale_wp the packed
	// Shoup companion of scale_w
	uint64_t scale_w, scale_wp;
	uint32_t bshift;  // bits - 2
	uint32_t logn;
};

// Device copy of one vkhel_ntt_tables: twiddles stored as {w, w'} pairs in
// the reference's bit-reversed index order (roots[m + i] of stage m).
struct DevTables {
	uint64_t n = 0;
	const ulonglong2 *fwd = nullptr;
	const ulonglong2 *inv = nullptr;
	NttConsts c{};
	void *alloc = nullptr;
};

}  // namespace vk

// Field order of the first seven members follows the reference's private
// struct (include/priv/ntt_tables.h:6-15) so code written against it keeps
// working; `id` keys the per-context device cache.
struct vkhel_ntt_tables {
	uint64_t n;
	uint64_t q;
	uint64_t w;
	uint64_t *roots_of_unity;
	uint64_t *inv_roots_of_unity;
	uint64_t *roots_barrett_factors;
	uint64_t *inv_roots_barrett_factors;
	uint64_t id;
	// 1 if w is a primitive 2n-th root of unity mod q (w^n = -1), which the
	// reference never checks; set by the create calls.
	int root_ok;
};

struct vkhel_ctx {
	int device = 0;
	hipStream_t stream = nullptr;
	// Private stream for table generation: never captured, so generating
	// a table is safe while another thread captures on `stream`.
	hipStream_t table_stream = nullptr;
	std::string device_name;
	int cu_count = 0;
	int ntt_algo = VKHEL_NTT_AUTO;
	// Tuned tile size (log2 words) per log2 n; 0 = default.
	std::vector<int> tile_bits = std::vector<int>(64, 0);
	// Tuned tile of a two-pass size's row pass (0: tile_bits), and stage
	// split (0: the default split, 1: the alternative, ntt_fused.h
	// alt_k1), per log2 n.
	std::vector<int> tile_bits2 = std::vector<int>(64, 0);
	std::vector<int> split = std::vector<int>(64, 0);
	// Tuned pass mode per log2 n: 0 = by batch size, 1 = whole polynomial
	// per workgroup, 2 = one grid per pass (one workgroup per tile), 3 = both
	// passes in one persistent launch (batches only).
	std::vector<int> pass_mode = std::vector<int>(64, 0);
	// Look-ahead of the one-launch two-pass schedule in polynomials
	// (handoff.h; 0: from the occupancy).
	int handoff_ahead = 0;

	std::mutex mu;  // guards `tables` (batch and serve threads share it)
	std::unordered_map<uint64_t, vk::DevTables *> tables;
	// Page-locked staging buffers of vkhel_vector_map, reused: a map takes
	// the smallest free buffer that fits (allocating only when none does),
	// unmap returns it.  staging_pool 0: free on unmap instead.
	struct Staging {
		void *p;
		size_t bytes;
		bool busy;
	};
	std::mutex staging_mu;
	std::vector<Staging> staging;
	int staging_pool = 1;
};

struct vkhel_vector {
	struct vkhel_ctx *ctx;
	uint64_t length;
	uint64_t *dev;
	void *staging;  // pinned host buffer while mapped
	size_t staging_words;
	bool view;      // aliases another vector's memory (vkhel_ext_vector_view)
};

namespace vk {

// Device tables of `t` on `ctx`, generated on the GPU on first use.
// nullptr on failure (error recorded).
const DevTables *device_tables(vkhel_ctx *ctx, const vkhel_ntt_tables *t);

// Launch arguments of the table generator: the modulus constants and the
// binary powers w^(2^j) and w^-(2^j), j < log2 n.
constexpr int kMaxTableLog = 28;
struct TableGenArgs {
	uint64_t q, mu, r64, qinv;
	uint32_t shift, logn;
	uint64_t wpow[kMaxTableLog], wipow[kMaxTableLog];
};
void table_gen_args(uint64_t n, uint64_t q, uint64_t w, TableGenArgs &g);
// Writes both interleaved {w, Shoup companion} tables of n = 2^g.logn
// entries (the layout tw_pos describes) on stream `st`.
hipError_t launch_gen_tables(const TableGenArgs &g, ulonglong2 *fwd,
		ulonglong2 *inv, hipStream_t st);
uint64_t inv_mod_2_64(uint64_t q);

// One polynomial transform for the fused kernel: which words, which
// direction's twiddles, the per-(n, q) constants and the pass split
// (k1 = stages in pass A; 0 for n <= 2^13).
// Device twiddle tables keep the reference's per-stage blocks (stage with m
// groups reads roots[m + i], src/vector.c:536-547) but store each block of m >= 4
// entries 4-way interleaved: entry 4j + g of the block at g * (m/4) + j.
// In the fused kernels' innermost stages a thread uses four consecutive
// entries, so with this order each wave-wide load reads one contiguous run
// instead of 16 bytes out of every 64.
__host__ __device__ __forceinline__ uint32_t tw_pos(uint32_t m, uint32_t i) {
	return m < 4 ? m + i : m + (i & 3) * (m >> 2) + (i >> 2);
}

// Transforms over moduli below this bound use the small-modulus kernels:
// forward butterflies without input reduction and with products below 4q,
// so values entering stage s >= 1 stay below (4s + 1) q, below 81q < 2^63
// after the last stage of n = 2^20; the inverse defers its sum reductions
// (bounds below 64q).  The 63-bit Shoup product accepts all of these.
// (2^55 rather than 2^56: the high word of such a q is below 2^23, which
// the final reduction's 24-bit multiply needs, modarith.h
// sub_small_multiple.)
constexpr uint64_t kNoReduceQ = 1ull << 55;
// Below this bound (and from kNoReduceQ up) the mid class: forward values
// kept below 8q with one reduction per butterfly, inverse below 4q, both
// with products below 4q (8q < 2^63).
constexpr uint64_t kMidQ = 1ull << 60;
enum { kClassGeneral = 0, kClassMid = 1, kClassSmall = 2 };
inline int modulus_class(uint64_t q) {
	return q < kNoReduceQ ? kClassSmall : q < kMidQ ? kClassMid : kClassGeneral;
}

struct NttTask {
	const uint64_t *in;
	uint64_t *out;
	const ulonglong2 *tw;
	NttConsts c;
	// Fused kernels only.  Inverse transforms: when set, the transform's
	// input is the pointwise product in[i] * in2[i] mod q of two canonical
	// spectra, formed as they are loaded (a polynomial product's last step).
	// Forward transforms: the operand of the c.scale_* epilogue.
	const uint64_t *in2;
	// Forward twiddles of a fused polynomial-product task (tw: inverse).
	const ulonglong2 *tw2;
};

// Launchers (ntt.hip).  `batch` polynomials of n words, poly p at
// in + p*n / out + p*n, all with the same tables.  Stream-ordered.
int launch_ntt(vkhel_ctx *ctx, const DevTables &t, const uint64_t *in,
		uint64_t *out, uint32_t batch, bool inverse, hipStream_t s);
NttTask make_task(const DevTables &t, const uint64_t *in, uint64_t *out,
		bool inverse);
// Device-resident tasks that share log2 n (which selects the compiled
// kernel), direction and modulus class (mc, modulus_class() of every q);
// the context supplies the tuned tile size.
// Pass mode 3 (one launch for both passes, k_ntt_handoff) needs a state
// buffer of handoff_state_bytes(ntasks); without one the passes run split.
int launch_fused_tasks(const vkhel_ctx *ctx, const NttTask *dev_tasks,
		uint32_t ntasks, uint32_t logn, bool inverse, int mc, hipStream_t st,
		uint32_t *handoff_state = nullptr, uint64_t handoff_bytes = 0);
uint64_t handoff_state_bytes(uint32_t ntasks);
// True when transforms of log2 n run both passes in one launch on ctx
// (pass mode 3 at a two-pass size).
bool uses_handoff(const vkhel_ctx *ctx, uint32_t logn);
// After a one-launch transform of npoly polynomials of log2 n on ctx (its
// tile) has completed: its state shows every ticket done and no wait given
// up.  Reads the state back synchronously.
bool handoff_state_ok(const vkhel_ctx *ctx, const uint32_t *state,
		uint32_t npoly, uint32_t logn);
// Stages of pass 1 for log2 n on 2^tb tiles with split choice s (0 for
// one-tile sizes).
int fused_split_k1(uint32_t logn, int tb, int s);
// Fused polynomial products (k_ntt_pm): the tile size (log2 words) they
// run on for log2 n on this context, 0 when that size and tile do not allow
// them.
int polymul_fused_tile(const vkhel_ctx *ctx, uint32_t logn);
// Launches `ntasks` fused polynomial products on 2^tile_bits tiles:
// pm_tasks {in = a, out, tw = inverse twiddles, tw2 = forward twiddles,
// in2 = b}.  Two-pass sizes also take col_tasks, two per product: tasks 2i
// and 2i + 1 the forward column passes of product i's a (into its out) and
// b (into scratch), which the pm task then names as in and in2; one-tile
// sizes take col_tasks null and read a and b themselves.
int launch_polymul_tasks(const NttTask *col_tasks, const NttTask *pm_tasks,
		uint32_t ntasks, uint32_t logn, int tile_bits, int mc, hipStream_t st);
// One fused polynomial product without task arrays: the column tasks of a
// and b (two-pass sizes, cols true) and the product task passed by value.
int launch_polymul_one(const NttTask &col_a, const NttTask &col_b,
		const NttTask &pm, bool cols, uint32_t logn, int tile_bits, int mc,
		hipStream_t st);
// One fused transform of task k (its in2 / epilogue fields honoured), by
// value: reads no task array.
int launch_task_fused(const vkhel_ctx *ctx, const NttTask &k, bool inverse,
		hipStream_t st);
// Tile size (log2 words) used for log2 n on this context.
int tile_bits_for(const vkhel_ctx *ctx, uint32_t logn);
bool ntt_fused_supported(uint32_t logn);
// Radix-2 path for one task (sizes the fused kernels do not cover).
int launch_task_ntt(const NttTask &k, bool inverse, hipStream_t s);

// Looks up this device's static scratch array (noalloc.hip) with the
// device current; called once per context creation.
int noalloc_init(int device);

// Pointwise launchers (elementwise.hip).
int launch_elemmul(const uint64_t *a, const uint64_t *b, uint64_t *out,
		uint64_t len, uint64_t q, hipStream_t s);
int launch_addmod(const uint64_t *a, const uint64_t *b, uint64_t *out,
		uint64_t len, uint64_t q, hipStream_t s);
// vkhel_vector_elemfma's kernel (elemfma.comp:36-55), stream-ordered.
int launch_elemfma(const uint64_t *a, const uint64_t *b, uint64_t *out,
		uint64_t len, uint64_t m, uint64_t q, hipStream_t s);

}  // namespace vk
