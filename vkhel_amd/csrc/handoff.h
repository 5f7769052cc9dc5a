// Work order of the single-launch two-pass transform (k_ntt_handoff in
// ntt_fused.h).  Plain constexpr integer functions, so the device kernel
// and the host-side model of the protocol (tests/c/handoff_model.cpp) run
// the same code.
//
// A two-pass transform of a batch runs as one persistent grid.  Each XCD
// (its own L2) serves its own queue: the polynomials p with p % 8 == xcd,
// local index j = p / 8.  Every polynomial is 2 * T tickets: T tiles of its
// first pass (the producer pass: columns for a forward transform, rows for
// an inverse) and T tiles of its second.  Queue order, in blocks of T
// tickets, runs the producers A polynomials ahead of the consumers:
//
//   P(0), ..., P(A-1), P(A), C(0), P(A+1), C(1), ..., P(J-1), C(J-1-A),
//   C(J-A), ..., C(J-1)
//
// (P = producer block, C = consumer block; A = 1: P(0), P(1), C(0), P(2),
// C(1), ...).  With A about the number of polynomials whose tickets the
// XCD's resident workgroups hold at once, consumers seldom wait.  A workgroup claims the next
// ticket of its XCD's queue with an atomic increment; a consumer ticket of
// polynomial j waits until all T producer tiles of j are done.  Those have
// smaller ticket numbers, so they were claimed earlier by workgroups that
// are running or finished, and producer tiles never wait: every wait ends.
// Producer and consumer of a polynomial run on one XCD, so the
// intermediate is handed over through that XCD's L2.
#pragma once

#include <cstdint>

namespace vk {

constexpr uint32_t kXcds = 8;

// Polynomials in XCD x's queue for a batch of npoly.
constexpr uint32_t handoff_queue_polys(uint32_t npoly, uint32_t xcd) {
	return xcd < npoly ? (npoly - xcd + kXcds - 1) / kXcds : 0;
}

// Tickets in XCD x's queue: 2 * T per polynomial.
constexpr uint32_t handoff_queue_tickets(uint32_t npoly, uint32_t xcd,
		uint32_t tiles) {
	return 2 * tiles * handoff_queue_polys(npoly, xcd);
}

// State words of a launch (zeroed before it): per XCD x, at
// x * kHandoffStride, its ticket counter and its count of finished
// tickets; the error flag at kHandoffErr; the producer counts of XCD x's
// polynomials from kHandoffCounts + x * handoff_count_stride(npoly).  Each
// XCD's words sit in 256-byte lines of their own.
constexpr uint32_t kHandoffStride = 64;
constexpr uint32_t kHandoffErr = kXcds * kHandoffStride;
constexpr uint32_t kHandoffCounts = kHandoffErr + kHandoffStride;

constexpr uint32_t handoff_done_word(uint32_t xcd) {
	return xcd * kHandoffStride + 1;
}

constexpr uint32_t handoff_err_word() {
	return kHandoffErr;
}

constexpr uint32_t handoff_count_stride(uint32_t npoly) {
	return (handoff_queue_polys(npoly, 0) + kHandoffStride - 1) /
		kHandoffStride * kHandoffStride;
}

constexpr uint32_t handoff_state_words(uint32_t npoly) {
	return kHandoffCounts + kXcds * handoff_count_stride(npoly);
}

constexpr uint32_t handoff_tiles(uint32_t logn, int tile_bits) {
	return 1u << (logn - (uint32_t) tile_bits);
}

struct HandoffWork {
	uint32_t poly;      // global polynomial index
	uint32_t tile;      // tile within the pass
	bool producer;      // first pass of the polynomial's direction
};

// Ticket k (< handoff_queue_tickets) of XCD x's queue, look-ahead A >= 1.
constexpr HandoffWork handoff_decode(uint32_t k, uint32_t xcd, uint32_t npoly,
		uint32_t tiles, uint32_t ahead) {
	const uint32_t J = handoff_queue_polys(npoly, xcd);
	const uint32_t A = ahead < J ? ahead : J;  // producer blocks up front
	const uint32_t M = J - A;                  // P(j + A), C(j) pairs
	const uint32_t b = k / tiles, tile = k % tiles;
	uint32_t j = 0;
	bool producer = true;
	if (b < A) {
		j = b;
	} else if (b - A < 2 * M) {
		j = (b - A) / 2;
		if ((b - A) & 1)
			producer = false;
		else
			j += A;
	} else {
		j = M + (b - A - 2 * M);
		producer = false;
	}
	return HandoffWork{j * kXcds + xcd, tile, producer};
}

}  // namespace vk
