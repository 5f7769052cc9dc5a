// Host-thread model of the single-launch two-pass protocol
// (vkhel_amd/csrc/handoff.h, kernel k_ntt_handoff): per-XCD ticket queues;
// each XCD runs at most `resident` workgroups at a time, and launches
// `launches` of them in all, each claiming one ticket (the kernel's grid
// gives every XCD as many workgroups as the longest queue has tickets);
// producer tiles publish with a release increment of a per-polynomial
// count, consumer tiles wait for the count with a bounded spin.  It runs
// the same handoff_decode and state layout as the kernel.
//
//   handoff_model            every configuration must finish with no
//                            spin giving up, every handoff visible and
//                            every ticket done
//   handoff_model --planted  two planted mistakes must be caught: a
//                            consumer block moved before its producers
//                            (the bounded spin gives up), and too few
//                            workgroups for a queue (tickets left undone;
//                            nothing waits on them)
//
// Prints one line per configuration and "model ok" / "planted caught".
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <random>
#include <thread>
#include <vector>

#include "handoff.h"

using namespace vk;

struct Run {
	uint32_t npoly, tiles, resident;
	bool planted;
	uint32_t launches;  // workgroups per XCD; 0: the kernel's grid
	uint32_t ahead;     // schedule look-ahead in polynomials
};

static uint32_t launches_of(const Run &r) {
	return r.launches ? r.launches : handoff_queue_tickets(r.npoly, 0, r.tiles);
}

// The planted mistake: consumer block of local polynomial 1 swapped with
// the producer block of local polynomial 1 (P(0), C(1)... ).
static HandoffWork decode(const Run &r, uint32_t k, uint32_t x) {
	HandoffWork w = handoff_decode(k, x, r.npoly, r.tiles, r.ahead);
	if (r.planted && handoff_queue_polys(r.npoly, x) >= 3) {
		const uint32_t j = w.poly / kXcds;
		if (j == 1)
			w.producer = !w.producer;
	}
	return w;
}

// The state layout: every XCD's ticket and done words and its counts lie
// inside the state, apart from each other and from the error word.
static bool check_layout(uint32_t npoly) {
	const uint32_t words = handoff_state_words(npoly);
	std::vector<int> owner(words, -1);
	auto claim = [&](uint32_t w, int who) {
		if (w >= words || owner[w] != -1)
			return false;
		owner[w] = who;
		return true;
	};
	for (uint32_t x = 0; x < kXcds; x++) {
		if (!claim(x * kHandoffStride, (int) x) || !claim(handoff_done_word(x), (int) x))
			return false;
		for (uint32_t j = 0; j < handoff_queue_polys(npoly, x); j++)
			if (!claim(kHandoffCounts + x * handoff_count_stride(npoly) + j, (int) x))
				return false;
	}
	if (!claim(handoff_err_word(), 99))
		return false;
	// words of different XCDs never share a 256-byte line
	for (uint32_t w = 0; w < words; w++)
		for (uint32_t v = w / 64 * 64; v < w / 64 * 64 + 64 && v < words; v++)
			if (owner[w] >= 0 && owner[v] >= 0 && owner[w] != owner[v])
				return false;
	return true;
}

// Static checks of the schedule: every ticket maps to a distinct
// (poly, pass, tile), every polynomial of the queue appears, and each
// polynomial's producer tickets all come before its consumer tickets.
static bool check_schedule(const Run &r) {
	for (uint32_t x = 0; x < kXcds; x++) {
		const uint32_t Q = handoff_queue_tickets(r.npoly, x, r.tiles);
		const uint32_t J = handoff_queue_polys(r.npoly, x);
		std::vector<int> seen(2 * r.tiles * J, 0);
		std::vector<int64_t> last_p(J, -1), first_c(J, 1 << 30);
		for (uint32_t k = 0; k < Q; k++) {
			const HandoffWork w = handoff_decode(k, x, r.npoly, r.tiles, r.ahead);
			if (w.poly % kXcds != x || w.poly >= r.npoly || w.tile >= r.tiles)
				return false;
			const uint32_t j = w.poly / kXcds;
			seen[(2 * j + (w.producer ? 0 : 1)) * r.tiles + w.tile]++;
			if (w.producer)
				last_p[j] = std::max<int64_t>(last_p[j], k);
			else
				first_c[j] = std::min<int64_t>(first_c[j], k);
		}
		for (int s : seen)
			if (s != 1)
				return false;
		for (uint32_t j = 0; j < J; j++)
			if (!(last_p[j] < first_c[j]))
				return false;
	}
	return true;
}

struct Shared {
	std::atomic<uint32_t> launched[kXcds];
	std::atomic<uint32_t> ticket[kXcds];
	std::atomic<uint32_t> done[kXcds];
	std::atomic<int> gave_up{0};
	std::atomic<int> stale{0};
	std::atomic<uint64_t> max_wait_us{0};
	std::vector<std::atomic<uint32_t>> count;
	std::vector<uint8_t> inter;   // producer tile markers (plain stores)
	std::vector<uint8_t> out;     // consumer tile markers
	explicit Shared(uint32_t npoly, uint32_t tiles)
		: count(npoly), inter(npoly * tiles, 0), out(npoly * tiles, 0) {
		for (uint32_t x = 0; x < kXcds; x++) {
			launched[x] = 0;
			ticket[x] = 0;
			done[x] = 0;
		}
		for (auto &c : count)
			c = 0;
	}
};

// A wait gives up after this long.  The kernel bounds its polls by count;
// host threads share few CPUs with many others, so the model bounds them
// by time: long enough that a correct schedule never reaches it even on a
// loaded machine, short enough for the planted deadlock to end the run.
static std::chrono::milliseconds g_wait_limit(30000);

// One resident slot of XCD x: runs workgroups one after another while the
// XCD has workgroups left to launch; each claims one ticket.
static void worker(const Run &r, Shared &s, uint32_t x, uint32_t seed) {
	std::mt19937 rng(seed);
	const uint32_t Q = handoff_queue_tickets(r.npoly, x, r.tiles);
	while (s.launched[x].fetch_add(1) < launches_of(r)) {
		const uint32_t k = s.ticket[x].fetch_add(1, std::memory_order_relaxed);
		if (k >= Q)
			continue;  // a workgroup without a ticket exits
		const HandoffWork w = decode(r, k, x);
		// uneven work: a few microseconds, sometimes a yield
		for (uint32_t i = rng() % 2000; i; i--)
			std::atomic_signal_fence(std::memory_order_seq_cst);
		if (rng() % 4 == 0)
			std::this_thread::yield();
		if (w.producer) {
			s.inter[w.poly * r.tiles + w.tile] = 1;
			s.count[w.poly].fetch_add(1, std::memory_order_release);
		} else {
			const auto t0 = std::chrono::steady_clock::now();
			uint64_t spins = 0;
			// (as in the kernel, once any wait has given up none waits again)
			while (!s.gave_up.load() &&
					s.count[w.poly].load(std::memory_order_acquire) < r.tiles) {
				if ((++spins & 255) == 0) {
					// let the producers' threads run; now and then check the clock
					std::this_thread::yield();
					if ((spins & 16383) == 0 &&
							std::chrono::steady_clock::now() - t0 > g_wait_limit) {
						s.gave_up++;
						break;
					}
				}
			}
			const uint64_t us = (uint64_t) std::chrono::duration_cast<
				std::chrono::microseconds>(std::chrono::steady_clock::now() - t0).count();
			uint64_t m = s.max_wait_us.load();
			while (us > m && !s.max_wait_us.compare_exchange_weak(m, us)) {
			}
			for (uint32_t t = 0; t < r.tiles; t++)
				if (!s.inter[w.poly * r.tiles + t])
					s.stale++;
			s.out[w.poly * r.tiles + w.tile] = 1;
		}
		s.done[x].fetch_add(1, std::memory_order_relaxed);
	}
}

// Returns 0 if the run finished cleanly, 1 if a spin gave up or a
// handoff was stale or a ticket was left undone.
static int simulate(const Run &r, uint32_t seed) {
	Shared s(r.npoly, r.tiles);
	std::vector<std::thread> th;
	for (uint32_t x = 0; x < kXcds; x++)
		for (uint32_t i = 0; i < r.resident; i++)
			th.emplace_back(worker, std::cref(r), std::ref(s), x, seed * 131 + x * 17 + i);
	for (auto &t : th)
		t.join();
	bool all = true;
	for (uint32_t x = 0; x < kXcds; x++)
		all &= s.done[x] == handoff_queue_tickets(r.npoly, x, r.tiles);
	for (uint32_t p = 0; p < r.npoly; p++)
		for (uint32_t t = 0; t < r.tiles; t++)
			all &= s.out[p * r.tiles + t] == 1;
	printf("npoly=%u tiles=%u resident=%u%s: gave_up=%d stale=%d complete=%d "
			"max_wait_us=%llu\n", r.npoly, r.tiles, r.resident,
			r.planted ? " (planted)" : "", s.gave_up.load(), s.stale.load(),
			(int) all, (unsigned long long) s.max_wait_us.load());
	return (s.gave_up || s.stale || !all) ? 1 : 0;
}

int main(int argc, char **argv) {
	const bool planted = argc > 1 && !strcmp(argv[1], "--planted");
	if (planted) {
		g_wait_limit = std::chrono::milliseconds(1000);
		// one resident workgroup per XCD: the moved consumer block waits for
		// producers that only a later workgroup would run
		const Run moved{64, 16, 1, true, 0, 1};
		// fewer workgroups than XCD 0's queue has tickets
		const Run short_grid{64, 16, 4, false, 100, 3};
		if (simulate(moved, 1) == 0 || simulate(short_grid, 1) == 0) {
			printf("planted mistake NOT caught\n");
			return 1;
		}
		printf("planted caught\n");
		return 0;
	}
	const Run runs[] = {
		{1, 16, 1, false, 0, 1}, {7, 16, 2, false, 0, 4}, {8, 32, 1, false, 0, 2},
		{64, 16, 1, false, 0, 1}, {64, 16, 3, false, 0, 5}, {512, 16, 4, false, 0, 4},
		{513, 8, 2, false, 0, 9}, {100, 2, 5, false, 0, 3}, {40, 64, 8, false, 0, 100},
		{9, 16, 2, false, 5000, 1}, {512, 16, 8, false, 0, 8},
	};
	int bad = 0;
	for (uint32_t npoly = 1; npoly <= 1100; npoly += 37)
		if (!check_layout(npoly)) {
			printf("layout check failed: npoly=%u\n", npoly);
			bad++;
		}
	for (const Run &r : runs) {
		if (!check_schedule(r)) {
			printf("schedule check failed: npoly=%u tiles=%u\n", r.npoly, r.tiles);
			bad++;
			continue;
		}
		for (uint32_t seed = 1; seed <= 3; seed++)
			bad += simulate(r, seed);
	}
	if (bad)
		return 1;
	printf("model ok\n");
	return 0;
}
