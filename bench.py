"""Benchmark: batched negacyclic NTT throughput on one MI355X.

Workload (the headline configuration): 8 RNS primes q_i < 2^50 with
q_i = 1 mod 2n, 64 polynomials per prime (512 polynomials), n = 2^16 uint64
coefficients uniform in [0, q_i).  One step = forward NTT of all 512
polynomials then inverse NTT of all 512 (in place, so every step sees the
same data), issued as one replayed hipGraph on the library's stream.

Reported `value` is transform-coefficients per second (2 * 512 * n per
step) in Gcoeff/s with inputs resident in HBM.  Also reported:
  * roofline of the dominant kernel (the fused forward NTT), timed with HIP
    events on the library stream;
  * the plain per-call path (vkhel_vector_forward_transform /
    _inverse_transform, one synchronous call per polynomial) for the same
    work;
  * cpu_baseline: the oracle's scalar C NTT (same arithmetic) on the host
    cores, over a bounded sample.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--logn 16]
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

HBM_PEAK_GBS = 8000.0

M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64_stream(seed, count):
    """splitmix64 outputs x_1..x_count of the stream seeded with `seed`."""
    with np.errstate(over="ignore"):
        st = np.uint64(seed) + np.arange(1, count + 1, dtype=np.uint64) * \
            np.uint64(0x9E3779B97F4A7C15)
        z = st
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def uniform_mod(q, count, seed):
    """`count` words uniform in [0, q): the top bitlen(q) bits of successive
    splitmix64 outputs, rejecting values >= q (same stream as the oracle's
    or_fill_uniform)."""
    bits = int(q).bit_length()
    parts, have, pos = [], 0, 0
    while have < count:
        need = int((count - have) * 1.3) + 64
        with np.errstate(over="ignore"):
            start = np.uint64(seed) + np.uint64(pos) * np.uint64(0x9E3779B97F4A7C15)
        x = splitmix64_stream(int(start), need) >> np.uint64(64 - bits)
        pos += need
        keep = x[x < np.uint64(q)]
        parts.append(keep)
        have += keep.size
    return np.concatenate(parts)[:count]


def is_prime(n):
    if n < 2:
        return False
    for p in (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37):
        if n % p == 0:
            return n == p
    d, s = n - 1, 0
    while d % 2 == 0:
        d //= 2
        s += 1
    for a in (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37):
        x = pow(a, d, n)
        if x in (1, n - 1):
            continue
        for _ in range(s - 1):
            x = x * x % n
            if x == n - 1:
                break
        else:
            return False
    return True


def primes_ntt(bound, m, count):
    out, c = [], ((bound - 2) // m) * m + 1
    while len(out) < count and c > m:
        if is_prime(c):
            out.append(c)
        c -= m
    return out


def root_2n(q, n):
    """psi = g^((q-1)/2n), g the smallest generator of (Z/qZ)^*."""
    m, fs, p = q - 1, [], 2
    while p * p <= m:
        if m % p == 0:
            fs.append(p)
            while m % p == 0:
                m //= p
        p += 1 if p == 2 else 2
    if m > 1:
        fs.append(m)
    g = 2
    while any(pow(g, (q - 1) // f, q) == 1 for f in fs):
        g += 1
    return pow(g, (q - 1) // (2 * n), q)


class HipEvents:
    """hipEvent timing on a given stream (the library's, not torch's)."""

    def __init__(self):
        self.h = ctypes.CDLL("libamdhip64.so")
        self.h.hipEventCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
        self.h.hipEventRecord.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        self.h.hipEventSynchronize.argtypes = [ctypes.c_void_p]
        self.h.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float),
                                               ctypes.c_void_p, ctypes.c_void_p]
        self.h.hipEventDestroy.argtypes = [ctypes.c_void_p]

    def new(self):
        e = ctypes.c_void_p()
        assert self.h.hipEventCreate(ctypes.byref(e)) == 0
        return e

    def record(self, e, stream):
        assert self.h.hipEventRecord(e, stream) == 0

    def ms(self, a, b):
        assert self.h.hipEventSynchronize(b) == 0
        out = ctypes.c_float()
        assert self.h.hipEventElapsedTime(ctypes.byref(out), a, b) == 0
        return float(out.value)


def cpu_baseline(logn, qs, ws, target_cpu_s=10.0):
    """Oracle scalar NTT (fwd+inv) on host cores over a bounded sample."""
    from oracle import oracle as O
    n = 1 << logn
    threads = max(1, min(os.cpu_count() or 1, 16))
    probe = np.stack([uniform_mod(qs[i % len(qs)], n, 7 + i) for i in range(4)])
    t1 = O.bench_ntt_batch(probe, n, qs, ws, 1)
    per_poly = max(t1 / 4, 1e-6)
    npolys = int(min(max(threads, target_cpu_s / per_poly), 4096))
    npolys = max(len(qs), npolys - npolys % len(qs))
    polys = np.stack([uniform_mod(qs[i % len(qs)], n, 100 + i) for i in range(npolys)])
    secs = O.bench_ntt_batch(polys, n, qs, ws, threads)
    return {"value": round(2 * npolys * n / secs / 1e9, 4), "unit": "Gcoeff/s",
            "cores": threads, "kind": "port",
            "sample": f"{npolys} polynomials of n=2^{logn} ({len(qs)} primes <2^50), "
                      f"forward+inverse, scalar C (oracle/vkhel_oracle.c) "
                      f"over {threads} threads, {secs:.2f} s wall"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--logn", type=int, default=16)
    ap.add_argument("--primes", type=int, default=8)
    ap.add_argument("--per-prime", type=int, default=64)
    ap.add_argument("--plain-polys", type=int, default=64,
                    help="polynomials timed on the plain per-call path")
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on exactly one GPU; "
                          "--gpus must be 1 (multi-GPU sharding is design-only)"}))
        sys.exit(2)

    import vkhel_amd as V

    n = 1 << args.logn
    qs = primes_ntt(1 << 50, 2 * n, args.primes)
    ws = [root_2n(q, n) for q in qs]
    ctx = V.Context()
    tabs = [V.NttTables(n, q, w) for q, w in zip(qs, ws)]
    polys, data = [], []
    for i in range(args.primes * args.per_prime):
        x = uniform_mod(qs[i % args.primes], n, 1000 + i)
        data.append(x)
        polys.append(ctx.from_array(x))
    both, fwd = V.Batch(ctx), V.Batch(ctx)
    for i, v in enumerate(polys):
        both.add(V.JOB_NTT_FWD, v, None, v, [tabs[i % args.primes]])
    for i, v in enumerate(polys):
        both.add(V.JOB_NTT_INV, v, None, v, [tabs[i % args.primes]])
    for i, v in enumerate(polys):
        fwd.add(V.JOB_NTT_FWD, v, None, v, [tabs[i % args.primes]])

    for _ in range(args.warmup):
        both.run()
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        both.launch()
    ctx.synchronize()
    secs = time.perf_counter() - t0
    npoly = len(polys)
    coeffs_per_step = 2 * npoly * n
    gcoeff = coeffs_per_step * args.steps / secs / 1e9

    # Results of the timed steps: every step is NTT then INTT in place.
    for i in (0, npoly - 1):
        assert np.array_equal(polys[i].read(), data[i]), "round trip broken"

    # Dominant kernel: fused forward NTT over all 512 polynomials.
    ev = HipEvents()
    stream = ctx.stream
    reps = max(3, args.steps // 2)
    e0, e1 = ev.new(), ev.new()
    fwd.run()  # build + warm
    ev.record(e0, stream)
    for _ in range(reps):
        fwd.launch()
    ev.record(e1, stream)
    fwd_ms = ev.ms(e0, e1) / reps
    # restore data (fwd-only runs changed it); not timed
    inv = V.Batch(ctx)
    for i, v in enumerate(polys):
        inv.add(V.JOB_NTT_INV, v, None, v, [tabs[i % args.primes]])
    for _ in range(reps + 1):
        inv.run()
    alg_bytes = 16 * npoly * n + 16 * n * args.primes
    achieved = alg_bytes / (fwd_ms * 1e-3) / 1e9

    # Plain per-call path: one synchronous call per polynomial.
    k = min(args.plain_polys, npoly)
    ctx.synchronize()
    t0 = time.perf_counter()
    for i in range(k):
        polys[i].forward_transform(polys[i], tabs[i % args.primes])
        polys[i].inverse_transform(polys[i], tabs[i % args.primes])
    plain_secs = time.perf_counter() - t0
    plain = 2 * k * n / plain_secs / 1e9

    cpu = None if args.no_cpu else cpu_baseline(args.logn, qs, ws)

    out = {
        "metric": "batched NTT Gcoeff/s & %HBM peak at N=2^16; served p50/p99 vs plain per-call",
        "value": round(gcoeff, 3),
        "unit": "Gcoeff/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(secs / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic (splitmix64, uniform in [0,q))",
        "config": {"workload": f"batched NTT fwd+inv, n=2^{args.logn}, "
                               f"{args.primes} primes <2^50 x {args.per_prime} polys",
                   "n": n, "polys": npoly, "primes": qs},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "k_ntt_fused<fwd>", "kernel_ms": round(fwd_ms, 4),
                     "alg_bytes_per_launch": alg_bytes},
        "plain_per_call": {"value": round(plain, 4), "unit": "Gcoeff/s",
                           "polys": k, "speedup": round(gcoeff / plain, 2)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))
    for v in polys:
        v.close()


if __name__ == "__main__":
    main()
