"""Benchmark: batched negacyclic NTT throughput on one MI355X.

Headline workload (`value`): 8 RNS primes q_i < 2^50 with q_i = 1 mod 2n,
64 polynomials per prime (512 polynomials), n = 2^16 uint64 coefficients
uniform in [0, q_i).  One step = forward NTT of all 512 polynomials then
inverse NTT of all 512 (in place, so every step sees the same data), issued
as one batch launch on the library's stream (its few kernel launches are
enqueued directly; batches of many launches replay a captured hipGraph).
`value` is transform-coefficients per second (2 * 512 * n per step) with
inputs resident in HBM.

Also reported, each with its own timing:
  * roofline of the dominant kernel, the forward NTT of the batch at
    n = 2^16 (two pass dispatches of k_ntt_pass), timed with HIP events on
    the library stream, with the HBM traffic and VALU occupancy of the
    committed rocprofv3 summary of the same workload;
  * plain_per_call: the same transforms through the synchronous reference
    API (vkhel_vector_forward_transform / _inverse_transform), one call per
    polynomial;
  * batched_n14: the same batch at n = 2^14;
  * polymul: 256 pairs x 16 primes < 2^55 at n = 2^15 (NTT a, NTT b,
    pointwise, INTT) as one batch, vs per call;
  * intt_center: INTT + centered lift into Z_p of the headline batch,
    fused in the inverse kernel vs as two batches;
  * rescale_ntt: RNS rescaling of NTT-form polynomials, one job vs the
    three-batch composition;
  * pointwise: the reference's pointwise ops as single calls on 1 GiB;
  * uneven_batch: 1000 seeded jobs, n = 2^10..2^17, 1..24 primes < 2^60,
    ops NTT / INTT / polymul / add / mul, as one batch and per call;
  * serving: the same 1000 jobs through the serving engine at Poisson
    arrival rates from 10% to 150% of the batched job rate (three runs per
    point), p50/p99 latency beside a plain per-call server at the same
    rates; every run checked against the batch's words;
  * map_unmap: vkhel_vector_map + unmap latency of a 2^16-word vector,
    pooled staging vs a staging buffer allocated per map;
  * pcie_inclusive: the headline batch with inputs and outputs in host
    memory (per-call reference API, and a pipelined batch with transfers
    on copy streams) -- reported beside `value`, never as it;
  * cpu_baseline: the oracle's scalar C NTT on the host cores (bounded
    sample), the polymul workload's CPU figure, and one n = 4096 polymul
    on one CPU thread against the per-call GPU path (the only leg that
    runs the oracle).

    python bench.py [--gpus 1] [--steps K] [--warmup W]
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
    """Up to `count` primes q < bound with q = 1 mod m, descending."""
    out, c = [], ((bound - 2) // m) * m + 1
    while len(out) < count and c > m:
        if is_prime(c):
            out.append(c)
        c -= m
    return out


_GEN = {}


def smallest_generator(q):
    if q not in _GEN:
        try:
            import sympy
            _GEN[q] = int(sympy.primitive_root(q))
        except ImportError:
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
            _GEN[q] = g
    return _GEN[q]


def root_2n(q, n):
    """psi = g^((q-1)/2n), g the smallest generator of (Z/qZ)^*."""
    return pow(smallest_generator(q), (q - 1) // (2 * n), q)


class HipEvents:
    """hipEvent timing on a given stream (the library's, not torch's)."""

    def __init__(self):
        self.h = ctypes.CDLL("libamdhip64.so")
        self.h.hipEventCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
        self.h.hipEventRecord.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        self.h.hipEventSynchronize.argtypes = [ctypes.c_void_p]
        self.h.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float),
                                               ctypes.c_void_p, ctypes.c_void_p]

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


def timed_launches(ctx, batch, reps, ev):
    """Average ms of `reps` replays of `batch`, HIP events on ctx.stream."""
    batch.run()
    e0, e1 = ev.new(), ev.new()
    ev.record(e0, ctx.stream)
    for _ in range(reps):
        batch.launch()
    ev.record(e1, ctx.stream)
    return ev.ms(e0, e1) / reps


def pct(xs, p):
    return float(np.percentile(np.asarray(xs, dtype=np.float64), p))


# ---------------------------------------------------------------- legs

def batched_ntt(V, ctx, logn, primes, per_prime, steps, warmup, ev):
    """Batched forward + inverse of primes*per_prime polynomials."""
    n = 1 << logn
    qs = primes_ntt(1 << 50, 2 * n, primes)
    tabs = [V.NttTables(n, q, root_2n(q, n)) for q in qs]
    polys, data = [], []
    for i in range(primes * per_prime):
        v = ctx.vector(n, False)
        v.fill_uniform(qs[i % primes], 1000 + i)
        polys.append(v)
    for i in (0, len(polys) - 1):
        data.append((i, polys[i].read()))
    both, fwd, inv = V.Batch(ctx), V.Batch(ctx), V.Batch(ctx)
    for i, v in enumerate(polys):
        both.add(V.JOB_NTT_FWD, v, None, v, [tabs[i % primes]])
        fwd.add(V.JOB_NTT_FWD, v, None, v, [tabs[i % primes]])
        inv.add(V.JOB_NTT_INV, v, None, v, [tabs[i % primes]])
    for i, v in enumerate(polys):
        both.add(V.JOB_NTT_INV, v, None, v, [tabs[i % primes]])
    for _ in range(warmup):
        both.run()
    ctx.synchronize()
    e0, e1 = ev.new(), ev.new()
    t0 = time.perf_counter()
    ev.record(e0, ctx.stream)
    for _ in range(steps):
        both.launch()
    ev.record(e1, ctx.stream)
    ctx.synchronize()
    secs = time.perf_counter() - t0
    gpu_ms = ev.ms(e0, e1)
    for i, x in data:  # every step is NTT then INTT in place
        assert np.array_equal(polys[i].read(), x), "round trip broken"
    # dominant kernel: the fused forward NTT over all polynomials
    reps = max(3, steps // 2)
    fwd_ms = timed_launches(ctx, fwd, reps, ev)
    inv_ms = timed_launches(ctx, inv, reps, ev)  # restores the data too
    npoly = len(polys)
    res = {"qs": qs, "tabs": tabs, "polys": polys, "n": n, "npoly": npoly,
           "secs": secs, "gpu_ms": gpu_ms, "steps": steps, "fwd_ms": fwd_ms,
           "inv_ms": inv_ms,
           "gcoeff": 2 * npoly * n * steps / secs / 1e9,
           "batches": (both, fwd, inv)}
    return res


def plain_per_call(V, ctx, r, k):
    """The same transforms, one synchronous reference-API call each."""
    tabs, polys, n = r["tabs"], r["polys"], r["n"]
    primes = len(tabs)
    k = min(k, len(polys))
    ctx.synchronize()
    t0 = time.perf_counter()
    for i in range(k):
        polys[i].forward_transform(polys[i], tabs[i % primes])
        polys[i].inverse_transform(polys[i], tabs[i % primes])
    secs = time.perf_counter() - t0
    return {"value": round(2 * k * n / secs / 1e9, 4), "unit": "Gcoeff/s",
            "polys": k, "us_per_call": round(secs / (2 * k) * 1e6, 2)}


def polymul_leg(V, ctx, ev, pairs_per_prime=256, primes=16, logn=15, reps=5):
    """Polynomial products a*b mod (X^n + 1, q_i), 256 pairs x 16 primes
    < 2^55 at n = 2^15."""
    n = 1 << logn
    qs = primes_ntt(1 << 55, 2 * n, primes)
    tabs = [V.NttTables(n, q, root_2n(q, n)) for q in qs]
    words = primes * n
    A, B, R = [], [], []
    for p in range(pairs_per_prime):
        a, b, c = ctx.vector(words, False), ctx.vector(words, False), ctx.vector(words)
        for i, q in enumerate(qs):
            a.fill_uniform(q, 7 * p + i, i * n, n)
            b.fill_uniform(q, 7 * p + 3 + i, i * n, n)
        A.append(a)
        B.append(b)
        R.append(c)
    bt = V.Batch(ctx)
    for a, b, c in zip(A, B, R):
        bt.add(V.JOB_POLYMUL, a, b, c, tabs)
    ms = timed_launches(ctx, bt, reps, ev)
    out_coeffs = pairs_per_prime * words
    # per call on 8 pairs (limb views through the reference API)
    sample = 8
    tmp = ctx.vector(n)
    ctx.synchronize()
    t0 = time.perf_counter()
    for p in range(sample):
        for i, (q, t) in enumerate(zip(qs, tabs)):
            va, vb, vc = A[p].view(i * n, n), B[p].view(i * n, n), R[p].view(i * n, n)
            va.forward_transform(vc, t)
            vb.forward_transform(tmp, t)
            vc.elemmul(tmp, vc, q)
            vc.inverse_transform(vc, t)
    plain_s = (time.perf_counter() - t0) / (sample * primes)
    res = {"workload": f"{pairs_per_prime} pairs x {primes} primes <2^55, n=2^{logn}",
           "ms": round(ms, 4),
           "output_gcoeff_s": round(out_coeffs / ms / 1e6, 3),
           "transform_gcoeff_s": round(3 * out_coeffs / ms / 1e6, 3),
           "alg_GBps": round(24 * out_coeffs / ms / 1e6, 1),
           "plain_per_call_output_gcoeff_s": round(n / plain_s / 1e9, 4)}
    res["speedup_vs_plain"] = round(res["output_gcoeff_s"] /
                                    res["plain_per_call_output_gcoeff_s"], 1)
    for v in A + B + R + [tmp]:
        v.close()
    bt.close()
    return res


def uneven_ops(V):
    """The five ops of the uneven mix, in draw order: NTT, INTT, polymul,
    add (the reference's vector add, elemfma with multiplier 1, as a batch
    job: JOB_ADD_REF), mul."""
    return [V.JOB_NTT_FWD, V.JOB_NTT_INV, V.JOB_POLYMUL, V.JOB_ADD_REF, V.JOB_MUL]


def make_uneven_jobs(V, ctx, njobs=1000, seed=20261202):
    """Uneven batch: n = 2^10..2^17, L = 1..24 primes < 2^60 (q = 1 mod
    2^18), op uniform over NTT, INTT, polymul, add (elemfma m = 1), mul."""
    ops = uneven_ops(V)
    rng = np.random.default_rng(seed)
    qs = primes_ntt(1 << 60, 1 << 18, 24)
    tables = {}
    jobs = []
    for k in range(njobs):
        logn = int(rng.integers(10, 18))
        L = int(rng.integers(1, 25))
        op = ops[int(rng.integers(0, 5))]
        n = 1 << logn
        for q in qs[:L]:
            if (n, q) not in tables:
                tables[(n, q)] = V.NttTables(n, q, root_2n(q, n))
        tabs = [tables[(n, q)] for q in qs[:L]]
        a = ctx.vector(L * n, False)
        b = ctx.vector(L * n, False) if op >= V.JOB_POLYMUL else None
        r = ctx.vector(L * n)
        for i, q in enumerate(qs[:L]):
            a.fill_uniform(q, (k << 8) + i, i * n, n)
            if b is not None:
                b.fill_uniform(q, (k << 8) + 128 + i, i * n, n)
        jobs.append((op, a, b, r, tabs))
    return jobs, tables


def job_coeffs(V, job):
    op, a, b, r, tabs = job
    words = len(tabs) * tabs[0].n
    return {V.JOB_NTT_FWD: words, V.JOB_NTT_INV: words,
            V.JOB_POLYMUL: 3 * words}.get(op, 0)


def per_call_service(V, ctx, jobs):
    """Service time (s) of each job through synchronous reference-API calls
    on limb views (NTT/INTT: 1 call per limb; polymul: 4; mul: elemmul;
    add: elemfma with multiplier 1, the function JOB_ADD_REF batches)."""
    times = []
    tmp = ctx.vector(1 << 17)
    for op, a, b, r, tabs in jobs:
        n = tabs[0].n
        views = []
        for i in range(len(tabs)):
            views.append((a.view(i * n, n), b.view(i * n, n) if b is not None else None,
                          r.view(i * n, n)))
        tv = tmp.view(0, n)
        ctx.synchronize()
        t0 = time.perf_counter()
        for (va, vb, vr), t in zip(views, tabs):
            if op == V.JOB_NTT_FWD:
                va.forward_transform(vr, t)
            elif op == V.JOB_NTT_INV:
                va.inverse_transform(vr, t)
            elif op == V.JOB_POLYMUL:
                va.forward_transform(vr, t)
                vb.forward_transform(tv, t)
                vr.elemmul(tv, vr, t.q)
                vr.inverse_transform(vr, t)
            elif op == V.JOB_MUL:
                va.elemmul(vb, vr, t.q)
            elif op == V.JOB_ADD_REF:
                va.elemfma(vb, vr, 1, t.q)
            else:
                raise ValueError(f"no per-call form for op {op}")
        times.append(time.perf_counter() - t0)
    tmp.close()
    return times


def poisson_arrivals(njobs, rate, seed):
    """Arrival times matching vkhel_server_replay's exponential gaps."""
    st = seed
    t, out = 0.0, []
    M = (1 << 64) - 1
    for _ in range(njobs):
        st = (st + 0x9E3779B97F4A7C15) & M
        z = st
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
        z ^= z >> 31
        u = ((z >> 11) + 1.0) / 9007199254740992.0
        t += -np.log(u) / rate
        out.append(t)
    return out


def job_sample(job):
    """The words of a job's last limb (the per-job sample the uneven and
    serving legs compare)."""
    op, a, b, r, tabs = job
    n = tabs[0].n
    return r.read((len(tabs) - 1) * n, n)


def clear_samples(jobs, seed):
    """Overwrites every job's sampled limb, so a run that skipped a job
    cannot pass the comparison."""
    for k, (op, a, b, r, tabs) in enumerate(jobs):
        n = tabs[0].n
        r.fill_uniform(3, seed + k, (len(tabs) - 1) * n, n)


def samples_equal(jobs, want):
    return all(np.array_equal(job_sample(j), w) for j, w in zip(jobs, want))


SERVE_FRACTIONS = (0.1, 0.25, 0.5, 0.75, 1.0, 1.25, 1.5)


def uneven_and_serving(V, ctx, ev, quick=False, repeats=3):
    """The 1000-job uneven batch (one batch vs the same jobs as per-call
    reference calls) and its serving at Poisson arrival rates from 10% to
    150% of the batched job rate, each point `repeats` times, beside a plain
    per-call server (one request at a time through the synchronous single
    calls) at the same rates.  Every run is checked against the batch's
    words on each job's last limb."""
    jobs, tables = make_uneven_jobs(V, ctx, 200 if quick else 1000)
    coeffs = sum(job_coeffs(V, j) for j in jobs)
    bt = V.Batch(ctx)
    for j in jobs:
        bt.add(*j)
    batch_ms = timed_launches(ctx, bt, 3, ev)
    bt.close()
    want = [job_sample(j) for j in jobs]
    clear_samples(jobs, 1 << 20)
    service = per_call_service(V, ctx, jobs)
    assert samples_equal(jobs, want), "per-call results differ from the batch's"
    plain_total = sum(service)
    uneven = {"workload": f"{len(jobs)} jobs, n=2^10..2^17, L=1..24 primes <2^60, "
                          "ops NTT/INTT/polymul/add(elemfma m=1)/mul (seeded)",
              "transform_coeffs": coeffs,
              "batch_ms": round(batch_ms, 3),
              "batched_gcoeff_s": round(coeffs / batch_ms / 1e6, 3),
              "plain_per_call_gcoeff_s": round(coeffs / plain_total / 1e9, 4),
              "plain_per_call_ms": round(plain_total * 1e3, 2),
              "speedup_vs_plain": round(plain_total * 1e3 / batch_ms, 1),
              "checked": "batch and per-call words equal on every job's last limb"}
    capacity = len(jobs) / (batch_ms * 1e-3)
    fracs = (0.5,) if quick else SERVE_FRACTIONS
    reps = 1 if quick else repeats

    def replay(srv, rate, seed):
        clear_samples(jobs, seed << 12)
        st, infos = srv.replay(jobs, rate_per_s=rate, seed=seed)
        assert st == 0, st
        assert samples_equal(jobs, want), f"served results differ at rate {rate}"
        lat = [(i.complete_ns - i.submit_ns) / 1e6 for i in infos]
        span = (max(i.complete_ns for i in infos) -
                min(i.submit_ns for i in infos)) / 1e9
        return (pct(lat, 50), pct(lat, 99), coeffs / span / 1e9,
                float(np.mean([i.batch_size for i in infos])))

    def summary(runs, k):
        xs = sorted(r[k] for r in runs)
        return round(xs[len(xs) // 2], 3), [round(xs[0], 3), round(xs[-1], 3)]

    points = []
    srv = V.Server(ctx, streams=4)
    plain = V.Server(ctx, per_call=True)
    for frac in fracs:
        rate = capacity * frac
        runs = [replay(srv, rate, 77 + r) for r in range(reps)]
        pruns = [replay(plain, rate, 177)]
        p = {"frac_of_capacity": frac, "rate_jobs_s": round(rate, 1), "runs": reps}
        for k, name in enumerate(("p50_ms", "p99_ms", "gcoeff_s")):
            p[name], p[name + "_range"] = summary(runs, k)
        p["mean_batch"] = summary(runs, 3)[0]
        p["plain_p50_ms"] = round(pruns[0][0], 3)
        p["plain_p99_ms"] = round(pruns[0][1], 3)
        p["plain_gcoeff_s"] = round(pruns[0][2], 3)
        points.append(p)
    st = srv.stats()
    srv.close()
    plain.close()
    serving = {"streams": 4, "hbm_budget_bytes": 200 * 10**9,
               "reserved_arena_bytes": int(st.reserved_bytes),
               "batched_capacity_jobs_s": round(capacity, 1),
               "plain_service_p50_ms": round(pct(service, 50) * 1e3, 3),
               "plain_service_p99_ms": round(pct(service, 99) * 1e3, 3),
               "points": points, "peak_inflight_bytes": int(st.peak_inflight_bytes),
               "plain": "a served FIFO: vkhel_server per_call mode, one request at a "
                        "time through the synchronous single calls",
               "checked": "every replay's words equal the batch's on each job's "
                          "last limb"}
    for op, a, b, r, tabs in jobs:
        a.close()
        r.close()
        if b is not None:
            b.close()
    return uneven, serving


def intt_center_leg(V, ctx, ev, logn=16, primes=8, per_prime=64, p=65537,
                    reps=5):
    """The RNS decryption-side step: INTT of every limb, then the centered
    lift into Z_p, for the headline batch shape; fused (INTT_CENTER, the
    lift applied by the inverse kernel as it stores) against the same work
    as two batches (INTT, then CENTER over the RNS polynomials)."""
    n = 1 << logn
    qs = primes_ntt(1 << 50, 2 * n, primes)
    tabs = [V.NttTables(n, q, root_2n(q, n)) for q in qs]
    words = primes * n
    A, R, T = [], [], []
    for k in range(per_prime):
        a = ctx.vector(words, False)
        for i, q in enumerate(qs):
            a.fill_uniform(q, 11 * k + i, i * n, n)
        A.append(a)
        R.append(ctx.vector(words))
        T.append(ctx.vector(words))
    fused, inv, center = V.Batch(ctx), V.Batch(ctx), V.Batch(ctx)
    for a, r, t in zip(A, R, T):
        fused.add_rns(V.JOB_INTT_CENTER, a, r, tabs, [p])
        inv.add(V.JOB_NTT_INV, a, None, t, tabs)
        center.add_rns(V.JOB_CENTER, t, t, tabs, [p])
    f_ms = timed_launches(ctx, fused, reps, ev)
    i_ms = timed_launches(ctx, inv, reps, ev)
    c_ms = timed_launches(ctx, center, reps, ev)
    same = all(np.array_equal(r.read(), t.read()) for r, t in zip(R[:4], T[:4]))
    coeffs = per_prime * words
    res = {"workload": f"INTT + centered lift into Z_{p}, {per_prime} polys x "
                       f"{primes} primes <2^50, n=2^{logn}",
           "fused_ms": round(f_ms, 4), "two_step_ms": round(i_ms + c_ms, 4),
           "fused_gcoeff_s": round(coeffs / f_ms / 1e6, 3),
           "speedup": round((i_ms + c_ms) / f_ms, 3),
           "fused_alg_GBps": round(16 * coeffs / f_ms / 1e6, 1),
           "results_equal": bool(same)}
    for v in A + R + T:
        v.close()
    for b in (fused, inv, center):
        b.close()
    return res


def rescale_ntt_leg(V, ctx, ev, logn=16, limbs=8, polys=64, reps=5):
    """RNS rescaling of NTT-form polynomials (drop the last of 8 limbs <
    2^50 at n = 2^16, 64 polynomials): one RESCALE_NTT job per polynomial
    (INTT of the dropped limb, its lift into the other limbs, their NTTs,
    one subtract-and-scale pass) against the textbook composition as
    three batches (INTT of every limb, coefficient-form RESCALE, NTT of the
    remaining limbs)."""
    n = 1 << logn
    qs = primes_ntt(1 << 50, 2 * n, limbs)
    tabs = [V.NttTables(n, q, root_2n(q, n)) for q in qs]
    words = limbs * n
    A, R, C, S, T = [], [], [], [], []
    for k in range(polys):
        a = ctx.vector(words, False)
        for i, q in enumerate(qs):
            a.fill_uniform(q, 13 * k + i, i * n, n)
        A.append(a)
        R.append(ctx.vector(words - n))
        C.append(ctx.vector(words))
        S.append(ctx.vector(words - n))
        T.append(ctx.vector(words - n))
    fused, inv, resc, fwd = V.Batch(ctx), V.Batch(ctx), V.Batch(ctx), V.Batch(ctx)
    for a, r, c, s_, t in zip(A, R, C, S, T):
        fused.add_rns(V.JOB_RESCALE_NTT, a, r, tabs)
        inv.add(V.JOB_NTT_INV, a, None, c, tabs)
        resc.add_rns(V.JOB_RESCALE, c, s_, tabs)
        fwd.add(V.JOB_NTT_FWD, s_, None, t, tabs[:-1])
    f_ms = timed_launches(ctx, fused, reps, ev)
    three_ms = sum(timed_launches(ctx, b, reps, ev) for b in (inv, resc, fwd))
    same = all(np.array_equal(r.read(), t.read()) for r, t in zip(R[:4], T[:4]))
    res = {"workload": f"rescale by the last of {limbs} primes <2^50 in NTT form, "
                       f"{polys} polys, n=2^{logn}",
           "fused_ms": round(f_ms, 4), "three_batches_ms": round(three_ms, 4),
           "speedup": round(three_ms / f_ms, 3),
           # read a (limbs x n words), write the result ((limbs - 1) x n)
           "fused_alg_GBps": round(8 * (2 * limbs - 1) * n * polys / f_ms / 1e6, 1),
           "results_equal": bool(same)}
    for v in A + R + C + S + T:
        v.close()
    for b in (fused, inv, resc, fwd):
        b.close()
    return res


def pointwise_leg(V, ctx, ev, words=1 << 27, reps=5):
    """The reference's pointwise ops as single calls on 1 GiB vectors (one
    launch each, synchronous): algorithmic bytes (24 per element for the
    two-input ops, 16 for the one-input ones) over the HIP-event time of
    the call, against the 8 TB/s peak."""
    q = primes_ntt(1 << 60, 1 << 17, 1)[0]
    a, b, r = (ctx.vector(words, False) for _ in range(3))
    a.fill_uniform(q, 1)
    b.fill_uniform(q, 2)
    ops = {"elemmul": (24, lambda: a.elemmul(b, r, q)),
           "elemfma_m1": (24, lambda: a.elemfma(b, r, 1, q)),
           "elemmod_p65537": (16, lambda: a.elemmod(r, 65537, q)),
           "elemgtadd": (16, lambda: a.elemgtadd(r, q // 2, 12345)),
           "elemgtsub": (16, lambda: a.elemgtsub(r, q // 2, 12345, 65537))}
    out = {"workload": f"single calls on {words} words (q < 2^60)"}
    for name, (bpe, fn) in ops.items():
        fn()
        e0, e1 = ev.new(), ev.new()
        ev.record(e0, ctx.stream)
        for _ in range(reps):
            fn()
        ev.record(e1, ctx.stream)
        ms = ev.ms(e0, e1) / reps
        gbs = bpe * words / ms / 1e6
        out[name] = {"ms": round(ms, 4), "alg_GBps": round(gbs, 1),
                     "frac_of_peak": round(gbs / HBM_PEAK_GBS, 3)}
    for v in (a, b, r):
        v.close()
    return out


def map_unmap_leg(V, ctx, logn=16, cycles=50):
    """vkhel_vector_map + vkhel_vector_unmap of one 2^16-word vector
    (src/vector.c:270-296): with the context's staging pool (default; the
    page-locked buffer is reused) and with it off (a buffer allocated and
    freed per cycle, as the reference stages per map).  Median microseconds
    per cycle over `cycles` cycles after a warm-up."""
    import ctypes
    L = V.load()
    n = 1 << logn
    v = ctx.vector(n, False)
    v.fill_uniform(1 << 60, 3)
    want = v.read()
    out = {"workload": f"map+unmap of 2^{logn} words", "cycles": cycles}
    for name, pool in (("pooled_us", True), ("per_call_alloc_us", False)):
        ctx.set_staging_pool(pool)
        times = []
        for k in range(cycles + 1):
            p = ctypes.c_void_p()
            t0 = time.perf_counter()
            L.vkhel_vector_map(v.handle, ctypes.byref(p), 8 * n)
            L.vkhel_vector_unmap(v.handle)
            dt = time.perf_counter() - t0
            if k:
                times.append(dt)
        out[name] = round(pct(times, 50) * 1e6, 2)
    ctx.set_staging_pool(True)
    assert np.array_equal(v.map_read(), want)
    v.close()
    out["speedup"] = round(out["per_call_alloc_us"] / out["pooled_us"], 2)
    return out


def pcie_inclusive(V, ctx, logn=16, primes=8, per_prime=64, chunks=16,
                   per_call_sample=32):
    """Host-resident inputs and outputs (the headline batch, n = 2^16,
    512 polynomials, forward then inverse): transfer time included.
    (a) reference API, one polynomial at a time: copy_from_host, forward,
        inverse, read back (a sample of polynomials);
    (b) pipelined: page-locked buffers, the batch in one device vector with
        a view per polynomial, in `chunks` parts, each part's upload and
        download one copy on its own copy stream, overlapping the
        transforms of the neighbouring parts on the library stream.  (One
        copy per polynomial instead, 512 KiB each, reached 32-34 GB/s
        against 55-57.5.)"""
    from vkhel_amd import hipstream as H
    n = 1 << logn
    npoly = primes * per_prime
    qs = primes_ntt(1 << 50, 2 * n, primes)
    tabs = [V.NttTables(n, q, root_2n(q, n)) for q in qs]
    up, down = V.HostBuffer(npoly * n), V.HostBuffer(npoly * n)
    for i in range(npoly):
        up.array[i * n:(i + 1) * n] = uniform_mod(qs[i % primes], n, 5000 + i)
    slab = ctx.vector(npoly * n, False)
    vecs = [slab.view(i * n, n) for i in range(npoly)]
    # (a) reference API per polynomial, pageable host memory
    host = [np.array(up.array[i * n:(i + 1) * n]) for i in range(per_call_sample)]
    ctx.synchronize()
    t0 = time.perf_counter()
    for i in range(per_call_sample):
        v = vecs[i]
        v.copy_from_host(host[i])
        v.forward_transform(v, tabs[i % primes])
        v.inverse_transform(v, tabs[i % primes])
        host[i] = v.read()
    per_call_s = time.perf_counter() - t0
    assert all(np.array_equal(host[i], up.array[i * n:(i + 1) * n])
               for i in range(per_call_sample)), "round trip broken"
    # (b) pipelined batch
    per = npoly // chunks
    batches = []
    for c in range(chunks):
        b = V.Batch(ctx)
        for i in range(c * per, (c + 1) * per):
            b.add(V.JOB_NTT_FWD, vecs[i], None, vecs[i], [tabs[i % primes]])
        for i in range(c * per, (c + 1) * per):
            b.add(V.JOB_NTT_INV, vecs[i], None, vecs[i], [tabs[i % primes]])
        b.run()
        batches.append(b)
    s_up, s_down = H.Stream(), H.Stream()
    ev_up = [H.Event() for _ in range(chunks)]
    ev_done = [H.Event() for _ in range(chunks)]
    start, end = H.Event(), H.Event()

    def run_pipeline():
        start.record(s_up)
        for c in range(chunks):
            part = slice(c * per * n, (c + 1) * per * n)
            slab.write_async(up.array[part], c * per * n, s_up)
            ev_up[c].record(s_up)
        for c in range(chunks):
            H.wait(ctx.stream, ev_up[c])
            batches[c].launch()
            ev_done[c].record(ctx.stream)
        for c in range(chunks):
            s_down.wait(ev_done[c])
            part = slice(c * per * n, (c + 1) * per * n)
            slab.read_async(down.array[part], c * per * n, s_down)
        end.record(s_down)
        s_down.synchronize()
        return start.elapsed_ms(end)

    run_pipeline()  # warm-up
    ms = min(run_pipeline() for _ in range(3))
    ok = bool(np.array_equal(down.array, up.array))
    for e in ev_up + ev_done + [start, end]:
        e.close()
    s_up.close()
    s_down.close()
    for b in batches:
        b.close()
    for v in vecs:
        v.close()
    slab.close()
    up.close()
    down.close()
    for t in tabs:
        t.close()
    bytes_moved = 2 * npoly * n * 8
    return {"workload": f"{npoly} x 2^{logn} forward+inverse, inputs and outputs "
                        f"in host memory",
            "pipelined_gcoeff_s": round(2 * npoly * n / (ms * 1e-3) / 1e9, 3),
            "pipelined_ms": round(ms, 3), "chunks": chunks,
            "pcie_GBps": round(bytes_moved / (ms * 1e-3) / 1e9, 1),
            "per_call_gcoeff_s": round(2 * per_call_sample * n / per_call_s / 1e9, 4),
            "per_call_sample": per_call_sample, "round_trip_exact": ok}


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(logn, qs, ws, V=None, ctx=None, target_cpu_s=16.0,
                 min_runs=5):
    """Oracle scalar NTT (fwd+inv) on host cores over a bounded sample: up
    to 4096 resident polynomials, at least `min_runs` timed passes after a
    warm-up pass (median reported) and about target_cpu_s thread-seconds
    of work.  Also the polymul workload's CPU figure and, given a context,
    the single n = 4096 polymul on one CPU thread against the per-call GPU
    path.  The only leg that runs the oracle."""
    from oracle import oracle as O
    n = 1 << logn
    threads = max(1, min(os.cpu_count() or 1, 16))
    probe = np.stack([uniform_mod(qs[i % len(qs)], n, 7 + i) for i in range(4)])
    O.bench_ntt_batch(probe, n, qs, ws, 1)  # warm-up
    t1 = O.bench_ntt_batch(probe, n, qs, ws, 1)
    per_poly = max(t1 / 4, 1e-6)
    npolys = int(min(max(threads, target_cpu_s / min_runs / per_poly), 4096))
    npolys = max(len(qs), npolys - npolys % len(qs))
    polys = np.stack([uniform_mod(qs[i % len(qs)], n, 100 + i) for i in range(npolys)])
    O.bench_ntt_batch(polys, n, qs, ws, threads)  # warm-up
    runs = []
    while len(runs) < 16 and (len(runs) < min_runs or sum(runs) * threads < target_cpu_s):
        runs.append(O.bench_ntt_batch(polys, n, qs, ws, threads))
    med = sorted(runs)[len(runs) // 2]
    return {"value": round(2 * npolys * n / med / 1e9, 4),
            "unit": "Gcoeff/s", "cores": threads, "kind": "port",
            "sample": f"median of {len(runs)} passes over {npolys} polynomials of "
                      f"n=2^{logn} ({len(qs)} primes <2^50), forward+inverse, scalar "
                      f"C (oracle/vkhel_oracle.c, -O3 -march=x86-64-v3) over "
                      f"{threads} threads, {sum(runs):.2f} s wall",
            "one_core_gcoeff_s": round(2 * 4 * n / t1 / 1e9, 5),
            "nproc": os.cpu_count(), "cpu_model": cpu_model(),
            "polymul": cpu_polymul_baseline(threads),
            "plain_polymul_n4096": (plain_polymul_n4096(V, ctx)
                                    if ctx is not None else None)}


def cpu_polymul_baseline(threads, logn=15, primes=16, target_cpu_s=8.0):
    """The polymul workload's CPU figure (n = 2^15, 16 primes < 2^55): the
    oracle's scalar forward, forward, pointwise product and inverse per
    pair on `threads` host threads over a bounded sample of pairs, output
    coefficients per second (median of 3 passes)."""
    from oracle import oracle as O
    n = 1 << logn
    qs = primes_ntt(1 << 55, 2 * n, primes)
    ws = [root_2n(q, n) for q in qs]
    probe_a = np.stack([uniform_mod(qs[i], n, 300 + i) for i in range(2)])
    probe_b = np.stack([uniform_mod(qs[i], n, 400 + i) for i in range(2)])
    t1, _ = O.bench_polymul_batch(probe_a, probe_b, n, qs[:2], ws[:2], 1)
    per_pair = max(t1 / 2, 1e-6)
    pairs = int(min(max(threads, target_cpu_s / 3 / per_pair), 1024))
    pairs = max(primes, pairs - pairs % primes)
    a = np.stack([uniform_mod(qs[i % primes], n, 500 + i) for i in range(pairs)])
    b = np.stack([uniform_mod(qs[i % primes], n, 9000 + i) for i in range(pairs)])
    runs = sorted(O.bench_polymul_batch(a, b, n, qs, ws, threads)[0] for _ in range(3))
    med = runs[1]
    return {"value": round(pairs * n / med / 1e9, 4), "unit": "Gcoeff/s (output)",
            "cores": threads, "kind": "port",
            "sample": f"median of 3 passes over {pairs} pairs of n=2^{logn} "
                      f"({primes} primes <2^55): forward, forward, product, inverse "
                      f"per pair, scalar C over {threads} threads",
            "one_core_value": round(2 * n / t1 / 1e9, 5)}


def plain_polymul_n4096(V, ctx, runs=7):
    """One n = 4096 polymul, q the largest 60-bit prime = 1 mod 8192: CPU
    oracle (one thread) vs the per-call GPU API (forward, forward, elemmul,
    inverse), each the median of `runs` timed runs after a warm-up run,
    bit-exact check of every GPU run."""
    from oracle import oracle as O
    n = 4096
    q = primes_ntt(1 << 60, 2 * n, 1)[0]
    w = root_2n(q, n)
    a, b = uniform_mod(q, n, 1), uniform_mod(q, n, 2)
    want = O.polymul_ntt(a, b, q, w)  # warm-up
    cpu = []
    for _ in range(runs):
        t0 = time.perf_counter()
        got = O.polymul_ntt(a, b, q, w)
        cpu.append(time.perf_counter() - t0)
        assert np.array_equal(got, want)
    t = V.NttTables(n, q, w)
    va, vb, vc = ctx.vector(n, False), ctx.vector(n, False), ctx.vector(n)
    ok, gpu = True, []
    for k in range(runs + 1):
        va.copy_from_host(a)
        vb.copy_from_host(b)
        ctx.synchronize()
        t0 = time.perf_counter()
        va.forward_transform(va, t)
        vb.forward_transform(vb, t)
        va.elemmul(vb, vc, q)
        vc.inverse_transform(vc, t)
        dt = time.perf_counter() - t0
        if k:  # the first run is the warm-up
            gpu.append(dt)
        ok = ok and bool(np.array_equal(vc.read(), want))
    for v in (va, vb, vc):
        v.close()
    t.close()
    return {"q": q, "psi": w, "runs": runs,
            "cpu_oracle_ms": round(sorted(cpu)[runs // 2] * 1e3, 3),
            "gpu_per_call_ms": round(sorted(gpu)[runs // 2] * 1e3, 3),
            "timing": f"median of {runs} runs after a warm-up run, each side",
            "bit_exact": ok}


def traffic_from_profile(logn, npoly, tile_bits):
    """HBM bytes per forward transform of the batch (both pass dispatches),
    from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary of this
    workload (scaled by the copy-kernel calibration recorded beside it), or
    None when no summary of the same tile size exists."""
    path = os.path.join(REPO, "profiles",
                        f"ntt_fwd_n{1 << logn}_x{npoly}_tile{tile_bits}.json")
    try:
        with open(path) as f:
            prof = json.load(f)
    except (OSError, ValueError):
        return None, None
    if prof.get("tile_bits") != tile_bits:
        return None, prof
    return prof.get("hbm_bytes_per_launch"), prof


def modulus_class(q):
    """The kernels' modulus class: 2 below 2^55, 1 below 2^60, else 0."""
    return 2 if q < (1 << 55) else 1 if q < (1 << 60) else 0


def _ntt_k1(logn, tb):
    # the library's default stage split (ntt_fused.h ntt_k1, split_ok, alt_k1)
    if logn <= tb:
        return 0
    k = logn - 9
    return k if k >= 1 and tb - k >= 4 and k < logn // 2 else logn // 2


def _split_ok(logn, tb, k1):
    return k1 >= 1 and tb - k1 >= 4 and logn - k1 <= tb


def _alt_k1(logn, tb):
    if logn <= tb:
        return -1
    k = _ntt_k1(logn, tb)
    return k + 1 if _split_ok(logn, tb, k + 1) else k - 1 if _split_ok(logn, tb, k - 1) else -1


def fwd_kernel_names(logn, tile_bits, pass_mode, mc, split=0, tile_bits2=0):
    """The dispatches of one forward transform at this size and tuning for
    moduli of class mc (the last template argument of a pass kernel is 1 for
    the alternative stage split where the size has one)."""
    if logn <= tile_bits or pass_mode == 1:
        return [f"k_ntt_fused<{logn}, false, {tile_bits}, {mc}>"]
    if pass_mode == 3:
        return [f"k_ntt_handoff<{logn}, false, {tile_bits}, {mc}>"]
    sp = 1 if split == 1 and _alt_k1(logn, tile_bits) > 0 else 0
    k1 = _alt_k1(logn, tile_bits) if sp else _ntt_k1(logn, tile_bits)
    tb2 = tile_bits2 or tile_bits
    if tb2 != tile_bits:
        k1b = (_alt_k1(logn, tb2) if split == 1 and _alt_k1(logn, tb2) > 0
               else _ntt_k1(logn, tb2))
        if logn <= tb2 or k1b != k1:
            tb2 = tile_bits
    sp2 = 1 if split == 1 and _alt_k1(logn, tb2) > 0 else 0
    return [f"k_ntt_pass<{logn}, 1, false, {tile_bits}, {mc}, {sp}>",
            f"k_ntt_pass<{logn}, 2, false, {tb2}, {mc}, {sp2}>"]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--logn", type=int, default=16)
    ap.add_argument("--primes", type=int, default=8)
    ap.add_argument("--per-prime", type=int, default=64)
    ap.add_argument("--plain-polys", type=int, default=64)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--headline-only", action="store_true",
                    help="skip the secondary legs (polymul, uneven, serving)")
    ap.add_argument("--tile-bits", type=int, default=0,
                    help="pin the headline size's tile (2^11/2^12/2^13 words, "
                         "split passes) instead of the tuner's choice, e.g. so "
                         "a profiled run uses the kernels an unprofiled run "
                         "picks (kernel tracing serialises dispatches and can "
                         "change the tuner's ranking)")
    ap.add_argument("--polymul-only", action="store_true",
                    help="only run the polymul leg (after tuning) and print "
                         "its JSON object")
    ap.add_argument("--no-tune", action="store_true",
                    help="skip the autotuner (library default tiles), e.g. so "
                         "that a trace run and a counter run of a profile use "
                         "the same kernels")
    ap.add_argument("--uneven-only", action="store_true",
                    help="only run the uneven-batch leg (1000 mixed jobs as "
                         "one batch, after tuning) and print its JSON object")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run 3 copies of 256 MiB (elemgtadd with bound "
                         "2^64-1) to calibrate FETCH_SIZE / WRITE_SIZE")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on exactly one GPU; "
                          "--gpus must be 1 (multi-GPU sharding is design-only)"}))
        sys.exit(2)

    import vkhel_amd as V

    ev = HipEvents()
    ctx = V.Context()
    if args.copy_calibration:
        words = 32 << 20
        a, b = ctx.vector(words, False), ctx.vector(words, False)
        a.fill_uniform(1 << 62, 1)
        for _ in range(3):
            a.elemgtadd(b, (1 << 64) - 1, 0)  # x > 2^64-1 never holds: a copy
        print(json.dumps({"copy_bytes_read": words * 8, "copy_bytes_written": words * 8,
                          "launches": 3}))
        return
    t_tune = time.perf_counter()
    if not args.no_tune:
        ctx.autotune(12, 17)
    if args.tile_bits:
        ctx.set_tile_bits(args.logn, args.tile_bits)
        ctx.set_pass_mode(args.logn, 2)
    if args.polymul_only:
        print(json.dumps({"polymul": polymul_leg(V, ctx, ev)}))
        return
    if args.uneven_only:
        jobs, _ = make_uneven_jobs(V, ctx, 1000)
        coeffs = sum(job_coeffs(V, j) for j in jobs)
        bt = V.Batch(ctx)
        for j in jobs:
            bt.add(*j)
        ms = timed_launches(ctx, bt, 3, ev)
        print(json.dumps({"uneven_batch": {"batch_ms": round(ms, 3),
                                           "transform_coeffs": coeffs,
                                           "batched_gcoeff_s": round(coeffs / ms / 1e6, 3)}}))
        bt.close()
        return
    tuned = {l: ctx.tile_bits(l) for l in range(12, 18)}
    modes = {l: ctx.pass_mode(l) for l in range(12, 18)}
    t_tune = time.perf_counter() - t_tune

    r = batched_ntt(V, ctx, args.logn, args.primes, args.per_prime, args.steps,
                    args.warmup, ev)
    n, npoly = r["n"], r["npoly"]
    alg_bytes = 16 * npoly * n + 16 * n * args.primes
    achieved = alg_bytes / (r["fwd_ms"] * 1e-3) / 1e9
    plain = plain_per_call(V, ctx, r, args.plain_polys)
    plain["speedup"] = round(r["gcoeff"] / plain["value"], 1)
    ws = [t.w for t in r["tabs"]]
    qs = r["qs"]
    for v in r["polys"]:
        v.close()
    for b in r["batches"]:
        b.close()

    extra = {}
    if not args.headline_only:
        r14 = batched_ntt(V, ctx, 14, 8, 64, args.steps, args.warmup, ev)
        extra["batched_n14"] = {"value": round(r14["gcoeff"], 3), "unit": "Gcoeff/s",
                                "fwd_kernel_ms": round(r14["fwd_ms"], 4),
                                "alg_GBps": round(16 * r14["npoly"] * r14["n"] /
                                                  (r14["fwd_ms"] * 1e-3) / 1e9, 1)}
        for v in r14["polys"]:
            v.close()
        for b in r14["batches"]:
            b.close()
        extra["polymul"] = polymul_leg(V, ctx, ev)
        extra["intt_center"] = intt_center_leg(V, ctx, ev)
        extra["rescale_ntt"] = rescale_ntt_leg(V, ctx, ev)
        extra["pointwise"] = pointwise_leg(V, ctx, ev)
        extra["uneven_batch"], extra["serving"] = uneven_and_serving(V, ctx, ev)
        extra["pcie_inclusive"] = pcie_inclusive(V, ctx)
        extra["map_unmap"] = map_unmap_leg(V, ctx)

    cpu = None if args.no_cpu else cpu_baseline(args.logn, qs, ws, V, ctx)
    tb = ctx.tile_bits(args.logn)
    traffic, prof = traffic_from_profile(args.logn, npoly, tb)
    out = {
        "metric": "batched NTT Gcoeff/s & %HBM peak at N=2^16; served p50/p99 vs plain per-call",
        "value": round(r["gcoeff"], 3),
        "unit": "Gcoeff/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r["secs"] / args.steps * 1e3, 4),
        "gpu_ms_per_step": round(r["gpu_ms"] / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic (device splitmix64, uniform in [0,q))",
        "config": {"workload": f"batched NTT fwd+inv, n=2^{args.logn}, "
                               f"{args.primes} primes <2^50 x {args.per_prime} polys",
                   "n": n, "polys": npoly, "primes": qs},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic,
                     "kernel": "forward NTT of the batch: " + " + ".join(
                         fwd_kernel_names(args.logn, tb, ctx.pass_mode(args.logn),
                                          modulus_class(max(qs)),
                                          ctx.split(args.logn),
                                          ctx.tile_bits2(args.logn))),
                     "kernel_ms": round(r["fwd_ms"], 4),
                     "inverse_kernel_ms": round(r["inv_ms"], 4),
                     "alg_bytes_per_launch": alg_bytes,
                     "profile": (f"profiles/ntt_fwd_n{n}_x{npoly}_tile{tb}.md"
                                 if prof else None),
                     "profiled_valu_busy": prof.get("valu_busy") if prof and traffic else None},
        "plain_per_call": plain,
        "tile_bits": tuned,
        "pass_mode": modes,
        "autotune_s": round(t_tune, 2),
        "cpu_baseline": cpu,
    }
    out.update(extra)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
