"""The uneven-size batch and its serving at their stated workload: 1000
seeded jobs, n = 2^10..2^17, 1..24 primes < 2^60, ops NTT / INTT / polymul /
add (the reference's elemfma with multiplier 1) / mul (bench.py
make_uneven_jobs, the mix bench.py times).

* The whole mix runs as one batch; a stratified sample -- every (log2 n, op)
  pair, first and last limb of its first job, and every limb of the first
  job with 24 limbs -- is checked against the oracle, which follows the
  reference's per-polynomial calls (src/vector.c:298-340, 388-427, 513-657).
* The same jobs are replayed through the serving engine, as a burst and at a
  Poisson arrival rate, into cleared result vectors; every served result
  must equal the batch's word for word.
"""
import numpy as np
import pytest

from oracle import oracle as O
import bench
import vkhel_amd as V

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mix():
    ctx = V.Context()
    jobs, tables = bench.make_uneven_jobs(V, ctx, 1000)
    bt = V.Batch(ctx)
    for j in jobs:
        bt.add(*j)
    bt.run()
    snapshot = [r.read(0, len(tabs) * tabs[0].n) for _, _, _, r, tabs in jobs]
    yield ctx, jobs, snapshot
    bt.close()
    for op, a, b, r, tabs in jobs:
        for v in (a, b, r):
            if v is not None:
                v.close()
    for t in tables.values():
        t.close()
    ctx.close()


def _logn(tabs):
    return tabs[0].n.bit_length() - 1


def _expected(op, a, b, q, w):
    if op == V.JOB_NTT_FWD:
        return O.ntt_forward(a, q, w)
    if op == V.JOB_NTT_INV:
        return O.ntt_inverse(a, q, w)
    if op == V.JOB_POLYMUL:
        return O.polymul_ntt(a, b, q, w)
    if op == V.JOB_ADD_REF:
        return O.elemfma(a, b, 1, q)
    if op == V.JOB_MUL:
        return O.elemmul(a, b, q)
    raise ValueError(op)


def test_mix_covers_the_stated_workload(mix):
    _, jobs, _ = mix
    sizes = {_logn(t) for _, _, _, _, t in jobs}
    ops = {op for op, *_ in jobs}
    limbs = [len(t) for *_, t in jobs]
    assert sizes == set(range(10, 18))
    assert ops == set(bench.uneven_ops(V))
    assert min(limbs) == 1 and max(limbs) == 24
    pairs = {(_logn(t), op) for op, _, _, _, t in jobs}
    assert len(pairs) == 8 * 5
    assert all(q < (1 << 60) for *_, t in jobs for q in (x.q for x in t))


def _check_job(job, got, limbs):
    op, a, b, r, tabs = job
    n = tabs[0].n
    for i in limbs:
        t = tabs[i]
        av = a.read(i * n, n)
        bv = b.read(i * n, n) if b is not None else None
        want = _expected(op, av, bv, t.q, t.w)
        assert np.array_equal(got[i * n:(i + 1) * n], want), \
            f"op {op} n=2^{_logn(tabs)} limb {i}/{len(tabs)} q={t.q}"


def test_uneven_batch_stratified_sample_vs_oracle(mix):
    _, jobs, snapshot = mix
    seen = set()
    checked = 0
    for k, job in enumerate(jobs):
        op, _, _, _, tabs = job
        key = (_logn(tabs), op)
        if key in seen:
            continue
        seen.add(key)
        print(f"job {k}: n=2^{key[0]} op={op} L={len(tabs)}", flush=True)
        _check_job(job, snapshot[k], sorted({0, len(tabs) - 1}))
        checked += 1
    assert checked == 40
    k24 = next(k for k, j in enumerate(jobs) if len(j[4]) == 24)
    print(f"job {k24}: every limb of a 24-limb job", flush=True)
    _check_job(jobs[k24], snapshot[k24], range(24))
    # the largest products of the mix: a 2^16 and a 2^17 polymul, every limb
    for logn in (16, 17):
        k = next(k for k, j in enumerate(jobs)
                 if j[0] == V.JOB_POLYMUL and _logn(j[4]) == logn)
        print(f"job {k}: polymul n=2^{logn}, every limb", flush=True)
        _check_job(jobs[k], snapshot[k], range(len(jobs[k][4])))


@pytest.mark.parametrize("rate", [0.0, 150000.0])
def test_uneven_jobs_served_equal_batch(mix, rate):
    ctx, jobs, snapshot = mix
    for k, (op, a, b, r, tabs) in enumerate(jobs):
        r.fill_uniform(3, 99 + k)  # words the served result must replace
    srv = V.Server(ctx, streams=4)
    try:
        st, infos = srv.replay(jobs, rate_per_s=rate, seed=77)
        assert st == 0, st
    finally:
        srv.close()
    for k, (op, a, b, r, tabs) in enumerate(jobs):
        got = r.read(0, len(tabs) * tabs[0].n)
        assert np.array_equal(got, snapshot[k]), f"served job {k} differs (rate {rate})"
        i = infos[k]
        assert i.status == 0 and i.submit_ns <= i.admit_ns <= i.complete_ns
    if rate:
        # arrivals were spread out: the replay took at least the Poisson span
        span = (max(i.submit_ns for i in infos) - min(i.submit_ns for i in infos)) / 1e9
        arrivals = bench.poisson_arrivals(len(jobs), rate, 77)
        assert span >= 0.9 * (arrivals[-1] - arrivals[0])
