"""GPU tests of the serving engine (vkhel_server_*, vkhel_ext.h).

Results must be bit-identical whether a request runs alone or mixed with
others, and equal to the oracle.  Admission, budget, queue, cancel and
timeout behaviour are forced from the host (budgets, queue sizes); no test
makes a kernel misbehave.
"""
import numpy as np
import pytest

from oracle import oracle as O
import vkhel_amd as V
from vkhel_amd import vkhel as VV

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = V.Context()
    yield c
    c.close()


_TABLE_CACHE = {}


def tables_for(logn, nlimbs):
    n = 1 << logn
    key = (logn, nlimbs)
    if key not in _TABLE_CACHE:
        qs = O.primes_ntt(1 << 60, 1 << 18, nlimbs)
        _TABLE_CACHE[key] = [V.NttTables(n, q, O.root_2n(q, n)) for q in qs]
    return _TABLE_CACHE[key]


def make_jobs(ctx, rng, count, logn_range=(10, 15), max_limbs=4, ops=None):
    """Random jobs; op drawn from NTT / INTT / polymul / mul / add, or
    ops[k % len(ops)] for job k when `ops` is given."""
    jobs = []
    for k in range(count):
        logn = int(rng.integers(*logn_range))
        L = int(rng.integers(1, max_limbs + 1))
        tabs = tables_for(logn, L)
        n = 1 << logn
        op = int(rng.integers(0, 5)) if ops is None else ops[k % len(ops)]
        binary = op >= V.JOB_POLYMUL
        a, r = ctx.vector(L * n, False), ctx.vector(L * n)
        b = ctx.vector(L * n, False) if binary else None
        for i, t in enumerate(tabs):
            a.fill_uniform(t.q, 1000 * k + i, i * n, n)
            if binary:
                b.fill_uniform(t.q, 1000 * k + 500 + i, i * n, n)
        jobs.append((op, a, b, r, tabs))
    return jobs


def run_alone(ctx, job):
    op, a, b, r, tabs = job
    alone = ctx.vector(r.length)
    bt = V.Batch(ctx)
    bt.add(op, a, b, alone, tabs)
    bt.run()
    out = alone.read()
    bt.close()
    alone.close()
    return out


def test_served_results_equal_alone_and_oracle(ctx):
    rng = np.random.default_rng(11)
    jobs = make_jobs(ctx, rng, 60)
    srv = V.Server(ctx, streams=4, max_batch=16)
    st, infos = srv.replay(jobs, rate_per_s=0)
    assert st == VV.OK
    assert all(i.status == VV.OK for i in infos)
    assert max(i.batch_size for i in infos) > 1  # requests were batched together
    for k, job in enumerate(jobs):
        op, a, b, r, tabs = job
        got = r.read()
        assert np.array_equal(got, run_alone(ctx, job)), k
        if k % 6 == 0:  # oracle spot checks across ops
            n = tabs[0].n
            xa = a.read()
            xb = b.read() if b is not None else None
            for i, t in enumerate(tabs):
                sl = slice(i * n, (i + 1) * n)
                if op == V.JOB_NTT_FWD:
                    want = O.ntt_forward(xa[sl], t.q, t.w)
                elif op == V.JOB_NTT_INV:
                    want = O.ntt_inverse(xa[sl], t.q, t.w)
                elif op == V.JOB_POLYMUL:
                    want = O.polymul_ntt(xa[sl], xb[sl], t.q, t.w)
                elif op == V.JOB_MUL:
                    want = O.elemmul(xa[sl], xb[sl], t.q)
                else:
                    want = O.addmul_mod(xa[sl], xb[sl], 1, t.q)
                assert np.array_equal(got[sl], want), (k, op, i)
    stats = srv.stats()
    assert stats.completed == len(jobs) and stats.inflight == 0
    srv.close()
    for _, a, b, r, _ in jobs:
        a.close()
        r.close()
        if b is not None:
            b.close()


def test_poisson_arrivals_latency_records(ctx):
    rng = np.random.default_rng(12)
    jobs = make_jobs(ctx, rng, 80, (10, 13), 3)
    srv = V.Server(ctx, streams=2)
    st, infos = srv.replay(jobs, rate_per_s=20000, seed=5)
    assert st == VV.OK
    for i in infos:
        assert i.status == VV.OK
        assert i.submit_ns <= i.admit_ns <= i.complete_ns
    subs = [i.submit_ns for i in infos]
    assert subs == sorted(subs)
    srv.close()


def test_budget_limits_inflight_bytes(ctx):
    rng = np.random.default_rng(13)
    jobs = make_jobs(ctx, rng, 24, (12, 14), 2)
    # admission cost: operands + result; the workers' scratch arenas are
    # reserved from the budget when the server is created
    def cost(b, words):
        return 8 * words * (3 if b is not None else 2)
    biggest = max(cost(b, r.length) for op, a, b, r, tabs in jobs)
    arena = 1 << 20
    budget = 4 * arena + int(biggest * 2)
    srv = V.Server(ctx, hbm_budget_bytes=budget, streams=4, scratch_bytes=arena)
    st, infos = srv.replay(jobs, 0)
    assert st == VV.OK
    s = srv.stats()
    assert s.reserved_bytes == 4 * arena
    assert 0 < s.peak_inflight_bytes <= budget - s.reserved_bytes
    srv.close()


def test_arenas_bound_micro_batches_and_requests(ctx):
    tabs = tables_for(13, 2)
    n = 1 << 13
    words = 2 * n
    vecs, jobs = [], []
    for k in range(6):
        a, b, r = (ctx.vector(words, False) for _ in range(3))
        for i, t in enumerate(tabs):
            a.fill_uniform(t.q, 40 + 10 * k + i, i * n, n)
            b.fill_uniform(t.q, 45 + 10 * k + i, i * n, n)
        vecs += [a, b, r]
        jobs.append((V.JOB_POLYMUL, a, b, r, tabs))
    # an arena that holds exactly one product's spectra (2 limbs x n words)
    srv = V.Server(ctx, streams=2, scratch_bytes=8 * words)
    st, infos = srv.replay(jobs, 0)
    assert st == VV.OK
    assert max(i.batch_size for i in infos) == 1
    for op, a, b, r, t in jobs:
        assert np.array_equal(r.read(), run_alone(ctx, (op, a, b, r, t)))
    srv.close()
    # a product whose spectra exceed one arena is refused at submit
    small = V.Server(ctx, streams=2, scratch_bytes=8 * n)
    with pytest.raises(V.VkhelError) as e:
        small.submit(*jobs[0])
    assert e.value.code == VV.ERR_BUDGET
    small.close()
    # arenas larger than the budget: the server is not created
    with pytest.raises(V.VkhelError) as e:
        V.Server(ctx, hbm_budget_bytes=1 << 20, streams=4, scratch_bytes=1 << 20)
    assert e.value.code == VV.ERR_BUDGET
    for v in vecs:
        v.close()


def test_per_call_server_matches_batch(ctx):
    """The plain per-call mode serves every request alone through the
    synchronous single calls; its words equal the batch path's."""
    rng = np.random.default_rng(19)
    ops = [V.JOB_NTT_FWD, V.JOB_NTT_INV, V.JOB_POLYMUL, V.JOB_ADD, V.JOB_MUL,
           V.JOB_ADD_REF]
    jobs = make_jobs(ctx, rng, 30, (10, 14), 3, ops=ops)
    srv = V.Server(ctx, per_call=True, streams=4, max_batch=64)
    st, infos = srv.replay(jobs, 50000, seed=3)
    assert st == VV.OK
    assert all(i.batch_size == 1 for i in infos)
    for job in jobs:
        assert np.array_equal(job[3].read(), run_alone(ctx, job)), job[0]
    srv.close()
    for _, a, b, r, _ in jobs:
        for v in (a, b, r):
            if v is not None:
                v.close()


def test_request_over_budget_rejected(ctx):
    tabs = tables_for(12, 1)
    a, r = ctx.vector(4096), ctx.vector(4096)
    srv = V.Server(ctx, hbm_budget_bytes=1000)
    with pytest.raises(V.VkhelError) as e:
        srv.submit(V.JOB_NTT_FWD, a, None, r, tabs)
    assert e.value.code == VV.ERR_BUDGET
    assert srv.stats().rejected == 1
    srv.close()


def test_queue_full_cancel_and_timeout(ctx):
    # One stream, one request in flight at a time, big polymul requests so
    # later ones are still waiting while we act on them.
    tabs = tables_for(16, 8)
    n = 1 << 16
    vecs, ops = [], []
    for k in range(8):
        a, b, r = (ctx.vector(8 * n, False) for _ in range(3))
        for i, t in enumerate(tabs):
            a.fill_uniform(t.q, k * 10 + i, i * n, n)
            b.fill_uniform(t.q, k * 10 + 5 + i, i * n, n)
        vecs += [a, b, r]
        ops.append((a, b, r))
    srv = V.Server(ctx, streams=1, max_inflight=1, queue_capacity=3)
    ids, busy = [], 0
    for a, b, r in ops:
        try:
            ids.append(srv.submit(V.JOB_POLYMUL, a, b, r, tabs))
        except V.VkhelError as e:
            assert e.code == VV.ERR_BUSY
            busy += 1
    assert busy >= 1
    last = ids[-1]
    st, _ = srv.wait(last, timeout_ns=1)
    assert st == VV.ERR_TIMEOUT
    srv.cancel(last)
    st, info = srv.wait(last)
    assert st == VV.ERR_CANCELLED and info.status == VV.ERR_CANCELLED
    for rid in ids[:-1]:
        st, _ = srv.wait(rid)
        assert st == VV.OK
    assert srv.stats().cancelled == 1
    srv.close()
    for v in vecs:
        v.close()


def test_server_rns_requests_match_batch(ctx):
    """RNS requests (centered lift, INTT_CENTER, both rescalings) submitted
    to the server among transform requests give the words the same jobs
    give in a batch."""
    rng = np.random.default_rng(17)
    n = 1 << 12
    tabs = tables_for(12, 4)
    srv = V.Server(ctx, streams=2, max_batch=8)
    reqs, vecs = [], []
    kinds = [(V.JOB_CENTER, [65537], 0), (V.JOB_INTT_CENTER, [2], 0),
             (V.JOB_RESCALE, None, 1), (V.JOB_RESCALE_NTT, None, 1)]
    for k in range(16):
        op, prm, drop = kinds[k % 4]
        a = ctx.vector(4 * n, False)
        for i, t in enumerate(tabs):
            a.fill_uniform(t.q, 7000 + 10 * k + i, i * n, n)
        r1, r2 = ctx.vector((4 - drop) * n), ctx.vector((4 - drop) * n)
        x, y = ctx.vector(4 * n, False), ctx.vector(4 * n)
        for i, t in enumerate(tabs):
            x.fill_uniform(t.q, 8000 + 10 * k + i, i * n, n)
        rid = srv.submit_rns(op, a, r1, tabs, prm)
        rid2 = srv.submit(V.JOB_NTT_FWD, x, None, y, tabs)
        reqs.append((rid, rid2, op, prm, a, r1, r2))
        vecs += [a, r1, r2, x, y]
    for rid, rid2, op, prm, a, r1, r2 in reqs:
        assert srv.wait(rid)[0] == VV.OK and srv.wait(rid2)[0] == VV.OK
        b = V.Batch(ctx)
        b.add_rns(op, a, r2, tabs, prm)
        b.run()
        b.close()
        assert np.array_equal(r1.read(), r2.read()), op
    with pytest.raises(V.VkhelError):
        srv.submit_rns(V.JOB_RESCALE, reqs[0][4], reqs[0][4], tabs)
    srv.close()
    for v in vecs:
        v.close()
