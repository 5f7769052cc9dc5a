"""bench.py's command-line contract, on the CPU: one GPU only, and the
helpers that size its workloads."""
import json
import os
import subprocess
import sys

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, REPO)

import bench  # noqa: E402
from oracle import oracle as O  # noqa: E402


def test_more_than_one_gpu_exits_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(REPO, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    msg = json.loads(r.stdout.strip().splitlines()[-1])
    assert "one GPU" in msg["error"]


def test_workload_primes_and_roots():
    # the headline's primes: the largest below 2^50 that are 1 mod 2n
    n = 1 << 16
    qs = bench.primes_ntt(1 << 50, 2 * n, 8)
    assert qs == O.primes_ntt(1 << 50, 2 * n, 8)
    assert all(q < 1 << 50 and q % (2 * n) == 1 and O.is_prime(q) for q in qs)
    assert qs == sorted(qs, reverse=True)
    for q in qs[:2]:
        w = bench.root_2n(q, n)
        assert pow(w, n, q) == q - 1  # a primitive 2n-th root


def test_uniform_mod_matches_the_oracle_stream():
    q = bench.primes_ntt(1 << 55, 1 << 16, 1)[0]
    assert np.array_equal(bench.uniform_mod(q, 5000, 9), O.fill_uniform(5000, q, 9))
