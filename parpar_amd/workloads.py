"""Shapes of the synthetic workloads the benchmark and the tests share.

batched_jobs: many independent PAR2 jobs run as one batch -- input count
uniform in [10, 3000], slice size log-uniform in [64 KiB, 4 MiB] (a multiple
of 4), the last input truncated to a random length, recovery count uniform in
[1, 500] with exponents 0..R-1.

serving_arrivals: files arriving as a Poisson process (20 per second) --
size log-uniform in [1 MiB, 8 GiB], slice size log-uniform in
[256 KiB, 16 MiB] (a multiple of 4, raised so that a file has at most 32768
slices), recovery count max(1, ceil(rho * slices)) with rho uniform in
[1 %, 30 %].
"""
import numpy as np


def batched_jobs(seed, count=64):
    """(inputs, slice bytes, last input bytes, recovery) per job."""
    rng = np.random.default_rng(seed)
    jobs = []
    for _ in range(count):
        n = int(rng.integers(10, 3001))
        s = int(np.exp(rng.uniform(np.log(64 << 10), np.log(4 << 20))))
        s -= s % 4
        last = int(rng.integers(1, s + 1))
        r = int(rng.integers(1, 501))
        jobs.append((n, s, last, r))
    return jobs


def serving_arrivals(seed, window_s, rate=20.0):
    """(arrival time s, file bytes, slice bytes, recovery) per file in the window."""
    rng = np.random.default_rng(seed)
    out = []
    t = 0.0
    while True:
        t += float(rng.exponential(1.0 / rate))
        if t >= window_s:
            break
        T = int(np.exp(rng.uniform(np.log(1 << 20), np.log(8 << 30))))
        S = int(np.exp(rng.uniform(np.log(256 << 10), np.log(16 << 20))))
        S -= S % 4
        if -(-T // S) > 32768:
            S = -(-T // 32768)
            S += (-S) % 4
        n = -(-T // S)
        rho = float(rng.uniform(0.01, 0.30))
        out.append((t, T, S, max(1, int(np.ceil(rho * n)))))
    return out
