"""Seeded synthetic workloads of the measured configurations, as plain specs.

Each spec names an instance by its generator seed and shape (the instance
itself is `generate_instance(seed, n, m, 1, 99)`: per job a Fisher-Yates
machine order and uniform durations from a splitmix64 stream), so the bench,
the GPU tests and the fixture script build the same inputs without sharing
anything but these draws.  Pure Python: no library needed.
"""
from __future__ import annotations

import random

DMIN, DMAX = 1, 99
SOLVE_SEED = 12345

# 50 jobs x 20 machines, K = 4096: the headline solve (instance seed 1).
HEADLINE = dict(gen_seed=1, n=50, m=20, K=4096, seed=SOLVE_SEED)
# 16 machine-focused beams sharing one bound: solve seeds 1..16 on the
# headline instance.
MULTISEED = dict(gen_seed=1, n=50, m=20, K=4096, seeds=list(range(1, 17)))


def batch_specs(count: int = 256, gen_seed: int = 2026) -> list[dict]:
    """The batched launch: n in {10,15,20,30,50}, m in {5,10,15,20},
    K in {256,1024}, uniform draws from random.Random(gen_seed)."""
    rng = random.Random(gen_seed)
    out = []
    for _ in range(count):
        n = rng.choice([10, 15, 20, 30, 50])
        m = rng.choice([5, 10, 15, 20])
        K = rng.choice([256, 1024])
        out.append(dict(gen_seed=rng.randrange(1, 2**31), n=n, m=m, K=K, seed=SOLVE_SEED))
    return out


def serve_specs(count: int = 64, seed: int = 7) -> tuple[list[dict], list[float]]:
    """The serving mix: n in [10,100], m in [5,20], K in {256,1024,4096}
    (random.Random(seed)), and unit-rate exponential inter-arrival gaps (a
    Poisson process, scaled to the arrival rate by the caller)."""
    rng = random.Random(seed)
    specs = []
    for _ in range(count):
        n = rng.randint(10, 100)
        m = rng.randint(5, 20)
        K = rng.choice([256, 1024, 4096])
        specs.append(dict(gen_seed=rng.randrange(1, 2**31), n=n, m=m, K=K, seed=SOLVE_SEED))
    gaps = [rng.expovariate(1.0) for _ in range(count)]
    return specs, gaps


# Members of the batch and of the serving mix whose oracle results are
# committed as fixtures (tests/golden/oracle_full_size.json).
BATCH_GOLDEN_MEMBERS = list(range(0, 256, 16))


def serve_golden_members(specs: list[dict], count: int = 10) -> list[int]:
    """The first `count` requests of the serving mix with K >= 1024 (the
    wide requests whose beams the K <= 512 serving tests never reach)."""
    return [i for i, s in enumerate(specs) if s["K"] >= 1024][:count]
