"""Synthetic shard workloads for the benchmarks and the engine tests.

``mixed_batch`` draws shards of uneven sizes from families of different
widths, the shape of a block's shard stream (shards of 2^8..2^22 rows across
23 families, crates/constants/src/lib.rs:1325-1359): by default 207 shards from
23 families of 4-24 columns, log2(rows) 60% in 8-14, 30% in 15-19, 10% in
20-22, witness-like scalars, degree-3 sumcheck. Scalars are not stored here:
a shard's column c is the counter-mode ChaCha20 stream with seed
``shard.seed + c``, generated straight into device memory.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass(frozen=True)
class ShardSpec:
    index: int
    family: int
    num_vars: int
    columns: int
    degree: int
    seed: int

    @property
    def rows(self) -> int:
        return 1 << self.num_vars

    @property
    def elements(self) -> int:
        return self.rows * self.columns


def mixed_batch(seed: int = 20261203, shards: int = 207, families: int = 23, min_cols: int = 4,
                max_cols: int = 24, classes=((0.6, 8, 14), (0.3, 15, 19), (0.1, 20, 22)),
                degree: int = 3) -> list:
    rng = np.random.default_rng(seed)
    fam_cols = rng.integers(min_cols, max_cols + 1, size=families)
    probs = np.array([c[0] for c in classes], dtype=float)
    probs /= probs.sum()
    out = []
    for i in range(shards):
        fam = int(rng.integers(0, families))
        cls = int(rng.choice(len(classes), p=probs))
        lo, hi = classes[cls][1], classes[cls][2]
        nv = int(rng.integers(lo, hi + 1))
        out.append(ShardSpec(i, fam, nv, int(fam_cols[fam]), degree, seed * 1000003 + i * 4096))
    return out


def poisson_arrivals(rate_per_s: float, count: int, seed: int) -> np.ndarray:
    """Arrival times (s) of a seeded Poisson process."""
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / rate_per_s, size=count))
