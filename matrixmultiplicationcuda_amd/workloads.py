"""Synthetic workloads of the batched and learned-dispatch paths.

A batched problem is one BCSR x dense product drawn from these
distributions: rows and cols log-uniform on [lo, hi] rounded to a multiple of
64, block shape from {16x16, 32x16, 64x16}, block density U(0.5%, 25%), the
"blockrandom_rho" pattern (Bernoulli blocks, each entry of a kept block
nonzero with probability rho ~ U(0.3, 1.0), zeros kept inside stored
blocks), N from {32, 128, 512}, fp16 values U(-1, 1).  Problem i of a set
uses generator seed base_seed + i for A and base_seed + i + 500000 for B.
"""
from dataclasses import dataclass

import numpy as np

from . import HostBCSR, generate, generate_bcsr

BLOCK_SHAPES = ((16, 16), (32, 16), (64, 16))
N_CHOICES = (32, 128, 512)


@dataclass
class Problem:
    rows: int
    cols: int
    bh: int
    bw: int
    density: float
    rho: float
    N: int
    seed: int

    def bcsr(self, dtype=np.float16) -> HostBCSR:
        """The problem's A (block values fp16, or fp32 for the quantised modes)."""
        h = generate_bcsr(self.rows, self.cols, 1.0 - self.density, "blockrandom_rho", bh=self.bh,
                          bw=self.bw, seed=self.seed, inblock_density=self.rho)
        if dtype == np.float16:
            return HostBCSR(h.hdr, h.idx, h.data.astype(np.float16), h.block_rows, h.block_cols,
                            self.bh, self.bw)
        return h

    def B(self) -> np.ndarray:
        return generate(self.cols, self.N, 0.0, "random", bh=16, seed=self.seed + 500000)


def _round64(x):
    return max(64, int(round(x / 64.0)) * 64)


def draw_problems(count, base_seed=1000, lo=512, hi=16384, rng_seed=None):
    """count problems from the batched distributions (parameters from
    numpy's PCG64 seeded with rng_seed, default base_seed)."""
    rng = np.random.default_rng(base_seed if rng_seed is None else rng_seed)
    out = []
    for i in range(count):
        rows = _round64(np.exp(rng.uniform(np.log(lo), np.log(hi))))
        cols = _round64(np.exp(rng.uniform(np.log(lo), np.log(hi))))
        bh, bw = BLOCK_SHAPES[rng.integers(len(BLOCK_SHAPES))]
        density = float(rng.uniform(0.005, 0.25))
        rho = float(rng.uniform(0.3, 1.0))
        N = int(N_CHOICES[rng.integers(len(N_CHOICES))])
        out.append(Problem(rows, cols, bh, bw, density, rho, N, base_seed + i))
    return out


def flops(h: HostBCSR, N: int) -> float:
    return 2.0 * h.nnzb * h.bh * h.bw * N
