"""Shared case construction for the tests.

make_case mirrors prepare_buffers (nvbench_main.cu:139-175): A from the
generator with seed 123 (block size = bh), B = generate_matrix(K, N, 0.0,
"random", bh, 456), both rounded double -> float -> half; BCSR of A with
bh x bw blocks.  Products are checked against the oracle's BCSR-order fp32
product (bit-identical to cpu_matmul_ref) with the per-entry bound
|C_gpu - C_cpu| <= 1e-3 * (|A|.|B|)_ij, exact zeros where (|A|.|B|)_ij = 0.
"""
from dataclasses import dataclass

import numpy as np

import matrixmultiplicationcuda_amd as mm
import oracle

REL_TOL = 1e-3  # per entry, relative to (|A|.|B|)_ij
REF_ABS_TOL = 1e-2  # the reference's own check, nvbench_main.cu:413-420


@dataclass
class Case:
    A16: np.ndarray
    B16: np.ndarray
    h: mm.HostBCSR
    bh: int
    bw: int

    @property
    def M(self):
        return self.A16.shape[0]

    @property
    def K(self):
        return self.A16.shape[1]

    @property
    def N(self):
        return self.B16.shape[1]


def make_case(M, K, N, sparsity=0.9, pattern="blockrandom", bh=16, bw=16, gen_bh=None,
              gen_bw=None, seed_a=123, seed_b=456, **gen_kw):
    gbh = bh if gen_bh is None else gen_bh
    gbw = gbh if gen_bw is None else gen_bw
    A16 = mm.generate(M, K, sparsity, pattern, bh=gbh, bw=gbw, seed=seed_a, **gen_kw)
    B16 = mm.generate(K, N, 0.0, "random", bh=bh, seed=seed_b)
    h = mm.bcsr_from_dense(A16, bh, bw)
    return Case(A16, B16, h, bh, bw)


def cpu_reference(case, br0=0, br1=None):
    """(C_cpu fp32, |A|.|B| fp64) from the oracle for block-rows [br0, br1)."""
    Bf = case.B16.astype(np.float32)
    hdr, idx = case.h.hdr, case.h.idx
    data16 = case.h.data.view(np.uint16)
    C = oracle.bcsr_spmm_f32(hdr, idx, data16, case.bh, case.bw, Bf, br0, br1, threads=8)
    data32 = case.h.data.astype(np.float32)
    _, Cabs = oracle.bcsr_spmm_f64(hdr, idx, data32, case.bh, case.bw, Bf, br0, br1, threads=8)
    return C, Cabs


def assert_close(C_gpu, C_cpu, Cabs, rows=None, what=""):
    if rows is not None:
        C_gpu, C_cpu, Cabs = C_gpu[rows], C_cpu[rows], Cabs[rows]
    err = np.abs(C_gpu.astype(np.float64) - C_cpu.astype(np.float64))
    zero = Cabs == 0
    assert np.all(C_gpu[zero] == 0), f"{what}: nonzero output where |A||B| = 0"
    bound = REL_TOL * Cabs
    bad = err > bound
    if bad.any():
        i = np.argwhere(bad)[0]
        raise AssertionError(f"{what}: {bad.sum()} entries exceed 1e-3*|A||B|; first at {tuple(i)}:"
                             f" gpu={C_gpu[tuple(i)]} cpu={C_cpu[tuple(i)]} bound={bound[tuple(i)]}")
    assert err.max(initial=0.0) <= REF_ABS_TOL, f"{what}: reference 1e-2 check failed"
    nz = ~zero
    return float((err[nz] / Cabs[nz]).max(initial=0.0))
