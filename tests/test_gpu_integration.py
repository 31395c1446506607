"""The drop-in call site, compiled C++: bcsr_integration_example is the
reference's sparseMatrixMulTensor_v2 call site (nvbench_main.cu:393-402)
rewritten for this library as INTEGRATION.md shows it -- the reference's
generator calls and seeds, BCSRMatrix(A, 16, 16).copyToDevice, hipMalloc'd
B and C, and the entry point called with the reference's arguments and no
stream.  Its C is checked against the oracle's literal cpu_matmul_ref
(nvbench_main.cu:78-91) at 1e-3 of (|A||B|)_ij, exact zeros where that is
0, and the reference's own 1e-2 absolute check."""
import os
import subprocess
import sys

import numpy as np
import pytest

import matrixmultiplicationcuda_amd as mm
import oracle
from helpers import assert_close

pytestmark = pytest.mark.gpu

EXE = os.path.join(os.path.dirname(mm.LIB_PATH), "bcsr_integration_example")


@pytest.mark.parametrize("M,K,N,sp", [(1024, 1024, 256, 0.9), (512, 768, 128, 0.8)])
def test_reference_call_site(tmp_path, M, K, N, sp):
    print(f"integration example M={M} K={K} N={N} sparsity={sp}", flush=True)
    sys.stdout.flush()
    assert mm.device_count() > 0
    out = tmp_path / "C.bin"
    r = subprocess.run([EXE, str(out), str(M), str(K), str(N), str(sp)], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    A16 = mm.generate(M, K, sp, "blockrandom", bh=16, seed=123)
    B16 = mm.generate(K, N, 0.0, "random", bh=16, seed=456)
    h = mm.bcsr_from_dense(A16, 16, 16)
    assert r.stdout.split() == [str(M), str(K), str(N), str(h.nnzb)]
    C = np.fromfile(out, dtype=np.float32).reshape(M, N)
    ref = oracle.matmul_dense_ref(A16.view(np.uint16), B16.view(np.uint16))
    _, Cabs = oracle.bcsr_spmm_f64(h.hdr, h.idx, h.data.astype(np.float32), 16, 16,
                                   B16.astype(np.float32), threads=8)
    assert_close(C, ref, Cabs, what="integration example")
