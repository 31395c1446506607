"""The reference's ".mat" argv runner on this library (bcsr_legacy_driver):
argument handling on the CPU; on the GPU, one CSV line per product in the
reference's schema name,time,diff,maxrelativeerr,size,pattern,sparsity,
matrixApath (main.cu:81-82), with every product close to the dense one."""
import os
import subprocess

import numpy as np
import pytest

import matrixmultiplicationcuda_amd as mm

DRIVER = os.path.join(os.path.dirname(mm.__file__), "bcsr_legacy_driver")


def test_driver_usage_and_missing_file(tmp_path):
    r = subprocess.run([DRIVER], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "usage" in r.stderr
    r = subprocess.run([DRIVER, str(tmp_path / "nope.mat"), str(tmp_path / "nope2.mat")],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "error" in r.stderr


def test_driver_rejects_shapes(tmp_path):
    a = np.ones((24, 32), np.float16)
    b = np.ones((32, 16), np.float16)
    mm.mat_write(str(tmp_path / "a.mat"), a)
    mm.mat_write(str(tmp_path / "b.mat"), b)
    r = subprocess.run([DRIVER, str(tmp_path / "a.mat"), str(tmp_path / "b.mat")],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "multiples of 16" in r.stderr


@pytest.mark.gpu
def test_driver_csv_rows(tmp_path):
    n = 256
    A = mm.generate(n, n, 0.7, "blockrandom", bh=16, seed=5)
    B = mm.generate(n, n, 0.0, "random", seed=6)
    pa, pb = str(tmp_path / "A_256.mat"), str(tmp_path / "B_256.mat")
    mm.mat_write(pa, A)
    mm.mat_write(pb, B)
    r = subprocess.run([DRIVER, pa, pb, "blockrandom"], capture_output=True, text=True, timeout=300)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    lines = r.stdout.strip().splitlines()
    assert lines[0].startswith("Matrix A: 256 x 256, nonZeros: ")
    rows = [ln.split(",") for ln in lines[1:]]
    names = [x[0] for x in rows]
    for want in ["rocBLAS GeMM", "SpMM 1 Co", "SpMM with Tensors", "SpMM with Tensors Op",
                 "rocSPARSE CSR", "rocSPARSE BSR", "SpMM BF16", "SpMM FP8 E4M3", "SpMM hi+lo"] + \
            [f"Hybrid CSR {t:.6f}" for t in (0.0, 0.2, 0.4, 0.6, 0.8)]:
        assert want in names, want
    for x in rows:
        assert len(x) == 8 and x[4] == "256" and x[5] == "blockrandom" and x[7] == pa
        t, diff = float(x[1]), float(x[2])
        assert t > 0
        bound = {"SpMM BF16": 0.5, "SpMM FP8 E4M3": 2.0}.get(x[0], 0.05)  # |C| ~ up to 80
        assert diff <= bound, (x[0], diff)
    sparsity = float(rows[0][6])
    assert abs(sparsity - np.count_nonzero(A) / A.size) < 1e-3
