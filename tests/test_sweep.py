"""bcsr_sweep, the axis-sweep runner standing in for the reference's NVBench
harness (nvbench_main.cu:1417-1451): benchmark names and axes on the CPU; on
the GPU a small sweep whose every state passes its check and whose CSV and
JSON records agree."""
import csv
import json
import os
import subprocess

import pytest

import matrixmultiplicationcuda_amd as mm

SWEEP = os.path.join(os.path.dirname(mm.__file__), "bcsr_sweep")

# the reference's registered benchmark names (nvbench_main.cu:1417-1424)
REGISTERED = ["cuBLAS_GEMM_TENSOR", "sparseMatrixMulTensor_v2", "sparseMatrixMulTensor32x16",
              "sparseMatrixMulTensor32x16_v2", "sparseMatrixMulTensorAdaptive_16x16",
              "sparseMatrixMulTensorAdaptive_32x16", "sparseMatrixMulTensor32x16_shared_vectorized",
              "sparseMatrixMulTensor32x16_v2_shared_vectorized"]


def run(*args, timeout=60):
    return subprocess.run([SWEEP, *args], capture_output=True, text=True, timeout=timeout)


def test_list_names_every_registered_benchmark():
    r = run("--list")
    assert r.returncode == 0, r.stderr
    rows = [ln.split("|")[1:-1] for ln in r.stdout.splitlines()[2:]]
    names = {c[1].strip(): c[3].strip() for c in rows}
    for n in REGISTERED:
        assert names.get(n) == "yes", n
    for n in ["sparseMatrixMulTensor64x16", "sparseMatrixMult1", "rocSPARSE_BSR", "bcsr_spmm_fp8_e4m3"]:
        assert names.get(n) == "no", n


@pytest.mark.parametrize("args,msg", [(["-a", "X=1"], "unknown axis"), (["-b", "nope"], "unknown benchmark"),
                                      (["-a", "N=[1:0]"], "empty range"), (["-a", "N=a"], "not an integer"),
                                      (["--frobnicate"], "unknown option"), (["-b"], "missing value"),
                                      (["--min-samples", "0"], "min-samples")])
def test_usage_errors(args, msg):
    r = run(*args)
    assert r.returncode == 2 and msg in r.stderr, r.stderr


@pytest.mark.gpu
def test_small_sweep_passes_and_records_agree(tmp_path):
    out_csv, out_json = str(tmp_path / "s.csv"), str(tmp_path / "s.json")
    benches = ["sparseMatrixMulTensor_v2", "sparseMatrixMulTensor32x16", "sparseMatrixMulTensorAdaptive_16x16",
               "sparseMatrixMulTensor64x16", "cuBLAS_GEMM_TENSOR", "sparseMatrixMult1Co", "sparseMatrixMult1",
               "rocSPARSE_CSR", "rocSPARSE_BSR", "bcsr_spmm_bf16", "bcsr_spmm_fp8_e4m3", "bcsr_spmm_hilo",
               "sparseMatrixMulTensor1"]
    args = []
    for b in benches:
        args += ["-b", b]
    args += ["-a", "M=256", "-a", "K=256", "-a", "N=[32,256]", "-a", "SPARS=70", "-a", "PAT=[0:6:2]",
             "--min-samples", "3", "--min-time", "0", "--flush-mib", "64", "--csv", out_csv, "--json", out_json]
    r = run(*args, timeout=600)
    print(r.stdout[-3000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0
    with open(out_csv) as f:
        rows = list(csv.DictReader(f))
    with open(out_json) as f:
        js = json.load(f)["states"]
    assert len(rows) == len(js) == len(benches) * 2 * 4
    for row, st in zip(rows, js):
        assert row["benchmark"] == st["benchmark"]
        assert int(row["N"]) == st["axes"]["N"] and int(row["PAT"]) == st["axes"]["PAT"]
        assert row["pattern"] == ["random", "checkerboard", "diagonal", "blockdiagonal", "blockrandom",
                                  "pattern64by16", "pattern32by16"][int(row["PAT"]) % 7]
        if row["benchmark"] == "sparseMatrixMulTensor1" and row["N"] != "256":
            assert row["check"].startswith("skip"), row
            continue
        assert row["check"] == "pass", row
        assert int(row["samples"]) >= 3 and float(row["gpu_time_mean_ms"]) > 0
        assert float(row["gpu_time_min_ms"]) <= float(row["gpu_time_mean_ms"]) <= float(row["gpu_time_max_ms"])
