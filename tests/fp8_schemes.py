"""Error of FP8 E4M3 storage schemes against the FP64 product, on the
precision workload's values (16 x 16 blocks, blockrandom, block magnitudes
log-uniform on [1e-3, 1e2] with random sign; B = U(-1, 1) fp16, seed 456).

Schemes (A storage / B operand):
  shipped        E8M0 scale per 16 x 32 tile of A (two device blocks, the
                 scale the CDNA4 scaled MFMA applies) / FP8 B with one E8M0
                 scale per 16-row slab and 256-column chunk (the library's
                 mode; for this U(-1, 1) B every slab of a chunk gets the
                 same scale)
  per_block_f32  FP32 absmax scale per 16 x 16 block of A / the same FP8 B
  a_only         the shipped A / B kept fp16
  a_only_f32     the per-block FP32 A / B kept fp16
Each product is taken in fp64 on the dequantised operands (the kernels add
in fp32; the difference is far below the quantisation error) and compared
with the fp64 product of the unquantised values: the maximum and mean of
|C - C64|_ij / (|A||B|)_ij over entries with (|A||B|)_ij > 0, and the
normwise ||C - C64||_F / ||C64||_F.

    python tests/fp8_schemes.py [M] [density] [N]
prints the table (the full precision workload is 16384 0.05 512).
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

import matrixmultiplicationcuda_amd as mm  # noqa: E402
import oracle  # noqa: E402

SCHEMES = ("shipped", "per_block_f32", "a_only", "a_only_f32")


def workload(M, density, N, seed=321):
    h = mm.generate_bcsr(M, M, 1.0 - density, "blockrandom", bh=16, bw=16, seed=seed,
                         value_dist="loguniform", lo=1e-3, hi=1e2)
    B16 = mm.generate(M, N, 0.0, "random", bh=16, seed=456)
    return h, B16


def errors(C, C64, Cabs):
    nz = Cabs > 0
    rel = np.abs(C - C64)[nz] / Cabs[nz]
    return {"max": float(rel.max(initial=0.0)), "mean": float(rel.mean()) if rel.size else 0.0,
            "normwise": float(np.linalg.norm(C - C64) / max(np.linalg.norm(C64), 1e-300))}


def measure(h, B16, threads=8, schemes=SCHEMES):
    Bf = B16.astype(np.float32)
    data = np.ascontiguousarray(h.data, dtype=np.float32)
    C64, Cabs = oracle.bcsr_spmm_f64(h.hdr, h.idx, data, 16, 16, Bf, threads=threads)
    a_tiles = oracle.fp8_tiles_dequant(h.hdr, h.idx, data, 16, 16)
    a_f32 = oracle.fp8_blocks_f32scale_dequant(data, 256)
    b8 = oracle.quantized_B("fp8", B16.view(np.uint16))
    ops = {"shipped": (a_tiles, b8), "per_block_f32": (a_f32, b8), "a_only": (a_tiles, Bf),
           "a_only_f32": (a_f32, Bf)}
    out = {}
    for name in schemes:
        Aq, Bq = ops[name]
        C, _ = oracle.bcsr_spmm_f64(h.hdr, h.idx, Aq, 16, 16, Bq, threads=threads, want_abs=False)
        out[name] = errors(C, C64, Cabs)
    return out


if __name__ == "__main__":
    M = int(sys.argv[1]) if len(sys.argv) > 1 else 4096
    d = float(sys.argv[2]) if len(sys.argv) > 2 else 0.05
    N = int(sys.argv[3]) if len(sys.argv) > 3 else 512
    h, B16 = workload(M, d, N)
    print(f"M = K = {M}, {d:g} blocks (nnzb {h.nnzb}), N = {N}")
    for k, v in measure(h, B16, threads=os.cpu_count() or 1).items():
        print(f"{k:14s} max {v['max']:.3e}  mean {v['mean']:.3e}  normwise {v['normwise']:.3e}")
