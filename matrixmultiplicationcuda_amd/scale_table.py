"""The plain fp16 product at scale over every block shape, pattern, density
and N of the at-scale workload, one row per cell.

    python -m matrixmultiplicationcuda_amd.scale_table --out table.json [--md table.md]

A = 16384 x 16384 fp16 BCSR, B = 16384 x N fp16 (seed 456), C fp32:
  * block shapes 16 x 16, 32 x 16, 64 x 16;
  * patterns: "blockrandom" at 1 %, 5 % and 20 % of the blocks (seed 123),
    "blockband" (half-width 8 block-columns) and "blockdiagonal" (256 x 256
    dense diagonal blocks, stored as the BCSR shape's blocks); the last two
    have the block density their structure fixes;
  * N in {64, 256, 512}.
Each cell times the automatic launch (hipEvents, warm-up first) and reports
effective TFLOPS 2 nnzb bh bw N / t, the roofline time max(F / 2.5 PFLOP/s,
Q / 8 TB/s) with Q the algorithmic bytes, its fraction, and the largest
difference over 64 sampled rows against rocBLAS's dense product of the same
rows (as a check of the run, not a tolerance test).  For square blocks it
also times rocSPARSE's BSR SpMM on the same operands (a comparison row).
"""
import argparse
import ctypes
import json
import sys
import time

import numpy as np

from . import DeviceArray, Event, HostBCSR, Plan, Stream, bcsr_from_dense, check, generate, generate_bcsr, lib

HBM = 8.0e12
MFMA_F16 = 2.5e15


def algorithmic_bytes(nnzb, bh, bw, M, K, N):
    return nnzb * bh * bw * 2 + 4 * nnzb + 4 * (M // bh + 1) + K * N * 2 + M * N * 4


def matrices(M, K, bh):
    """(pattern label, block density label, HostBCSR fp16) for one block shape."""
    bw = 16
    for d in (0.01, 0.05, 0.20):
        h = generate_bcsr(M, K, 1.0 - d, "blockrandom", bh=bh, bw=bw, seed=123)
        yield "blockrandom", d, HostBCSR(h.hdr, h.idx, h.data.astype(np.float16), h.block_rows, h.block_cols,
                                          bh, bw)
    # the band is defined on bh x bw blocks: block-columns within 8 of the
    # ones the block-row's diagonal crosses
    h = generate_bcsr(M, K, 0.0, "blockband", bh=bh, bw=bw, seed=123, band_halfwidth=8)
    yield "blockband", None, HostBCSR(h.hdr, h.idx, h.data.astype(np.float16), h.block_rows, h.block_cols, bh, bw)
    A = generate(M, K, 0.0, "blockdiagonal", bh=256, seed=123)
    yield "blockdiagonal", None, bcsr_from_dense(A, bh, bw)


def dense_rows(h, rows):
    """Rows `rows` of the BCSR matrix as a dense fp16 array."""
    K = h.block_cols * h.bw
    out = np.zeros((len(rows), K), np.float16)
    for i, r in enumerate(rows):
        br, rr = divmod(int(r), h.bh)
        for k in range(h.hdr[br], h.hdr[br + 1]):
            c = h.idx[k] * h.bw
            out[i, c:c + h.bw] = h.data[k, rr]
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawTextHelpFormatter)
    ap.add_argument("--M", type=int, default=16384)
    ap.add_argument("--K", type=int, default=16384)
    ap.add_argument("--N", default="64,256,512")
    ap.add_argument("--bh", default="16,32,64")
    ap.add_argument("--reps", type=int, default=20)
    ap.add_argument("--out", required=True)
    ap.add_argument("--md", default=None)
    args = ap.parse_args(argv)
    Ns = [int(x) for x in args.N.split(",")]
    M, K = args.M, args.K
    check(lib().bcsr_set_device(0))
    stream = Stream()
    rng = np.random.default_rng(7)
    sample = np.sort(rng.choice(M, 64, replace=False))
    Bs = {}
    for N in Ns:
        B16 = generate(K, N, 0.0, "random", bh=16, seed=456)
        Bs[N] = (B16, DeviceArray.from_numpy(B16))
    rows = []
    t_start = time.time()
    for bh in [int(x) for x in args.bh.split(",")]:
        for pattern, dens, h in matrices(M, K, bh):
            plan = Plan(h, "f16")
            plain = None
            if h.bh == h.bw:  # rocSPARSE BSR takes square blocks
                plain = [DeviceArray.from_numpy(np.ascontiguousarray(h.hdr, np.int32)),
                         DeviceArray.from_numpy(np.ascontiguousarray(h.idx, np.int32)),
                         DeviceArray.from_numpy(np.ascontiguousarray(h.data, np.float16))]
            Ad = dense_rows(h, sample)
            dA = DeviceArray.from_numpy(Ad)
            density = h.nnzb / float(h.block_rows * h.block_cols)
            for N in Ns:
                B16, dB = Bs[N]
                C = DeviceArray((M, N), np.float32)
                ms = plan.time_ptr(dB.ptr, C.ptr, N, stream, warmup=5, reps=args.reps)
                Cs = DeviceArray((len(sample), N), np.float32)
                check(lib().bcsr_dense_gemm_f16(dA.ptr, dB.ptr, Cs.ptr, len(sample), K, N, stream.handle))
                stream.synchronize()
                diff = float(np.abs(C.numpy()[sample] - Cs.numpy()).max(initial=0.0))
                sp_us = None
                if plain is not None:
                    ref = ctypes.c_void_p()
                    check(lib().bcsr_rocsparse_create(plain[0].ptr, plain[1].ptr, plain[2].ptr, h.block_rows,
                                                      h.block_cols, h.nnzb, h.bh, dB.ptr, C.ptr, N,
                                                      ctypes.byref(ref)), "rocsparse bsr")
                    run = lambda: check(lib().bcsr_rocsparse_run(ref, stream.handle))
                    for _ in range(3):
                        run()
                    stream.synchronize()
                    e0, e1 = Event(), Event()
                    e0.record(stream)
                    for _ in range(args.reps):
                        run()
                    e1.record(stream)
                    stream.synchronize()
                    sp_us = e0.elapsed_ms(e1) / args.reps * 1e3
                    lib().bcsr_rocsparse_destroy(ref)
                F = 2.0 * h.nnzb * h.bh * h.bw * N
                Q = algorithmic_bytes(h.nnzb, h.bh, h.bw, M, K, N)
                t_star = max(F / MFMA_F16, Q / HBM)
                t = ms * 1e-3
                row = {"bh": bh, "bw": 16, "pattern": pattern, "block_density": round(density, 5),
                       "density_axis": dens, "N": N, "nnzb": int(h.nnzb), "us": round(ms * 1e3, 3),
                       "effective_tflops": round(F / t / 1e12, 2), "gbs": round(Q / t / 1e9, 1),
                       "bound": "mfma" if F / MFMA_F16 > Q / HBM else "hbm",
                       "roofline_us": round(t_star * 1e6, 3), "roofline_frac": round(t_star / t, 4),
                       "max_abs_diff_sampled_rows_vs_rocblas": diff, "kernel": plan.kernel_name(N),
                       "rocsparse_bsr_us": None if sp_us is None else round(sp_us, 3),
                       "speedup_vs_rocsparse_bsr": None if sp_us is None else round(sp_us / (ms * 1e3), 2)}
                rows.append(row)
                print(json.dumps(row), flush=True)
                C.free()
                Cs.free()
            dA.free()
            for a in plain or []:
                a.free()
            plan.close()
    for _, dB in Bs.values():
        dB.free()
    stream.destroy()
    with open(args.out, "w") as f:
        json.dump({"M": M, "K": K, "rows": rows, "seconds": round(time.time() - t_start, 1)}, f, indent=1)
    if args.md:
        with open(args.md, "w") as f:
            f.write("| block | pattern | block density | N | time us | eff. TFLOPS | GB/s | bound | roofline frac | "
                    "rocSPARSE BSR us (speedup) | kernel |\n")
            f.write("|---|---|---|---|---|---|---|---|---|---|---|\n")
            for r in rows:
                sp = "-" if r["rocsparse_bsr_us"] is None else f"{r['rocsparse_bsr_us']} ({r['speedup_vs_rocsparse_bsr']}x)"
                f.write(f"| {r['bh']}x16 | {r['pattern']} | {r['block_density'] * 100:.2f}% | {r['N']} | {r['us']} | "
                        f"{r['effective_tflops']} | {r['gbs']} | {r['bound']} | {r['roofline_frac']} | {sp} | "
                        f"{r['kernel']} |\n")
    return 0


if __name__ == "__main__":
    sys.exit(main())
