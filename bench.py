#!/usr/bin/env python3
"""Benchmark of the MI355X BCSR x dense engine on one GPU.

Default workload: A = 16384 x 16384 fp16 BCSR with 16 x 16 blocks, 5% of the
blocks nonzero ("blockrandom", seed 123), B = 16384 x 512 fp16 (seed 456),
C fp32 -- the shape the headline metric is quoted on.  A step is one
C = A . B product (one kernel launch) with A, B and C resident in HBM.

Prints one JSON line: effective TFLOPS = 2 * nnzb * bh * bw * N / t, the
dominant kernel's roofline position (algorithmic bytes / time vs 8 TB/s
HBM) and the oracle's CPU product timed on this host as cpu_baseline.
"""
import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "BCSR SpMM effective TFLOPS + % roofline; FP8/batched/learned vs plain FP16"
HBM_PEAK_GBS = 8000.0
MFMA_PEAK_TFLOPS = {"f16": 2500.0, "bf16": 2500.0, "fp8": 5000.0, "hilo": 2500.0}


def parse():
    p = argparse.ArgumentParser(description=__doc__)
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=200)
    p.add_argument("--warmup", type=int, default=500,
                   help="untimed launches first (lets the GPU reach its steady clocks)")
    p.add_argument("--M", type=int, default=16384)
    p.add_argument("--K", type=int, default=16384)
    p.add_argument("--N", type=int, default=512)
    p.add_argument("--density", type=float, default=0.05, help="fraction of nonzero blocks")
    p.add_argument("--bh", type=int, default=16)
    p.add_argument("--seed", type=int, default=123, help="generator seed of A (the reference's 123)")
    p.add_argument("--seed-b", type=int, default=456, help="generator seed of B (the reference's 456)")
    p.add_argument("--pattern", default=None,
                   help="generator pattern (default: blockrandom / pattern32by16 / pattern64by16 by bh)")
    p.add_argument("--tile-n", type=int, default=0)
    p.add_argument("--variant", type=int, default=-1,
                   help="kernel variant (-1 automatic, 0 plain, 1 pipelined, 2 64-bit addressing)")
    p.add_argument("--calib-copy-mb", type=int, default=0,
                   help="after timing, run the plain copy kernel over this many MiB once "
                        "(reference for profiler byte counters)")
    p.add_argument("--xcd-map", type=int, default=-1)
    p.add_argument("--waves-per-block", type=int, default=0)
    p.add_argument("--precision", default="f16", choices=["f16", "bf16", "fp8", "hilo"],
                   help="block storage: fp16 (the reference's), bf16, FP8 E4M3 with per-block "
                        "scales, or hi+lo fp16 pairs")
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--no-extensions", action="store_true",
                   help="skip the precision-mode and batched/learned measurements")
    p.add_argument("--batch-count", type=int, default=96)
    p.add_argument("--cpu-threads", type=int, default=0, help="0 = all host cores (max 16)")
    return p.parse_args()


A_BYTES = {"f16": 2, "bf16": 2, "fp8": 1, "hilo": 4}


def algorithmic_bytes(nnzb, bh, bw, M, K, N, precision="f16"):
    """A values (+ FP8 block scales) + idx + hdr + B + C, each read or written once."""
    scales = 4 * nnzb if precision == "fp8" else 0
    return (nnzb * bh * bw * A_BYTES[precision] + scales + 4 * nnzb + 4 * (M // bh + 1)
            + K * N * 2 + M * N * 4)


def time_ms(stream, fn, warmup=2, reps=10):
    import matrixmultiplicationcuda_amd as mm
    for _ in range(warmup):
        fn()
    stream.synchronize()
    e0, e1 = mm.Event(), mm.Event()
    e0.record(stream)
    for _ in range(reps):
        fn()
    e1.record(stream)
    stream.synchronize()
    return e0.elapsed_ms(e1) / reps


def precision_modes(mm, np, h16, h32, B, N, stream, ref_ms):
    """The target matrix in bf16, FP8 and hi+lo storage (automatic launch)."""
    out = {}
    flops = 2.0 * h16.nnzb * h16.bh * h16.bw * N
    M = h16.block_rows * h16.bh
    for prec in ("bf16", "fp8", "hilo"):
        plan = mm.Plan(h32, prec)
        C = mm.DeviceArray((M, N), np.float32)
        ms = time_ms(stream, lambda: plan.spmm_ptr(B.ptr, C.ptr, N, stream), reps=20)
        out[prec] = {"ms": round(ms, 5), "tflops": round(flops / ms / 1e9, 2),
                     "vs_f16": round(ref_ms / ms, 3), "kernel": plan.kernel_name(N)}
        if prec == "fp8":
            # B quantised once (bcsr_quantize_b_e4m3), then only the product
            # (bcsr_spmm_b8): the cost per product when B is reused
            K = h16.block_cols * h16.bw
            chunks = (N + 255) // 256
            d8 = mm.DeviceArray((K * N,), np.uint8)
            ds = mm.DeviceArray((chunks,), np.uint8)
            work = mm.DeviceArray((K // 16 * chunks,), np.float32)
            mm.check(mm.lib().bcsr_quantize_b_e4m3(B.ptr, K, N, d8.ptr, ds.ptr, work.ptr, stream.handle))
            fn = lambda: mm.check(mm.lib().bcsr_spmm_b8(plan._handle, d8.ptr, ds.ptr, C.ptr, N, None,
                                                        stream.handle))
            ms8 = time_ms(stream, fn, reps=20)
            out["fp8_prequantised_b"] = {"ms": round(ms8, 5), "tflops": round(flops / ms8 / 1e9, 2),
                                         "vs_f16": round(ref_ms / ms8, 3), "kernel": plan.kernel_name(N)}
            for a in (d8, ds, work):
                a.free()
        C.free()
        plan.close()
    return out


def library_rows(mm, np, h, B, N, stream, ref_ms, C_ref):
    """The library comparison rows at the target workload: rocBLAS's dense fp16
    GEMM on the densified A (the reference's cuBLAS rows) and rocSPARSE's BSR
    SpMM on the plain BCSR arrays, each timed like the kernel above and
    checked against its output."""
    M, K = h.block_rows * h.bh, h.block_cols * h.bw
    flops_sparse = 2.0 * h.nnzb * h.bh * h.bw * N
    C = mm.DeviceArray((M, N), np.float32)
    out = {}

    def check(name):
        got = C.numpy().astype(np.float64)
        ref = C_ref.astype(np.float64)
        out[name]["max_abs_diff_vs_bcsr"] = float(np.abs(got - ref).max(initial=0.0))

    # dense: A densified on the host (M x K fp16)
    A = np.zeros((h.block_rows, h.bh, h.block_cols, h.bw), np.float16)
    rows = np.repeat(np.arange(h.block_rows), np.diff(h.hdr))
    A[rows, :, h.idx, :] = h.data.astype(np.float16)
    dA = mm.DeviceArray.from_numpy(A.reshape(M, K))
    del A
    fn = lambda: mm.check(mm.lib().bcsr_dense_gemm_f16(dA.ptr, B.ptr, C.ptr, M, K, N, stream.handle))
    ms = time_ms(stream, fn, warmup=3, reps=20)
    out["rocblas_dense_f16"] = {"ms": round(ms, 5), "dense_tflops": round(2.0 * M * K * N / ms / 1e9, 2),
                                "effective_tflops": round(flops_sparse / ms / 1e9, 2),
                                "bcsr_speedup": round(ms / ref_ms, 3)}
    check("rocblas_dense_f16")
    dA.free()
    # rocSPARSE BSR (square blocks only)
    if h.bh == h.bw:
        dh = mm.DeviceArray.from_numpy(np.ascontiguousarray(h.hdr, np.int32))
        di = mm.DeviceArray.from_numpy(np.ascontiguousarray(h.idx, np.int32))
        dd = mm.DeviceArray.from_numpy(np.ascontiguousarray(h.data, np.float16))
        ref = ctypes.c_void_p()
        mm.check(mm.lib().bcsr_rocsparse_create(dh.ptr, di.ptr, dd.ptr, h.block_rows, h.block_cols,
                                                h.nnzb, h.bh, B.ptr, C.ptr, N, ctypes.byref(ref)),
                 "rocsparse bsr")
        fn = lambda: mm.check(mm.lib().bcsr_rocsparse_run(ref, stream.handle))
        ms = time_ms(stream, fn, warmup=3, reps=20)
        out["rocsparse_bsr_f16"] = {"ms": round(ms, 5), "effective_tflops": round(flops_sparse / ms / 1e9, 2),
                                    "bcsr_speedup": round(ms / ref_ms, 3)}
        check("rocsparse_bsr_f16")
        mm.lib().bcsr_rocsparse_destroy(ref)
        for a in (dh, di, dd):
            a.free()
    C.free()
    return out


# The reference's published medians (ms) of its legacy "SpMM with Tensors"
# kernel: square n x n A with 16 x 16 blocks ("blockrandom", fraction of zero
# blocks as given), B n x n, N = n; column time_med of avg_results.csv:86-113
# (hardware not recorded by the reference).
PUBLISHED_SPMM_WITH_TENSORS_MS = {
    5376: {0.7: 17.9727365, 0.8: 11.217856, 0.4: 36.245167, 0.9: 5.721632, 0.6: 26.1274395,
           0.5: 31.784544, 0.3: 40.057983},
    1280: {0.7: 0.192896, 0.8: 0.150608, 0.4: 0.314752, 0.9: 0.111984, 0.6: 0.235488,
           0.5: 0.275552, 0.3: 0.356448},
    4096: {0.7: 4.937264, 0.8: 3.390464, 0.4: 13.035104, 0.9: 1.917968, 0.6: 6.68472,
           0.5: 9.60736, 0.3: 15.07904},
    8192: {0.7: 80.295952, 0.8: 52.0944155, 0.4: 158.6035995, 0.9: 24.801424, 0.6: 105.1238405,
           0.5: 130.032112, 0.3: 252.054352},
}


def published_shapes(mm, np, stream):
    """This library's fp16 kernel on the reference's published shapes, each
    checked against rocBLAS's dense product of the same operands."""
    rows = []
    for n, by_sp in sorted(PUBLISHED_SPMM_WITH_TENSORS_MS.items()):
        B16 = mm.generate(n, n, 0.0, "random", bh=16, seed=456)
        dB = mm.DeviceArray.from_numpy(B16)
        del B16
        C = mm.DeviceArray((n, n), np.float32)
        Cd = mm.DeviceArray((n, n), np.float32)
        for sp, ref_ms in sorted(by_sp.items()):
            h32 = mm.generate_bcsr(n, n, sp, "blockrandom", bh=16, bw=16, seed=123)
            h = mm.HostBCSR(h32.hdr, h32.idx, h32.data.astype(np.float16), h32.block_rows, h32.block_cols,
                            16, 16)
            del h32
            plan = mm.Plan(h, "f16")
            reps = 20 if n <= 4096 else 5
            ms = time_ms(stream, lambda: plan.spmm_ptr(dB.ptr, C.ptr, n, stream), warmup=2, reps=reps)
            A = np.zeros((h.block_rows, 16, h.block_cols, 16), np.float16)
            A[np.repeat(np.arange(h.block_rows), np.diff(h.hdr)), :, h.idx, :] = h.data
            dA = mm.DeviceArray.from_numpy(A.reshape(n, n))
            del A
            mm.check(mm.lib().bcsr_dense_gemm_f16(dA.ptr, dB.ptr, Cd.ptr, n, n, n, stream.handle))
            stream.synchronize()
            diff = float(np.abs(C.numpy() - Cd.numpy()).max(initial=0.0))
            dA.free()
            flops = 2.0 * h.nnzb * 256 * n
            rows.append({"n": n, "zero_block_fraction": sp, "nnzb": int(h.nnzb), "ms": round(ms, 5),
                         "effective_tflops": round(flops / ms / 1e9, 2),
                         "dense_equivalent_tflops": round(2.0 * n ** 3 / ms / 1e9, 2),
                         "reference_ms": ref_ms, "speedup_vs_reference": round(ref_ms / ms, 2),
                         "max_abs_diff_vs_rocblas": diff, "kernel": plan.kernel_name(n)})
            plan.close()
        for a in (dB, C, Cd):
            a.free()
    sp_list = [r["speedup_vs_reference"] for r in rows]
    return {"workload": "square n x n, 16 x 16 blockrandom blocks, N = n (the reference's published runs, "
                        "avg_results.csv:86-113, its hardware not recorded)",
            "geomean_speedup_vs_reference": round(float(np.exp(np.mean(np.log(sp_list)))), 2),
            "min_speedup_vs_reference": min(sp_list), "rows": rows}


def batched(mm, np, count, stream):
    """The batched uneven workload: plain fp16 per-matrix launches (the
    library's static choice each) against the batched persistent kernel
    (graph replays), fp16 and FP8.  The batch's tile widths come from the
    autotuner, which measures the policies -- the batch's automatic widths,
    the learned dispatcher's per-matrix picks and (FP8) each matrix's own
    exhaustive autotune -- each with its grid/item-size sweep."""
    import concurrent.futures as cf

    from matrixmultiplicationcuda_amd import autotune, learned, workloads
    probs = workloads.draw_problems(count, base_seed=1000)
    with cf.ThreadPoolExecutor(max_workers=min(12, os.cpu_count() or 1)) as ex:
        built = list(ex.map(lambda p: (p.bcsr(np.float32), p.B()), probs))
    h32s = [b[0] for b in built]
    h16s = [mm.HostBCSR(h.hdr, h.idx, h.data.astype(np.float16), h.block_rows, h.block_cols,
                        h.bh, h.bw) for h in h32s]
    Ns = [p.N for p in probs]
    flops = sum(workloads.flops(h, p.N) for h, p in zip(h16s, probs))
    Bs = [mm.DeviceArray.from_numpy(b[1]) for b in built]
    Cs = [mm.DeviceArray((p.rows, p.N), np.float32) for p in probs]
    res = {"workload": f"{count} problems: rows/cols log-uniform 512..16384 (x64), blocks "
                       "16/32/64 x 16, block density U(0.5%, 25%), blockrandom_rho U(0.3, 1), "
                       "N in {32, 128, 512}; seeds 1000 + i",
           "gflop": round(flops / 1e9, 3)}
    out = {}
    for prec, hosts in (("f16", h16s), ("fp8", h32s)):
        plans = [mm.Plan(h, prec) for h in hosts]

        def seq():
            for pl, B, C in zip(plans, Bs, Cs):
                pl.spmm_ptr(B.ptr, C.ptr, B.shape[1], stream)
        t_seq = time_ms(stream, seq, reps=5)
        disp = learned.Dispatcher(prec)
        policies = {"auto": None, "learned": disp.batch_tiles(plans, h16s, Ns)}
        # per-matrix launches with the learned dispatcher's configurations
        cfgs = [autotune.cfg_of(disp.choose(pl, h, N)[0]) for pl, h, N in zip(plans, h16s, Ns)]
        disp.close()

        def seq_learned():
            for pl, B, C, cfg in zip(plans, Bs, Cs, cfgs):
                pl.spmm_ptr(B.ptr, C.ptr, B.shape[1], stream, cfg)
        t_seq_l = time_ms(stream, seq_learned, reps=5)
        if prec == "fp8":
            # per-problem exhaustive autotune of the tiles (each matrix alone)
            configs = autotune.configs_for("fp8")
            times = [autotune.time_configs(p, B, C, B.shape[1], configs, stream, reps=3)
                     for p, B, C in zip(plans, Bs, Cs)]
            policies["per_matrix_autotuned"] = autotune.batch_tiles_from_times(
                plans, Ns, times, configs, 2)
        best, table = autotune.tune_batch_policies(plans, Bs, Cs, policies, stream)
        batch = mm.Batch(plans, Bs, Cs, tiles=best["tiles"], waves_per_cu=best["waves_per_cu"],
                         bin_work=best["bin_work"])
        t_b = batch.time(stream, graph=True, warmup=3, reps=20)
        entry = {"per_matrix_ms": round(t_seq, 4), "per_matrix_tflops": round(flops / t_seq / 1e9, 2),
                 "per_matrix_learned_ms": round(t_seq_l, 4),
                 "batched_ms": round(t_b, 4), "batched_tflops": round(flops / t_b / 1e9, 2),
                 "tile_policy": best["policy"], "waves_per_cu": best["waves_per_cu"],
                 "bin_work": best["bin_work"],
                 "policies_ms": {k: round(v["ms"], 4) for k, v in table.items()}}
        if prec == "fp8":
            # error of the batched FP8 results against the FP64 product (3 problems)
            batch.run(stream)
            stream.synchronize()
            errs = []
            for i in sorted(range(count), key=lambda i: probs[i].rows * probs[i].N)[:3]:
                h = h32s[i]
                Cg = Cs[i].numpy().astype(np.float64)
                # FP64 product of the generator's values (dense, numpy)
                A = np.zeros((h.block_rows, h.bh, h.block_cols, h.bw), np.float64)
                A[np.repeat(np.arange(h.block_rows), np.diff(h.hdr)), :, h.idx, :] = h.data
                A = A.reshape(h.block_rows * h.bh, h.block_cols * h.bw)
                B64 = built[i][1].astype(np.float64)
                C64, Ca = A @ B64, np.abs(A) @ np.abs(B64)
                nz = Ca > 0
                e = np.abs(Cg - C64)[nz] / Ca[nz]
                errs.append((float(e.max()), float(e.mean())))
            entry["error_vs_fp64_per_entry"] = {"max": max(e[0] for e in errs),
                                                "mean": float(np.mean([e[1] for e in errs])),
                                                "problems": 3}
        batch.close()
        for pl in plans:
            pl.close()
        out[prec] = entry
    for a in Bs + Cs:
        a.free()
    res.update(out)
    res["speedup_fp8_batched_vs_fp16_per_matrix"] = round(
        out["f16"]["per_matrix_ms"] / out["fp8"]["batched_ms"], 3)
    return res


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU; sharding over several GPUs (by "
                                   "block-rows, or by columns of B and C) is not launched",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    import numpy as np

    import matrixmultiplicationcuda_amd as mm

    pattern = args.pattern or {16: "blockrandom", 32: "pattern32by16",
                               64: "pattern64by16"}[args.bh]
    bh, bw = args.bh, 16
    M, K, N = args.M, args.K, args.N
    t0 = time.time()
    # the BCSR arrays of generate_matrix(M, K, 1 - density, pattern, bh, 123)
    # rounded to fp16, produced block by block (no dense M x K matrix)
    h32 = mm.generate_bcsr(M, K, 1.0 - args.density, pattern, bh=bh, bw=bw, seed=args.seed)
    h = mm.HostBCSR(h32.hdr, h32.idx, h32.data.astype(np.float16), h32.block_rows,
                    h32.block_cols, bh, bw)
    B16 = mm.generate(K, N, 0.0, "random", bh=bh, seed=args.seed_b)
    mm.check(mm.lib().bcsr_set_device(0))
    # fp16 storage takes the fp16 matrix (as the reference does); the other
    # modes quantise the generator's fp32 values
    plan = mm.Plan(h if args.precision == "f16" else h32, args.precision)
    t_setup = time.time() - t0
    print(f"# setup {t_setup:.1f}s nnzb={h.nnzb}", file=sys.stderr, flush=True)

    B = mm.DeviceArray.from_numpy(B16)
    C = mm.DeviceArray((M, N), np.float32)
    cfg = mm.make_cfg(tile_n=args.tile_n, xcd_map=args.xcd_map, variant=args.variant,
                      waves_per_block=args.waves_per_block)
    kname = plan.kernel_name(N, cfg)
    stream = mm.Stream()
    plan.spmm_repeat_ptr(B.ptr, C.ptr, N, args.warmup, stream, cfg)
    stream.synchronize()
    mm.synchronize()
    start, end = mm.Event(), mm.Event()
    start.record(stream)
    plan.spmm_repeat_ptr(B.ptr, C.ptr, N, args.steps, stream, cfg)
    end.record(stream)
    stream.synchronize()
    mm.synchronize()
    ms = start.elapsed_ms(end) / max(args.steps, 1)
    t = ms * 1e-3
    print(f"# timed {args.steps} steps: {ms * 1e3:.2f} us/step ({kname})", file=sys.stderr,
          flush=True)
    if args.calib_copy_mb:
        nb = args.calib_copy_mb << 20
        src, dst = mm.DeviceArray((nb,), np.uint8), mm.DeviceArray((nb,), np.uint8)
        mm.check(mm.lib().bcsr_copy_kernel(dst.ptr, src.ptr, nb, stream.handle), "copy kernel")
        stream.synchronize()
        src.free()
        dst.free()

    flops = 2.0 * h.nnzb * bh * bw * N
    q = algorithmic_bytes(h.nnzb, bh, bw, M, K, N, args.precision)
    tflops = flops / t / 1e12
    gbs = q / t / 1e9
    t_star = max(q / (HBM_PEAK_GBS * 1e9), flops / (MFMA_PEAK_TFLOPS[args.precision] * 1e12))
    bound = "hbm" if q / (HBM_PEAK_GBS * 1e9) >= flops / (MFMA_PEAK_TFLOPS[args.precision] * 1e12) else "mfma"
    if bound == "hbm":
        roof = {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": None}
    else:
        pk = MFMA_PEAK_TFLOPS[args.precision]
        roof = {"bound": "mfma", "achieved": round(tflops, 2), "peak": pk, "unit": "TFLOP/s",
                "frac": round(tflops / pk, 4), "traffic": None}
    # the matrix cores' share of their dense peak at this throughput
    roof["mfma_utilisation"] = round(tflops / MFMA_PEAK_TFLOPS[args.precision], 4)
    roof["kernel"] = kname
    roof["roofline_time_us"] = round(t_star * 1e6, 3)
    roof["algorithmic_bytes"] = int(q)
    traffic_file = os.path.join(ROOT, "profiles", "traffic.json")
    wl_key = f"{M}x{K}_bh{bh}_d{args.density}_N{N}_{args.precision}"
    if os.path.exists(traffic_file):
        with open(traffic_file) as f:
            tr = json.load(f).get(wl_key)
        if tr and tr.get("kernel") == kname:
            roof["traffic"] = tr["bytes_per_launch"]
            # the roofline at the arithmetic intensity the counters measured
            # (flops over the HBM bytes a launch really moves), next to the
            # algorithmic one above
            ai = flops / tr["bytes_per_launch"]
            ceiling = min(MFMA_PEAK_TFLOPS[args.precision], ai * HBM_PEAK_GBS / 1e3)
            roof["at_measured_traffic"] = {"arith_intensity": round(ai, 2), "ceiling_tflops": round(ceiling, 1),
                                           "frac": round(tflops / ceiling, 4)}

    # correctness of this very run against the CPU reference product
    C_gpu = C.numpy()
    cpu = None
    if not args.no_cpu_baseline:
        import oracle

        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        Bf = B16.astype(np.float32)
        data16 = h.data.view(np.uint16)
        times = []
        C_cpu = None
        for rep in range(3):
            c0 = time.perf_counter()
            C_cpu = oracle.bcsr_spmm_f32(h.hdr, h.idx, data16, bh, bw, Bf, threads=threads, C=C_cpu)
            times.append(time.perf_counter() - c0)
            if sum(times) > 20.0:
                break
        tc = float(np.median(times))
        # the same loop on one core (one run)
        c0 = time.perf_counter()
        C_one = oracle.bcsr_spmm_f32(h.hdr, h.idx, data16, bh, bw, Bf, threads=1)
        t1 = time.perf_counter() - c0
        assert np.array_equal(C_one, C_cpu), "the oracle's product depends on its thread count"
        del C_one
        C64, Cabs = oracle.bcsr_spmm_f64(h.hdr, h.idx, h.data.astype(np.float32), bh, bw, Bf,
                                         threads=threads)
        err = np.abs(C_gpu.astype(np.float64) - C_cpu)
        nz = Cabs > 0
        rel = float((err[nz] / Cabs[nz]).max(initial=0.0))
        exact_zero = bool(np.all(C_gpu[~nz] == 0))
        # FP64 products: of the stored fp16 values (the kernel's arithmetic
        # error) and of the generator's values before fp16 rounding (storage
        # included)
        rel64 = float((np.abs(C_gpu - C64)[nz] / Cabs[nz]).max(initial=0.0))
        del C64
        C64g, Cabsg = oracle.bcsr_spmm_f64(h32.hdr, h32.idx, h32.data, bh, bw, Bf, threads=threads)
        nzg = Cabsg > 0
        rel64g = float((np.abs(C_gpu - C64g)[nzg] / Cabsg[nzg]).max(initial=0.0))
        mean64g = float((np.abs(C_gpu - C64g)[nzg] / Cabsg[nzg]).mean()) if nzg.any() else 0.0
        del C64g, Cabsg
        # fp16 storage is held to 1e-3 of |A||B| per entry; the other modes
        # store different values and are held to their own bounds in the tests
        bound = {"f16": 1e-3, "bf16": 2e-2, "hilo": 1e-3, "fp8": 0.25}[args.precision]
        cpu = {"value": round(flops / tc / 1e12, 5), "unit": "TFLOPS", "cores": threads,
               "kind": "port",
               "sample": f"full product (all {h.block_rows} block-rows), oracle BCSR-order fp32 "
                         f"loop over {threads} threads, median of {len(times)} runs "
                         f"({tc:.3f} s each)",
               "seconds": round(tc, 4),
               "single_core": {"value": round(flops / t1 / 1e12, 5), "unit": "TFLOPS", "cores": 1,
                               "seconds": round(t1, 4)},
               "gpu_vs_cpu": {"max_err_over_absAabsB": rel, "bound": bound,
                              "exact_zero_where_absAabsB_zero": exact_zero,
                              "ok": rel <= bound and exact_zero},
               "gpu_vs_fp64": {"max_err_over_absAabsB_stored_values": rel64,
                               "max_err_over_absAabsB_generator_values": rel64g,
                               "mean_err_over_absAabsB_generator_values": mean64g}}

    extensions = None
    if not args.no_extensions and args.precision == "f16":
        t1 = time.time()
        extensions = {"precision_modes_at_this_workload":
                      precision_modes(mm, np, h, h32, B, N, stream, ms)}
        extensions["library_rows_at_this_workload"] = library_rows(mm, np, h, B, N, stream, ms, C_gpu)
        extensions["reference_published_shapes"] = published_shapes(mm, np, stream)
        extensions["batched"] = batched(mm, np, args.batch_count, stream)
        print(f"# extensions {time.time() - t1:.1f}s", file=sys.stderr, flush=True)

    out = {
        "metric": METRIC,
        "value": round(tflops, 3),
        "unit": "TFLOPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 6),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": args.precision,
        "data": f"synthetic (reference generator, seeds {args.seed}/{args.seed_b})",
        "config": {"workload": f"bcsr_spmm {M}x{K} {bh}x{bw} blocks {pattern} "
                               f"{args.density * 100:g}% blocks N={N} {args.precision}",
                   "M": M, "K": K, "N": N, "bh": bh, "bw": bw, "pattern": pattern,
                   "block_density": args.density, "nnzb": int(h.nnzb), "seed_a": args.seed,
                   "seed_b": args.seed_b,
                   "tile_n": args.tile_n, "xcd_map": args.xcd_map, "variant": args.variant,
                   "waves_per_block": args.waves_per_block,
                   "parallelism": "single-gpu"},
        "gbs": round(gbs, 1),
        "roofline": roof,
        "cpu_baseline": cpu,
        "extensions": extensions,
    }
    print(json.dumps(out), flush=True)
    stream.destroy()
    B.free()
    C.free()
    plan.close()


if __name__ == "__main__":
    main()
