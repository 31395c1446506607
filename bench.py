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
    p.add_argument("--cpu-threads", type=int, default=0,
                   help="0 = every CPU this process may run on (its affinity set)")
    p.add_argument("--timing", default="cold", choices=["cold", "warm"],
                   help="the headline: cold = before every timed launch, outside its events, a "
                        "copy over 2 x 512 MiB evicts A, B and C from L2 and the 256 MB Infinity "
                        "Cache (the reference's NVBench cold sampling); warm = back-to-back "
                        "launches.  Both are measured and reported; this picks `value`.")
    p.add_argument("--flush-mb", type=int, default=512)
    return p.parse_args()


A_BYTES = {"f16": 2, "bf16": 2, "fp8": 1, "hilo": 4}


def algorithmic_bytes(nnzb, bh, bw, M, K, N, precision="f16"):
    """A values (+ FP8 block scales) + idx + hdr + B + C, each read or written once."""
    scales = 4 * nnzb if precision == "fp8" else 0
    return (nnzb * bh * bw * A_BYTES[precision] + scales + 4 * nnzb + 4 * (M // bh + 1)
            + K * N * 2 + M * N * 4)


def timed_launches(mm, np, fn, stream, steps, flush=None):
    """Per-launch milliseconds of `steps` launches of fn (HIP events around
    each launch on its stream).  flush(), when given, runs on the same stream
    before each launch, outside that launch's events.  Returns the list of
    times and the wall-clock seconds of the whole loop (between two
    synchronisations)."""
    ev = [(mm.Event(), mm.Event()) for _ in range(steps)]
    stream.synchronize()
    mm.synchronize()
    w0 = time.perf_counter()
    for e0, e1 in ev:
        if flush is not None:
            flush()
        e0.record(stream)
        fn()
        e1.record(stream)
    stream.synchronize()
    mm.synchronize()
    wall = time.perf_counter() - w0
    ms = [e0.elapsed_ms(e1) for e0, e1 in ev]
    for e0, e1 in ev:
        e0.destroy()
        e1.destroy()
    return ms, wall


def host_cpu():
    """CPU model, logical CPUs, the CPUs this process may run on and the
    cgroup CPU quota (None if unlimited or unreadable)."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    quota = None
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            q, period = f.read().split()
            if q != "max":
                quota = round(int(q) / int(period), 2)
    except (OSError, ValueError):
        pass
    return {"model": model, "logical_cpus": os.cpu_count(),
            "affinity_cpus": len(os.sched_getaffinity(0)), "cgroup_cpu_quota": quota}


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


def fp64_reference(np, h, B16):
    """(C, |A||B|) in float64 by scipy's BSR product (an independent check,
    not the oracle)."""
    import scipy.sparse as sps
    M, K = h.block_rows * h.bh, h.block_cols * h.bw
    d = h.data.astype(np.float64)
    Bd = B16.astype(np.float64)
    A = sps.bsr_matrix((d, h.idx, h.hdr), shape=(M, K))
    Aa = sps.bsr_matrix((np.abs(d), h.idx, h.hdr), shape=(M, K))
    return np.asarray(A @ Bd), np.asarray(Aa @ np.abs(Bd))


def error_vs(np, C, C64, Cabs):
    nz = Cabs > 0
    rel = np.abs(C.astype(np.float64) - C64)[nz] / Cabs[nz]
    return {"max": float(rel.max(initial=0.0)), "mean": float(rel.mean()) if rel.size else 0.0,
            "normwise": float(np.linalg.norm(C - C64) / max(np.linalg.norm(C64), 1e-300))}


def precision_modes(mm, np, N, stream, ref, flush, steps=50):
    """The precision workload: 16384^2, 16 x 16 blocks, 5 % ("blockrandom",
    seed 321), block magnitudes log-uniform on [1e-3, 1e2] with random sign,
    B = U(-1, 1) fp16 (seed 456), N = 512; fp16, bf16, hi+lo and FP8 storage.
    Each mode: cold and warm median launch times, TFLOPS, speed over the
    fp16 headline, and the error of its output against the FP64 product of
    the unquantised values (max and mean of |C - C64| / (|A||B|) per entry,
    and normwise ||C - C64|| / ||C64||)."""
    M = K = 16384
    h32 = mm.generate_bcsr(M, K, 0.95, "blockrandom", bh=16, bw=16, seed=321, value_dist="loguniform",
                           lo=1e-3, hi=1e2)
    B16 = mm.generate(K, N, 0.0, "random", bh=16, seed=456)
    C64, Cabs = fp64_reference(np, h32, B16)
    B = mm.DeviceArray.from_numpy(B16)
    C = mm.DeviceArray((M, N), np.float32)
    flops = 2.0 * h32.nnzb * 256 * N
    out = {"workload": "16384^2, 16 x 16 blocks, 5 % blockrandom (seed 321), |a| log-uniform on "
                       "[1e-3, 1e2] with random sign, B U(-1, 1) fp16 (seed 456), N = " + str(N),
           "nnzb": int(h32.nnzb), "fp64_reference": "scipy.sparse BSR product in float64"}
    for prec in ("f16", "bf16", "fp8", "hilo"):
        plan = mm.Plan(h32, prec)
        fn = lambda: plan.spmm_ptr(B.ptr, C.ptr, N, stream)
        fn()
        cold, _ = timed_launches(mm, np, fn, stream, steps, flush=flush)
        warm, _ = timed_launches(mm, np, fn, stream, steps)
        got = C.numpy()
        cm, wm = float(np.median(cold)), float(np.median(warm))
        out[prec] = {"cold_ms": round(cm, 5), "warm_ms": round(wm, 5),
                     "tflops_cold": round(flops / cm / 1e9, 2), "tflops_warm": round(flops / wm / 1e9, 2),
                     "vs_f16_headline_cold": round(ref["cold"] / cm, 3),
                     "vs_f16_headline_warm": round(ref["warm"] / wm, 3),
                     "kernel": plan.kernel_name(N), "error_vs_fp64": error_vs(np, got, C64, Cabs)}
        if prec == "fp8":
            # B quantised once (bcsr_quantize_b_e4m3), then only the product
            # (bcsr_spmm_b8): the cost per product when B is reused
            d8 = mm.DeviceArray((K * N,), np.uint8)
            ds = mm.DeviceArray((mm.lib().bcsr_b8_scale_bytes(K, N),), np.uint8)
            mm.check(mm.lib().bcsr_quantize_b_e4m3(B.ptr, K, N, d8.ptr, ds.ptr, None, stream.handle))
            fn8 = lambda: mm.check(mm.lib().bcsr_spmm_b8(plan._handle, d8.ptr, ds.ptr, C.ptr, N, None,
                                                         stream.handle))
            fn8()
            cold8, _ = timed_launches(mm, np, fn8, stream, steps, flush=flush)
            warm8, _ = timed_launches(mm, np, fn8, stream, steps)
            c8, w8 = float(np.median(cold8)), float(np.median(warm8))
            out["fp8_prequantised_b"] = {"cold_ms": round(c8, 5), "warm_ms": round(w8, 5),
                                         "tflops_cold": round(flops / c8 / 1e9, 2),
                                         "vs_f16_headline_cold": round(ref["cold"] / c8, 3),
                                         "vs_f16_headline_warm": round(ref["warm"] / w8, 3)}
            for a in (d8, ds):
                a.free()
        plan.close()
    B.free()
    C.free()
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


def batched(mm, np, count, stream, headline_tflops):
    """The batched uneven workload: plain fp16 per-matrix launches (the
    library's static choice each) against the batched persistent kernel
    (graph replays), fp16 and FP8.  The batch's tile widths come from the
    autotuner, which measures the policies -- the batch's automatic widths,
    the learned dispatcher's per-matrix picks and each matrix's own
    exhaustive autotune -- each with its grid/item-size sweep.  Per matrix,
    the learned dispatcher's live regret against exhaustive autotuning and
    its gain over the library's automatic choice are measured on the same
    problems."""
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
        # each problem on its own, as the timing database records a matrix:
        # every configuration and the library's automatic choice timed on
        # this box (exhaustive autotune = the fastest of them); the learned
        # dispatcher's pick is read from the same table (regret 0 where it
        # picks the fastest), as the held-out evaluation does
        configs = autotune.configs_for(prec)
        times, t_auto = [], []
        for p, B, C in zip(plans, Bs, Cs):  # a problem's configurations and automatic choice together
            times.append(autotune.time_configs(p, B, C, B.shape[1], configs, stream, reps=5))
            t_auto.append(autotune.time_auto(p, B, C, B.shape[1], stream, reps=5))
        times, t_auto = np.array(times), np.array(t_auto)
        t_best = np.fmin(np.nanmin(times, axis=1), t_auto)
        pc, pa = disp.predict(list(zip(h16s, Ns)))
        masks = np.array([autotune.supported_mask(p, N, configs) for p, N in zip(plans, Ns)], np.float32)
        rep = learned.regret_report(pc if pa is None else np.concatenate([pc, pa[:, None]], axis=1),
                                    masks, times, t_auto, margin=disp.margin)
        picks = [disp.choose(pl, h, N, B, C, stream) for pl, h, N, B, C in zip(plans, h16s, Ns, Bs, Cs)]
        cfgs = [autotune.cfg_of(c) for c, _ in picks]
        t_pick = np.array([t_auto[i] if c is None else times[i, configs.index(tuple(c))]
                           for i, (c, _) in enumerate(picks)])
        live = {"problems": count, "configs": len(configs),
                "learned_mean_regret": round(rep["dispatch_mean_regret"], 4),
                "learned_p90_regret": round(float(np.percentile(t_pick / t_best - 1.0, 90)), 4),
                "static_mean_regret": round(rep["static_mean_regret"], 4),
                "static_p90_regret": round(rep["static_p90_regret"], 4),
                "autotune_fallbacks": int(sum(fb for _, fb in picks)),
                "kept_automatic": int(sum(c is None for c, _ in picks)),
                "margin": disp.margin,
                "learned_gain_over_static_sum_ms": round(float(t_auto.sum() / t_pick.sum()), 4),
                "model_trained_on": disp.meta.get("db"), "model_configs": len(disp.configs)}
        disp.close()
        policies["per_matrix_autotuned"] = autotune.batch_tiles_from_times(
            plans, Ns, times, configs, mm.PRECISIONS[prec])

        def seq_learned():
            for pl, B, C, cfg in zip(plans, Bs, Cs, cfgs):
                pl.spmm_ptr(B.ptr, C.ptr, B.shape[1], stream, cfg)
        t_seq_l = time_ms(stream, seq_learned, reps=5)
        best, table = autotune.tune_batch_policies(plans, Bs, Cs, policies, stream)
        batch = mm.Batch(plans, Bs, Cs, tiles=best["tiles"], waves_per_cu=best["waves_per_cu"],
                         bin_work=best["bin_work"])
        t_b = batch.time(stream, graph=True, warmup=3, reps=20)
        entry = {"per_matrix_ms": round(t_seq, 4), "per_matrix_tflops": round(flops / t_seq / 1e9, 2),
                 "per_matrix_learned_ms": round(t_seq_l, 4),
                 "batched_ms": round(t_b, 4), "batched_tflops": round(flops / t_b / 1e9, 2),
                 "tile_policy": best["policy"], "waves_per_cu": best["waves_per_cu"],
                 "bin_work": best["bin_work"],
                 "policies_ms": {k: round(v["ms"], 4) for k, v in table.items()},
                 "per_matrix_dispatch_live": live}
        # error of every batched result against the FP64 product of the
        # generator's values (scipy's BSR product in float64)
        batch.run(stream)
        stream.synchronize()
        errs = [error_vs(np, Cs[i].numpy(), *fp64_reference(np, h32s[i], built[i][1]))
                for i in range(count)]
        entry["error_vs_fp64"] = {"max": max(e["max"] for e in errs),
                                  "mean_of_means": float(np.mean([e["mean"] for e in errs])),
                                  "max_normwise": max(e["normwise"] for e in errs),
                                  "problems": count}
        batch.close()
        for pl in plans:
            pl.close()
        out[prec] = entry
    for a in Bs + Cs:
        a.free()
    res.update(out)
    f16, f8 = out["f16"], out["fp8"]
    res["speedup_fp8_batched_vs_fp16_per_matrix"] = round(f16["per_matrix_ms"] / f8["batched_ms"], 3)
    # FP8 + batched + learned TFLOPS over the plain fp16 kernel's headline
    # TFLOPS, and the factors that ratio is made of on this workload
    res["fp8_batched_over_fp16_headline"] = {
        "fp8_batched_tflops_over_fp16_headline_cold": round(f8["batched_tflops"] / headline_tflops["cold"], 3),
        "fp8_batched_tflops_over_fp16_headline_warm": round(f8["batched_tflops"] / headline_tflops["warm"], 3),
        "factors_on_this_workload": {
            "batching_fp16": round(f16["per_matrix_ms"] / f16["batched_ms"], 3),
            "fp8_over_fp16_both_batched": round(f16["batched_ms"] / f8["batched_ms"], 3),
            "learned_tiles_over_automatic_fp16": round(f16["policies_ms"]["auto"] / f16["policies_ms"]["learned"], 3),
            "learned_tiles_over_automatic_fp8": round(f8["policies_ms"]["auto"] / f8["policies_ms"]["learned"], 3),
            "per_matrix_learned_over_static_fp16": round(f16["per_matrix_ms"] / f16["per_matrix_learned_ms"], 3),
            "per_matrix_learned_over_static_fp8": round(f8["per_matrix_ms"] / f8["per_matrix_learned_ms"], 3)},
        "policy_used": {"f16": f16["tile_policy"], "fp8": f8["tile_policy"]}}
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
    launch = lambda: plan.spmm_ptr(B.ptr, C.ptr, N, stream, cfg)
    # cold: a copy over 2 x flush-mb MiB between launches (outside their
    # events) leaves none of A, B, C in L2 or the Infinity Cache
    nb = args.flush_mb << 20
    fsrc, fdst = mm.DeviceArray((nb,), np.uint8), mm.DeviceArray((nb,), np.uint8)
    flush = lambda: mm.check(mm.lib().bcsr_copy_kernel(fdst.ptr, fsrc.ptr, nb, stream.handle), "flush")
    cold_ms, cold_wall = timed_launches(mm, np, launch, stream, args.steps, flush=flush)
    warm_ms, warm_wall = timed_launches(mm, np, launch, stream, args.steps)
    fsrc.free()
    fdst.free()
    timing = {"cold": {"median_ms": float(np.median(cold_ms)), "mean_ms": float(np.mean(cold_ms)),
                       "min_ms": float(np.min(cold_ms)), "max_ms": float(np.max(cold_ms)),
                       "wall_s_incl_flushes": round(cold_wall, 4),
                       "flush": f"copy of {args.flush_mb} MiB to a second {args.flush_mb} MiB "
                                f"buffer before each launch, outside its events"},
              "warm": {"median_ms": float(np.median(warm_ms)), "mean_ms": float(np.mean(warm_ms)),
                       "min_ms": float(np.min(warm_ms)), "max_ms": float(np.max(warm_ms)),
                       "wall_s": round(warm_wall, 4),
                       "note": "back-to-back launches: A, B and C stay in the 256 MB Infinity Cache"}}
    ms = timing[args.timing]["median_ms"]
    t = ms * 1e-3
    print(f"# timed {args.steps} steps: cold median {timing['cold']['median_ms'] * 1e3:.2f} us, "
          f"warm median {timing['warm']['median_ms'] * 1e3:.2f} us ({kname})", file=sys.stderr,
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
        other = {k: round(q / (v["median_ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4) for k, v in timing.items()}
    else:
        pk = MFMA_PEAK_TFLOPS[args.precision]
        roof = {"bound": "mfma", "achieved": round(tflops, 2), "peak": pk, "unit": "TFLOP/s",
                "frac": round(tflops / pk, 4), "traffic": None}
        other = {k: round(flops / (v["median_ms"] * 1e-3) / 1e12 / pk, 4) for k, v in timing.items()}
    roof["timing"] = args.timing
    roof["frac_cold"], roof["frac_warm"] = other["cold"], other["warm"]
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
            # the launch the headline times (cold or warm) where both were counted
            tb = tr.get("bytes_per_launch_" + args.timing, tr["bytes_per_launch"])
            roof["traffic"] = tb
            # the roofline at the arithmetic intensity the counters measured
            # (flops over the HBM bytes a launch really moves), next to the
            # algorithmic one above
            ai = flops / tb
            ceiling = min(MFMA_PEAK_TFLOPS[args.precision], ai * HBM_PEAK_GBS / 1e3)
            roof["at_measured_traffic"] = {"arith_intensity": round(ai, 2), "ceiling_tflops": round(ceiling, 1),
                                           "frac": round(tflops / ceiling, 4)}

    # correctness of this very run against the CPU reference product
    C_gpu = C.numpy()
    cpu = None
    if not args.no_cpu_baseline:
        import oracle

        host = host_cpu()
        threads = args.cpu_threads or host["affinity_cpus"]
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
        # the reference's literal dense cpu_matmul_ref (nvbench_main.cu:78-91),
        # one thread, on the small case (1024^2, 10 % of 16 x 16 blocks,
        # N = 256): O(M K N), so only there
        A0 = mm.generate(1024, 1024, 0.9, "blockrandom", bh=16, seed=123)
        B0 = mm.generate(1024, 256, 0.0, "random", bh=16, seed=456)
        h0 = mm.bcsr_from_dense(A0, 16, 16)
        c0 = time.perf_counter()
        C0 = oracle.matmul_dense_ref(A0.view(np.uint16), B0.view(np.uint16))
        td = time.perf_counter() - c0
        c0 = time.perf_counter()
        C0b = oracle.bcsr_spmm_f32(h0.hdr, h0.idx, h0.data.view(np.uint16), 16, 16, B0.astype(np.float32),
                                   threads=1)
        tb = time.perf_counter() - c0
        assert np.array_equal(C0.astype(np.float32), C0b), "dense and BCSR-order CPU products differ"
        f0 = 2.0 * h0.nnzb * 256 * 256
        dense_small = {"workload": "1024 x 1024, 10 % of 16 x 16 blocks (nnzb "
                                   f"{h0.nnzb}), N = 256, seeds 123 / 456",
                       "dense_literal": {"seconds": round(td, 4), "cores": 1,
                                         "effective_tflops": round(f0 / td / 1e12, 6),
                                         "dense_equivalent_tflops": round(2.0 * 1024 * 1024 * 256 / td / 1e12, 6)},
                       "bcsr_order": {"seconds": round(tb, 5), "cores": 1,
                                      "effective_tflops": round(f0 / tb / 1e12, 6)},
                       "bit_identical": True}
        del A0, B0, C0, C0b
        cpu = {"value": round(flops / tc / 1e12, 5), "unit": "TFLOPS", "cores": threads,
               "kind": "port",
               "sample": f"full product (all {h.block_rows} block-rows), oracle BCSR-order fp32 "
                         f"loop (bit-identical to cpu_matmul_ref) over {threads} threads, median "
                         f"of {len(times)} runs ({tc:.3f} s each)",
               "seconds": round(tc, 4),
               "host": host,
               "single_core": {"value": round(flops / t1 / 1e12, 5), "unit": "TFLOPS", "cores": 1,
                               "seconds": round(t1, 4)},
               "reference_dense_loop_small_case": dense_small,
               "gpu_vs_cpu": {"max_err_over_absAabsB": rel, "bound": bound,
                              "exact_zero_where_absAabsB_zero": exact_zero,
                              "ok": rel <= bound and exact_zero},
               "gpu_vs_fp64": {"max_err_over_absAabsB_stored_values": rel64,
                               "max_err_over_absAabsB_generator_values": rel64g,
                               "mean_err_over_absAabsB_generator_values": mean64g}}

    extensions = None
    if not args.no_extensions and args.precision == "f16":
        t1 = time.time()
        nb = args.flush_mb << 20
        fsrc, fdst = mm.DeviceArray((nb,), np.uint8), mm.DeviceArray((nb,), np.uint8)
        flush = lambda: mm.check(mm.lib().bcsr_copy_kernel(fdst.ptr, fsrc.ptr, nb, stream.handle), "flush")
        ref = {k: timing[k]["median_ms"] for k in ("cold", "warm")}
        extensions = {"precision_modes": precision_modes(mm, np, N, stream, ref, flush)}
        extensions["library_rows_at_this_workload"] = library_rows(mm, np, h, B, N, stream, ms, C_gpu)
        extensions["reference_published_shapes"] = published_shapes(mm, np, stream)
        extensions["batched"] = batched(mm, np, args.batch_count, stream,
                                        {k: flops / (v * 1e-3) / 1e12 for k, v in ref.items()})
        fsrc.free()
        fdst.free()
        print(f"# extensions {time.time() - t1:.1f}s", file=sys.stderr, flush=True)

    out = {
        "metric": METRIC,
        "value": round(tflops, 3),
        "unit": "TFLOPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 6),
        "timing": timing,
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
