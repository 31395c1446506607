"""The batched uneven workload on its own: one persistent launch over many
BCSR x dense problems, replayed from a hipGraph, with a fixed policy.

    python -m matrixmultiplicationcuda_amd.batch_bench [--precision fp8] [--count 96]
        [--waves-per-cu 4] [--bin-work 128] [--queues 2] [--reps 20]

Problems come from workloads.draw_problems (rows and cols log-uniform on
[512, 16384] rounded to 64, 16/32/64 x 16 blocks, block density U(0.5%, 25%),
in-block density U(0.3, 1), N in {32, 128, 512}; seeds 1000 + i).  Tiles are
the batch's automatic widths.  One line of JSON: milliseconds per run (FP8:
the two B quantisers, the counter reset and the persistent kernel), TFLOPS,
and the per-matrix launches of the same problems for comparison.  Meant to
be run under rocprofv3 --kernel-trace --stats for the per-kernel split.
"""
import argparse
import concurrent.futures as cf
import json
import os
import sys

import numpy as np

from . import Batch, DeviceArray, HostBCSR, Plan, Stream, Event, check, lib, workloads


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawTextHelpFormatter)
    ap.add_argument("--precision", choices=("f16", "bf16", "fp8", "hilo"), default="fp8")
    ap.add_argument("--count", type=int, default=96)
    ap.add_argument("--waves-per-cu", type=int, default=0, help="0 = the library's choice")
    ap.add_argument("--bin-work", type=int, default=128)
    ap.add_argument("--queues", type=int, default=0,
                    help="0 = automatic (one global queue), 1 = global, 2 = one per XCD")
    ap.add_argument("--reps", type=int, default=20)
    args = ap.parse_args(argv)
    check(lib().bcsr_set_device(0))
    stream = Stream()
    probs = workloads.draw_problems(args.count, base_seed=1000)
    with cf.ThreadPoolExecutor(max_workers=min(12, os.cpu_count() or 1)) as ex:
        built = list(ex.map(lambda p: (p.bcsr(np.float32), p.B()), probs))
    hosts = []
    for h, _ in built:
        if args.precision == "f16":
            h = HostBCSR(h.hdr, h.idx, h.data.astype(np.float16), h.block_rows, h.block_cols, h.bh, h.bw)
        hosts.append(h)
    flops = sum(workloads.flops(h, p.N) for h, p in zip(hosts, probs))
    Bs = [DeviceArray.from_numpy(b) for _, b in built]
    Cs = [DeviceArray((p.rows, p.N), np.float32) for p in probs]
    plans = [Plan(h, args.precision) for h in hosts]
    print(f"# {args.count} problems built, {flops / 1e9:.1f} GFLOP", file=sys.stderr, flush=True)

    # per-matrix launches of the same problems (automatic choice each)
    for _ in range(2):
        for pl, B, C in zip(plans, Bs, Cs):
            pl.spmm_ptr(B.ptr, C.ptr, B.shape[1], stream)
    stream.synchronize()
    e0, e1 = Event(), Event()
    e0.record(stream)
    for _ in range(5):
        for pl, B, C in zip(plans, Bs, Cs):
            pl.spmm_ptr(B.ptr, C.ptr, B.shape[1], stream)
    e1.record(stream)
    stream.synchronize()
    per_matrix_ms = e0.elapsed_ms(e1) / 5

    batch = Batch(plans, Bs, Cs, waves_per_cu=args.waves_per_cu, bin_work=args.bin_work,
                  queues=args.queues)
    ms = batch.time(stream, graph=True, warmup=3, reps=args.reps)
    print(json.dumps({"precision": args.precision, "problems": args.count, "gflop": round(flops / 1e9, 3),
                      "batched_ms": round(ms, 4), "batched_tflops": round(flops / ms / 1e9, 2),
                      "per_matrix_ms": round(per_matrix_ms, 4),
                      "per_matrix_tflops": round(flops / per_matrix_ms / 1e9, 2),
                      "waves_per_cu": args.waves_per_cu, "bin_work": args.bin_work,
                      "queues": args.queues}), flush=True)
    batch.close()
    for pl in plans:
        pl.close()
    for a in Bs + Cs:
        a.free()
    stream.destroy()
    return 0


if __name__ == "__main__":
    sys.exit(main())
