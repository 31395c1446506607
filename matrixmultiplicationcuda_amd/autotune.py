"""Autotuner: exhaustive timing of launch configurations, per matrix and per
batch, with a persistent timing database.

Per matrix the candidate set is CONFIGS (output tile width x kernel variant
x XCD mapping for the 16-bit modes, tile width x XCD mapping for FP8); a
configuration the library does not build for a plan (bcsr_spmm_kernel_name
refuses it) is recorded as missing.  Per batch the tuner measures tile-width policies (the batch's
automatic widths, the learned dispatcher's picks, each problem's own fastest
width restricted to those the persistent kernel is compiled for), each with a
sweep of the persistent grid (waves per CU) and the work-item size, and keeps
the fastest.

The database is a JSON-lines file, one record per tuned matrix: its
features, N, the measured milliseconds of every configuration (null where
unsupported) and of the library's automatic choice.  Appending is the
checkpoint: a run that stops can be resumed from the records already there.
"""
import json
import os

import numpy as np

from . import Batch, check, lib, make_cfg

TILES = (32, 64, 128, 256)
XCD_MAPS = (0, 1, 2, 3, 4)  # linear, tiles pinned, rows pinned, chunked runs, 2 x 4 grid
# (tile_n, variant, xcd_map); a combination the library refuses for a plan
# (say tiles pinned to XCDs for a tile count 8 does not divide) is missing
CONFIGS_16 = tuple((t, v, m) for t in TILES for v in (0, 1) for m in XCD_MAPS)
CONFIGS_FP8 = tuple((t, -1, m) for t in TILES for m in XCD_MAPS)
BATCH_WAVES = (4, 6, 8, 12, 16)
BATCH_BINS = (32, 64, 128, 256)


def configs_for(precision):
    return CONFIGS_FP8 if precision == "fp8" else CONFIGS_16


def cfg_of(config):
    """Launch configuration of a (tile_n, variant[, xcd_map]) tuple (no map:
    the library's automatic mapping); None stays None (the automatic choice)."""
    if config is None:
        return None
    tile, variant = config[0], config[1]
    xmap = config[2] if len(config) > 2 else -1
    return make_cfg(tile_n=tile, variant=variant, xcd_map=xmap, waves_per_block=1)


def supported_mask(plan, N, configs):
    return np.array([plan.supports(N, cfg_of(c)) for c in configs], dtype=bool)


def time_configs(plan, B, C, N, configs, stream=None, warmup=2, reps=5):
    """Milliseconds per launch of every configuration (nan if unsupported)."""
    out = np.full(len(configs), np.nan)
    for k, c in enumerate(configs):
        cfg = cfg_of(c)
        if cfg is None or plan.supports(N, cfg):
            out[k] = plan.time_ptr(B.ptr, C.ptr, N, stream, cfg, warmup, reps)
    return out


def time_auto(plan, B, C, N, stream=None, warmup=2, reps=5):
    return plan.time_ptr(B.ptr, C.ptr, N, stream, None, warmup, reps)


class TimingDB:
    """JSON-lines store of per-matrix timings (append-only)."""

    def __init__(self, path):
        self.path = path

    def load(self):
        if not self.path or not os.path.exists(self.path):
            return []
        with open(self.path) as f:
            return [json.loads(line) for line in f if line.strip()]

    def keys(self):
        return {r["key"] for r in self.load()}

    def append(self, record):
        if not self.path:
            return
        d = os.path.dirname(self.path)
        if d:
            os.makedirs(d, exist_ok=True)
        with open(self.path, "a") as f:
            f.write(json.dumps(record) + "\n")


def record(key, feats, N, precision, configs, times, auto_ms):
    return {"key": key, "precision": precision, "N": int(N),
            "features": [float(x) for x in feats],
            "configs": [list(c) for c in configs],
            "ms": [None if not np.isfinite(t) else float(t) for t in times],
            "auto_ms": float(auto_ms)}


def batch_variant(precision_id, dev_bh, tile):
    """The kernel variant the batched persistent kernel runs for a device
    block height and tile width: the prefetching one (1) where its two load
    sets fit, else the plain one (0), as batch_prefetch in kernels_batch.hip.
    FP8 has a single kernel, which its configurations name as variant -1
    (CONFIGS_FP8)."""
    if precision_id == 2:
        return -1
    h = 2 if precision_id == 3 else 1
    return 1 if (dev_bh // 16) * (tile // 16) * h <= 16 else 0


def best_batch_tile(configs, values, precision_id, dev_bh):
    """The tile width of the smallest value (a time or a predicted log time)
    among the configurations the persistent kernel runs for this device
    block height: the variant batch_variant names, a tile width the batch is
    compiled for, a finite value.  0 (the batch's automatic width) if none."""
    best, best_v = 0, np.inf
    for c, v in zip(configs, values):
        tile, variant = c[0], c[1]  # the batch has its own queue: any XCD map
        if v is None or not np.isfinite(v) or variant != batch_variant(precision_id, dev_bh, tile):
            continue
        if not lib().bcsr_batch_tile_supported(precision_id, dev_bh, tile):
            continue
        if v < best_v:
            best, best_v = tile, v
    return best


def batch_tiles_from_times(plans, Ns, times, configs, precision_id):
    """Each problem's fastest tile width among those the persistent kernel
    supports, judged by the timings of the variant the batch runs there."""
    return [best_batch_tile(configs, t, precision_id, p.info.dev_bh) for p, t in zip(plans, times)]


def tune_batch(plans, Bs, Cs, tiles, stream, waves=BATCH_WAVES, bins=BATCH_BINS, graph=True,
               warmup=2, reps=5):
    """Sweep the persistent grid and item size for fixed per-problem tiles.
    Returns (best dict, list of all (waves, bin, ms))."""
    trials = []
    best = None
    for w in waves:
        for b in bins:
            batch = Batch(plans, Bs, Cs, tiles=tiles, waves_per_cu=w, bin_work=b)
            tiles_used = list(batch.tiles)
            try:
                ms = batch.time(stream, graph=graph, warmup=warmup, reps=reps)
            finally:
                batch.close()
            trials.append((w, b, ms))
            if best is None or ms < best["ms"]:
                best = {"waves_per_cu": w, "bin_work": b, "ms": ms, "tiles": tiles_used}
    return best, trials


def tune_batch_policies(plans, Bs, Cs, policies, stream, waves=BATCH_WAVES, bins=BATCH_BINS,
                        graph=True, warmup=2, reps=3):
    """Tile widths by policy (name -> per-problem tiles, None = the batch's
    automatic widths), each with its own grid/item sweep; returns the best
    {policy, tiles, waves_per_cu, bin_work, ms} and the ms of every policy."""
    table = {}
    best = None
    for name, tiles in policies.items():
        b, _ = tune_batch(plans, Bs, Cs, tiles, stream, waves, bins, graph, warmup, reps)
        table[name] = {"ms": b["ms"], "waves_per_cu": b["waves_per_cu"], "bin_work": b["bin_work"]}
        if best is None or b["ms"] < best["ms"]:
            best = dict(b, policy=name)
    return best, table


def precision_id(precision):
    from . import PRECISIONS
    return PRECISIONS[precision]


__all__ = ["CONFIGS_16", "CONFIGS_FP8", "XCD_MAPS", "TimingDB", "time_configs", "time_auto", "tune_batch",
           "batch_variant", "best_batch_tile", "batch_tiles_from_times", "tune_batch_policies",
           "configs_for", "cfg_of", "supported_mask", "record", "check"]
