"""Learned dispatch: an MLP cost model that maps a matrix's sparsity features
to the predicted log time of every launch configuration and of the library's
automatic choice, trained on the autotuner's own timings with this library's
HIP kernels (bcsr_cm_*), and a dispatcher that picks the predicted-fastest
configuration -- or falls back to exhaustive autotuning when the predicted
spread between the fastest and the slowest candidate is under 5%, and keeps
the automatic choice unless a configuration is predicted faster than it by
more than a margin chosen on validation matrices.

Features (24, in this order): log2 rows, log2 cols, log2 N, bh / 64,
bw / 64, block density, log2(nnzb + 1), log2(mean blocks per block-row + 1),
coefficient of variation of blocks per block-row, log2(max blocks per
block-row + 1), fraction of empty block-rows, bandwidth (max distance of a
block from the scaled diagonal / cols), mean distance / cols, fraction of
blocks within one block of the diagonal, in-block density (nonzero fraction
of stored values), log2 flops, log2 algorithmic bytes, log2 arithmetic
intensity, N / 512, ceil(N / 128), one-hot bh in {16, 32, 64}, log2 block-rows.
"""
import ctypes
import json
import math

import numpy as np

from . import DeviceArray, HostBCSR, check, lib, _ptr, stream_handle

N_FEATURES = 24
HIDDEN = 64
FALLBACK_SPREAD = 0.05
AUTO = -1  # choose(): the library's automatic choice
# margins (predicted gain over the automatic choice) the trainer tries on
# its validation matrices
MARGINS = (0.0, 0.01, 0.02, 0.03, 0.05, 0.075, 0.1, 0.15, 0.2, 0.3)


def features(h: HostBCSR, N: int) -> np.ndarray:
    brows, bcols = h.block_rows, h.block_cols
    rows, cols = brows * h.bh, bcols * h.bw
    cnt = np.diff(np.asarray(h.hdr, dtype=np.int64)).astype(np.float64)
    nnzb = float(cnt.sum())
    mean = float(cnt.mean()) if brows else 0.0
    cv = float(cnt.std() / mean) if mean > 0 else 0.0
    mx = float(cnt.max()) if brows else 0.0
    empty = float((cnt == 0).mean()) if brows else 1.0
    if nnzb > 0:
        bi = np.repeat(np.arange(brows, dtype=np.float64), cnt.astype(np.int64))
        bj = np.asarray(h.idx, dtype=np.float64)
        # distance of each block centre from the diagonal scaled to the shape
        dist = np.abs((bj + 0.5) * h.bw - (bi + 0.5) * h.bh * (cols / max(rows, 1)))
        band = float(dist.max() / cols)
        mdist = float(dist.mean() / cols)
        diag = float((dist <= max(h.bh, h.bw)).mean())
        inblock = float(np.count_nonzero(h.data) / h.data.size)
    else:
        band = mdist = diag = inblock = 0.0
    density = nnzb / max(brows * bcols, 1)
    flops = 2.0 * nnzb * h.bh * h.bw * N
    nbytes = nnzb * h.bh * h.bw * 2 + 4 * nnzb + 4 * (brows + 1) + cols * N * 2 + rows * N * 4
    f = [math.log2(max(rows, 1)), math.log2(max(cols, 1)), math.log2(max(N, 1)), h.bh / 64.0,
         h.bw / 64.0, density, math.log2(nnzb + 1), math.log2(mean + 1), cv, math.log2(mx + 1),
         empty, band, mdist, diag, inblock, math.log2(flops + 1), math.log2(max(nbytes, 1)),
         math.log2(flops / max(nbytes, 1) + 1e-9), N / 512.0, math.ceil(N / 128.0),
         float(h.bh == 16), float(h.bh == 32), float(h.bh == 64), math.log2(max(brows, 1))]
    assert len(f) == N_FEATURES
    return np.array(f, dtype=np.float32)


class CostModel:
    """The MLP on the GPU (bcsr_cm_*): fit, predict, save, load."""

    def __init__(self, n_in=N_FEATURES, hidden=HIDDEN, n_out=8, seed=0):
        self._lib = lib()
        self.n_in, self.hidden, self.n_out = n_in, hidden, n_out
        self._h = ctypes.c_void_p()
        check(self._lib.bcsr_cm_create(n_in, hidden, n_out, seed, ctypes.byref(self._h)),
              "bcsr_cm_create")
        n = ctypes.c_int32()
        check(self._lib.bcsr_cm_num_params(self._h, ctypes.byref(n)), "cm params")
        self.n_params = n.value
        self.mu = np.zeros(n_in, np.float32)
        self.sigma = np.ones(n_in, np.float32)

    # parameters
    def get_params(self):
        out = np.empty(self.n_params, np.float32)
        check(self._lib.bcsr_cm_get_params(self._h, _ptr(out), self.n_params), "cm get")
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        check(self._lib.bcsr_cm_set_params(self._h, _ptr(p), self.n_params), "cm set")

    def get_grads(self):
        out = np.empty(self.n_params, np.float32)
        check(self._lib.bcsr_cm_get_grads(self._h, _ptr(out), self.n_params), "cm grads")
        return out

    # device-array level
    def forward_dev(self, X: DeviceArray, n, P: DeviceArray, stream=None):
        check(self._lib.bcsr_cm_forward(self._h, X.ptr, n, P.ptr, stream_handle(stream)), "cm fwd")

    def backward_dev(self, X, Y, M, n, inv_count, stream=None):
        loss = ctypes.c_float()
        check(self._lib.bcsr_cm_backward(self._h, X.ptr, Y.ptr, M.ptr, n, inv_count,
                                         ctypes.byref(loss), stream_handle(stream)), "cm bwd")
        return loss.value

    def train_dev(self, X, Y, M, n, inv_count, steps, lr, stream=None, b1=0.9, b2=0.999,
                  eps=1e-8):
        loss = ctypes.c_float()
        check(self._lib.bcsr_cm_train(self._h, X.ptr, Y.ptr, M.ptr, n, inv_count, steps, lr, b1,
                                      b2, eps, ctypes.byref(loss), stream_handle(stream)),
              "cm train")
        return loss.value

    # host level (standardised features)
    def fit(self, F, Y, M, steps=3000, lr=3e-3, log_every=0):
        F = np.asarray(F, np.float32)
        self.mu = F.mean(0)
        self.sigma = np.maximum(F.std(0), 1e-6).astype(np.float32)
        X = ((F - self.mu) / self.sigma).astype(np.float32)
        Ym = np.where(M > 0, Y, 0.0).astype(np.float32)
        n = X.shape[0]
        dX, dY = DeviceArray.from_numpy(X), DeviceArray.from_numpy(Ym)
        dM = DeviceArray.from_numpy(np.asarray(M, np.float32))
        inv = 1.0 / float(np.sum(M))
        losses = []
        chunk = log_every or steps
        done = 0
        while done < steps:
            k = min(chunk, steps - done)
            losses.append(self.train_dev(dX, dY, dM, n, inv, k, lr))
            done += k
        for a in (dX, dY, dM):
            a.free()
        return losses

    def predict(self, F):
        F = np.atleast_2d(np.asarray(F, np.float32))
        X = ((F - self.mu) / self.sigma).astype(np.float32)
        n = X.shape[0]
        dX = DeviceArray.from_numpy(X)
        dP = DeviceArray((n, self.n_out), np.float32)
        self.forward_dev(dX, n, dP)
        from . import synchronize
        synchronize()
        P = dP.numpy()
        dX.free()
        dP.free()
        return P

    def save(self, path, configs, meta=None):
        with open(path, "w") as f:
            json.dump({"n_in": self.n_in, "hidden": self.hidden, "n_out": self.n_out,
                       "configs": [list(c) for c in configs], "mu": self.mu.tolist(),
                       "sigma": self.sigma.tolist(), "params": self.get_params().tolist(),
                       "meta": meta or {}}, f)

    @classmethod
    def load(cls, path):
        with open(path) as f:
            d = json.load(f)
        m = cls(d["n_in"], d["hidden"], d["n_out"])
        m.set_params(np.array(d["params"], np.float32))
        m.mu = np.array(d["mu"], np.float32)
        m.sigma = np.array(d["sigma"], np.float32)
        return m, [tuple(c) for c in d["configs"]], d.get("meta", {})

    def close(self):
        if self._h:
            check(self._lib.bcsr_cm_destroy(self._h), "cm destroy")
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def dataset(records):
    """(features, log-ms targets, support mask, ms, auto_ms) from DB records.
    Targets and mask have one column per configuration and a last one for
    the automatic choice (mask 0 where a record has no automatic time)."""
    F = np.array([r["features"] for r in records], np.float32)
    ms = np.array([[np.nan if t is None else t for t in r["ms"]] for r in records], np.float64)
    auto = np.array([np.nan if r.get("auto_ms") is None else r["auto_ms"] for r in records], np.float64)
    full = np.concatenate([ms, auto[:, None]], axis=1)
    M = np.isfinite(full).astype(np.float32)
    Y = np.where(M > 0, np.log(np.where(M > 0, full, 1.0)), 0.0).astype(np.float32)
    return F, Y, M, ms, auto


def choose(pred_log, mask, spread=FALLBACK_SPREAD, pred_auto=None, margin=0.0):
    """(k, fallback): k the index of the predicted-fastest supported config,
    or AUTO when pred_auto (the predicted log time of the library's
    automatic choice) is given and no supported config is predicted faster
    than it by more than `margin`; fallback whether the predicted spread of
    the supported configs is under `spread` (then the dispatcher autotunes)."""
    p = np.where(mask > 0, pred_log, np.inf)
    k = int(np.argmin(p))
    sup = pred_log[mask > 0]
    if sup.size and not np.all(np.isfinite(sup)):
        return k, True  # a prediction that is not a number: autotune instead
    # spread of predicted times exp(max - min) - 1, compared in log space
    # (a badly trained model can predict log-time gaps far beyond exp's range)
    fallback = bool(sup.size > 1 and float(sup.max() - sup.min()) < math.log1p(spread))
    if (pred_auto is not None and not fallback and np.isfinite(pred_auto)
            and not float(p[k]) < float(pred_auto) - math.log1p(margin)):
        return AUTO, False
    return k, fallback


def split_prediction(pred_log, n_configs):
    """(per-config predictions, automatic-choice predictions or None)."""
    pred_log = np.atleast_2d(pred_log)
    if pred_log.shape[1] == n_configs + 1:
        return pred_log[:, :n_configs], pred_log[:, n_configs]
    return pred_log[:, :n_configs], None


def regret_report(pred_log, M, ms, auto_ms, spread=FALLBACK_SPREAD, margin=0.0):
    """Regret of learned dispatch (with and without the autotune fallback)
    and of the library's static choice, against exhaustive autotuning (the
    fastest of the configurations and the automatic choice).
    pred_log has one column per configuration and, for a model with the
    automatic choice's output, one more: then the dispatcher keeps the
    automatic choice unless a config is predicted faster by > margin."""
    T = ms.shape[1]
    pc, pa = split_prediction(pred_log, T)
    # the fastest of everything timed: every configuration and the automatic choice
    best = np.fmin(np.nanmin(ms, axis=1), auto_ms)
    learned, learned_fb, fallbacks, kept_auto = [], [], 0, 0
    for i in range(ms.shape[0]):
        k, fb = choose(pc[i], M[i, :T], spread, None if pa is None else pa[i], margin)
        t = auto_ms[i] if k == AUTO else ms[i, k]
        kept_auto += k == AUTO
        r = t / best[i] - 1.0
        learned.append(r)
        if fb:
            fallbacks += 1
            learned_fb.append(0.0)
        else:
            learned_fb.append(r)
    static = auto_ms / best - 1.0
    return {"n": int(ms.shape[0]), "margin": margin,
            "learned_mean_regret": float(np.mean(learned)),
            "learned_p90_regret": float(np.percentile(learned, 90)),
            "dispatch_mean_regret": float(np.mean(learned_fb)),
            "fallback_fraction": fallbacks / max(ms.shape[0], 1),
            "kept_automatic_fraction": kept_auto / max(ms.shape[0], 1),
            "static_mean_regret": float(np.nanmean(static)),
            "static_p90_regret": float(np.nanpercentile(static, 90))}


def tune_margin(pred_log, M, ms, auto_ms, margins=MARGINS):
    """The margin with the least mean dispatch regret on (validation)
    records, the smaller one on ties."""
    scores = [(regret_report(pred_log, M, ms, auto_ms, margin=m)["dispatch_mean_regret"], m) for m in margins]
    return min(scores)[1]


def default_model_path(precision):
    import os
    return os.path.join(os.path.dirname(os.path.abspath(__file__)), "data",
                        f"costmodel_{precision}.json")


class Dispatcher:
    """Picks a launch configuration per matrix from the trained cost model;
    when the predicted spread is under FALLBACK_SPREAD it times every
    candidate instead (exhaustive autotune) if operands are given."""

    def __init__(self, precision="f16", path=None):
        self.precision = precision
        self.model, self.configs, self.meta = CostModel.load(path or default_model_path(precision))
        self.margin = float(self.meta.get("margin", 0.0))

    def predict(self, hosts_and_N):
        """(per-config log-time predictions, automatic-choice predictions or None)."""
        F = np.stack([features(h, N) for h, N in hosts_and_N])
        return split_prediction(self.model.predict(F), len(self.configs))

    def choose(self, plan, h, N, B=None, C=None, stream=None):
        """(config, used_fallback) for one product; config None = the
        library's automatic choice."""
        from . import autotune
        pc, pa = self.predict([(h, N)])
        mask = autotune.supported_mask(plan, N, self.configs).astype(np.float32)
        k, fb = choose(pc[0], mask, pred_auto=None if pa is None else pa[0], margin=self.margin)
        if fb and B is not None and C is not None:
            t = autotune.time_configs(plan, B, C, N, self.configs, stream)
            k = int(np.nanargmin(t))
        return (None if k == AUTO else self.configs[k]), fb

    def batch_tiles(self, plans, hosts, Ns):
        """Per-problem tile widths for a batch: the predicted-fastest
        configuration among those the persistent kernel runs (the variant
        autotune.batch_variant names for each tile width)."""
        from . import PRECISIONS, autotune
        prec = PRECISIONS[self.precision]
        pred, _ = self.predict(list(zip(hosts, Ns)))
        return [autotune.best_batch_tile(self.configs, row, prec, p.info.dev_bh)
                for p, row in zip(plans, pred)]

    def close(self):
        self.model.close()
