"""Train the learned dispatcher's cost model on the GPU from a timing
database and report its regret on the held-out matrices.

    python -m matrixmultiplicationcuda_amd.train_costmodel \\
        --db data/timings_f16.jsonl --out data/costmodel_f16.json

Records whose generator seed is below --split train the model (fit by the
library's HIP forward/backward/Adam kernels), except the last --validation
of them, on which the dispatcher's margin is chosen (how much faster than
the library's automatic choice a configuration must be predicted before the
dispatcher takes it); the others are held out and only evaluated: regret of
the dispatcher (predicted-fastest configuration, autotune fallback when the
predicted spread is under 5%, the automatic choice kept within the margin)
against exhaustive autotuning, next to the regret of the library's static
choice and of the dispatcher without the margin.
"""
import argparse
import json
import sys

from . import check, lib
from . import autotune, learned


def split(records, seed_split):
    train = [r for r in records if int(r["key"].split(":")[1]) < seed_split]
    test = [r for r in records if int(r["key"].split(":")[1]) >= seed_split]
    return train, test


def train(records_train, steps=3000, lr=3e-3, seed=0):
    F, Y, M, _, _ = learned.dataset(records_train)
    cm = learned.CostModel(learned.N_FEATURES, learned.HIDDEN, Y.shape[1], seed=seed)
    losses = cm.fit(F, Y, M, steps=steps, lr=lr, log_every=max(steps // 6, 1))
    return cm, losses


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawTextHelpFormatter)
    ap.add_argument("--db", required=True)
    ap.add_argument("--out", required=True)
    ap.add_argument("--split", type=int, default=200000)
    ap.add_argument("--validation", type=int, default=500)
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--lr", type=float, default=3e-3)
    ap.add_argument("--seed", type=int, default=0)
    args = ap.parse_args(argv)
    check(lib().bcsr_set_device(0))
    recs = autotune.TimingDB(args.db).load()
    tr, te = split(recs, args.split)
    tr.sort(key=lambda r: int(r["key"].split(":")[1]))
    nval = min(args.validation, len(tr) // 2)
    fit, val = tr[:len(tr) - nval], tr[len(tr) - nval:]
    cm, losses = train(fit, args.steps, args.lr, args.seed)
    configs = [tuple(c) for c in tr[0]["configs"]]
    margin = 0.0
    meta = {"train": len(fit), "validation": len(val), "held_out": len(te), "steps": args.steps,
            "lr": args.lr, "losses": losses, "db": args.db, "auto_output": True}
    if val:
        Fv, _, Mv, msv, autov = learned.dataset(val)
        margin = learned.tune_margin(cm.predict(Fv), Mv, msv, autov)
    meta["margin"] = margin
    if te:
        Ft, _, Mt, mst, autot = learned.dataset(te)
        pt = cm.predict(Ft)
        meta["regret"] = learned.regret_report(pt, Mt, mst, autot, margin=margin)
        meta["regret_without_margin"] = learned.regret_report(pt, Mt, mst, autot, margin=0.0)
        meta["regret_without_automatic_output"] = learned.regret_report(pt[:, :mst.shape[1]], Mt, mst, autot)
    cm.save(args.out, configs, meta)
    print(json.dumps(meta), flush=True)
    cm.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
