"""Train the learned dispatcher's cost model on the GPU from a timing
database and report its regret on the held-out matrices.

    python -m matrixmultiplicationcuda_amd.train_costmodel \\
        --db data/timings_f16.jsonl --out data/costmodel_f16.json

Records whose generator seed is below --split train the model (fit by the
library's HIP forward/backward/Adam kernels); the others are held out and
only evaluated: regret of the dispatcher (predicted-fastest configuration,
autotune fallback when the predicted spread is under 5%) against exhaustive
autotuning, next to the regret of the library's static choice.
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
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--lr", type=float, default=3e-3)
    ap.add_argument("--seed", type=int, default=0)
    args = ap.parse_args(argv)
    check(lib().bcsr_set_device(0))
    recs = autotune.TimingDB(args.db).load()
    tr, te = split(recs, args.split)
    cm, losses = train(tr, args.steps, args.lr, args.seed)
    configs = [tuple(c) for c in tr[0]["configs"]]
    meta = {"train": len(tr), "held_out": len(te), "steps": args.steps, "lr": args.lr,
            "losses": losses, "db": args.db}
    if te:
        Ft, _, Mt, mst, autot = learned.dataset(te)
        meta["regret"] = learned.regret_report(cm.predict(Ft), Mt, mst, autot)
    cm.save(args.out, configs, meta)
    print(json.dumps(meta), flush=True)
    cm.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
