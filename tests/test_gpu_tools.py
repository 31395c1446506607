"""The command-line tools of the package on small inputs, on the GPU:
scale_table (the at-scale table, checked rows), batch_bench (the batched
workload on its own), collect_timings (the autotuner's timing database) and
train_costmodel (training on that database and the resumed collection)."""
import json
import sys

import pytest

from matrixmultiplicationcuda_amd import batch_bench, collect_timings, learned, scale_table, train_costmodel

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_scale_table_small(tmp_path):
    out, md = tmp_path / "t.json", tmp_path / "t.md"
    log("scale_table 512 x 512")
    assert scale_table.main(["--M", "512", "--K", "512", "--N", "64,128", "--bh", "16,32", "--reps", "2",
                             "--out", str(out), "--md", str(md)]) == 0
    rows = json.loads(out.read_text())["rows"]
    # 2 block shapes x (3 random densities + band + block-diagonal) x 2 N
    assert len(rows) == 20
    for r in rows:
        assert r["max_abs_diff_sampled_rows_vs_rocblas"] <= 1e-3, r
        assert r["us"] > 0 and r["effective_tflops"] > 0
        assert (r["rocsparse_bsr_us"] is not None) == (r["bh"] == 16)
    assert md.read_text().count("\n") == 22  # header, rule, 20 rows


def test_batch_bench_small(capsys):
    log("batch_bench 6 problems")
    assert batch_bench.main(["--count", "6", "--precision", "fp8", "--reps", "2"]) == 0
    line = [x for x in capsys.readouterr().out.splitlines() if x.startswith("{")][-1]
    r = json.loads(line)
    assert r["problems"] == 6 and r["batched_ms"] > 0 and r["per_matrix_ms"] > 0


def test_collect_timings_then_train(tmp_path, capsys):
    db = tmp_path / "db.jsonl"
    log("collect_timings 3 problems")
    args = ["--count", "3", "--base-seed", "424242", "--out", str(db), "--threads", "2", "--reps", "1"]
    assert collect_timings.main(args) == 0
    recs = [json.loads(x) for x in db.read_text().splitlines()]
    assert len(recs) == 3
    for r in recs:
        assert len(r["configs"]) == len(r["ms"]) and r["auto_ms"] > 0
    # a second run resumes: nothing left to time, nothing appended
    assert collect_timings.main(args) == 0
    assert len(db.read_text().splitlines()) == 3
    capsys.readouterr()
    log("train_costmodel on 2 + 1 held out")
    model = tmp_path / "cm.json"
    assert train_costmodel.main(["--db", str(db), "--out", str(model), "--split", "424244",
                                 "--steps", "50"]) == 0
    meta = json.loads([x for x in capsys.readouterr().out.splitlines() if x.startswith("{")][-1])
    # of the 2 training records one fits the model and one chooses its margin
    assert meta["train"] == 1 and meta["validation"] == 1 and meta["held_out"] == 1
    assert "regret" in meta and meta["margin"] in learned.MARGINS
    assert model.exists()
