"""GPU parity of the batched persistent kernel.

The persistent kernel runs the per-matrix kernels' work functions over a
device work queue, so each problem's C must equal, bit for bit, the
per-matrix launch with the same tile width (automatic variant: the
prefetching kernel where it fits, as in the batch); the fp16
batch is also checked against the oracle's CPU product (1e-3 of |A||B|).
Graph replays are checked one by one with eager work in between.
"""
import sys

import numpy as np
import pytest

import matrixmultiplicationcuda_amd as mm
import oracle
from matrixmultiplicationcuda_amd import workloads

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def small_problems(count, seed):
    return workloads.draw_problems(count, base_seed=seed, lo=64, hi=1536)


def setup(problems, precision):
    plans, Bs, Cs, hosts, Bh = [], [], [], [], []
    for pr in problems:
        h = pr.bcsr(np.float16 if precision == "f16" else np.float32)
        B16 = pr.B()
        plans.append(mm.Plan(h, precision))
        Bs.append(mm.DeviceArray.from_numpy(B16))
        C = mm.DeviceArray((pr.rows, pr.N), np.float32)
        C.fill_bytes(0xFF)  # NaN: every element must be written
        Cs.append(C)
        hosts.append(h)
        Bh.append(B16)
    return plans, Bs, Cs, hosts, Bh


def per_matrix(plans, Bs, tiles):
    out = []
    for p, B, t in zip(plans, Bs, tiles):
        # automatic variant: the prefetching kernel where it fits, else the
        # plain one (64-row blocks at 128 columns), as the batch runs them
        cfg = mm.make_cfg(tile_n=int(t), variant=-1)
        out.append(p.matmul(B.numpy(), cfg=cfg))
    return out


def teardown(plans, Bs, Cs):
    for x in plans + Bs + Cs:
        x.close() if hasattr(x, "close") else x.free()


@pytest.mark.parametrize("queues", [1, 2])
@pytest.mark.parametrize("precision", ["f16", "fp8", "bf16", "hilo"])
def test_batch_equals_per_matrix(precision, queues):
    """One global queue (1) or one per XCD (2): each problem's C is bit-equal
    to its own per-matrix launch at the same tile width."""
    problems = small_problems(12, seed=7)
    log(f"batch {precision} queues {queues}: "
        + ", ".join(f"{p.rows}x{p.cols}/{p.bh}/N{p.N}" for p in problems))
    mm.check(mm.lib().bcsr_set_device(0))
    plans, Bs, Cs, hosts, Bh = setup(problems, precision)
    batch = mm.Batch(plans, Bs, Cs, queues=queues)
    tiles = batch.tiles
    log(f"  items {batch.info.items} grid {batch.info.grid} tiles {list(tiles)}")
    batch.run()
    mm.synchronize()
    want = per_matrix(plans, Bs, tiles)
    for i, (C, w) in enumerate(zip(Cs, want)):
        got = C.numpy()
        assert np.array_equal(got, w), f"problem {i} ({problems[i]}) differs from its own launch"
    if precision == "f16":
        for i in (0, 5, 11):
            h = hosts[i]
            Bf = Bh[i].astype(np.float32)
            ref = oracle.bcsr_spmm_f32(h.hdr, h.idx, h.data.view(np.uint16), h.bh, h.bw, Bf,
                                       threads=8)
            _, cabs = oracle.bcsr_spmm_f64(h.hdr, h.idx, h.data.astype(np.float32), h.bh, h.bw,
                                           Bf, threads=8)
            err = np.abs(Cs[i].numpy().astype(np.float64) - ref)
            assert np.all(err <= 1e-3 * cabs + 1e-30)
    batch.close()
    teardown(plans, Bs, Cs)


def test_batch_explicit_tiles_and_queue_shapes():
    """Every tile width the kernel is compiled for, tiny and large work
    items, several persistent grid sizes: same bits every time."""
    problems = small_problems(9, seed=21)
    mm.check(mm.lib().bcsr_set_device(0))
    plans, Bs, Cs, _, _ = setup(problems, "f16")
    ref = None
    for tiles, wpc, bin_work, queues in [(None, 0, 0, 0), ("min", 1, 1, 1), ("max", 4, 100000, 2),
                                         ("min", 16, 64, 2), ("min", 1, 1, 2)]:
        if tiles == "min":
            t = [32] * len(problems)
        elif tiles == "max":
            t = [128 if mm.lib().bcsr_batch_tile_supported(0, p.bh, 128) else 64
                 for p in problems]
        else:
            t = None
        for C in Cs:
            C.fill_bytes(0xFF)
        batch = mm.Batch(plans, Bs, Cs, tiles=t, waves_per_cu=wpc, bin_work=bin_work, queues=queues)
        log(f"  tiles {list(batch.tiles)} wpc {wpc} bin {bin_work} queues {queues} "
            f"items {batch.info.items}")
        batch.run()
        mm.synchronize()
        want = per_matrix(plans, Bs, batch.tiles)
        for i, C in enumerate(Cs):
            assert np.array_equal(C.numpy(), want[i])
        got = [C.numpy() for C in Cs]
        if ref is not None:
            for a, b in zip(got, ref):
                # different tile widths change only the MFMA column grouping
                np.testing.assert_allclose(a, b, rtol=0, atol=1e-5 * max(1.0, np.abs(b).max()))
        ref = got
        batch.close()
    teardown(plans, Bs, Cs)


@pytest.mark.parametrize("precision", ["f16", "fp8"])
def test_batch_graph_replays(precision):
    problems = small_problems(10, seed=33)
    mm.check(mm.lib().bcsr_set_device(0))
    plans, Bs, Cs, _, _ = setup(problems, precision)
    batch = mm.Batch(plans, Bs, Cs)
    stream = mm.Stream()
    batch.run(stream)
    stream.synchronize()
    want = [C.numpy() for C in Cs]
    for rep in range(4):
        for C in Cs:
            C.fill_bytes(0xFF)  # eager work between replays
        mm.synchronize()
        batch.run(stream, graph=True)
        stream.synchronize()
        for i, C in enumerate(Cs):
            assert np.array_equal(C.numpy(), want[i]), f"replay {rep} problem {i}"
        log(f"  replay {rep} ok")
    batch.close()
    stream.destroy()
    teardown(plans, Bs, Cs)


def test_batch_errors():
    mm.check(mm.lib().bcsr_set_device(0))
    problems = small_problems(3, seed=44)
    plans, Bs, Cs, hosts, _ = setup(problems, "f16")
    h32 = problems[0].bcsr(np.float32)
    p8 = mm.Plan(h32, "fp8")
    with pytest.raises(mm.BcsrError) as e:  # mixed precisions
        mm.Batch([plans[0], p8], Bs[:2], Cs[:2])
    assert e.value.status == -7
    big = [256 if p.bh == 16 else 128 for p in problems]
    with pytest.raises(mm.BcsrError) as e:  # tile width not compiled into the batch kernel
        mm.Batch(plans, Bs, Cs, tiles=[256, 32, 32])
    assert e.value.status == -7
    del big
    with pytest.raises(mm.BcsrError) as e:  # queue mode out of range
        mm.Batch(plans, Bs, Cs, queues=3)
    assert e.value.status == -1
    batch = mm.Batch(plans, Bs, Cs)
    with pytest.raises(mm.BcsrError) as e:  # the null stream cannot be captured
        batch.run(None, graph=True)
    assert e.value.status == -1
    batch.close()
    empty = mm.Batch([], [], [])
    assert empty.info.items == 0
    empty.run()
    mm.synchronize()
    empty.close()
    p8.close()
    teardown(plans, Bs, Cs)


def test_fp8_with_empty_inner_dimension():
    """FP8 products with K = 0 (no block-columns): C = 0, per matrix and in a
    batch beside an ordinary problem; no quantiser runs and no B scale is
    read."""
    mm.check(mm.lib().bcsr_set_device(0))
    N, rows = 64, 64
    empty = mm.HostBCSR(np.zeros(rows // 16 + 1, np.int32), np.zeros(0, np.int32),
                        np.zeros((0, 16, 16), np.float32), rows // 16, 0, 16, 16)
    log("fp8 K = 0, per matrix")
    p0 = mm.Plan(empty, "fp8")
    Bdummy = mm.DeviceArray((1, N), np.float16)
    C0 = mm.DeviceArray((rows, N), np.float32)
    C0.fill_bytes(0xFF)
    p0.spmm_ptr(Bdummy.ptr, C0.ptr, N)
    mm.synchronize()
    assert np.all(C0.numpy() == 0.0)
    log("fp8 K = 0 in a batch")
    other = small_problems(1, seed=11)[0]
    h1 = other.bcsr(np.float32)
    p1 = mm.Plan(h1, "fp8")
    B1 = mm.DeviceArray.from_numpy(other.B())
    C1 = mm.DeviceArray((other.rows, other.N), np.float32)
    C0.fill_bytes(0xFF)
    batch = mm.Batch([p0, p1], [Bdummy, B1], [C0, C1])
    batch.run()
    mm.synchronize()
    assert np.all(C0.numpy() == 0.0)
    want = p1.matmul(B1.numpy(), cfg=mm.make_cfg(tile_n=int(batch.tiles[1]), variant=-1))
    assert np.array_equal(C1.numpy(), want)
    batch.close()
    for x in (p0, p1):
        x.close()
    for a in (Bdummy, C0, B1, C1):
        a.free()


def test_fp8_batch_with_slab_scales_differ():
    """FP8 batch whose B matrices have 16-row slabs of different magnitudes
    (each slab keeps its own scale, the two-MFMA path): every C bit-equal to
    its own per-matrix launch, over two runs."""
    problems = small_problems(8, seed=23)
    log("fp8 batch, B slabs scaled: " + ", ".join(f"{p.rows}x{p.cols}/{p.bh}/N{p.N}" for p in problems))
    mm.check(mm.lib().bcsr_set_device(0))
    rng = np.random.default_rng(23)
    plans, Bs, Cs = [], [], []
    for pr in problems:
        plans.append(mm.Plan(pr.bcsr(np.float32), "fp8"))
        B = pr.B().astype(np.float32)
        B *= np.repeat(2.0 ** rng.integers(-6, 6, pr.cols // 16), 16)[:, None].astype(np.float32)
        Bs.append(mm.DeviceArray.from_numpy(B.astype(np.float16)))
        C = mm.DeviceArray((pr.rows, pr.N), np.float32)
        C.fill_bytes(0xFF)
        Cs.append(C)
    batch = mm.Batch(plans, Bs, Cs)
    want = per_matrix(plans, Bs, batch.tiles)
    for run in range(2):
        batch.run()
        mm.synchronize()
        for i, (C, w) in enumerate(zip(Cs, want)):
            assert np.array_equal(C.numpy(), w), f"run {run}: problem {i} ({problems[i]})"
    batch.close()
    teardown(plans, Bs, Cs)
