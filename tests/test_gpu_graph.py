"""HIP graphs of the library's launches (bcsr_graph_*): a captured product,
replayed several times with eager work and synchronisations in between,
equals the eager product bit for bit on every replay.  FP8 plans: a graph
captured at one N stays valid after a larger N grows the plan's workspace
(the outgrown workspace is kept until the plan is destroyed), and a capture
that would need a larger workspace is refused with BCSR_ERR_UNSUPPORTED.
"""
import sys

import numpy as np
import pytest

import matrixmultiplicationcuda_amd as mm
import oracle

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def plan_for(precision, M=512, K=1024, bh=16, seed=11):
    A32 = mm.generate(M, K, 0.85, "blockrandom", bh=bh, bw=bh, seed=seed, dtype=np.float32)
    h = mm.bcsr_from_dense(A32.astype(np.float16), bh, 16)
    if precision != "f16":
        h = mm.HostBCSR(h.hdr, h.idx, oracle.bcsr_values_f32(A32, h.hdr, h.idx, bh, 16), h.block_rows,
                        h.block_cols, bh, 16)
    return mm.Plan(h, precision), M, K


def eager(plan, dB, M, N, stream):
    C = mm.DeviceArray((M, N), np.float32)
    C.fill_bytes(0xFF, stream)
    plan.spmm_ptr(dB.ptr, C.ptr, N, stream)
    stream.synchronize()
    out = C.numpy()
    C.free()
    return out


@pytest.mark.parametrize("precision", ["f16", "bf16", "fp8", "hilo"])
def test_graph_replays_equal_eager(precision):
    s = mm.Stream()
    plan, M, K = plan_for(precision)
    N = 256
    dB = mm.DeviceArray.from_numpy(mm.generate(K, N, 0.0, "random", seed=5))
    want = eager(plan, dB, M, N, s)  # also allocates an FP8 plan's workspace
    C = mm.DeviceArray((M, N), np.float32)
    g = mm.Graph.capture(s, lambda: plan.spmm_ptr(dB.ptr, C.ptr, N, s))
    other = mm.DeviceArray((M, N), np.float32)
    for rep in range(3):
        log(f"{precision} replay {rep}")
        C.fill_bytes(0xFF, s)
        s.synchronize()
        g.launch(s)
        plan.spmm_ptr(dB.ptr, other.ptr, N, s)  # eager work between replays
        s.synchronize()
        assert np.array_equal(C.numpy(), want), rep
        assert np.array_equal(other.numpy(), want), rep
    g.destroy()
    for a in (C, other, dB):
        a.free()
    plan.close()
    s.destroy()


def test_fp8_graph_survives_workspace_growth():
    s = mm.Stream()
    plan, M, K = plan_for("fp8")
    small, large = 64, 512
    dB_s = mm.DeviceArray.from_numpy(mm.generate(K, small, 0.0, "random", seed=1))
    dB_l = mm.DeviceArray.from_numpy(mm.generate(K, large, 0.0, "random", seed=2))
    want_s = eager(plan, dB_s, M, small, s)
    C = mm.DeviceArray((M, small), np.float32)
    g = mm.Graph.capture(s, lambda: plan.spmm_ptr(dB_s.ptr, C.ptr, small, s))
    want_l = eager(plan, dB_l, M, large, s)  # grows the workspace
    fresh, _, _ = plan_for("fp8")
    assert np.array_equal(want_l, eager(fresh, dB_l, M, large, s))
    fresh.close()
    for rep in range(2):
        log(f"fp8 replay after growth {rep}")
        C.fill_bytes(0xFF, s)
        g.launch(s)
        s.synchronize()
        assert np.array_equal(C.numpy(), want_s), rep
    g.destroy()
    for a in (C, dB_s, dB_l):
        a.free()
    plan.close()
    s.destroy()


def test_fp8_capture_needing_a_larger_workspace_is_refused():
    s = mm.Stream()
    plan, M, K = plan_for("fp8")
    N = 128
    dB = mm.DeviceArray.from_numpy(mm.generate(K, N, 0.0, "random", seed=3))
    C = mm.DeviceArray((M, N), np.float32)
    with pytest.raises(mm.BcsrError) as e:
        mm.Graph.capture(s, lambda: plan.spmm_ptr(dB.ptr, C.ptr, N, s))  # no eager call first
    assert e.value.status == -7
    # the stream is usable again and the plan still computes the product
    got = eager(plan, dB, M, N, s)
    fresh, _, _ = plan_for("fp8")
    assert np.array_equal(got, eager(fresh, dB, M, N, s))
    fresh.close()
    for a in (C, dB):
        a.free()
    plan.close()
    s.destroy()


def test_graph_begin_rejects_the_null_stream():
    assert mm.lib().bcsr_graph_begin(None) == -1
