"""Host threads calling the library at once, each on its own stream.

The C ABI's concurrency contract: calls of a 16-bit plan only read the plan
and may overlap on any streams and threads; the registry behind
bcsr_copy_to_device / the sparseMatrixMulTensor* entry points and
bcsr_release_device is shared and locked; launch variant 3 builds its panel
schedule on the first launch (under the plan's lock); an FP8 plan keeps one
quantised-B workspace, so each thread here owns its FP8 plan.  Every result
must equal the same call made alone, bit for bit.  (ctypes releases the GIL
for the duration of each call, so the threads do run the library at the same
time.)
"""
import ctypes
import sys
import threading

import numpy as np
import pytest

import matrixmultiplicationcuda_amd as mm

pytestmark = pytest.mark.gpu

THREADS = 8


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def run_threads(target, n=THREADS):
    """Run target(k, barrier) in n threads; re-raise the first failure."""
    barrier = threading.Barrier(n)
    errors = []

    def body(k):
        try:
            mm.check(mm.lib().bcsr_set_device(0))
            target(k, barrier)
        except BaseException as e:  # noqa: BLE001 -- reported below
            errors.append((k, e))
            barrier.abort()

    ts = [threading.Thread(target=body, args=(k,)) for k in range(n)]
    for t in ts:
        t.start()
    for t in ts:
        t.join(timeout=120)
        assert not t.is_alive(), "a thread did not finish"
    if errors:
        k, e = errors[0]
        raise AssertionError(f"thread {k}: {e!r}") from e


def f16_host(M, K, bh, sparsity, seed):
    pattern = {16: "blockrandom", 32: "pattern32by16", 64: "pattern64by16"}[bh]
    h = mm.generate_bcsr(M, K, sparsity, pattern, bh=bh, bw=16, seed=seed)
    return mm.HostBCSR(h.hdr, h.idx, h.data.astype(np.float16), h.block_rows, h.block_cols, bh, 16)


def test_threads_share_a_16bit_plan():
    """Eight threads, one fp16 plan, ten calls each on their own streams with
    their own B and C: every C equals the plan's result for that B alone."""
    M = K = 2048
    N = 256
    mm.check(mm.lib().bcsr_set_device(0))
    plan = mm.Plan(f16_host(M, K, 16, 0.9, seed=5), "f16")
    Bs = [mm.generate(K, N, 0.0, "random", bh=16, seed=100 + k) for k in range(THREADS)]
    want = []
    for k in range(THREADS):
        B = mm.DeviceArray.from_numpy(Bs[k])
        C = mm.DeviceArray((M, N), np.float32)
        plan.spmm_ptr(B.ptr, C.ptr, N)
        mm.synchronize()
        want.append(C.numpy())
        B.free()
        C.free()
    log(f"shared plan: nnzb {plan.info.nnzb}, {THREADS} threads x 10 calls")
    got = [None] * THREADS

    def work(k, barrier):
        st = mm.Stream()
        B = mm.DeviceArray.from_numpy(Bs[k])
        C = mm.DeviceArray((M, N), np.float32)
        barrier.wait()
        for _ in range(10):
            plan.spmm_ptr(B.ptr, C.ptr, N, st)
        st.synchronize()
        got[k] = C.numpy()
        B.free()
        C.free()
        st.destroy()

    run_threads(work)
    for k in range(THREADS):
        assert np.array_equal(got[k], want[k]), f"thread {k}"
    plan.close()


def test_threads_register_call_and_release_drop_in_matrices():
    """Each thread copies its own matrix to the device (the registry behind
    the reference-named entry points), calls its reference kernel name five
    times and releases it, all at once; each C equals a plan built from the
    same arrays."""
    M = K = 1024
    N = 128
    shapes = [16, 32, 64, 16, 32, 64, 16, 16]
    names = {16: "sparseMatrixMulTensor_v2", 32: "sparseMatrixMulTensor32x16", 64: "sparseMatrixMulTensor64x16"}
    hosts = [f16_host(M, K, shapes[k], 0.85, seed=200 + k) for k in range(THREADS)]
    Bs = [mm.generate(K, N, 0.0, "random", bh=16, seed=300 + k) for k in range(THREADS)]
    mm.check(mm.lib().bcsr_set_device(0))
    want = []
    for k in range(THREADS):
        plan = mm.Plan(hosts[k], "f16")
        want.append(plan.matmul(Bs[k]))
        plan.close()
    log(f"drop-in registry: {THREADS} threads, block heights {shapes}")
    got = [None] * THREADS
    lib = mm.lib()

    def work(k, barrier):
        h = hosts[k]
        st = mm.Stream()
        B = mm.DeviceArray.from_numpy(Bs[k])
        C = mm.DeviceArray((M, N), np.float32)
        data = np.ascontiguousarray(h.data)
        ghdr, gidx, gdata = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
        barrier.wait()
        mm.check(lib.bcsr_copy_to_device(h.hdr.ctypes.data_as(ctypes.c_void_p),
                                         h.idx.ctypes.data_as(ctypes.c_void_p),
                                         data.ctypes.data_as(ctypes.c_void_p), h.block_rows, h.block_cols,
                                         h.bh, 16, ctypes.byref(ghdr), ctypes.byref(gidx), ctypes.byref(gdata)))
        fn = getattr(lib, names[h.bh])
        for _ in range(5):
            mm.check(fn(ghdr, gidx, gdata, B.ptr, C.ptr, M, N, st.handle), names[h.bh])
        st.synchronize()
        got[k] = C.numpy()
        mm.check(lib.bcsr_release_device(ghdr))
        assert lib.bcsr_release_device(ghdr) == -6  # gone from the registry
        B.free()
        C.free()
        st.destroy()

    run_threads(work)
    for k in range(THREADS):
        assert np.array_equal(got[k], want[k]), f"thread {k} ({shapes[k]} x 16)"


def test_threads_race_the_first_panel_launch():
    """Launch variant 3 builds its panel schedule on the plan's first such
    launch; eight threads make that first launch together.  The schedule is
    built once and every C equals a later variant-3 launch alone."""
    M = K = 2048
    N = 256
    mm.check(mm.lib().bcsr_set_device(0))
    plan = mm.Plan(f16_host(M, K, 16, 0.9, seed=7), "f16")
    cfg = mm.make_cfg(tile_n=128, variant=3)
    assert plan.supports(N, cfg)
    Bs = [mm.generate(K, N, 0.0, "random", bh=16, seed=400 + k) for k in range(THREADS)]
    log(f"first panel launch from {THREADS} threads, nnzb {plan.info.nnzb}")
    got = [None] * THREADS

    def work(k, barrier):
        st = mm.Stream()
        B = mm.DeviceArray.from_numpy(Bs[k])
        C = mm.DeviceArray((M, N), np.float32)
        barrier.wait()
        plan.spmm_ptr(B.ptr, C.ptr, N, st, cfg)
        st.synchronize()
        got[k] = C.numpy()
        B.free()
        C.free()
        st.destroy()

    run_threads(work)
    for k in range(THREADS):
        B = mm.DeviceArray.from_numpy(Bs[k])
        C = mm.DeviceArray((M, N), np.float32)
        plan.spmm_ptr(B.ptr, C.ptr, N, None, cfg)
        mm.synchronize()
        assert np.array_equal(got[k], C.numpy()), f"thread {k}"
        B.free()
        C.free()
    plan.close()


def test_threads_each_with_an_fp8_plan():
    """Each thread owns an FP8 plan of the same matrix and calls it with a
    rising N (its workspace grows twice); every C equals the same calls made
    by one thread alone."""
    M = K = 1024
    Ns = (64, 256, 512)
    h32 = mm.generate_bcsr(M, K, 0.9, "blockrandom", bh=16, bw=16, seed=9)
    Bs = {N: mm.generate(K, N, 0.0, "random", bh=16, seed=500 + N) for N in Ns}
    mm.check(mm.lib().bcsr_set_device(0))
    ref = mm.Plan(h32, "fp8")
    want = {N: ref.matmul(Bs[N]) for N in Ns}
    ref.close()
    log(f"FP8 plans: {THREADS} threads, N {Ns}")
    got = [None] * THREADS
    plans = [mm.Plan(h32, "fp8") for _ in range(THREADS)]

    def work(k, barrier):
        st = mm.Stream()
        outs = {}
        barrier.wait()
        for N in Ns:
            B = mm.DeviceArray.from_numpy(Bs[N])
            C = mm.DeviceArray((M, N), np.float32)
            for _ in range(3):
                plans[k].spmm_ptr(B.ptr, C.ptr, N, st)
            st.synchronize()
            outs[N] = C.numpy()
            B.free()
            C.free()
        got[k] = outs
        st.destroy()

    run_threads(work)
    for k in range(THREADS):
        for N in Ns:
            assert np.array_equal(got[k][N], want[N]), f"thread {k}, N = {N}"
    for p in plans:
        p.close()
