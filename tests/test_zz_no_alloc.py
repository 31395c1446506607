"""bcsr_spmm_noalloc, the entry point that allocates nothing, through the C
ABI: its C equals bcsr_spmm_cfg's (the entry point that allocates the FP8
workspace) bit for bit, on two calls in a row, the second with a smaller
input; an FP8 product larger than the fixed scratch is refused with
BCSR_ERR_SCRATCH_LIMIT before any launch; 16-bit plans run unchanged."""
import ctypes
import sys

import numpy as np
import pytest

import matrixmultiplicationcuda_amd as mm

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def product(M, K, N, precision, seed, sparsity=0.9):
    h = mm.generate_bcsr(M, K, sparsity, "blockrandom", bh=16, bw=16, seed=seed)
    if precision == "f16":
        h = mm.HostBCSR(h.hdr, h.idx, h.data.astype(np.float16), h.block_rows, h.block_cols, 16, 16)
    B16 = mm.generate(K, N, 0.0, "random", bh=16, seed=seed + 1)
    return h, B16


def run_both(plan, B16, M, N, cfg=None):
    """(C of bcsr_spmm_cfg, C of bcsr_spmm_noalloc) on the same B."""
    lib = mm.lib()
    B = mm.DeviceArray.from_numpy(B16)
    outs = []
    for fn in (lib.bcsr_spmm_cfg, lib.bcsr_spmm_noalloc):
        C = mm.DeviceArray((M, N), np.float32)
        C.fill_bytes(0xFF)  # NaN: every element must be written
        rc = fn(plan._handle, B.ptr, C.ptr, N, ctypes.byref(cfg) if cfg is not None else None, None)
        assert rc == 0, mm.STATUS.get(rc)
        mm.synchronize()
        outs.append(C.numpy())
        C.free()
    B.free()
    return outs


@pytest.mark.parametrize("precision,slab_mags", [("fp8", False), ("fp8", True), ("f16", False)])
def test_noalloc_equals_spmm_on_two_calls(precision, slab_mags):
    """slab_mags: B's 16-row slabs scaled by 2^-6 .. 2^5, so FP8 chunks carry
    different slab scales (the two-MFMA path)."""
    assert mm.device_count() > 0
    mm.check(mm.lib().bcsr_set_device(0))
    # the second call is smaller (fewer rows, K and N), so stale scratch
    # contents of the first would show
    for (M, K, N, seed) in [(1024, 4096, 512, 11), (512, 1024, 96, 12)]:
        log(f"noalloc {precision} slab magnitudes {slab_mags}: M={M} K={K} N={N}")
        h, B16 = product(M, K, N, precision, seed)
        if slab_mags:
            rng = np.random.default_rng(seed)
            B16 = (B16.astype(np.float32) * np.repeat(2.0 ** rng.integers(-6, 6, K // 16), 16)[:, None]).astype(
                np.float16)
        plan = mm.Plan(h, precision)
        want, got = run_both(plan, B16, M, N)
        assert np.isfinite(got).all()
        assert np.array_equal(want, got), f"{precision} M={M} K={K} N={N}"
        plan.close()


def test_noalloc_refuses_beyond_its_scratch():
    mm.check(mm.lib().bcsr_set_device(0))
    limit = 16 << 20
    K, N = 16384, 1024 + 16  # K * N just over 16 MiB of quantised B
    assert K * N > limit
    log(f"noalloc limit: K={K} N={N}")
    h, _ = product(64, K, 16, "fp8", 5, sparsity=0.999)
    plan = mm.Plan(h, "fp8")
    B = mm.DeviceArray((16,), np.float16)  # never read: the call is refused first
    C = mm.DeviceArray((64, N), np.float32)
    rc = mm.lib().bcsr_spmm_noalloc(plan._handle, B.ptr, C.ptr, N, None, None)
    assert rc == -9 and mm.STATUS[rc] == "BCSR_ERR_SCRATCH_LIMIT"
    # the largest FP8 product it takes at this K still runs and matches
    N2 = 1024
    B2 = mm.generate(K, N2, 0.0, "random", bh=16, seed=6)
    want, got = run_both(plan, B2, 64, N2)
    assert np.array_equal(want, got)
    for a in (B, C):
        a.free()
    plan.close()
