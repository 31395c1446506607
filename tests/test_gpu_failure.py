"""Failure paths at the boundary on the GPU: an allocation the device cannot
satisfy comes back as BCSR_ERR_OUT_OF_MEMORY (with the HIP error kept for
bcsr_last_hip_error), and the library keeps working after it."""
import ctypes

import numpy as np
import pytest

import matrixmultiplicationcuda_amd as mm
import oracle

pytestmark = pytest.mark.gpu


def test_failed_allocation_is_reported_and_recoverable():
    lib = mm.lib()
    mm.check(lib.bcsr_set_device(0))
    p = ctypes.c_void_p()
    # 1 PiB: far beyond the device's memory, refused by the runtime
    assert lib.bcsr_dev_malloc(ctypes.byref(p), 1 << 50) == -5
    assert not p.value
    assert lib.bcsr_last_hip_error() != 0
    # the device and the library still work
    A = mm.generate(256, 256, 0.8, "blockrandom", bh=16, seed=3)
    B16 = mm.generate(256, 64, 0.0, "random", seed=4)
    h = mm.bcsr_from_dense(A, 16, 16)
    plan = mm.Plan(h, "f16")
    dB = mm.DeviceArray.from_numpy(B16)
    dC = mm.DeviceArray((256, 64), np.float32)
    plan.spmm_ptr(dB.ptr, dC.ptr, 64)
    mm.synchronize()
    want = oracle.matmul_dense_ref(A.view(np.uint16), B16.view(np.uint16))
    absAB = np.abs(A.astype(np.float64)) @ np.abs(B16.astype(np.float64))
    assert np.all(np.abs(dC.numpy() - want) <= 1e-3 * absAB + 1e-30)
    plan.close()
