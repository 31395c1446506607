"""GPU parity of the CSR product (sparseMatrixMult1Co's computation), the
hybrid HCSR sequence of the reference's legacy driver and the dense rocBLAS
comparison row."""
import sys

import numpy as np
import pytest

import matrixmultiplicationcuda_amd as mm
import oracle

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def dev(a):
    return mm.DeviceArray.from_numpy(np.ascontiguousarray(a))


def csr_dev(A16):
    hdr, idx, data = oracle.csr_from_dense(A16.view(np.uint16))
    d_idx = dev(idx if len(idx) else np.zeros(1, np.int32))
    d_data = dev(data if len(data) else np.zeros(1, np.uint16))
    return hdr, idx, data, dev(hdr), d_idx, d_data


@pytest.mark.parametrize("M,K,N,sp", [(64, 64, 16, 0.9), (96, 160, 48, 0.7), (256, 128, 272, 0.95),
                                      (128, 512, 512, 0.99), (32, 48, 256, 1.0)])
def test_csr_kernel_bit_exact(M, K, N, sp):
    mm.check(mm.lib().bcsr_set_device(0))
    A16 = mm.generate(M, K, sp, "random", seed=M + N)
    B16 = mm.generate(K, N, 0.0, "random", seed=7)
    hdr, idx, data, dh, di, dd = csr_dev(A16)
    log(f"csr M={M} K={K} N={N} nnz={len(idx)}")
    dB = dev(B16)
    dC = mm.DeviceArray((M, N), np.float32)
    dC.fill_bytes(0xFF)
    mm.check(mm.lib().bcsr_csr_spmm(dh.ptr, di.ptr, dd.ptr, dB.ptr, dC.ptr, M, K, N, 0, None))
    mm.synchronize()
    want = oracle.csr_spmm_f32(hdr, idx, data, B16.astype(np.float32))
    assert np.array_equal(dC.numpy(), want)
    # accumulate mode adds onto what C holds
    X = np.random.default_rng(1).standard_normal((M, N)).astype(np.float32)
    dX = dev(X)
    mm.check(mm.lib().bcsr_csr_spmm(dh.ptr, di.ptr, dd.ptr, dB.ptr, dX.ptr, M, K, N, 1, None))
    mm.synchronize()
    assert np.array_equal(dX.numpy(), X + want)


def test_reference_names_and_errors():
    mm.check(mm.lib().bcsr_set_device(0))
    n = 128
    A16 = mm.generate(n, n, 0.9, "random", seed=3)
    B16 = mm.generate(n, n, 0.0, "random", seed=4)
    hdr, idx, data, dh, di, dd = csr_dev(A16)
    dB = dev(B16)
    dC = mm.DeviceArray((n, n), np.float32)
    mm.check(mm.lib().sparseMatrixMult1Co(dh.ptr, di.ptr, dd.ptr, dB.ptr, dC.ptr, n, None))
    P = np.ones((n, n), np.float32)
    dP = dev(P)
    mm.check(mm.lib().addMatrices(dC.ptr, dP.ptr, n, None))
    mm.synchronize()
    want = oracle.csr_spmm_f32(hdr, idx, data, B16.astype(np.float32)) + np.float32(1.0)
    assert np.array_equal(dC.numpy(), want)
    # sparseMatrixMult1: the running sum starts from what C holds
    X = np.random.default_rng(2).standard_normal((n, n)).astype(np.float32) * np.float32(1e3)
    dX = dev(X)
    mm.check(mm.lib().sparseMatrixMult1(dh.ptr, di.ptr, dd.ptr, dB.ptr, dX.ptr, n, None))
    mm.synchronize()
    assert np.array_equal(dX.numpy(), oracle.csr_spmm_f32(hdr, idx, data, B16.astype(np.float32), C0=X))
    assert mm.lib().bcsr_csr_spmm(dh.ptr, di.ptr, dd.ptr, dB.ptr, dC.ptr, n, n, 16, 3, None) == -1
    assert mm.lib().bcsr_csr_spmm(dh.ptr, di.ptr, dd.ptr, dB.ptr, dC.ptr, n, n, 40, 0, None) == -3
    assert mm.lib().bcsr_csr_spmm(dh.ptr, di.ptr, dd.ptr, dB.ptr, dC.ptr, -1, n, n, 0, None) == -1


@pytest.mark.parametrize("threshold", [0.0, 0.2, 0.4, 0.6, 0.8])
def test_hybrid_hcsr(threshold):
    """16 x 16 blocks with element density >= threshold go to the BCSR plan,
    the rest to CSR (formats/HCSRMatrix.cpp:10-36); C = BCSR part + CSR part."""
    mm.check(mm.lib().bcsr_set_device(0))
    M, K, N = 256, 384, 128
    # blocks of varied density: blockrandom_rho with in-block density 0.5
    A32 = mm.generate(M, K, 0.6, "blockrandom_rho", bh=16, bw=16, seed=11, inblock_density=0.5,
                      dtype=np.float32)
    A32 += mm.generate(M, K, 0.97, "random", seed=12, dtype=np.float32)  # scattered singletons
    A16 = A32.astype(np.float16)
    bits = A16.view(np.uint16)
    dens = ((bits & 0x7FFF) != 0).reshape(M // 16, 16, K // 16, 16).mean(axis=(1, 3))
    dense_part = np.repeat(np.repeat(dens >= threshold, 16, 0), 16, 1)
    A_b = np.where(dense_part, A16, np.float16(0))
    A_c = np.where(dense_part, np.float16(0), A16)
    h = mm.bcsr_from_dense(A_b, 16, 16)
    plan = mm.Plan(h, "f16")
    hdr, idx, data, dh, di, dd = csr_dev(A_c)
    log(f"threshold {threshold}: bcsr nnzb {h.nnzb}, csr nnz {len(idx)}")
    B16 = mm.generate(K, N, 0.0, "random", seed=13)
    dB = dev(B16)
    dC = mm.DeviceArray((M, N), np.float32)
    dC.fill_bytes(0xFF)
    mm.check(mm.lib().bcsr_hybrid_spmm(plan._handle, dh.ptr, di.ptr, dd.ptr, dB.ptr, dC.ptr, N, None))
    mm.synchronize()
    C = dC.numpy().astype(np.float64)
    ref = oracle.matmul_dense_ref(bits, B16.view(np.uint16)).astype(np.float64)
    absAB = np.abs(A16.astype(np.float64)) @ np.abs(B16.astype(np.float64))
    assert np.all(np.abs(C - ref) <= 1e-3 * absAB + 1e-30)
    assert np.all(C[absAB == 0] == 0)
    plan.close()


@pytest.mark.parametrize("M,K,N", [(256, 512, 128), (1024, 1024, 256)])
def test_dense_gemm_row(M, K, N):
    mm.check(mm.lib().bcsr_set_device(0))
    A16 = mm.generate(M, K, 0.0, "random", seed=21)
    B16 = mm.generate(K, N, 0.0, "random", seed=22)
    dA, dB = dev(A16), dev(B16)
    dC = mm.DeviceArray((M, N), np.float32)
    mm.check(mm.lib().bcsr_dense_gemm_f16(dA.ptr, dB.ptr, dC.ptr, M, K, N, None))
    mm.synchronize()
    ref = A16.astype(np.float64) @ B16.astype(np.float64)
    absAB = np.abs(A16.astype(np.float64)) @ np.abs(B16.astype(np.float64))
    err = np.abs(dC.numpy() - ref)
    log(f"dense {M}x{K}x{N}: max err / |A||B| {np.max(err / absAB):.2e}")
    assert np.all(err <= 1e-3 * absAB)
