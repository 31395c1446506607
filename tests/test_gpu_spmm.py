"""GPU parity of the fp16 BCSR x dense kernels against the oracle.

Every case goes through the C ABI (libbcsr_mi355x.so); the oracle's BCSR-order
fp32 product is the reference CPU result (bit-identical to cpu_matmul_ref,
nvbench_main.cu:78-91).  Bound per entry: 1e-3 * (|A|.|B|)_ij, exact zero
where (|A|.|B|)_ij = 0, plus the reference's own |diff| <= 1e-2
(nvbench_main.cu:413-420).
"""
import ctypes
import functools
import sys

import numpy as np
import pytest

import matrixmultiplicationcuda_amd as mm
import oracle
from helpers import Case, assert_close, cpu_reference, make_case

pytestmark = pytest.mark.gpu


def _gpu():
    assert mm.device_count() > 0, "GPU tests need a visible MI355X"


def run_plan(case, cfg=None, precision="f16"):
    _gpu()
    plan = mm.Plan(case.h, precision)
    B = mm.DeviceArray.from_numpy(case.B16)
    C = mm.DeviceArray((case.M, case.N), np.float32)
    C.fill_bytes(0xFF)  # NaN pattern: every entry must be written
    plan.spmm(B, C, cfg=cfg)
    mm.synchronize()
    out = C.numpy()
    plan.close()
    B.free()
    C.free()
    return out


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_small_case_against_dense_cpu_matmul_ref():
    """The small reference case: 1024^2 blockrandom sp 0.9 16x16 seed 123, B 1024x256 seed 456."""
    case = make_case(1024, 1024, 256, 0.9, "blockrandom", 16, 16)
    assert case.h.nnzb == 400
    log("1024^2 blockrandom 16x16 N=256: nnzb=400")
    C = run_plan(case)
    C_dense = oracle.matmul_dense_ref(case.A16.view(np.uint16), case.B16.view(np.uint16))
    C_cpu, Cabs = cpu_reference(case)
    assert np.array_equal(C_dense, C_cpu)  # the oracle's two forms agree bit for bit
    rel = assert_close(C, C_cpu, Cabs, what="1024^2 blockrandom")
    log(f"1024^2 blockrandom max err / (|A||B|) = {rel:.3e}")


@pytest.mark.parametrize("bh,pattern", [(16, "blockrandom"), (32, "pattern32by16"),
                                        (64, "pattern64by16")])
@pytest.mark.parametrize("N", [16, 48, 64, 128, 512])
def test_block_shapes_and_widths(bh, pattern, N):
    case = make_case(512, 768, N, 0.8, pattern, bh, 16)
    log(f"shape bh={bh} N={N} nnzb={case.h.nnzb}")
    C = run_plan(case)
    C_cpu, Cabs = cpu_reference(case)
    assert_close(C, C_cpu, Cabs, what=f"bh={bh} N={N}")


@functools.lru_cache(maxsize=None)
def launch_case(bh, long_rows):
    """1024 x 2048, N = 512; long_rows (sparsity 0.3) gives block-rows of
    ~45-90 pairs, exercising long pipelines and odd tails."""
    pattern = {16: "blockrandom", 32: "pattern32by16", 64: "pattern64by16"}[bh]
    case = make_case(1024, 2048, 512, 0.3 if long_rows else 0.85, pattern, bh, 16)
    C_cpu, Cabs = cpu_reference(case)
    return case, C_cpu, Cabs


@pytest.mark.parametrize("wpb", [1, 4])
@pytest.mark.parametrize("variant", [0, 1, 2])
@pytest.mark.parametrize("tile_n", [32, 64, 128, 256])
@pytest.mark.parametrize("xcd_map", [0, 1, 2, 3, 4])
@pytest.mark.parametrize("bh", [16, 32, 64])
def test_launch_configs(tile_n, xcd_map, bh, variant, wpb):
    case, C_cpu, Cabs = launch_case(bh, variant > 0)
    log(f"cfg tile={tile_n} map={xcd_map} bh={bh} variant={variant} wpb={wpb}")
    rs, epl = bh // 16, tile_n // 16
    if (epl == 16 and rs != 1) or (variant == 1 and rs * epl > 16):
        # shapes whose register tiles do not fit are not instantiated
        with pytest.raises(mm.BcsrError) as e:
            run_plan(case, cfg=mm.make_cfg(tile_n, xcd_map, variant, wpb))
        assert e.value.status == -7
        return
    C = run_plan(case, cfg=mm.make_cfg(tile_n, xcd_map, variant, wpb))
    assert_close(C, C_cpu, Cabs, what=f"tile={tile_n} map={xcd_map} variant={variant} wpb={wpb}")


@pytest.mark.parametrize("pattern", mm.REFERENCE_PATTERNS + mm.EXTRA_PATTERNS)
def test_all_patterns(pattern):
    case = make_case(640, 512, 96, 0.7, pattern, 16, 16, gen_bh=32 if pattern == "blockdiagonal" else 16,
                     inblock_density=0.4)
    log(f"pattern {pattern} nnzb={case.h.nnzb}")
    C = run_plan(case)
    C_cpu, Cabs = cpu_reference(case)
    assert_close(C, C_cpu, Cabs, what=pattern)


@pytest.mark.parametrize("bh,bw", [(32, 32), (48, 16), (128, 32), (16, 64)])
def test_wide_and_tall_blocks(bh, bw):
    """Blocks other than bh x 16 are cut into device blocks of 16 columns and
    16/32/64 rows."""
    case = make_case(384, 512, 64, 0.75, "blockrandom", bh, bw, gen_bh=16)
    log(f"blocks {bh}x{bw} nnzb={case.h.nnzb}")
    C = run_plan(case)
    C_cpu, Cabs = cpu_reference(case)
    assert_close(C, C_cpu, Cabs, what=f"{bh}x{bw}")


def test_empty_matrix_writes_zeros():
    case = make_case(256, 256, 64, 1.0, "blockrandom", 16, 16)
    assert case.h.nnzb == 0
    C = run_plan(case)
    assert np.all(C == 0)


def test_single_block_and_empty_rows():
    A = np.zeros((128, 256), dtype=np.float16)
    A[16:32, 48:64] = np.arange(256, dtype=np.float16).reshape(16, 16) / 256
    case = make_case(128, 256, 32, 0.9)
    case.A16 = A
    case.h = mm.bcsr_from_dense(A, 16, 16)
    assert case.h.nnzb == 1
    C = run_plan(case)
    C_cpu, Cabs = cpu_reference(case)
    assert_close(C, C_cpu, Cabs, what="single block")
    assert np.all(C[:16] == 0) and np.all(C[32:] == 0)


def test_identity_blocks_select_rows_of_B():
    """Closed form: A = identity made of 16x16 identity blocks -> C = B exactly."""
    _gpu()
    n = 512
    A = np.eye(n, dtype=np.float16)
    h = mm.bcsr_from_dense(A, 16, 16)
    B = mm.generate(n, 128, 0.0, "random", seed=7)
    plan = mm.Plan(h)
    assert np.array_equal(plan.matmul(B), B.astype(np.float32))
    plan.close()


def test_drop_in_entry_points():
    """bcsr_copy_to_device + every sparseMatrixMulTensor* name of the reference."""
    _gpu()
    lib = mm.lib()
    for bh, pattern in [(16, "blockrandom"), (32, "pattern32by16"), (64, "pattern64by16")]:
        case = make_case(512, 512, 128, 0.8, pattern, bh, 16)
        h = case.h
        ghdr, gidx, gdata = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
        data = np.ascontiguousarray(h.data)
        mm.check(lib.bcsr_copy_to_device(h.hdr.ctypes.data_as(ctypes.c_void_p),
                                         h.idx.ctypes.data_as(ctypes.c_void_p),
                                         data.ctypes.data_as(ctypes.c_void_p), h.block_rows,
                                         h.block_cols, bh, 16, ctypes.byref(ghdr),
                                         ctypes.byref(gidx), ctypes.byref(gdata)))
        # the device hdr is the reference prefix array, idx the block-columns
        hdr_back = np.empty_like(h.hdr)
        idx_back = np.empty_like(h.idx)
        mm.check(lib.bcsr_memcpy_dtoh(hdr_back.ctypes.data_as(ctypes.c_void_p), ghdr, hdr_back.nbytes))
        mm.check(lib.bcsr_memcpy_dtoh(idx_back.ctypes.data_as(ctypes.c_void_p), gidx, idx_back.nbytes))
        assert np.array_equal(hdr_back, h.hdr) and np.array_equal(idx_back, h.idx)
        B = mm.DeviceArray.from_numpy(case.B16)
        C_cpu, Cabs = cpu_reference(case)
        stream = mm.Stream()
        s = stream.handle
        for name, ebh in list(mm.ENTRY_POINTS.items()) + list(mm.ADAPTIVE_ENTRY_POINTS.items()):
            C = mm.DeviceArray((case.M, case.N), np.float32)
            C.fill_bytes(0xFF)
            mm.synchronize()
            fn = getattr(lib, name)
            args = [ghdr, gidx, gdata, B.ptr, C.ptr, case.M, case.N]
            if name in mm.ADAPTIVE_ENTRY_POINTS:
                args.append(2)
            rc = fn(*args, s)
            if ebh != bh:
                assert rc == -2, f"{name} accepted a {bh}x16 matrix"
                continue
            assert rc == 0, f"{name}: {mm.STATUS.get(rc)}"
            stream.synchronize()
            assert_close(C.numpy(), C_cpu, Cabs, what=name)
            C.free()
        C = mm.DeviceArray((case.M, case.N), np.float32)
        if bh == 16:
            # sparseMatrixMulTensor1 is square n x n: M = N = n
            rc = lib.sparseMatrixMulTensor1(ghdr, gidx, gdata, B.ptr, C.ptr, case.N, s)
            assert rc == -3  # n = 128 != M = 512
        # wrong M, unregistered hdr, misaligned B, N % 16 != 0
        name = {16: "sparseMatrixMulTensor", 32: "sparseMatrixMulTensor32x16",
                64: "sparseMatrixMulTensor64x16"}[bh]
        fn = getattr(lib, name)
        assert fn(ghdr, gidx, gdata, B.ptr, C.ptr, case.M + 16, case.N, s) == -3
        assert fn(gidx, gidx, gdata, B.ptr, C.ptr, case.M, case.N, s) == -6
        assert fn(ghdr, gidx, gdata, B.offset(2), C.ptr, case.M, case.N, s) == -3
        assert fn(ghdr, gidx, gdata, B.ptr, C.ptr, case.M, case.N - 8, s) == -3
        stream.synchronize()
        stream.destroy()
        C.free()
        B.free()
        mm.check(lib.bcsr_release_device(ghdr))
        assert lib.bcsr_release_device(ghdr) == -6
    mm.synchronize()


def test_target_scale_sampled_rows_and_checksum():
    """16384^2, 16x16, 5% blocks, N=512: sampled block-rows against the oracle
    and the row-sum identity C.1 = A.(B.1) for every row."""
    _gpu()
    M = K = 16384
    N = 512
    case = make_case(M, K, N, 0.95, "blockrandom", 16, 16)
    log(f"target nnzb={case.h.nnzb}")
    plan = mm.Plan(case.h)
    Cg = plan.matmul(case.B16)
    rng = np.random.default_rng(5)
    brs = sorted(set(rng.integers(0, M // 16, 48).tolist()) | {0, M // 16 - 1})
    for br in brs:
        C_cpu, Cabs = cpu_reference(case, br, br + 1)
        sl = slice(br * 16, br * 16 + 16)
        assert_close(Cg[sl], C_cpu[sl], Cabs[sl], what=f"block-row {br}")
    # row sums: sum_j C[i, j] = sum_k A[i, k] * (sum_j B[k, j]) in fp64
    bsum = case.B16.astype(np.float64).sum(axis=1)
    babs = np.abs(case.B16.astype(np.float64)).sum(axis=1)
    h = case.h
    rows_ref = np.zeros(M)
    rows_abs = np.zeros(M)
    data = h.data.astype(np.float64)
    for br in range(h.block_rows):
        for k in range(h.hdr[br], h.hdr[br + 1]):
            c = h.idx[k]
            rows_ref[br * 16:(br + 1) * 16] += data[k] @ bsum[c * 16:(c + 1) * 16]
            rows_abs[br * 16:(br + 1) * 16] += np.abs(data[k]) @ babs[c * 16:(c + 1) * 16]
    got = Cg.astype(np.float64).sum(axis=1)
    assert np.all(np.abs(got - rows_ref) <= 1e-3 * rows_abs + 1e-6)
    plan.close()


@pytest.mark.parametrize("bh,density,N,prec,want", [
    (16, 0.05, 512, "f16", "bcsr_spmm_pair<1, 8, true, f16>"),     # the target: 128-column tiles
    (16, 0.20, 512, "f16", "bcsr_spmm_pair<1, 16, true, f16>"),    # dense: 256-column tiles
    (16, 0.20, 512, "bf16", "bcsr_spmm_pair<1, 16, true, bf16>"),
    (16, 0.20, 512, "hilo", "bcsr_spmm_pair<1, 8, true, hilo>"),   # hi+lo keeps 128
    (64, 0.05, 64, "f16", "bcsr_spmm_pair<4, 2, true, f16>"),      # 64-row blocks, N <= 128: 32 columns
    (64, 0.05, 256, "f16", "bcsr_spmm_pair<4, 4, true, f16>"),
    (64, 0.05, 512, "f16", "bcsr_spmm_pair<4, 8, false, f16>"),
    (32, 0.20, 128, "fp8", "bcsr_spmm_fp8<2, 4>"),                 # FP8, 32-row blocks, N = 128: 64 columns
    (16, 0.20, 512, "fp8", "bcsr_spmm_fp8<1, 16>"),                # FP8 dense: 256 columns
    (16, 0.05, 512, "fp8", "bcsr_spmm_fp8<1, 8>"),
    (32, 0.20, 512, "fp8", "bcsr_spmm_fp8<2, 16>"),                # FP8 dense 32-row blocks: 256 columns
    (16, 0.05, 384, "f16", "bcsr_spmm_pair<1, 8, true, f16>"),     # three tiles: chunked XCD mapping
    (16, 0.20, 1280, "f16", "bcsr_spmm_pair<1, 8, true, f16>"),    # N > 1024: 128 columns
])
def test_automatic_choices(bh, density, N, prec, want):
    """The automatic launch picks the tile widths measured fastest (kernel
    names carry the block sub-tiles and 16-column groups per wave), and a
    product with few block-rows keeps 128-column tiles."""
    pattern = {16: "blockrandom", 32: "pattern32by16", 64: "pattern64by16"}[bh]
    M, K = 16384, 2048  # >= 16384 rows: the wide-tile rules apply
    h = mm.generate_bcsr(M, K, 1.0 - density, pattern, bh=bh, bw=16, seed=5)
    if prec == "f16":
        h = mm.HostBCSR(h.hdr, h.idx, h.data.astype(np.float16), h.block_rows, h.block_cols, bh, 16)
    plan = mm.Plan(h, prec)
    assert plan.kernel_name(N) == want
    plan.close()
    if want.endswith("16, true, f16>"):
        few = mm.generate_bcsr(2048, K, 1.0 - density, pattern, bh=bh, bw=16, seed=5)
        few = mm.HostBCSR(few.hdr, few.idx, few.data.astype(np.float16), few.block_rows, few.block_cols, bh, 16)
        plan = mm.Plan(few, prec)
        # 128 block-rows: at 256 or 128 columns fewer than 640 work items,
        # so the automatic choice halves the tile to 64 columns
        assert plan.kernel_name(N) == "bcsr_spmm_pair<1, 4, true, f16>"
        plan.close()


@pytest.mark.parametrize("rows,N,prec,want", [
    (512, 512, "f16", "bcsr_spmm_pair<1, 2, true, f16>"),    # 32 block-rows: 32 columns (512 items)
    (2560, 256, "f16", "bcsr_spmm_pair<1, 4, true, f16>"),   # 160 block-rows x 4 tiles of 64 = 640 items
    (2560, 512, "f16", "bcsr_spmm_pair<1, 8, true, f16>"),   # 160 x 4 tiles of 128: enough items
    (1024, 512, "fp8", "bcsr_spmm_fp8<1, 2>"),     # 64 block-rows: 32 columns (1024 items)
])
def test_automatic_tile_for_few_work_items(rows, N, prec, want):
    """Products with fewer than 640 (device block-row, column tile) items at
    the usual width take narrower tiles, down to 32 columns; the results
    stay equal to the oracle's."""
    h = mm.generate_bcsr(rows, 1024, 0.9, "blockrandom", bh=16, bw=16, seed=7)
    if prec == "f16":
        h = mm.HostBCSR(h.hdr, h.idx, h.data.astype(np.float16), h.block_rows, h.block_cols, 16, 16)
    plan = mm.Plan(h, prec)
    assert plan.kernel_name(N) == want
    if prec == "f16":
        B16 = mm.generate(1024, N, 0.0, "random", bh=16, seed=8)
        A16 = np.zeros((rows, 1024), np.float16)
        case = Case(A16, B16, h, 16, 16)
        C_cpu, Cabs = cpu_reference(case)
        assert_close(plan.matmul(B16), C_cpu, Cabs, what=want)
    plan.close()


def test_b_beyond_2gib_uses_64bit_addressing():
    """B of 1048704 x 1024 fp16 (2.0 GiB + 800 KiB): 32-bit buffer offsets no
    longer reach it, so the automatic choice takes the 64-bit global-load
    kernel.  Four block-rows hold blocks at block-columns near both ends of K
    (the last ones lie beyond byte offset 2^31 of B); the product is checked
    against the oracle on the compacted slabs those blocks use."""
    K, N, M = 1048704, 1024, 64
    bc = K // 16
    rng = np.random.default_rng(9)
    cols = [sorted({0, r + 1, 40000 + r, bc - 1 - r}) for r in range(M // 16)]
    hdr = np.zeros(M // 16 + 1, np.int32)
    hdr[1:] = np.cumsum([len(c) for c in cols])
    idx = np.concatenate([np.asarray(c, np.int32) for c in cols])
    data = rng.uniform(-1, 1, (len(idx), 16, 16)).astype(np.float16)
    h = mm.HostBCSR(hdr, idx, data, M // 16, bc, 16, 16)
    used = np.unique(idx)
    B = np.zeros((K, N), np.float16)
    slabs = rng.uniform(-1, 1, (len(used), 16, N)).astype(np.float16)
    for s, c in zip(slabs, used):
        B[16 * c:16 * c + 16] = s
    assert K * N * 2 > 2**31 - 1
    print(f"K={K} N={N} B bytes={K * N * 2} nnzb={len(idx)}", flush=True)
    dB = mm.DeviceArray.from_numpy(B)
    del B
    plan = mm.Plan(h, "f16")
    name = plan.kernel_name(N)
    assert name.startswith("bcsr_spmm_pair_global"), name
    C = mm.DeviceArray((M, N), np.float32)
    C.fill_bytes(0xFF)
    plan.spmm_ptr(dB.ptr, C.ptr, N)
    mm.synchronize()
    got = C.numpy()
    # the same product on the compacted slabs
    remap = {int(c): i for i, c in enumerate(used)}
    small = Case(np.zeros((M, 16 * len(used)), np.float16), slabs.reshape(-1, N),
                 mm.HostBCSR(hdr, np.asarray([remap[int(c)] for c in idx], np.int32), data, M // 16,
                             len(used), 16, 16), 16, 16)
    C_cpu, Cabs = cpu_reference(small)
    assert_close(got, C_cpu, Cabs, what="B > 2 GiB")
    for a in (C, dB):
        a.free()
    plan.close()
