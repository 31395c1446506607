"""GPU parity of the storage-precision modes (bf16, hi+lo fp16, FP8 E4M3).

These modes have no reference counterpart (the reference is fp16 only).  Each
is checked twice:
  1. against the exact product of the values the mode stores (the oracle's
     restatement of the quantisers, dequantised, multiplied in fp64): the
     16-bit modes may differ only by fp32 accumulation, bound
     1e-4 * (|Aq|.|Bq|)_ij.  FP8 gets 2^-11: the scaled 16x16x128 FP8 MFMA
     does not add its products in IEEE fp32 (on MI355X its result differs
     from the exact sum of the exact products by up to 1.2e-4 of their
     magnitude sum; no fixed-alignment truncation model reproduces it);
  2. against the FP64 product of the unquantised A (fp32 generator values)
     and B (the fp16 input), with each mode's error bound per entry relative
     to (|A|.|B|)_ij: fp16 2^-11, bf16 2^-8 + 2^-8 (A and the B operand are
     rounded), hi+lo 5e-5; FP8 (2^-4 per operand above E4M3's subnormal
     range) max <= 0.095, mean <= 0.0126 and normwise <= 0.054 on
     log-uniform values (1.5x the measured errors).
The measured maxima are printed per mode.
"""
import sys

import numpy as np
import pytest

import matrixmultiplicationcuda_amd as mm
import oracle

pytestmark = pytest.mark.gpu

BOUND_VS_FP64 = {"f16": 2.0**-11 + 1e-6, "bf16": 2.0**-8 * 2 + 2.0**-16 + 1e-6, "hilo": 5e-5}
PATTERN = {16: "blockrandom", 32: "pattern32by16", 64: "pattern64by16"}


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def case(M, K, N, bh, sparsity=0.9, dist="loguniform", seed=321):
    A32 = mm.generate(M, K, sparsity, PATTERN[bh], bh=bh, bw=bh, seed=seed, value_dist=dist,
                      lo=1e-3, hi=1e2, dtype=np.float32)
    h = mm.bcsr_from_dense(A32.astype(np.float16), bh, 16)
    data32 = oracle.bcsr_values_f32(A32, h.hdr, h.idx, bh, 16)
    h32 = mm.HostBCSR(h.hdr, h.idx, data32, h.block_rows, h.block_cols, bh, 16)
    B16 = mm.generate(K, N, 0.0, "random", bh=16, seed=456)
    return h32, data32, B16


def run_mode(h32, B16, mode, cfg=None):
    assert mm.device_count() > 0
    plan = mm.Plan(h32, mode)
    C = plan.matmul(B16, cfg=cfg)
    plan.close()
    return C


@pytest.mark.parametrize("mode", ["f16", "bf16", "hilo", "fp8"])
@pytest.mark.parametrize("bh,N", [(16, 256), (16, 96), (32, 128), (64, 64), (16, 512)])
def test_precision_mode(mode, bh, N):
    M, K = 512, 1024
    h32, data32, B16 = case(M, K, N, bh)
    log(f"mode={mode} bh={bh} N={N} nnzb={h32.nnzb}")
    C = run_mode(h32, B16, mode).astype(np.float64)
    Aq = oracle.dequantized_blocks(mode, data32, bh, 16, h32.hdr, h32.idx)
    Aq32 = np.asarray(Aq, dtype=np.float32)
    assert np.array_equal(Aq32.astype(np.float64), np.asarray(Aq, dtype=np.float64))
    Bq = oracle.quantized_B(mode, B16.view(np.uint16))
    Cq, Cqabs = oracle.bcsr_spmm_f64(h32.hdr, h32.idx, Aq32, bh, 16, Bq, threads=8)
    err_q = np.abs(C - Cq)
    tight = 2.0**-11 if mode == "fp8" else 1e-4
    log(f"  {mode}: max |C - Cq| / (|Aq||Bq|) {np.max(err_q / np.maximum(Cqabs, 1e-30)):.3e}")
    assert np.all(err_q <= tight * Cqabs + 1e-30), f"{mode}: kernel differs from its stored values"
    Bf = B16.astype(np.float32)
    C64, Cabs = oracle.bcsr_spmm_f64(h32.hdr, h32.idx, data32, bh, 16, Bf, threads=8)
    nz = Cabs > 0
    rel = np.abs(C - C64)[nz] / Cabs[nz]
    normwise = np.linalg.norm(C - C64) / np.linalg.norm(C64)
    log(f"  {mode}: max rel |A||B| {rel.max():.3e} mean {rel.mean():.3e} normwise {normwise:.3e}")
    if mode == "fp8":
        # E4M3 keeps 2^-4 relative precision per operand (A and B are both
        # quantised); with 16-64 blocks per row an entry sums few terms, so
        # its error reaches 5-6e-2 (the full 16384^2 workload, ~51 blocks
        # per row, averages it down to 1.9e-2).  Bounds: 1.5x the measured
        # 6.3e-2 / 8.4e-3 / 3.6e-2.
        assert rel.max() <= 0.095, f"fp8 max error vs FP64 {rel.max():.3e}"
        assert rel.mean() <= 0.0126, f"fp8 mean error vs FP64 {rel.mean():.3e}"
        assert normwise <= 0.054, f"fp8 normwise error vs FP64 {normwise:.3e}"
    else:
        assert rel.max() <= BOUND_VS_FP64[mode], f"{mode}: error vs FP64 above bound ({rel.max():.3e})"
    assert np.all(C[~nz] == 0)


def test_fp8_b_quantiser_matches_oracle():
    """bcsr_quantize_b_e4m3 == the oracle's quantiser, byte for byte: one
    scale per (16-row slab, 256-column chunk), chunk rows 16-byte aligned."""
    lib = mm.lib()
    K, N = 4176, 640  # 261 slabs (an odd last pair, scale rows padded to 272), 3 chunks, the last partial
    B16 = mm.generate(K, N, 0.0, "random", seed=11)
    B16[5, 7] = np.float16(0.0)
    B16[17, :] *= np.float16(1e-3)  # small values (subnormal codes in slab 1)
    B16[4000:4016, 300:310] = np.float16(3.0)  # one slab of chunk 1 with a larger scale
    B16[4160:4176, :256] = np.float16(0.0)  # an all-zero patch (scale 2^0)
    dB = mm.DeviceArray.from_numpy(B16)
    d8 = mm.DeviceArray((K * N,), np.uint8)
    stride = oracle.b8_scale_stride(K)
    assert stride == 272 and lib.bcsr_b8_scale_bytes(K, N) == 3 * stride
    ds = mm.DeviceArray((3 * stride,), np.uint8)
    want_q, want_s = oracle.quant_e4m3_B(B16.astype(np.float32))
    assert len(set(want_s[1].tolist())) > 1 and want_s[0, 260] == 127
    for call in range(3):  # repeated calls
        d8.fill_bytes(0xFF)
        ds.fill_bytes(0)
        mm.check(lib.bcsr_quantize_b_e4m3(dB.ptr, K, N, d8.ptr, ds.ptr, None, None))
        mm.synchronize()
        q_k_minor = d8.numpy().reshape(K // 16, N, 16)
        got = q_k_minor.transpose(0, 2, 1).reshape(K, N)
        got_s = ds.numpy().reshape(3, stride)
        assert np.array_equal(got_s[:, :K // 16], want_s), f"call {call}"
        assert np.array_equal(got, want_q), f"call {call}"
    assert lib.bcsr_quantize_b_e4m3(dB.ptr, K, N, d8.ptr, ds.offset(8), None, None) == -3  # misaligned scales


@pytest.mark.parametrize("bh,N,tile,K", [(16, 512, 0, 1024), (16, 512, 64, 1024), (16, 512, 256, 1024),
                                         (16, 96, 32, 1024), (32, 384, 0, 1024), (64, 272, 0, 1024),
                                         (64, 128, 64, 1024), (16, 256, 0, 20480)])
def test_fp8_slab_scales_differ(bh, N, tile, K):
    """B whose 16-row slabs differ in magnitude (2^-6 .. 2^5) in its first
    256 columns only: there each slab keeps its own E8M0 scale and the
    product takes two MFMAs per octet (even slots, odd slots), the other
    chunk stays on the one-MFMA path; K = 20480 gives 1280 slabs, more than
    one read of the chunk's scales.  Against the exact product of the
    stored values (2^-11, as every FP8 product) and, per entry, within the
    uniform-B bounds of the FP64 product (the slab scales follow the
    magnitudes, so the relative error stays that of uniform B)."""
    M = 512 if K <= 1024 else 128
    h32, data32, B16 = case(M, K, N, bh, dist="uniform")
    rng = np.random.default_rng(bh * 1000 + N)
    mags = (2.0 ** rng.integers(-6, 6, K // 16)).astype(np.float32)
    B = B16.astype(np.float32)
    B[:, :256] *= np.repeat(mags, 16)[:, None]
    B16 = B.astype(np.float16)
    _, sc = oracle.quant_e4m3_B(B16.astype(np.float32))
    assert len(set(sc[0].tolist())) > 4
    assert N <= 256 or len(set(sc[1].tolist())) == 1
    log(f"bh={bh} N={N} tile={tile} K={K} nnzb={h32.nnzb}: chunk 0 slab scales {sorted(set(sc[0].tolist()))}")
    C = run_mode(h32, B16, "fp8", mm.make_cfg(tile_n=tile) if tile else None).astype(np.float64)
    Aq = np.asarray(oracle.dequantized_blocks("fp8", data32, bh, 16, h32.hdr, h32.idx), dtype=np.float32)
    Bq = oracle.quantized_B("fp8", B16.view(np.uint16))
    Cq, Cqabs = oracle.bcsr_spmm_f64(h32.hdr, h32.idx, Aq, bh, 16, Bq, threads=8)
    err_q = np.abs(C - Cq)
    log(f"  max |C - Cq| / (|Aq||Bq|) {np.max(err_q / np.maximum(Cqabs, 1e-30)):.3e}")
    assert np.all(err_q <= 2.0**-11 * Cqabs + 1e-30), "kernel differs from its stored values"
    C64, Cabs = oracle.bcsr_spmm_f64(h32.hdr, h32.idx, data32, bh, 16, B16.astype(np.float32), threads=8)
    nz = Cabs > 0
    rel = np.abs(C - C64)[nz] / Cabs[nz]
    log(f"  max rel |A||B| {rel.max():.3e} mean {rel.mean():.3e}")
    assert rel.max() <= 0.095 and rel.mean() <= 0.0126
    assert np.all(C[~nz] == 0)


def test_fp8_64_row_blocks_have_no_256_column_kernel():
    """FP8 builds no kernel for 64-row blocks at 256-column tiles (its 256
    accumulator registers leave no room for the operands): the plan reports
    that configuration unsupported and a call with it fails with
    BCSR_ERR_UNSUPPORTED, while 128-column tiles run and match the automatic
    choice."""
    h32, _, B16 = case(256, 1024, 512, 64, dist="uniform")
    plan = mm.Plan(h32, "fp8")
    try:
        assert not plan.supports(512, mm.make_cfg(tile_n=256))
        assert plan.supports(512, mm.make_cfg(tile_n=128))
        with pytest.raises(mm.BcsrError, match="BCSR_ERR_UNSUPPORTED"):
            plan.matmul(B16, cfg=mm.make_cfg(tile_n=256))
        C128 = plan.matmul(B16, cfg=mm.make_cfg(tile_n=128))
        assert np.array_equal(C128, plan.matmul(B16))
    finally:
        plan.close()


@pytest.mark.parametrize("slab_mags", [False, True])
@pytest.mark.parametrize("bh,N", [(16, 512), (32, 96), (64, 272)])
def test_spmm_on_prequantised_b_equals_spmm(bh, N, slab_mags):
    """bcsr_spmm_b8 on bcsr_quantize_b_e4m3's output == bcsr_spmm (which
    quantises B itself) bit for bit, over two calls with different N, for B
    of one magnitude and for B whose 16-row slabs differ in magnitude (two
    MFMAs per 16 columns); a 16-bit plan and a misaligned B8 are refused."""
    lib = mm.lib()
    M, K = 512, 1024
    A32 = mm.generate(M, K, 0.8, "blockrandom", bh=bh, seed=bh, dtype=np.float32)
    h = mm.bcsr_from_dense(A32.astype(np.float16), bh, 16)
    h32 = mm.HostBCSR(h.hdr, h.idx, oracle.bcsr_values_f32(A32, h.hdr, h.idx, bh, 16), h.block_rows,
                      h.block_cols, bh, 16)
    plan = mm.Plan(h32, "fp8")
    for n in (N, 48):
        B16 = mm.generate(K, n, 0.0, "random", seed=n)
        if slab_mags:
            mags = 2.0 ** np.random.default_rng(n).integers(-6, 6, K // 16)
            B16 = (B16.astype(np.float32) * np.repeat(mags, 16)[:, None]).astype(np.float16)
        dB = mm.DeviceArray.from_numpy(B16)
        d8 = mm.DeviceArray((K * n,), np.uint8)
        ds = mm.DeviceArray((lib.bcsr_b8_scale_bytes(K, n),), np.uint8)
        C1 = mm.DeviceArray((M, n), np.float32)
        C2 = mm.DeviceArray((M, n), np.float32)
        C1.fill_bytes(0xFF)
        mm.check(lib.bcsr_quantize_b_e4m3(dB.ptr, K, n, d8.ptr, ds.ptr, None, None))
        mm.check(lib.bcsr_spmm_b8(plan._handle, d8.ptr, ds.ptr, C1.ptr, n, None, None))
        plan.spmm_ptr(dB.ptr, C2.ptr, n)
        mm.synchronize()
        assert np.array_equal(C1.numpy(), C2.numpy()), (bh, n)
    assert lib.bcsr_spmm_b8(plan._handle, d8.offset(8), ds.ptr, C1.ptr, 48, None, None) == -3
    assert lib.bcsr_spmm_b8(plan._handle, d8.ptr, ds.offset(4), C1.ptr, 48, None, None) == -3
    p16 = mm.Plan(h, "f16")
    assert lib.bcsr_spmm_b8(p16._handle, d8.ptr, ds.ptr, C1.ptr, 48, None, None) == -7
    p16.close()
    plan.close()


@pytest.mark.parametrize("mode", ["bf16", "hilo", "fp8"])
def test_precision_uniform_values_target_shape_rows(mode):
    """16384^2, 5% blocks, N = 512 (U(-1,1) values): sampled block-rows."""
    M = K = 16384
    N = 512
    A32 = mm.generate(M, K, 0.95, "blockrandom", bh=16, seed=123, dtype=np.float32)
    h = mm.bcsr_from_dense(A32.astype(np.float16), 16, 16)
    data32 = oracle.bcsr_values_f32(A32, h.hdr, h.idx, 16, 16)
    del A32
    h32 = mm.HostBCSR(h.hdr, h.idx, data32, h.block_rows, h.block_cols, 16, 16)
    B16 = mm.generate(K, N, 0.0, "random", seed=456)
    C = run_mode(h32, B16, mode).astype(np.float64)
    Bf = B16.astype(np.float32)
    rows = list(range(0, 1024, 37))
    worst = 0.0
    for br in rows:
        C64, Cabs = oracle.bcsr_spmm_f64(h32.hdr, h32.idx, data32, 16, 16, Bf, br, br + 1)
        sl = slice(br * 16, br * 16 + 16)
        nz = Cabs[sl] > 0
        rel = (np.abs(C[sl] - C64[sl])[nz] / Cabs[sl][nz]).max(initial=0.0)
        worst = max(worst, rel)
    log(f"  {mode} target rows: max rel |A||B| {worst:.3e}")
    assert worst <= {"bf16": 2e-2, "hilo": 5e-5, "fp8": 2e-2}[mode]


@pytest.mark.parametrize("xcd_map", [0, 1, 2, 3])
@pytest.mark.parametrize("bh,N", [(16, 512), (64, 256), (32, 384)])
def test_fp8_xcd_mappings_agree(bh, N, xcd_map):
    """The FP8 kernel under every work-to-XCD mapping gives the automatic
    launch's C bit for bit (the mapping only moves work items)."""
    M, K = 1024, 1024
    A32 = mm.generate(M, K, 0.8, "blockrandom", bh=bh, seed=bh + 1, dtype=np.float32)
    h = mm.bcsr_from_dense(A32.astype(np.float16), bh, 16)
    h32 = mm.HostBCSR(h.hdr, h.idx, oracle.bcsr_values_f32(A32, h.hdr, h.idx, bh, 16), h.block_rows,
                      h.block_cols, bh, 16)
    plan = mm.Plan(h32, "fp8")
    B16 = mm.generate(K, N, 0.0, "random", seed=7)
    dB = mm.DeviceArray.from_numpy(B16)
    C1 = mm.DeviceArray((M, N), np.float32)
    C2 = mm.DeviceArray((M, N), np.float32)
    C1.fill_bytes(0xFF)
    C2.fill_bytes(0xFF)
    plan.spmm_ptr(dB.ptr, C1.ptr, N)
    cfg = mm.make_cfg(xcd_map=xcd_map)
    if xcd_map == 1 and (N + 127) // 128 == 3:  # three tiles cannot be pinned
        with pytest.raises(mm.BcsrError) as e:
            plan.spmm_ptr(dB.ptr, C2.ptr, N, None, cfg)
        assert e.value.status == -1
    else:
        plan.spmm_ptr(dB.ptr, C2.ptr, N, None, cfg)
        mm.synchronize()
        assert np.array_equal(C1.numpy(), C2.numpy())
    plan.close()


def test_fp8_workspace_growth_keeps_results():
    """One FP8 plan called with N = 64, 512 (its workspace grows), then 64
    again gives, each time, the C of a fresh plan called once at that N."""
    M, K = 512, 1024
    h32, _, _ = case(M, K, 16, 16)
    plan = mm.Plan(h32, "fp8")
    for N in (64, 512, 64, 256):
        B16 = mm.generate(K, N, 0.0, "random", seed=N)
        log(f"fp8 workspace N={N}")
        C = plan.matmul(B16)
        fresh = mm.Plan(h32, "fp8")
        C_ref = fresh.matmul(B16)
        fresh.close()
        assert np.array_equal(C, C_ref), N
    plan.close()
