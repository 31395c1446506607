"""GPU parity of the slab-streaming panel kernel (launch variant 3,
kernels_panel.hip) against the oracle, through the C ABI: same bound as the
pair kernels (1e-3 * (|A|.|B|)_ij, exact zeros where (|A|.|B|)_ij = 0),
every entry written, plus the shapes it refuses."""
import sys

import numpy as np
import pytest

import matrixmultiplicationcuda_amd as mm
from helpers import assert_close, cpu_reference, make_case

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def run(case, cfg, precision="f16", h=None):
    assert mm.device_count() > 0, "GPU tests need a visible MI355X"
    plan = mm.Plan(case.h if h is None else h, precision)
    B = mm.DeviceArray.from_numpy(case.B16)
    C = mm.DeviceArray((case.M, case.N), np.float32)
    C.fill_bytes(0xFF)  # NaN pattern: every entry must be written
    try:
        plan.spmm(B, C, cfg=cfg)
        mm.synchronize()
        return C.numpy(), plan.kernel_name(case.N, cfg)
    finally:
        plan.close()
        B.free()
        C.free()


@pytest.mark.parametrize("M,K,N,sp,pattern,bw", [
    (1024, 2048, 512, 0.85, "blockrandom", 16),   # 4 panels, 4 column tiles
    (1024, 2048, 512, 0.3, "blockrandom", 16),    # dense rows: chunks cut by A capacity
    (1040, 768, 208, 0.75, "blockrandom", 16),    # partial last panel, ragged last tile
    (512, 1024, 256, 0.8, "blockrandom", 32),     # 16 x 32 blocks: two device blocks each
    (768, 768, 256, 0.0, "blockband", 16),        # band: few slabs per panel
    (1024, 1024, 128, 0.0, "blockdiagonal", 16),  # one column tile
    (512, 768, 64, 0.8, "blockrandom", 16),       # N below the 128-column tile
])
@pytest.mark.parametrize("xcd_map", [0, 1])
def test_panel_matches_oracle(M, K, N, sp, pattern, bw, xcd_map):
    kw = {"band_halfwidth": 4} if pattern == "blockband" else {}
    gen_bh = 64 if pattern == "blockdiagonal" else 16
    case = make_case(M, K, N, sp, pattern, 16, bw, gen_bh=gen_bh, gen_bw=gen_bh if bw == 16 else bw, **kw)
    log(f"panel {M}x{K} N={N} {pattern} sp={sp} bw={bw} map={xcd_map} nnzb={case.h.nnzb}")
    C, name = run(case, mm.make_cfg(variant=3, xcd_map=xcd_map))
    assert name == "bcsr_spmm_panel<f16>"
    C_cpu, Cabs = cpu_reference(case)
    assert_close(C, C_cpu, Cabs, what=f"panel {pattern} sp={sp}")


def test_panel_empty_rows_and_matrix():
    case = make_case(640, 512, 256, 0.9, "blockrandom", 16, 16)
    h = case.h
    # clear every third block-row
    keep = np.ones(h.nnzb, bool)
    for br in range(0, h.block_rows, 3):
        keep[h.hdr[br]:h.hdr[br + 1]] = False
    counts = np.diff(h.hdr) * [(br % 3) != 0 for br in range(h.block_rows)]
    hdr = np.concatenate([[0], np.cumsum(counts)]).astype(np.int32)
    h2 = mm.HostBCSR(hdr, h.idx[keep], h.data[keep], h.block_rows, h.block_cols, 16, 16)
    A = np.zeros_like(case.A16)
    for br in range(h2.block_rows):
        for k in range(h2.hdr[br], h2.hdr[br + 1]):
            c = h2.idx[k]
            A[16 * br:16 * br + 16, 16 * c:16 * c + 16] = h2.data[k]
    case2 = type(case)(A, case.B16, h2, 16, 16)
    log(f"panel empty rows nnzb={h2.nnzb}")
    C, _ = run(case2, mm.make_cfg(variant=3))
    C_cpu, Cabs = cpu_reference(case2)
    assert_close(C, C_cpu, Cabs, what="panel empty rows")
    empty = mm.HostBCSR(np.zeros(h.block_rows + 1, np.int32), np.zeros(0, np.int32),
                        np.zeros((0, 16, 16), np.float16), h.block_rows, h.block_cols, 16, 16)
    case3 = type(case)(np.zeros_like(case.A16), case.B16, empty, 16, 16)
    C, _ = run(case3, mm.make_cfg(variant=3))
    assert np.all(C == 0)


def test_panel_bf16_matches_pair_kernel():
    case = make_case(1024, 1024, 256, 0.8, "blockrandom", 16, 16)
    h32 = mm.HostBCSR(case.h.hdr, case.h.idx, case.h.data.astype(np.float32), case.h.block_rows,
                      case.h.block_cols, 16, 16)
    log("panel bf16")
    Cp, name = run(case, mm.make_cfg(variant=3), "bf16", h32)
    assert name == "bcsr_spmm_panel<bf16>"
    Cq, _ = run(case, mm.make_cfg(variant=1), "bf16", h32)
    # same bf16 operands, fp32 accumulation: within the bf16 mode's product bound
    Babs = np.abs(case.B16.astype(np.float64))
    Aabs = np.abs(case.A16.astype(np.float64))
    assert np.all(np.abs(Cp.astype(np.float64) - Cq) <= 1e-4 * (Aabs @ Babs) + 1e-30)


@pytest.mark.parametrize("what", ["bh32", "hilo", "fp8", "tile64"])
def test_panel_unsupported_shapes(what):
    bh = 32 if what == "bh32" else 16
    pattern = "pattern32by16" if bh == 32 else "blockrandom"
    N = 256
    case = make_case(512, 512, N, 0.8, pattern, bh, 16)
    prec = {"hilo": "hilo", "fp8": "fp8"}.get(what, "f16")
    h = case.h
    if prec != "f16":
        h = mm.HostBCSR(h.hdr, h.idx, h.data.astype(np.float32), h.block_rows, h.block_cols, bh, 16)
    cfg = mm.make_cfg(tile_n=64 if what == "tile64" else 0, variant=3)
    with pytest.raises(mm.BcsrError) as e:
        run(case, cfg, prec, h)
    assert e.value.status == -7


def test_panel_target_scale_sampled_rows():
    """16384^2, 16x16, 5% blocks, N = 512 through the panel kernel: sampled
    block-rows against the oracle."""
    M = K = 16384
    case = make_case(M, K, 512, 0.95, "blockrandom", 16, 16)
    log(f"panel target nnzb={case.h.nnzb}")
    C, _ = run(case, mm.make_cfg(variant=3))
    rng = np.random.default_rng(11)
    for br in sorted(set(rng.integers(0, M // 16, 24).tolist()) | {0, M // 16 - 1}):
        C_cpu, Cabs = cpu_reference(case, br, br + 1)
        sl = slice(16 * br, 16 * br + 16)
        assert_close(C[sl], C_cpu[sl], Cabs[sl], what=f"panel block-row {br}")
