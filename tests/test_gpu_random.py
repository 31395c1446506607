"""Seeded random parity sweep: BCSR matrices built directly from random
structure (block heights 16-128 including 48/80/96, block widths 16-48,
1-24 block-rows, 1-48 block-columns, empty rows, explicit zeros inside
stored blocks), random N (multiples of 16 up to 640) and random launch
configurations (tile width, work-to-XCD mapping, kernel variant, waves per
workgroup).  Every supported launch must match the oracle's CPU product
(1e-3 |A||B| per entry, exact zeros where |A||B| = 0); unsupported
combinations must be refused with BCSR_ERR_INVALID_ARG or
BCSR_ERR_UNSUPPORTED, never run."""
import sys

import numpy as np
import pytest

import matrixmultiplicationcuda_amd as mm
from helpers import Case, assert_close, cpu_reference

pytestmark = pytest.mark.gpu


def random_bcsr(rng):
    bh = int(rng.choice([16, 32, 48, 64, 80, 96, 128]))
    bw = int(rng.choice([16, 32, 48]))
    br, bc = int(rng.integers(1, 25)), int(rng.integers(1, 49))
    density = float(rng.uniform(0.0, 0.7))
    rows = []
    for _ in range(br):
        n = int(rng.binomial(bc, density))
        rows.append(np.sort(rng.choice(bc, size=n, replace=False)).astype(np.int32))
    hdr = np.zeros(br + 1, np.int32)
    hdr[1:] = np.cumsum([len(r) for r in rows])
    idx = np.concatenate(rows) if hdr[-1] else np.zeros(0, np.int32)
    data = rng.uniform(-1, 1, (int(hdr[-1]), bh, bw)).astype(np.float16)
    data[rng.random(data.shape) < 0.2] = 0  # explicit zeros inside stored blocks
    return mm.HostBCSR(hdr, idx, data, br, bc, bh, bw)


@pytest.mark.parametrize("seed", range(6))
def test_random_structures_and_launches(seed):
    rng = np.random.default_rng(1000 + seed)
    ran = refused = 0
    for case_no in range(12):
        h = random_bcsr(rng)
        M, K = h.shape
        N = 16 * int(rng.integers(1, 41))
        B16 = rng.uniform(-1, 1, (K, N)).astype(np.float16)
        case = Case(np.zeros((M, K), np.float16), B16, h, h.bh, h.bw)
        C_cpu, Cabs = cpu_reference(case)
        plan = mm.Plan(h, "f16")
        dB = mm.DeviceArray.from_numpy(B16)
        C = mm.DeviceArray((M, N), np.float32)
        for launch in range(3):
            cfg = None if launch == 0 else mm.make_cfg(
                tile_n=int(rng.choice([0, 32, 64, 128, 256])), xcd_map=int(rng.choice([-1, 0, 1, 2, 3, 4])),
                variant=int(rng.choice([-1, 0, 1, 2])), waves_per_block=int(rng.choice([0, 1, 2, 4])))
            what = (f"seed={seed} case={case_no} launch={launch} bh={h.bh} bw={h.bw} M={M} K={K} N={N} "
                    f"nnzb={h.nnzb}" + ("" if cfg is None else
                                        f" tile={cfg.tile_n} map={cfg.xcd_map} variant={cfg.variant} "
                                        f"wpb={cfg.waves_per_block}"))
            print(what, flush=True)
            sys.stdout.flush()
            C.fill_bytes(0xFF)
            try:
                plan.spmm_ptr(dB.ptr, C.ptr, N, None, cfg)
            except mm.BcsrError as e:
                assert cfg is not None and e.status in (-1, -7), what
                refused += 1
                continue
            mm.synchronize()
            assert_close(C.numpy(), C_cpu, Cabs, what=what)
            ran += 1
        for a in (C, dB):
            a.free()
        plan.close()
    assert ran >= 12, (ran, refused)


@pytest.mark.parametrize("seed", range(3))
def test_random_structures_precision_modes(seed):
    """bf16, hi+lo and FP8 plans of random structures (fp32 block values,
    automatic launch) against the exact product of the values each mode
    stores: 1e-4 |Aq||Bq| per entry for the 16-bit modes (fp32 accumulation
    only), 2^-11 for FP8 (its scaled MFMA does not add in IEEE fp32)."""
    import oracle

    rng = np.random.default_rng(2000 + seed)
    for case_no in range(5):
        bh = int(rng.choice([16, 32, 48, 64]))
        bw = int(rng.choice([16, 32]))
        br, bc = int(rng.integers(1, 9)), int(rng.integers(1, 17))
        rows = [np.sort(rng.choice(bc, size=int(rng.binomial(bc, 0.5)), replace=False)).astype(np.int32)
                for _ in range(br)]
        hdr = np.zeros(br + 1, np.int32)
        hdr[1:] = np.cumsum([len(r) for r in rows])
        idx = np.concatenate(rows) if hdr[-1] else np.zeros(0, np.int32)
        data32 = (rng.uniform(-1, 1, (int(hdr[-1]), bh, bw)) *
                  np.exp(rng.uniform(-4, 4, (int(hdr[-1]), 1, 1)))).astype(np.float32)
        h32 = mm.HostBCSR(hdr, idx, data32, br, bc, bh, bw)
        M, K = h32.shape
        N = 16 * int(rng.integers(1, 25))
        B16 = rng.uniform(-1, 1, (K, N)).astype(np.float16)
        for mode in ("bf16", "hilo", "fp8"):
            what = f"seed={seed} case={case_no} mode={mode} bh={bh} bw={bw} M={M} K={K} N={N} nnzb={h32.nnzb}"
            print(what, flush=True)
            plan = mm.Plan(h32, mode)
            C = plan.matmul(B16).astype(np.float64)
            plan.close()
            Aq = np.asarray(oracle.dequantized_blocks(mode, data32, bh, bw, hdr, idx), dtype=np.float64)
            Aq32 = Aq.astype(np.float32)
            assert np.array_equal(Aq32.astype(np.float64), Aq), what  # every stored value is an fp32
            Bq = oracle.quantized_B(mode, B16.view(np.uint16))
            Cq, Cqabs = oracle.bcsr_spmm_f64(hdr, idx, Aq32, bh, bw, Bq, threads=8)
            tight = 2.0**-11 if mode == "fp8" else 1e-4
            assert np.all(np.abs(C - Cq) <= tight * Cqabs + 1e-30), what
