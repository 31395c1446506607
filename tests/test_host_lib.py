"""CPU tests of the product's host side (libbcsr_mi355x.so without GPU calls).

- every function declared in include/bcsr_mi355x.h is exported;
- the C++ generator equals the oracle (and so the reference) draw for draw;
- BCSR/CSR construction equals the oracle's;
- argument validation happens before any HIP call;
- the .mat reader/writer round-trips and rejects malformed files;
- the extension patterns have the structure they are defined with;
- the C++ classes compile and run as the reference's callers use them.
"""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

import matrixmultiplicationcuda_amd as mm
import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "bcsr_mi355x.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    src = "\n".join(l for l in src.splitlines() if not l.lstrip().startswith(("#", "BCSR_ENTRY(")))
    names = set(re.findall(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*[\s*]+([A-Za-z_][A-Za-z0-9_]*)\s*\(",
                           src, flags=re.M))
    names |= set(re.findall(r"BCSR_ENTRY\((\w+)\);", src))
    names.discard("BCSR_ENTRY")
    names.discard("name")
    return sorted(names)


def test_library_exports_every_declared_symbol():
    lib = mm.lib()
    names = declared_functions()
    assert len(names) > 40
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", mm.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    assert not [n for n in names if n not in exported]


def test_status_strings():
    lib = mm.lib()
    for code in mm.STATUS:
        assert lib.bcsr_status_string(code)
    assert lib.bcsr_status_string(-6) == b"device hdr pointer not registered by copyToDevice"


@pytest.mark.parametrize("elem", ["float", "double", "int", "longlong"])
@pytest.mark.parametrize("pattern", mm.REFERENCE_PATTERNS)
def test_generator_equals_oracle(pattern, elem):
    for (r, c, bs, sp, seed) in [(40, 48, 8, 0.7, 42), (130, 70, 16, 0.5, 7), (8, 8, 2, 0.7, 42),
                                 (64, 96, 64, 0.9, 123)]:
        a = mm.generate_matrix(r, c, sp, pattern, bs, seed, elem)
        b = oracle.generate_matrix(r, c, sp, pattern, bs, seed, elem)
        assert np.array_equal(a, b)


@pytest.mark.parametrize("pattern", mm.REFERENCE_PATTERNS + mm.EXTRA_PATTERNS)
@pytest.mark.parametrize("dist", ["uniform", "loguniform"])
def test_extended_generator_equals_oracle(pattern, dist):
    for (r, c, bh, bw) in [(64, 96, 16, 16), (128, 64, 32, 16), (100, 90, 16, 32)]:
        kw = dict(bh=bh, bw=bw, seed=9, value_dist=dist, inblock_density=0.5, band_halfwidth=2)
        a = mm.generate(r, c, 0.6, pattern, dtype=np.float32, **kw)
        b = oracle.generate_ex(r, c, 0.6, pattern, **kw)
        assert np.array_equal(a, b)
        a16 = mm.generate(r, c, 0.6, pattern, **kw)
        assert np.array_equal(a16.view(np.uint16), oracle.f32_to_f16_bits(b))


def test_reference_path_equals_flat_path():
    """generate_matrix<float> (vector of vectors) == generate_into (flat)."""
    a = mm.generate_matrix(96, 80, 0.8, "blockrandom", 16, 5).astype(np.float32)
    b = mm.generate(96, 80, 0.8, "blockrandom", bh=16, seed=5, dtype=np.float32)
    assert np.array_equal(a, b)


def test_unknown_pattern():
    with pytest.raises(mm.BcsrError) as e:
        mm.generate(16, 16, 0.5, "largerandom")
    assert e.value.status == -8
    with pytest.raises(mm.BcsrError):
        mm.generate_matrix(16, 16, 0.5, "largerandom")


@pytest.mark.parametrize("bh,bw", [(16, 16), (32, 16), (64, 16), (16, 32), (48, 16)])
def test_bcsr_equals_oracle(bh, bw):
    A16 = mm.generate(192, 256, 0.7, "blockrandom", bh=16, seed=3)
    h = mm.bcsr_from_dense(A16, bh, bw)
    hdr, idx, data = oracle.bcsr_from_dense(A16.view(np.uint16), bh, bw)
    assert np.array_equal(h.hdr, hdr) and np.array_equal(h.idx, idx)
    assert np.array_equal(h.data.view(np.uint16), data)


@pytest.mark.parametrize("pattern,bh,gbh,kw", [
    ("blockrandom", 16, 16, {}), ("blockrandom", 32, 32, {"value_dist": "loguniform"}),
    ("blockrandom_rho", 16, 16, {"inblock_density": 0.3}), ("blockband", 16, 16, {"band_halfwidth": 2}),
    ("pattern32by16", 32, 16, {}), ("pattern64by16", 64, 16, {}), ("pattern64by16", 32, 16, {}),
    ("pattern32by16", 64, 16, {}), ("random", 16, 16, {}), ("diagonal", 16, 16, {}),
    ("checkerboard", 32, 32, {}), ("blockdiagonal", 16, 16, {})])
def test_generate_bcsr_equals_dense_path(pattern, bh, gbh, kw):
    """The block-by-block generator == the oracle's dense generator followed by
    the oracle's BCSR builder (same blocks, same fp32 values)."""
    bw = 16 if pattern.startswith("pattern") else bh
    h = mm.generate_bcsr(256, 320, 0.8, pattern, bh=bh, bw=bw, seed=17, **kw)
    A = oracle.generate_ex(256, 320, 0.8, pattern, bh=gbh, bw=gbh if bw == gbh else bw, seed=17,
                           **kw)
    hdr, idx, _ = oracle.bcsr_from_dense(oracle.f32_to_f16_bits(A), bh, bw)
    assert np.array_equal(h.hdr, hdr) and np.array_equal(h.idx, idx)
    assert np.array_equal(h.data, oracle.bcsr_values_f32(A, hdr, idx, bh, bw))


def test_generate_bcsr_rejects_unaligned():
    with pytest.raises(mm.BcsrError) as e:
        mm.generate_bcsr(64, 64, 0.5, "nosuchpattern", bh=16, bw=16)
    assert e.value.status == -8
    with pytest.raises(mm.BcsrError) as e:
        mm.generate_bcsr(72, 64, 0.5, "blockrandom", bh=16, bw=16)
    assert e.value.status == -3


@pytest.mark.parametrize("bh,bw,rows,cols,code", [(8, 16, 64, 64, -2), (16, 24, 64, 64, -2),
                                                  (16, 16, 40, 64, -3), (32, 16, 48, 64, -3)])
def test_bcsr_rejects_bad_shapes(bh, bw, rows, cols, code):
    A16 = np.ones((rows, cols), dtype=np.float16)
    with pytest.raises(mm.BcsrError) as e:
        mm.bcsr_from_dense(A16, bh, bw)
    assert e.value.status == code


def test_plan_validation_before_device_work():
    """Bad arguments are rejected on the host, before any HIP call."""
    lib = mm.lib()
    p = ctypes.c_void_p()
    hdr = np.array([0, 1], dtype=np.int32)
    idx = np.array([0], dtype=np.int32)
    data = np.zeros(256, dtype=np.float16)
    P = lambda a: a.ctypes.data_as(ctypes.c_void_p)  # noqa: E731
    assert lib.bcsr_plan_create(P(hdr), P(idx), P(data), 1, 1, 8, 16, ctypes.byref(p)) == -2
    assert lib.bcsr_plan_create(P(hdr), P(idx), P(data), 1, 1, 16, 20, ctypes.byref(p)) == -2
    bad_idx = np.array([3], dtype=np.int32)  # block-column out of range
    assert lib.bcsr_plan_create(P(hdr), P(bad_idx), P(data), 1, 1, 16, 16, ctypes.byref(p)) == -1
    bad_hdr = np.array([1, 1], dtype=np.int32)
    assert lib.bcsr_plan_create(P(bad_hdr), P(idx), P(data), 1, 1, 16, 16, ctypes.byref(p)) == -1
    hdr2 = np.array([0, 2], dtype=np.int32)
    unsorted = np.array([1, 0], dtype=np.int32)
    assert lib.bcsr_plan_create(P(hdr2), P(unsorted), P(data), 1, 2, 16, 16, ctypes.byref(p)) == -1
    assert lib.bcsr_plan_create(None, P(idx), P(data), 1, 1, 16, 16, ctypes.byref(p)) == -1
    assert lib.bcsr_plan_create_f32(P(hdr), P(idx), None, 1, 1, 16, 16, 9, ctypes.byref(p)) == -1
    # unregistered device pointer
    assert lib.bcsr_release_device(ctypes.c_void_p(0x1000)) == -6
    assert lib.sparseMatrixMulTensor(ctypes.c_void_p(0x1000), None, None, None, None, 16, 16,
                                     None) == -6
    # a batch whose plan list holds a null plan
    plans = (ctypes.c_void_p * 1)(None)
    ptrs = (ctypes.c_void_p * 1)(None)
    n = np.array([16], dtype=np.int32)
    batch = ctypes.c_void_p()
    assert lib.bcsr_batch_create(plans, ptrs, ptrs, n.ctypes.data_as(ctypes.c_void_p), None, 1, None,
                                 ctypes.byref(batch)) == -1


def test_mat_roundtrip(tmp_path):
    A = mm.generate(24, 40, 0.6, "random", seed=4)
    path = str(tmp_path / "a.mat")
    mm.mat_write(path, A)
    back, nz = mm.mat_read(path)
    assert np.array_equal(back.view(np.uint16), A.view(np.uint16))
    assert nz == int(np.count_nonzero(A.astype(np.float32)))


def test_mat_reader_format(tmp_path):
    """Hand-written file: header "rows cols", whitespace-separated values,
    nonZeros counted on the float before half rounding (Matrix.cu:24-31)."""
    path = tmp_path / "h.mat"
    path.write_text("2 3\n1 0 -2.5\n 0.1\t1e-9 65504\n")
    A, nz = mm.mat_read(str(path))
    assert A.shape == (2, 3)
    want = np.array([[1, 0, -2.5], [0.1, 1e-9, 65504]], dtype=np.float32).astype(np.float16)
    assert np.array_equal(A.view(np.uint16), want.view(np.uint16))
    assert nz == 5  # 1e-9 is nonzero as a float although it rounds to 0 in half


@pytest.mark.parametrize("text", ["", "3\n", "2 2\n1 2 3\n", "x y\n1 2\n", "-1 2\n"])
def test_mat_reader_rejects_malformed(tmp_path, text):
    path = tmp_path / "bad.mat"
    path.write_text(text)
    with pytest.raises(mm.BcsrError):
        mm.mat_read(str(path))


def test_mat_missing_file(tmp_path):
    with pytest.raises(mm.BcsrError):
        mm.mat_read(str(tmp_path / "nope.mat"))


# ---- extension patterns: structure ------------------------------------------------
@pytest.mark.parametrize("bh,expect", [(16, 17), (32, 18), (64, 20)])
def test_blockband_structure(bh, expect):
    n = 1024
    A = mm.generate(n, n, 0.0, "blockband", bh=bh, bw=16, band_halfwidth=8, seed=1,
                    dtype=np.float32)
    h = mm.bcsr_from_dense(A.astype(np.float16), bh, 16)
    counts = np.diff(h.hdr)
    nbr = n // bh
    # interior block-rows hold the diagonal block-columns plus 8 either side
    assert np.all(counts[nbr // 4: 3 * nbr // 4] == expect)
    for br in range(nbr):
        d0, d1 = br * bh // 16, (br * bh + bh - 1) // 16
        cols = h.idx[h.hdr[br]:h.hdr[br + 1]]
        assert np.array_equal(cols, np.arange(max(0, d0 - 8), min(n // 16, d1 + 9)))


def test_blockrandom_rho_in_block_density():
    A = mm.generate(1024, 1024, 0.9, "blockrandom_rho", bh=16, bw=16, inblock_density=0.4,
                    seed=2, dtype=np.float32)
    h = mm.bcsr_from_dense(A.astype(np.float16), 16, 16)
    frac = np.count_nonzero(h.data.astype(np.float32)) / h.data.size
    assert abs(frac - 0.4) < 0.02
    assert abs(h.nnzb / (64 * 64) - 0.1) < 0.02


def test_loguniform_values():
    from scipy import stats

    A = mm.generate(512, 512, 0.9, "blockrandom", bh=16, seed=3, value_dist="loguniform",
                    lo=1e-3, hi=1e2, dtype=np.float32)
    v = A[A != 0].astype(np.float64)
    assert np.all((np.abs(v) >= 1e-3 * (1 - 1e-6)) & (np.abs(v) <= 1e2 * (1 + 1e-6)))
    assert abs(np.mean(v > 0) - 0.5) < 0.02
    u = (np.log10(np.abs(v)) + 3) / 5
    assert stats.kstest(u, "uniform").pvalue > 1e-3


# ---- the C++ classes as the reference's callers use them ---------------------------
CPP_PROGRAM = r"""
#include <cstdio>
#include <stdexcept>
#include "BCSRMatrix.h"
#include "CSRMatrix.h"
#include "HCSRMatrix.h"
#include "Matrix.h"
#include "matrix_generator.h"
#include "miscutil.h"

int main() {
    // prepare_buffers (nvbench_main.cu:144-156) without the device copies
    auto genA = mg::generate_matrix<float>(64, 96, 0.7, "blockrandom", 16, 123);
    Matrix A(64, 96);
    for (int i = 0; i < 64; ++i)
        for (int j = 0; j < 96; ++j) A.data[i * 96 + j] = float2half_rn(genA[i][j]);
    BCSRMatrix b(A, 16, 16);
    CSRMatrix c(A);
    HCSRMatrix hy(A, 0.5f);
    float d = blockDensity(A, 0, 0);
    bool threw = false;
    try { BCSRMatrix bad(A, 8, 16); } catch (const std::invalid_argument &) { threw = true; }
    std::printf("%d %d %d %d %d %.4f %d\n", b.blockRows, b.blockCols, b.hdr[b.blockRows], c.nnz(),
                hy.bcsr->hdr[hy.bcsr->blockRows] + (hy.csr->nnz() > 0), d, threw);
    return 0;
}
"""


def test_cpp_api_compiles_and_runs(tmp_path):
    src = tmp_path / "api.cpp"
    src.write_text(CPP_PROGRAM)
    exe = tmp_path / "api"
    subprocess.run(["/opt/rocm/bin/hipcc", "-std=c++17", "-O1", f"-I{ROOT}/include", str(src),
                    "-o", str(exe), f"-L{os.path.dirname(mm.LIB_PATH)}", "-lbcsr_mi355x",
                    f"-Wl,-rpath,{os.path.dirname(mm.LIB_PATH)}"], check=True,
                   capture_output=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()
    A16 = mm.generate(64, 96, 0.7, "blockrandom", bh=16, seed=123)
    h = mm.bcsr_from_dense(A16, 16, 16)
    assert out[:3] == ["4", "6", str(h.nnzb)]
    assert int(out[3]) == int(np.count_nonzero(A16.astype(np.float32)))
    blk0 = A16[:16, :16].astype(np.float32)
    assert float(out[5]) == pytest.approx(np.count_nonzero(blk0) / 256, abs=1e-4)
    assert out[6] == "1"


def test_dispatcher_choice_edge_cases():
    """learned.choose: the predicted-fastest supported configuration, and the
    autotune fallback when the predicted spread is under 5 % -- also for
    predictions far outside exp's range and for predictions that are NaN."""
    import math

    from matrixmultiplicationcuda_amd import learned

    mask = np.array([1, 1, 0, 1], np.float32)
    k, fb = learned.choose(np.array([2.0, 1.0, -5.0, 3.0]), mask)
    assert k == 1 and not fb  # masked-out config 2 is never chosen
    k, fb = learned.choose(np.array([1.0, 1.0 + math.log(1.04), 0.0, 1.01]), mask)
    assert k == 0 and fb  # all within 4 %: autotune
    k, fb = learned.choose(np.array([1.0, 1.0 + math.log(1.06), 0.0, 1.0]), mask)
    assert not fb  # 6 % apart: trust the model
    k, fb = learned.choose(np.array([-800.0, 900.0, 0.0, 5.0]), mask)
    assert k == 0 and not fb  # no overflow for a 1700-nat spread
    k, fb = learned.choose(np.array([1.0, np.nan, 0.0, 2.0]), mask)
    assert fb
    k, fb = learned.choose(np.array([1.0, 2.0, 0.0, 3.0]), np.array([0, 1, 0, 0], np.float32))
    assert k == 1 and not fb  # a single supported configuration


def test_dispatcher_keeps_the_automatic_choice_within_its_margin():
    """With the automatic choice's predicted time, a configuration is taken
    only when predicted faster than it by more than the margin; the regret
    report charges the automatic choice's own time then, and tune_margin
    picks the margin with the least regret on the records it is given."""
    import math

    from matrixmultiplicationcuda_amd import learned

    mask = np.array([1, 1, 1], np.float32)
    pred = np.array([1.0, 1.2, 1.5])
    k, fb = learned.choose(pred, mask, pred_auto=1.0 + math.log(1.08), margin=0.05)
    assert k == 0 and not fb  # 8 % faster than automatic, margin 5 %: take config 0
    k, fb = learned.choose(pred, mask, pred_auto=1.0 + math.log(1.03), margin=0.05)
    assert k == learned.AUTO and not fb  # only 3 % faster: keep the automatic choice
    k, fb = learned.choose(pred, mask, pred_auto=0.5, margin=0.0)
    assert k == learned.AUTO  # automatic predicted fastest
    k, fb = learned.choose(np.array([1.0, 1.01, 1.02]), mask, pred_auto=2.0, margin=0.0)
    assert fb and k == 0  # predicted spread under 5 %: autotune, whatever the automatic time
    k, fb = learned.choose(pred, mask, pred_auto=np.nan, margin=0.0)
    assert k == 0  # no automatic prediction: the fastest configuration

    # two records, three configs; record 0: config 0 is best (1.0 ms), the
    # automatic choice takes 1.1 ms; record 1: automatic (1.0 ms) is best
    ms = np.array([[1.0, 2.0, 3.0], [1.3, 2.0, 3.0]])
    auto = np.array([1.1, 1.0])
    M = np.ones((2, 4), np.float32)
    # the model predicts config 0 about 4 % faster than automatic on both
    p = np.log(np.array([[1.0, 2.0, 3.0, 1.04], [1.0, 2.0, 3.0, 1.04]]))
    r0 = learned.regret_report(p, M, ms, auto, margin=0.0)
    assert r0["learned_mean_regret"] == pytest.approx((0.0 + 0.3) / 2)
    assert r0["static_mean_regret"] == pytest.approx((0.1 + 0.0) / 2)
    r5 = learned.regret_report(p, M, ms, auto, margin=0.05)
    assert r5["learned_mean_regret"] == pytest.approx((0.1 + 0.0) / 2)
    assert r5["kept_automatic_fraction"] == 1.0
    assert learned.tune_margin(p, M, ms, auto, margins=(0.0, 0.05, 0.1)) == 0.05
    # a model without the automatic output: the fastest configuration always
    r = learned.regret_report(p[:, :3], M, ms, auto)
    assert r["learned_mean_regret"] == pytest.approx(0.15) and r["kept_automatic_fraction"] == 0.0


class _StubInfo:
    def __init__(self, dev_bh):
        self.dev_bh = dev_bh


class _StubPlan:
    def __init__(self, dev_bh):
        self.info = _StubInfo(dev_bh)


@pytest.mark.parametrize("precision", ["f16", "fp8"])
def test_batch_tile_pickers_take_the_fastest_batch_tile(precision):
    """Both batch tile pickers (timings, cost-model predictions) pick the
    fastest configuration among those the persistent kernel runs: for FP8
    (whose configurations carry variant -1) as well as for fp16."""
    from matrixmultiplicationcuda_amd import autotune

    prec = mm.PRECISIONS[precision]
    configs = autotune.configs_for(precision)
    for dev_bh in (16, 32, 64):
        batch_tiles = [t for t in autotune.TILES
                       if mm.lib().bcsr_batch_tile_supported(prec, dev_bh, t)]
        assert batch_tiles
        for want in batch_tiles:
            # every configuration 1 ms except the batch's variant of `want`;
            # other variants of the same tile are faster still but not run
            # by the batch, and 256 columns (never batched) fastest of all
            t = np.ones(len(configs))
            for k, c in enumerate(configs):
                tile, variant = c[0], c[1]
                if tile == want:
                    t[k] = 0.5 if variant == autotune.batch_variant(prec, dev_bh, tile) else 0.1
                if tile == 256:
                    t[k] = 0.01
            got = autotune.batch_tiles_from_times([_StubPlan(dev_bh)], [512], [t], configs, prec)
            assert got == [want], (precision, dev_bh, want)
            assert autotune.best_batch_tile(configs, np.log(t), prec, dev_bh) == want
        # nothing timed: the batch's automatic width
        assert autotune.best_batch_tile(configs, [np.nan] * len(configs), prec, dev_bh) == 0


def test_quantize_b_with_no_rows_is_a_no_op():
    """bcsr_quantize_b_e4m3 with K = 0 (or N = 0) returns before any HIP call."""
    assert mm.lib().bcsr_quantize_b_e4m3(None, 0, 64, None, None, None, None) == 0
    assert mm.lib().bcsr_quantize_b_e4m3(None, 64, 0, None, None, None, None) == 0
    assert mm.lib().bcsr_quantize_b_e4m3(None, 8, 64, None, None, None, None) == -3


def test_header_in_c_and_the_reference_call_site_in_cpp(tmp_path):
    """bcsr_mi355x.h compiles as C (every stream argument explicit) and as
    C++, where a reference call site keeps its arguments -- half pointers,
    unsigned M and N -- and drops only the launch configuration (the stream
    defaults to the null stream)."""
    c_src = tmp_path / "use.c"
    c_src.write_text('#include "bcsr_mi355x.h"\n'
                     'int f(const int *h, const int *i, const void *d, const void *B, float *C) {\n'
                     '  return sparseMatrixMulTensor_v2(h, i, d, B, C, 16u, 16u, (void *)0);\n}\n')
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-fsyntax-only", f"-I{ROOT}/include",
                    str(c_src)], check=True, capture_output=True)
    cpp_src = tmp_path / "use.cpp"
    cpp_src.write_text('#include "bcsr_mi355x.h"\n#include "half_type.h"\n'
                       'int f(int *h, int *i, half *d, half *B, float *C, unsigned M, unsigned N) {\n'
                       '  int rc = sparseMatrixMulTensor_v2(h, i, d, B, C, M, N);\n'
                       '  rc |= sparseMatrixMulTensorAdaptive64x16(h, i, d, B, C, M, N, 2u);\n'
                       '  return rc | sparseMatrixMulTensor1(h, i, d, B, C, N);\n}\n')
    subprocess.run(["/opt/rocm/llvm/bin/clang++", "-std=c++17", "-Wall", "-Werror", "-fsyntax-only",
                    f"-I{ROOT}/include", str(cpp_src)], check=True, capture_output=True)
