"""Pins the oracle (the CPU checker) before it is trusted.

Each check is independent of the oracle's own formulas: the C++11 known
answer and numpy's MT19937 for the engine, numpy's float16 for rounding,
the reference generator itself (oracle/_ref, compiled from the reference's
matrix_generator.cpp) and the committed golden fixtures it produced, a pure
Python BCSR construction, scipy's CSR, numpy float64 products and closed
forms (identity blocks, one-hot B).
"""
import hashlib
import json
import os

import numpy as np
import pytest

import oracle

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")
PATTERNS = ["random", "checkerboard", "diagonal", "blockdiagonal", "blockrandom",
            "pattern64by16", "pattern32by16"]
needs_ref = pytest.mark.skipif(not oracle.ref_available(),
                               reason="oracle/_ref not built (reference tree absent)")


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def golden():
    with open(os.path.join(GOLDEN, "golden.json")) as f:
        return json.load(f)


# ---- std::mt19937 ------------------------------------------------------------
def test_mt19937_known_answer():
    # C++11 [rand.predef]: the 10000th invocation of a default-constructed
    # mt19937 (seed 5489) produces 4123659995.
    assert oracle.mt_stream(5489, 10000)[-1] == 4123659995


@pytest.mark.parametrize("seed", [0, 1, 123, 456, 2**32 - 1])
def test_mt19937_matches_numpy(seed):
    rs = np.random.RandomState(seed)
    want = rs.randint(0, 2**32, size=2000, dtype=np.uint64).astype(np.uint32)
    assert np.array_equal(oracle.mt_stream(seed, 2000), want)


# ---- float <-> half ----------------------------------------------------------
def test_half_to_float_all_values():
    h = np.arange(65536, dtype=np.uint32).astype(np.uint16)
    got = oracle.f16_bits_to_f32(h)
    want = h.view(np.float16).astype(np.float32)
    nan = np.isnan(want)
    assert np.array_equal(np.isnan(got), nan)
    assert np.array_equal(got[~nan].view(np.uint32), want[~nan].view(np.uint32))


def test_float_to_half_rounding():
    rng = np.random.default_rng(0)
    halves = np.arange(65536, dtype=np.uint32).astype(np.uint16).view(np.float16)
    finite = halves[np.isfinite(halves)].astype(np.float32)
    # every half, the exact midpoints between neighbours (ties), and nearby
    mids = ((finite[:-1].astype(np.float64) + finite[1:].astype(np.float64)) / 2).astype(np.float32)
    x = np.concatenate([finite, mids, np.nextafter(mids, np.float32(np.inf)),
                        np.nextafter(mids, np.float32(-np.inf)),
                        rng.standard_normal(200000).astype(np.float32) * 1e4,
                        rng.uniform(-1, 1, 200000).astype(np.float32) * 1e-6,
                        np.array([65504, 65519.99, 65520, 1e9, -1e9, 2**-25, 2**-24, 3 * 2**-26,
                                  np.inf, -np.inf, 0.0, -0.0], dtype=np.float32)])
    got = oracle.f32_to_f16_bits(x)
    want = x.astype(np.float16).view(np.uint16)
    assert np.array_equal(got, want)


# ---- generator -----------------------------------------------------------------
@pytest.mark.parametrize("pattern", PATTERNS)
def test_generator_matches_golden_tiny(pattern):
    z = np.load(os.path.join(GOLDEN, "tiny_patterns.npz"))
    for (r, c, bs, sp, seed) in [(40, 48, 8, 0.7, 42), (37, 70, 16, 0.5, 7), (64, 64, 16, 0.9, 123)]:
        key = f"{pattern}_{r}x{c}_b{bs}_s{sp}_seed{seed}"
        got = oracle.generate_matrix(r, c, sp, pattern, bs, seed, "float").astype(np.float32)
        assert np.array_equal(got.view(np.uint32), z[key + "_float"].view(np.uint32)), key
        got_i = oracle.generate_matrix(r, c, sp, pattern, bs, seed, "int").astype(np.int32)
        assert np.array_equal(got_i, z[key + "_int"]), key


def test_generator_test_program_case():
    """matrix_generator_test.cpp:6 -- generate_matrix<float>(8,8,0.7,"checkerboard",2,42)."""
    want = np.array(golden()["generator_test_checkerboard"], dtype=np.uint32).view(np.float32)
    got = oracle.generate_matrix(8, 8, 0.7, "checkerboard", 2, 42).astype(np.float32)
    assert np.array_equal(got, want)
    # checkerboard structure: blocks (i, j) with i + j even are dense
    for i in range(8):
        for j in range(8):
            assert (got[i, j] != 0) == ((i // 2 + j // 2) % 2 == 0)


@needs_ref
@pytest.mark.parametrize("elem", ["float", "double", "int", "longlong"])
@pytest.mark.parametrize("pattern", PATTERNS)
def test_generator_matches_reference_build(pattern, elem):
    for (r, c, bs, sp, seed) in [(33, 100, 32, 0.3, 1), (130, 70, 16, 0.5, 7), (8, 8, 2, 0.7, 42),
                                 (96, 80, 64, 0.95, 99), (17, 17, 1, 0.0, 3)]:
        got = oracle.generate_matrix(r, c, sp, pattern, bs, seed, elem)
        want = oracle.ref_generate_matrix(r, c, sp, pattern, bs, seed, elem)
        assert np.array_equal(got, want), (pattern, elem, r, c, bs, sp, seed)


@needs_ref
def test_unknown_pattern_raises_in_both():
    with pytest.raises(ValueError):
        oracle.generate_matrix(8, 8, 0.5, "largerandom", 16, 1)
    with pytest.raises(ValueError):
        oracle.ref_generate_matrix(8, 8, 0.5, "largerandom", 16, 1)


def test_small_blockrandom_golden():
    g = golden()["blockrandom_1024x1024_bs16"]
    A = oracle.generate_matrix(1024, 1024, 0.9, "blockrandom", 16, 123).astype(np.float32)
    A16 = oracle.f32_to_f16_bits(A)
    assert sha(A16) == g["sha_dense_f16"]
    hdr, idx, data = oracle.bcsr_from_dense(A16, 16, 16)
    assert int(hdr[-1]) == g["nnzb"] == 400
    assert hdr.tolist() == g["hdr"] and idx.tolist() == g["idx"]
    assert sha(data) == g["sha_data"]


@pytest.mark.parametrize("name", ["B_1024x256", "B_16384x64"])
def test_B_golden(name):
    g = golden()["B"][name]
    B = oracle.generate_matrix(g["K"], g["N"], 0.0, "random", 16, g["seed"]).astype(np.float32)
    assert sha(oracle.f32_to_f16_bits(B)) == g["sha_f16"]


# ---- BCSR / CSR construction -----------------------------------------------------
def python_bcsr(A16, bh, bw):
    """Literal restatement of formats/BCSRMatrix.cu:36-71 in pure Python."""
    rows, cols = A16.shape
    hdr, idx, blocks = [0], [], []
    for i in range(0, rows, bh):
        n = 0
        for j in range(0, cols, bw):
            blk = A16[i:i + bh, j:j + bw]
            if np.any((blk & 0x7FFF) != 0):
                idx.append(j // bw)
                blocks.append(blk.copy())
                n += 1
        hdr.append(hdr[-1] + n)
    return np.array(hdr), np.array(idx), np.array(blocks).reshape(-1, bh, bw)


@pytest.mark.parametrize("bh,bw", [(16, 16), (32, 16), (16, 32), (64, 16)])
def test_bcsr_matches_python_restatement(bh, bw):
    A = oracle.generate_matrix(128, 192, 0.6, "blockrandom", 16, 5).astype(np.float32)
    A[3, 5] = -0.0  # a negative zero never makes a block nonzero
    A16 = oracle.f32_to_f16_bits(A)
    hdr, idx, data = oracle.bcsr_from_dense(A16, bh, bw)
    h2, i2, d2 = python_bcsr(A16, bh, bw)
    assert np.array_equal(hdr, h2) and np.array_equal(idx, i2) and np.array_equal(data, d2)


def test_bcsr_invariants_and_densify_roundtrip():
    A = oracle.generate_matrix(512, 640, 0.85, "blockrandom", 16, 11).astype(np.float32)
    A16 = oracle.f32_to_f16_bits(A)
    hdr, idx, data = oracle.bcsr_from_dense(A16, 16, 16)
    assert hdr[0] == 0 and np.all(np.diff(hdr) >= 0)
    for r in range(len(hdr) - 1):
        row = idx[hdr[r]:hdr[r + 1]]
        assert np.all(np.diff(row) > 0) and np.all(row < 640 // 16)
    assert np.all(np.any((data.reshape(len(idx), -1) & 0x7FFF) != 0, axis=1))
    back = oracle.bcsr_densify(hdr, idx, data, 512 // 16, 640 // 16, 16, 16)
    assert np.array_equal(back, A16)


def test_csr_matches_scipy():
    import scipy.sparse as sps

    A = oracle.generate_matrix(96, 160, 0.8, "random", 16, 3).astype(np.float32)
    A16 = oracle.f32_to_f16_bits(A)
    hdr, idx, data = oracle.csr_from_dense(A16)
    S = sps.csr_matrix(A16.view(np.float16).astype(np.float32))
    assert np.array_equal(hdr, S.indptr) and np.array_equal(idx, S.indices)
    assert np.array_equal(oracle.f16_bits_to_f32(data), S.data)


def test_csr_product_equals_dense_cpu_matmul_ref_bitwise():
    """CSR-order product == the literal dense cpu_matmul_ref loop, bit for bit."""
    A = oracle.generate_matrix(80, 112, 0.85, "random", 16, 4).astype(np.float32)
    B = oracle.generate_matrix(112, 48, 0.0, "random", 16, 5).astype(np.float32)
    A16, B16 = oracle.f32_to_f16_bits(A), oracle.f32_to_f16_bits(B)
    hdr, idx, data = oracle.csr_from_dense(A16)
    C = oracle.csr_spmm_f32(hdr, idx, data, oracle.f16_bits_to_f32(B16))
    assert np.array_equal(C, oracle.matmul_dense_ref(A16, B16))


def test_csr_running_sum_from_c():
    """sparseMatrixMult1's in-place order: C = 0 gives the plain product; a
    row of two unit products onto 2^24 rounds each step back to 2^24 (ties to
    even), where 2^24 + (1 + 1) would be exact."""
    A = oracle.generate_matrix(48, 64, 0.8, "random", 16, 8).astype(np.float32)
    B = oracle.generate_matrix(64, 32, 0.0, "random", 16, 9).astype(np.float32)
    A16, B16 = oracle.f32_to_f16_bits(A), oracle.f32_to_f16_bits(B)
    hdr, idx, data = oracle.csr_from_dense(A16)
    Bf = oracle.f16_bits_to_f32(B16)
    assert np.array_equal(oracle.csr_spmm_f32(hdr, idx, data, Bf, C0=np.zeros((48, 32), np.float32)),
                          oracle.csr_spmm_f32(hdr, idx, data, Bf))
    one = oracle.f32_to_f16_bits(np.array([1.0], np.float32))[0]
    hdr1 = np.array([0, 2], np.int32)
    idx1 = np.array([0, 1], np.int32)
    data1 = np.array([one, one], np.uint16)
    B1 = np.ones((2, 16), np.float32)
    C0 = np.full((1, 16), 2.0**24, np.float32)
    got = oracle.csr_spmm_f32(hdr1, idx1, data1, B1, C0=C0)
    assert np.all(got == np.float32(2.0**24))
    assert np.all(C0 + oracle.csr_spmm_f32(hdr1, idx1, data1, B1) == np.float32(2.0**24 + 2))


# ---- products ------------------------------------------------------------------------
def small_case(M=256, K=384, N=96, sp=0.8, seed=9):
    A = oracle.generate_matrix(M, K, sp, "blockrandom", 16, seed).astype(np.float32)
    B = oracle.generate_matrix(K, N, 0.0, "random", 16, seed + 1).astype(np.float32)
    return oracle.f32_to_f16_bits(A), oracle.f32_to_f16_bits(B)


def test_bcsr_order_equals_dense_cpu_matmul_ref_bitwise():
    A16, B16 = small_case()
    hdr, idx, data = oracle.bcsr_from_dense(A16, 16, 16)
    Cd = oracle.matmul_dense_ref(A16, B16)
    Cb = oracle.bcsr_spmm_f32(hdr, idx, data, 16, 16, oracle.f16_bits_to_f32(B16), threads=3)
    assert np.array_equal(Cd.view(np.uint32), Cb.view(np.uint32))


def test_products_against_numpy_float64():
    A16, B16 = small_case()
    A = oracle.f16_bits_to_f32(A16).astype(np.float64)
    B = oracle.f16_bits_to_f32(B16).astype(np.float64)
    exact = A @ B
    absb = np.abs(A) @ np.abs(B)
    hdr, idx, data = oracle.bcsr_from_dense(A16, 16, 16)
    Cb = oracle.bcsr_spmm_f32(hdr, idx, data, 16, 16, B.astype(np.float32))
    # fp32 sequential sum: |error| <= K * 2^-24 * (|A||B|) (first-order bound)
    assert np.all(np.abs(Cb - exact) <= A16.shape[1] * 2.0**-24 * absb + 1e-30)
    C64, Cabs = oracle.bcsr_spmm_f64(hdr, idx, oracle.f16_bits_to_f32(data), 16, 16,
                                     B.astype(np.float32))
    assert np.allclose(C64, exact, rtol=0, atol=1e-12 * absb.max())
    assert np.allclose(Cabs, absb, rtol=1e-12)


def test_closed_forms():
    # identity blocks: C = B exactly; one-hot B: C = a column of A exactly
    n = 128
    I16 = oracle.f32_to_f16_bits(np.eye(n, dtype=np.float32))
    B16 = oracle.f32_to_f16_bits(oracle.generate_matrix(n, 48, 0.0, "random", 16, 2))
    hdr, idx, data = oracle.bcsr_from_dense(I16, 16, 16)
    C = oracle.bcsr_spmm_f32(hdr, idx, data, 16, 16, oracle.f16_bits_to_f32(B16))
    assert np.array_equal(C, oracle.f16_bits_to_f32(B16))
    A16, _ = small_case(M=128, K=128, N=16)
    onehot = np.zeros((128, 16), dtype=np.float32)
    for j in range(16):
        onehot[(7 * j + 3) % 128, j] = 1.0
    hdr, idx, data = oracle.bcsr_from_dense(A16, 16, 16)
    C = oracle.bcsr_spmm_f32(hdr, idx, data, 16, 16, onehot)
    Af = oracle.f16_bits_to_f32(A16)
    for j in range(16):
        assert np.array_equal(C[:, j], Af[:, (7 * j + 3) % 128])


def test_planted_mistakes_are_caught():
    """A transposed operand, a dropped block or a wrong index all change the
    product beyond the tolerance the GPU tests use (1e-3 of |A||B|)."""
    A16, B16 = small_case(M=128, K=128, N=128)
    Bf = oracle.f16_bits_to_f32(B16)
    hdr, idx, data = oracle.bcsr_from_dense(A16, 16, 16)
    good = oracle.bcsr_spmm_f32(hdr, idx, data, 16, 16, Bf)
    _, Cabs = oracle.bcsr_spmm_f64(hdr, idx, oracle.f16_bits_to_f32(data), 16, 16, Bf)

    def caught(bad):
        return np.any(np.abs(bad - good) > 1e-3 * Cabs)

    assert caught(oracle.bcsr_spmm_f32(hdr, idx, data.transpose(0, 2, 1).copy(), 16, 16, Bf))
    idx2 = idx.copy()
    idx2[0] = (idx2[0] + 1) % 8
    assert caught(oracle.bcsr_spmm_f32(hdr, idx2, data, 16, 16, Bf))
    hdr2 = hdr.copy()
    hdr2[1:] -= 1
    hdr2[0] = 0
    hdr2 = np.maximum(hdr2, 0)
    assert caught(oracle.bcsr_spmm_f32(hdr2, idx[1:], data[1:], 16, 16, Bf))


# ---- the 16384^2 workloads the bench and the GPU tests time -----------------------
@pytest.mark.slow
@pytest.mark.parametrize("name", ["c1_16x16_d1", "c1_16x16_d5", "c1_16x16_d20", "c1_32x16_d5",
                                  "c1_64x16_d5", "c1_blockdiag256"])
def test_library_generator_matches_large_golden(name):
    """The library's generator (mm.generate_bcsr, or the dense generator plus
    BCSRMatrix for the 256 x 256 block-diagonal) gives, at 16384^2, the
    arrays whose hashes were written from the reference's own
    matrix_generator.cpp (matrix_generator.cpp:28-140, formats/BCSRMatrix.cu:22-74)."""
    import matrixmultiplicationcuda_amd as mm

    g = golden()["large"][name]
    if g["pattern"] == "blockdiagonal":
        A = mm.generate(g["M"], g["K"], g["sparsity"], g["pattern"], bh=g["blocksize"],
                        seed=g["seed"])
        h = mm.bcsr_from_dense(A, g["bh"], g["bw"])
        del A
        data16 = np.ascontiguousarray(h.data, dtype=np.float16)
    else:
        assert g["blocksize"] == g["bh"]
        h = mm.generate_bcsr(g["M"], g["K"], g["sparsity"], g["pattern"], bh=g["bh"], bw=g["bw"],
                             seed=g["seed"])
        data16 = np.ascontiguousarray(h.data.astype(np.float16))
    assert h.nnzb == g["nnzb"]
    assert sha(np.ascontiguousarray(h.hdr, np.int32)) == g["sha_hdr"]
    assert sha(np.ascontiguousarray(h.idx, np.int32)) == g["sha_idx"]
    assert sha(data16.view(np.uint16)) == g["sha_data"]


@pytest.mark.slow
@pytest.mark.parametrize("name", ["B_1024x256", "B_16384x64", "B_16384x256", "B_16384x512"])
def test_library_B_matches_golden(name):
    """mm.generate's B (seed 456, fp16) at every N the bench and tests use."""
    import matrixmultiplicationcuda_amd as mm

    g = golden()["B"][name]
    B = mm.generate(g["K"], g["N"], 0.0, "random", bh=16, seed=g["seed"])
    assert sha(np.ascontiguousarray(B).view(np.uint16)) == g["sha_f16"]
