"""ctypes loader of the oracle (liboracle.so) and of oracle/_ref.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker.  The product never imports it.
See oracle.h for what each function restates and what pins it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libmg_ref.so")
REFERENCE_TREE = "/root/reference"

_o = None
_r = None
_P = ctypes.c_void_p


def build(with_ref: bool | None = None):
    """Compile liboracle.so (and oracle/_ref when the reference tree exists)."""
    subprocess.run(["make", "-s", "-C", HERE, "all"], check=True)
    if with_ref is None:
        with_ref = os.path.isdir(REFERENCE_TREE)
    if with_ref:
        subprocess.run(["make", "-s", "-C", HERE, "ref", f"REF={REFERENCE_TREE}"], check=True)


def lib():
    global _o
    if _o is None:
        if not os.path.exists(LIB):
            raise RuntimeError(f"{LIB} missing; run `make -C oracle`")
        o = ctypes.CDLL(LIB)
        i, d, u32 = ctypes.c_int, ctypes.c_double, ctypes.c_uint32
        sz = ctypes.c_size_t

        def f(name, res, *args):
            fn = getattr(o, name)
            fn.restype = res
            fn.argtypes = list(args)

        f("orc_mt_stream", None, u32, _P, sz)
        f("orc_f2h", ctypes.c_uint16, ctypes.c_float)
        f("orc_h2f", ctypes.c_float, ctypes.c_uint16)
        f("orc_f32_to_f16", None, _P, _P, sz)
        f("orc_f16_to_f32", None, _P, _P, sz)
        f("orc_generate_matrix", i, i, i, d, ctypes.c_char_p, i, u32, i, _P)
        f("orc_generate_ex", i, i, i, d, ctypes.c_char_p, i, i, u32, i, d, d, i, d, _P)
        f("orc_bcsr_count", ctypes.c_int64, _P, i, i, i, i, _P)
        f("orc_bcsr_fill", None, _P, i, i, i, i, _P, _P, _P)
        f("orc_bcsr_fill_f32", None, _P, i, i, i, i, _P, _P, _P)
        f("orc_bcsr_densify", None, _P, _P, _P, i, i, i, i, _P)
        f("orc_csr_count", ctypes.c_int64, _P, i, i, _P)
        f("orc_csr_fill", None, _P, i, i, _P, _P, _P)
        f("orc_matmul_dense_ref", None, _P, _P, i, i, i, _P)
        f("orc_bcsr_spmm_f32", None, _P, _P, _P, i, i, i, i, _P, i, _P, i)
        f("orc_bcsr_spmm_f64", None, _P, _P, _P, i, i, i, i, _P, i, _P, _P, i)
        f("orc_f2e4m3", ctypes.c_uint8, ctypes.c_float)
        f("orc_e4m32f", ctypes.c_float, ctypes.c_uint8)
        f("orc_e8m0_for_absmax", ctypes.c_uint8, ctypes.c_float)
        f("orc_quant_e4m3_groups", None, _P, ctypes.c_int64, ctypes.c_int64, _P, _P)
        f("orc_quant_e4m3_B", None, _P, i, i, _P, _P)
        f("orc_quant_e4m3_tiles", None, _P, _P, _P, i, i, i, _P)
        f("orc_quant_e4m3_blocks_f32scale", None, _P, ctypes.c_int64, ctypes.c_int64, _P)
        f("orc_csr_spmm_f32", None, _P, _P, _P, i, _P, i, _P)
        f("orc_csr_spmm_acc_f32", None, _P, _P, _P, i, _P, i, _P)
        f("orc_f2bf16", ctypes.c_uint16, ctypes.c_float)
        f("orc_bf162f", ctypes.c_float, ctypes.c_uint16)
        f("orc_hilo_split", None, ctypes.c_float, ctypes.POINTER(ctypes.c_uint16),
          ctypes.POINTER(ctypes.c_uint16))
        f("orc_f2bf16_n", None, _P, ctypes.c_int64, _P)
        f("orc_hilo_split_n", None, _P, ctypes.c_int64, _P, _P)
        _o = o
    return _o


def ref_available() -> bool:
    return os.path.exists(REF_LIB)


def ref_lib():
    global _r
    if _r is None:
        r = ctypes.CDLL(REF_LIB)
        r.ref_generate_matrix.restype = ctypes.c_int
        r.ref_generate_matrix.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                          ctypes.c_char_p, ctypes.c_int, ctypes.c_uint,
                                          ctypes.c_int, _P]
        r.ref_generate_matrix_f32.restype = ctypes.c_int
        r.ref_generate_matrix_f32.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                              ctypes.c_char_p, ctypes.c_int, ctypes.c_uint, _P]
        _r = r
    return _r


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


ELEM = {"float": 0, "double": 1, "int": 2, "longlong": 3}


def mt_stream(seed, n):
    out = np.empty(n, dtype=np.uint32)
    lib().orc_mt_stream(seed, _p(out), n)
    return out


def generate_matrix(rows, cols, sparsity, pattern, blocksize, seed, elem_type="float"):
    out = np.empty((rows, cols), dtype=np.float64)
    rc = lib().orc_generate_matrix(rows, cols, sparsity, pattern.encode(), blocksize, seed,
                                   ELEM[elem_type], _p(out))
    if rc != 0:
        raise ValueError(f"Unknown pattern: {pattern}")
    return out


def ref_generate_matrix(rows, cols, sparsity, pattern, blocksize, seed, elem_type="float"):
    out = np.empty((rows, cols), dtype=np.float64)
    rc = ref_lib().ref_generate_matrix(rows, cols, sparsity, pattern.encode(), blocksize, seed,
                                       ELEM[elem_type], _p(out))
    if rc != 0:
        raise ValueError(f"Unknown pattern: {pattern}")
    return out


def ref_generate_f32(rows, cols, sparsity, pattern, blocksize, seed):
    out = np.empty((rows, cols), dtype=np.float32)
    rc = ref_lib().ref_generate_matrix_f32(rows, cols, sparsity, pattern.encode(), blocksize,
                                           seed, _p(out))
    if rc != 0:
        raise ValueError(f"Unknown pattern: {pattern}")
    return out


def generate_ex(rows, cols, sparsity, pattern, bh=16, bw=None, seed=123, value_dist="uniform",
                lo=1e-3, hi=1e2, band_halfwidth=8, inblock_density=1.0):
    bw = bh if bw is None else bw
    out = np.empty((rows, cols), dtype=np.float32)
    vd = {"uniform": 0, "loguniform": 1}[value_dist]
    rc = lib().orc_generate_ex(rows, cols, sparsity, pattern.encode(), bh, bw, seed, vd, lo, hi,
                               band_halfwidth, inblock_density, _p(out))
    if rc != 0:
        raise ValueError(f"Unknown pattern: {pattern}")
    return out


def f32_to_f16_bits(a):
    a = np.ascontiguousarray(a, dtype=np.float32)
    out = np.empty(a.shape, dtype=np.uint16)
    lib().orc_f32_to_f16(_p(a), _p(out), a.size)
    return out


def f16_bits_to_f32(h):
    h = np.ascontiguousarray(h, dtype=np.uint16)
    out = np.empty(h.shape, dtype=np.float32)
    lib().orc_f16_to_f32(_p(h), _p(out), h.size)
    return out


def bcsr_from_dense(A16_bits, bh, bw):
    """(hdr, idx, data uint16 [nnzb, bh, bw]) of BCSRMatrix(A, bh, bw)."""
    A = np.ascontiguousarray(A16_bits, dtype=np.uint16)
    rows, cols = A.shape
    hdr = np.zeros(rows // bh + 1, dtype=np.int32)
    n = lib().orc_bcsr_count(_p(A), rows, cols, bh, bw, _p(hdr))
    idx = np.zeros(max(n, 1), dtype=np.int32)
    data = np.zeros((max(n, 1), bh, bw), dtype=np.uint16)
    lib().orc_bcsr_fill(_p(A), rows, cols, bh, bw, _p(hdr), _p(idx), _p(data))
    return hdr, idx[:n], data[:n]


def bcsr_values_f32(A32, hdr, idx, bh, bw):
    A = np.ascontiguousarray(A32, dtype=np.float32)
    rows, cols = A.shape
    n = int(hdr[-1])
    data = np.zeros((max(n, 1), bh, bw), dtype=np.float32)
    idx = np.ascontiguousarray(idx, dtype=np.int32) if n else np.zeros(1, np.int32)
    lib().orc_bcsr_fill_f32(_p(A), rows, cols, bh, bw, _p(np.ascontiguousarray(hdr, np.int32)),
                            _p(idx), _p(data))
    return data[:n]


def bcsr_densify(hdr, idx, data, block_rows, block_cols, bh, bw):
    out = np.empty((block_rows * bh, block_cols * bw), dtype=np.uint16)
    idx = np.ascontiguousarray(idx, dtype=np.int32) if len(idx) else np.zeros(1, np.int32)
    data = np.ascontiguousarray(data, dtype=np.uint16) if len(data) else np.zeros(1, np.uint16)
    lib().orc_bcsr_densify(_p(np.ascontiguousarray(hdr, np.int32)), _p(idx), _p(data), block_rows,
                           block_cols, bh, bw, _p(out))
    return out


def csr_from_dense(A16_bits):
    A = np.ascontiguousarray(A16_bits, dtype=np.uint16)
    rows, cols = A.shape
    hdr = np.zeros(rows + 1, dtype=np.int32)
    n = lib().orc_csr_count(_p(A), rows, cols, _p(hdr))
    idx = np.zeros(max(n, 1), dtype=np.int32)
    data = np.zeros(max(n, 1), dtype=np.uint16)
    lib().orc_csr_fill(_p(A), rows, cols, _p(hdr), _p(idx), _p(data))
    return hdr, idx[:n], data[:n]


def matmul_dense_ref(A16_bits, B16_bits):
    """cpu_matmul_ref (nvbench_main.cu:78-91), literal dense loop."""
    A = np.ascontiguousarray(A16_bits, dtype=np.uint16)
    B = np.ascontiguousarray(B16_bits, dtype=np.uint16)
    M, K = A.shape
    N = B.shape[1]
    C = np.empty((M, N), dtype=np.float32)
    lib().orc_matmul_dense_ref(_p(A), _p(B), M, K, N, _p(C))
    return C


def bcsr_spmm_f32(hdr, idx, data16, bh, bw, Bf32, br0=0, br1=None, threads=1, C=None):
    """BCSR-order fp32 product, bit-identical to cpu_matmul_ref.  Rows of
    block-rows [br0, br1) of C are written."""
    hdr = np.ascontiguousarray(hdr, dtype=np.int32)
    nbr = len(hdr) - 1
    br1 = nbr if br1 is None else br1
    Bf = np.ascontiguousarray(Bf32, dtype=np.float32)
    N = Bf.shape[1]
    if C is None:
        C = np.zeros((nbr * bh, N), dtype=np.float32)
    idx = np.ascontiguousarray(idx, dtype=np.int32) if len(idx) else np.zeros(1, np.int32)
    data = np.ascontiguousarray(data16, dtype=np.uint16) if len(idx) else np.zeros(1, np.uint16)
    lib().orc_bcsr_spmm_f32(_p(hdr), _p(idx), _p(data), bh, bw, br0, br1, _p(Bf), N, _p(C),
                            threads)
    return C


def bcsr_spmm_f64(hdr, idx, data32, bh, bw, Bf32, br0=0, br1=None, threads=1, want_abs=True):
    """FP64 product on fp32 block values and the |A|.|B| bound."""
    hdr = np.ascontiguousarray(hdr, dtype=np.int32)
    nbr = len(hdr) - 1
    br1 = nbr if br1 is None else br1
    Bf = np.ascontiguousarray(Bf32, dtype=np.float32)
    N = Bf.shape[1]
    C = np.zeros((nbr * bh, N), dtype=np.float64)
    Ca = np.zeros((nbr * bh, N), dtype=np.float64) if want_abs else None
    idx = np.ascontiguousarray(idx, dtype=np.int32) if len(idx) else np.zeros(1, np.int32)
    data = np.ascontiguousarray(data32, dtype=np.float32) if len(idx) else np.zeros(1, np.float32)
    lib().orc_bcsr_spmm_f64(_p(hdr), _p(idx), _p(data), bh, bw, br0, br1, _p(Bf), N, _p(C),
                            _p(Ca) if want_abs else None, threads)
    return C, Ca


# ---- FP8 E4M3 storage mode ------------------------------------------------
def f32_to_e4m3(x):
    x = np.ascontiguousarray(x, dtype=np.float32)
    return np.array([lib().orc_f2e4m3(float(v)) for v in x.ravel()], dtype=np.uint8).reshape(x.shape)


E4M3_TABLE = None


def e4m3_to_f32(q):
    global E4M3_TABLE
    if E4M3_TABLE is None:
        E4M3_TABLE = np.array([lib().orc_e4m32f(b) for b in range(256)], dtype=np.float32)
    return E4M3_TABLE[np.asarray(q, dtype=np.uint8)]


def quant_e4m3_groups(x, group_len):
    """(q, scale_exp): one E8M0 scale per consecutive group of group_len."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    n = x.size // group_len
    q = np.empty(x.shape, dtype=np.uint8)
    sc = np.empty(n, dtype=np.uint8)
    lib().orc_quant_e4m3_groups(_p(x), n, group_len, _p(q), _p(sc))
    return q, sc


def b8_scale_stride(K):
    """Bytes from one chunk's slab scales to the next (rows padded to 16)."""
    return (K // 16 + 15) & ~15


def quant_e4m3_B(B):
    """(q K x N, scale_exp [ceil(N/256), K/16]) of the B quantiser: one scale
    per (16-row slab, 256-column chunk)."""
    B = np.ascontiguousarray(B, dtype=np.float32)
    K, N = B.shape
    chunks, stride = (N + 255) // 256, b8_scale_stride(K)
    q = np.empty((K, N), dtype=np.uint8)
    sc = np.zeros((chunks, stride), dtype=np.uint8)
    if K and N:
        lib().orc_quant_e4m3_B(_p(B), K, N, _p(q), _p(sc))
    return q, sc[:, :K // 16]


def csr_spmm_f32(hdr, idx, data16, Bf32, C0=None):
    """CSR x dense in CSR order (fp32, k ascending per row), from +0 or, with
    C0, as the running sum started from C0 (sparseMatrixMult1's order)."""
    hdr = np.ascontiguousarray(hdr, dtype=np.int32)
    M = len(hdr) - 1
    Bf = np.ascontiguousarray(Bf32, dtype=np.float32)
    N = Bf.shape[1]
    idx = np.ascontiguousarray(idx, dtype=np.int32) if len(idx) else np.zeros(1, np.int32)
    data = np.ascontiguousarray(data16, dtype=np.uint16) if len(data16) else np.zeros(1, np.uint16)
    if C0 is None:
        C = np.zeros((M, N), dtype=np.float32)
        lib().orc_csr_spmm_f32(_p(hdr), _p(idx), _p(data), M, _p(Bf), N, _p(C))
    else:
        C = np.array(C0, dtype=np.float32, order="C", copy=True).reshape(M, N)
        lib().orc_csr_spmm_acc_f32(_p(hdr), _p(idx), _p(data), M, _p(Bf), N, _p(C))
    return C


def fp8_tiles_dequant(hdr, idx, data32, bh, bw):
    """A as the FP8 mode stores it: 16 x 32 tiles (consecutive 16-column
    device blocks of a 16-row slice) with one E8M0 scale each, dequantised."""
    hdr = np.ascontiguousarray(hdr, dtype=np.int32)
    data = np.ascontiguousarray(data32, dtype=np.float32)
    out = np.zeros_like(data)
    idx = np.ascontiguousarray(idx, dtype=np.int32) if len(idx) else np.zeros(1, np.int32)
    if data.size:
        lib().orc_quant_e4m3_tiles(_p(hdr), _p(idx), _p(data), len(hdr) - 1, bh, bw, _p(out))
    return out


def fp8_blocks_f32scale_dequant(data32, blk):
    """A's blocks quantised with one FP32 absmax scale each (the per-block
    FP32 form of FP8 storage), dequantised; blk = values per block."""
    data = np.ascontiguousarray(data32, dtype=np.float32)
    out = np.zeros_like(data)
    if data.size:
        lib().orc_quant_e4m3_blocks_f32scale(_p(data), data.size // blk, blk, _p(out))
    return out


def e8m0_to_f32(e):
    return np.ldexp(np.float32(1.0), np.asarray(e, dtype=np.int32) - 127).astype(np.float32)


def f32_to_bf16_bits(x):
    x = np.ascontiguousarray(x, dtype=np.float32)
    out = np.empty(x.shape, dtype=np.uint16)
    if x.size:
        lib().orc_f2bf16_n(_p(x), x.size, _p(out))
    return out


def bf16_bits_to_f32(b):
    return (np.asarray(b, dtype=np.uint32) << 16).astype(np.uint32).view(np.float32)


def hilo_split(x):
    x = np.ascontiguousarray(x, dtype=np.float32)
    hi = np.empty(x.shape, np.uint16)
    lo = np.empty(x.shape, np.uint16)
    if x.size:
        lib().orc_hilo_split_n(_p(x), x.size, _p(hi), _p(lo))
    return hi, lo


def dequantized_blocks(mode, data32, bh, bw, hdr=None, idx=None):
    """Block values as the storage mode keeps them (fp32-exact); FP8 needs
    the BCSR structure (hdr, idx) because its scales span block pairs."""
    data32 = np.ascontiguousarray(data32, dtype=np.float32)
    if mode == "f16":
        return f16_bits_to_f32(f32_to_f16_bits(data32))
    if mode == "bf16":
        return bf16_bits_to_f32(f32_to_bf16_bits(data32))
    if mode == "hilo":
        hi, lo = hilo_split(data32)
        return (f16_bits_to_f32(hi).astype(np.float64)
                + f16_bits_to_f32(lo).astype(np.float64) * 2.0**-11)
    if mode == "fp8":
        return fp8_tiles_dequant(hdr, idx, data32, bh, bw)
    raise ValueError(mode)


def quantized_B(mode, B16_bits):
    """B (given in fp16) as the mode's kernel multiplies it."""
    Bf = f16_bits_to_f32(B16_bits)
    if mode in ("f16", "hilo"):
        return Bf
    if mode == "bf16":
        return bf16_bits_to_f32(f32_to_bf16_bits(Bf))
    if mode == "fp8":
        q, sc = quant_e4m3_B(Bf)
        K, N = Bf.shape
        # scale of (row k, column n) = that of slab k // 16, chunk n // 256
        scale = np.repeat(np.repeat(e8m0_to_f32(sc).T, 16, axis=0), 256, axis=1)[:K, :N]
        return e4m3_to_f32(q) * scale
    raise ValueError(mode)
