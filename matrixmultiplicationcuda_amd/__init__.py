"""MI355X-native BCSR x dense matrix multiplication.

Python mirror of the C ABI in include/bcsr_mi355x.h (libbcsr_mi355x.so, built
from csrc/ by build.py).  The product path is the HIP library; nothing here
computes on the CPU except host-side data preparation (generation, BCSR
construction), which itself runs in the C++ library.

Reference counterparts:
  generate_matrix       mg::generate_matrix<T>      matrix_generator.cpp:28-140
  generate              extended generator (new patterns/value distributions)
  bcsr_from_dense       BCSRMatrix::BCSRMatrix      formats/BCSRMatrix.cu:22-74
  Plan / spmm           copyToDevice + sparseMatrixMulTensor* kernels
                        formats/BCSRMatrix.cu:106-132, src/cuda_kernels.cu
  mat_read / mat_write  Matrix(const string&)       Matrix.cu:14-32
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libbcsr_mi355x.so")

PREC_F16, PREC_BF16, PREC_FP8_E4M3, PREC_F16_HILO = 0, 1, 2, 3
PRECISIONS = {"f16": PREC_F16, "bf16": PREC_BF16, "fp8": PREC_FP8_E4M3, "hilo": PREC_F16_HILO}

STATUS = {
    0: "BCSR_OK",
    -1: "BCSR_ERR_INVALID_ARG",
    -2: "BCSR_ERR_BLOCK_SIZE",
    -3: "BCSR_ERR_SHAPE",
    -4: "BCSR_ERR_HIP",
    -5: "BCSR_ERR_OUT_OF_MEMORY",
    -6: "BCSR_ERR_NOT_REGISTERED",
    -7: "BCSR_ERR_UNSUPPORTED",
    -8: "BCSR_ERR_UNKNOWN_PATTERN",
    -9: "BCSR_ERR_SCRATCH_LIMIT",
}

# The reference's pattern list (nvbench_main.cu:56-64) plus this library's
# extensions.
REFERENCE_PATTERNS = ["random", "checkerboard", "diagonal", "blockdiagonal", "blockrandom",
                      "pattern64by16", "pattern32by16"]
EXTRA_PATTERNS = ["blockband", "blockrandom_rho"]

# Every drop-in entry point and its block height (src/cuda_kernels.cuh:39-114).
ENTRY_POINTS = {
    "sparseMatrixMulTensor": 16,
    "sparseMatrixMulTensor_v2": 16,
    "sparseMatrixMulTensor_v3": 16,
    "sparseMatrixMulTensor_v1_improved": 16,
    "sparseMatrixMulTensor_option2_ldmatrix_sm80": 16,
    "sparseMatrixMulTensor64x16": 64,
    "sparseMatrixMulTensor64x16_v2": 64,
    "sparseMatrixMulTensor32x16": 32,
    "sparseMatrixMulTensor32x16_shared": 32,
    "sparseMatrixMulTensor32x16_shared_vectorized": 32,
    "sparseMatrixMulTensor32x16_v2": 32,
    "sparseMatrixMulTensor32x16_v2_shared_vectorized": 32,
}
ADAPTIVE_ENTRY_POINTS = {
    "sparseMatrixMulTensorAdaptive": 16,
    "sparseMatrixMulTensorAdaptive32x16": 32,
    "sparseMatrixMulTensorAdaptive64x16": 64,
}


class BcsrError(RuntimeError):
    def __init__(self, status: int, what: str = ""):
        self.status = status
        name = STATUS.get(status, str(status))
        super().__init__(f"{what}: {name}" if what else name)


class bcsr_launch_cfg(ctypes.Structure):
    _fields_ = [("tile_n", ctypes.c_int32), ("xcd_map", ctypes.c_int32),
                ("variant", ctypes.c_int32), ("waves_per_block", ctypes.c_int32),
                ("reserved", ctypes.c_int32 * 4)]


class bcsr_plan_info(ctypes.Structure):
    _fields_ = [("rows", ctypes.c_int32), ("cols", ctypes.c_int32), ("bh", ctypes.c_int32),
                ("bw", ctypes.c_int32), ("block_rows", ctypes.c_int32),
                ("block_cols", ctypes.c_int32), ("precision", ctypes.c_int32),
                ("dev_bh", ctypes.c_int32), ("nnzb", ctypes.c_int64),
                ("dev_pairs", ctypes.c_int64), ("device_bytes", ctypes.c_int64)]


class bcsr_batch_cfg(ctypes.Structure):
    _fields_ = [("waves_per_cu", ctypes.c_int32), ("bin_work", ctypes.c_int32),
                ("queues", ctypes.c_int32), ("reserved", ctypes.c_int32 * 5)]


class bcsr_batch_info(ctypes.Structure):
    _fields_ = [("count", ctypes.c_int32), ("items", ctypes.c_int32), ("grid", ctypes.c_int32),
                ("precision", ctypes.c_int32), ("flops", ctypes.c_double),
                ("device_bytes", ctypes.c_int64)]


_lib_handle = None

_P = ctypes.c_void_p
_I32 = ctypes.c_int32
_I64P = ctypes.POINTER(ctypes.c_int64)
_I32P = ctypes.POINTER(ctypes.c_int32)


def _declare(lib):
    def f(name, res, *args):
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = list(args)

    i = ctypes.c_int
    f("bcsr_status_string", ctypes.c_char_p, i)
    f("bcsr_last_hip_error", i)
    f("bcsr_plan_create", i, _P, _P, _P, _I32, _I32, _I32, _I32, ctypes.POINTER(_P))
    f("bcsr_plan_create_f32", i, _P, _P, _P, _I32, _I32, _I32, _I32, _I32, ctypes.POINTER(_P))
    f("bcsr_plan_destroy", i, _P)
    f("bcsr_plan_get_info", i, _P, ctypes.POINTER(bcsr_plan_info))
    f("bcsr_spmm", i, _P, _P, _P, _I32, _P)
    f("bcsr_spmm_cfg", i, _P, _P, _P, _I32, ctypes.POINTER(bcsr_launch_cfg), _P)
    f("bcsr_spmm_repeat", i, _P, _P, _P, _I32, ctypes.POINTER(bcsr_launch_cfg), _I32, _P)
    f("bcsr_spmm_time", i, _P, _P, _P, _I32, ctypes.POINTER(bcsr_launch_cfg), _I32, _I32, _P,
      ctypes.POINTER(ctypes.c_float))
    f("bcsr_spmm_kernel_name", i, _P, _I32, ctypes.POINTER(bcsr_launch_cfg), ctypes.c_char_p,
      ctypes.c_size_t)
    f("bcsr_batch_tile_supported", i, _I32, _I32, _I32)
    f("bcsr_batch_create", i, _P, _P, _P, _P, _P, _I32, ctypes.POINTER(bcsr_batch_cfg),
      ctypes.POINTER(_P))
    f("bcsr_batch_run", i, _P, _P)
    f("bcsr_batch_run_graph", i, _P, _P)
    f("bcsr_batch_time", i, _P, _I32, _I32, _I32, _P, ctypes.POINTER(ctypes.c_float))
    f("bcsr_batch_get_info", i, _P, ctypes.POINTER(bcsr_batch_info))
    f("bcsr_batch_tiles", i, _P, _P)
    f("bcsr_batch_destroy", i, _P)
    f("bcsr_cm_create", i, _I32, _I32, _I32, ctypes.c_uint32, ctypes.POINTER(_P))
    f("bcsr_cm_destroy", i, _P)
    f("bcsr_cm_num_params", i, _P, _I32P)
    for name in ("bcsr_cm_set_params", "bcsr_cm_get_params", "bcsr_cm_get_grads"):
        f(name, i, _P, _P, _I32)
    f("bcsr_cm_forward", i, _P, _P, _I32, _P, _P)
    f("bcsr_cm_backward", i, _P, _P, _P, _P, _I32, ctypes.c_float,
      ctypes.POINTER(ctypes.c_float), _P)
    f("bcsr_cm_train", i, _P, _P, _P, _P, _I32, ctypes.c_float, _I32, ctypes.c_float,
      ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.POINTER(ctypes.c_float), _P)
    f("bcsr_csr_spmm", i, _P, _P, _P, _P, _P, _I32, _I32, _I32, _I32, _P)
    f("sparseMatrixMult1Co", i, _P, _P, _P, _P, _P, ctypes.c_uint, _P)
    f("sparseMatrixMult1", i, _P, _P, _P, _P, _P, ctypes.c_uint, _P)
    f("bcsr_add_into", i, _P, _P, ctypes.c_int64, _P)
    f("addMatrices", i, _P, _P, ctypes.c_uint, _P)
    f("bcsr_hybrid_spmm", i, _P, _P, _P, _P, _P, _P, _I32, _P)
    f("bcsr_dense_gemm_f16", i, _P, _P, _P, _I32, _I32, _I32, _P)
    f("bcsr_rocsparse_create", i, _P, _P, _P, _I32, _I32, ctypes.c_int64, _I32, _P, _P, _I32,
      ctypes.POINTER(ctypes.c_void_p))
    f("bcsr_rocsparse_csr_create", i, _P, _P, _P, _I32, _I32, ctypes.c_int64, _P, _P, _I32,
      ctypes.POINTER(ctypes.c_void_p))
    f("bcsr_rocsparse_run", i, _P, _P)
    f("bcsr_rocsparse_destroy", i, _P)
    f("bcsr_copy_to_device", i, _P, _P, _P, _I32, _I32, _I32, _I32, ctypes.POINTER(_P),
      ctypes.POINTER(_P), ctypes.POINTER(_P))
    f("bcsr_release_device", i, _P)
    f("bcsr_plan_from_device", i, _P, ctypes.POINTER(_P))
    for name in ENTRY_POINTS:
        f(name, i, _P, _P, _P, _P, _P, ctypes.c_uint, ctypes.c_uint, _P)
    f("sparseMatrixMulTensor1", i, _P, _P, _P, _P, _P, ctypes.c_uint, _P)
    for name in ADAPTIVE_ENTRY_POINTS:
        f(name, i, _P, _P, _P, _P, _P, ctypes.c_uint, ctypes.c_uint, ctypes.c_uint, _P)
    f("mg_generate_matrix", i, _I32, _I32, ctypes.c_double, ctypes.c_char_p, _I32, ctypes.c_uint32,
      _I32, _P)
    for name in ("mg_generate_f32", "mg_generate_f16"):
        f(name, i, _I32, _I32, ctypes.c_double, ctypes.c_char_p, _I32, _I32, _I32, ctypes.c_double,
          _I32, ctypes.c_double, ctypes.c_double, ctypes.c_uint32, _P)
    f("mg_generate_bcsr", i, _I32, _I32, ctypes.c_double, ctypes.c_char_p, _I32, _I32, _I32,
      ctypes.c_double, _I32, ctypes.c_double, ctypes.c_double, ctypes.c_uint32, _P, _I64P,
      ctypes.POINTER(_P), ctypes.POINTER(_P))
    f("mg_free", None, _P)
    f("bcsr_host_count", i, _P, _I32, _I32, _I32, _I32, _P, _I64P)
    f("bcsr_host_fill", i, _P, _I32, _I32, _I32, _I32, _P, _P, _P)
    f("mat_read_header", i, ctypes.c_char_p, _I32P, _I32P)
    f("mat_read", i, ctypes.c_char_p, _I32, _I32, _P, _I32P)
    f("mat_write", i, ctypes.c_char_p, _I32, _I32, _P)
    f("bcsr_device_count", i, _I32P)
    f("bcsr_set_device", i, _I32)
    f("bcsr_dev_malloc", i, ctypes.POINTER(_P), ctypes.c_size_t)
    f("bcsr_dev_free", i, _P)
    f("bcsr_memcpy_htod", i, _P, _P, ctypes.c_size_t)
    f("bcsr_memcpy_dtoh", i, _P, _P, ctypes.c_size_t)
    f("bcsr_memcpy_dtod", i, _P, _P, ctypes.c_size_t, _P)
    f("bcsr_memset_async", i, _P, _I32, ctypes.c_size_t, _P)
    f("bcsr_copy_kernel", i, _P, _P, ctypes.c_size_t, _P)
    f("bcsr_quantize_b_e4m3", i, _P, _I32, _I32, _P, _P, _P, _P)
    f("bcsr_b8_scale_bytes", ctypes.c_int64, _I32, _I32)
    f("bcsr_spmm_b8", i, _P, _P, _P, _P, _I32, _P, _P)
    f("bcsr_spmm_noalloc", i, _P, _P, _P, _I32, ctypes.POINTER(bcsr_launch_cfg), _P)
    f("bcsr_stream_create", i, ctypes.POINTER(_P))
    f("bcsr_stream_destroy", i, _P)
    f("bcsr_stream_synchronize", i, _P)
    f("bcsr_device_synchronize", i)
    f("bcsr_event_create", i, ctypes.POINTER(_P))
    f("bcsr_event_destroy", i, _P)
    f("bcsr_event_record", i, _P, _P)
    f("bcsr_event_elapsed_ms", i, _P, _P, ctypes.POINTER(ctypes.c_float))
    f("bcsr_graph_begin", i, _P)
    f("bcsr_graph_end", i, _P, ctypes.POINTER(_P))
    f("bcsr_graph_launch", i, _P, _P)
    f("bcsr_graph_destroy", i, _P)


def lib():
    """The loaded C ABI.  Raises if the HIP library has not been built: there
    is no fallback implementation."""
    global _lib_handle
    if _lib_handle is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; "
                f"g.build()'` (or make -C matrixmultiplicationcuda_amd/csrc)")
        h = ctypes.CDLL(LIB_PATH)
        _declare(h)
        _lib_handle = h
    return _lib_handle


def check(status: int, what: str = ""):
    if status != 0:
        raise BcsrError(status, what)


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


# ---- host-side generator and formats --------------------------------------
def generate_matrix(rows, cols, sparsity=0.7, pattern="random", blocksize=16, seed=123,
                    elem_type="float"):
    """mg::generate_matrix<T>; returns a float64 array holding the T values."""
    et = {"float": 0, "double": 1, "int": 2, "longlong": 3}[elem_type]
    out = np.empty((rows, cols), dtype=np.float64)
    check(lib().mg_generate_matrix(rows, cols, float(sparsity), pattern.encode(), blocksize,
                                   seed, et, _ptr(out)), f"generate_matrix({pattern})")
    return out


def generate(rows, cols, sparsity=0.7, pattern="random", bh=16, bw=None, seed=123,
             band_halfwidth=8, inblock_density=1.0, value_dist="uniform", lo=1e-3, hi=1e2,
             dtype=np.float16):
    """Extended generator (mg::generate_into).  dtype float16 rounds
    double -> float -> half as nvbench_main.cu:67-75 does."""
    bw = bh if bw is None else bw
    vd = {"uniform": 0, "loguniform": 1}[value_dist]
    if dtype == np.float16:
        out = np.empty((rows, cols), dtype=np.float16)
        fn = lib().mg_generate_f16
    else:
        out = np.empty((rows, cols), dtype=np.float32)
        fn = lib().mg_generate_f32
    check(fn(rows, cols, float(sparsity), pattern.encode(), bh, bw, band_halfwidth,
             float(inblock_density), vd, float(lo), float(hi), seed, _ptr(out)),
          f"generate({pattern})")
    return out


@dataclass
class HostBCSR:
    hdr: np.ndarray  # int32[block_rows + 1]
    idx: np.ndarray  # int32[nnzb]
    data: np.ndarray  # float16 or float32 [nnzb, bh, bw]
    block_rows: int
    block_cols: int
    bh: int
    bw: int

    @property
    def nnzb(self) -> int:
        return int(self.hdr[-1])

    @property
    def shape(self):
        return self.block_rows * self.bh, self.block_cols * self.bw


def generate_bcsr(rows, cols, sparsity, pattern, bh=16, bw=16, seed=123, band_halfwidth=8,
                  inblock_density=1.0, value_dist="uniform", lo=1e-3, hi=1e2) -> HostBCSR:
    """BCSR (float32 block values) of generate(...), any pattern, built one
    band of block-rows at a time without the dense matrix (mg::generate_bcsr)."""
    vd = {"uniform": 0, "loguniform": 1}[value_dist]
    hdr = np.zeros(rows // bh + 1, dtype=np.int32)
    nnzb = ctypes.c_int64(0)
    pidx, pdata = ctypes.c_void_p(), ctypes.c_void_p()
    check(lib().mg_generate_bcsr(rows, cols, float(sparsity), pattern.encode(), bh, bw,
                                 band_halfwidth, float(inblock_density), vd, float(lo), float(hi),
                                 seed, _ptr(hdr), ctypes.byref(nnzb), ctypes.byref(pidx),
                                 ctypes.byref(pdata)), f"generate_bcsr({pattern})")
    n = nnzb.value
    try:
        idx = np.ctypeslib.as_array(ctypes.cast(pidx, ctypes.POINTER(ctypes.c_int32)),
                                    shape=(max(n, 1),))[:n].copy()
        data = np.ctypeslib.as_array(ctypes.cast(pdata, ctypes.POINTER(ctypes.c_float)),
                                     shape=(max(n, 1) * bh * bw,))[:n * bh * bw].copy()
    finally:
        lib().mg_free(pidx)
        lib().mg_free(pdata)
    return HostBCSR(hdr, idx, data.reshape(n, bh, bw), rows // bh, cols // bw, bh, bw)


def bcsr_from_dense(A: np.ndarray, bh=16, bw=16) -> HostBCSR:
    """BCSRMatrix(Matrix, bh, bw) through the C++ host library."""
    A = np.ascontiguousarray(A, dtype=np.float16)
    rows, cols = A.shape
    if bh > 0 and rows % bh == 0:
        nbr = rows // bh
    else:
        nbr = max(rows // max(bh, 1), 0)
    hdr = np.zeros(nbr + 1, dtype=np.int32)
    nnzb = ctypes.c_int64(0)
    check(lib().bcsr_host_count(_ptr(A), rows, cols, bh, bw, _ptr(hdr), ctypes.byref(nnzb)),
          "bcsr_host_count")
    n = nnzb.value
    idx = np.zeros(max(n, 1), dtype=np.int32)
    data = np.zeros((max(n, 1), bh, bw), dtype=np.float16)
    check(lib().bcsr_host_fill(_ptr(A), rows, cols, bh, bw, _ptr(hdr), _ptr(idx), _ptr(data)),
          "bcsr_host_fill")
    return HostBCSR(hdr, idx[:n], data[:n], nbr, cols // bw, bh, bw)


def bcsr_values_f32(h: HostBCSR, A32: np.ndarray) -> np.ndarray:
    """Block values of the same structure taken from an fp32 matrix."""
    out = np.empty((h.nnzb, h.bh, h.bw), dtype=np.float32)
    for br in range(h.block_rows):
        for k in range(h.hdr[br], h.hdr[br + 1]):
            c = h.idx[k]
            out[k] = A32[br * h.bh:(br + 1) * h.bh, c * h.bw:(c + 1) * h.bw]
    return out


def mat_read(path: str):
    r, c = ctypes.c_int32(), ctypes.c_int32()
    check(lib().mat_read_header(path.encode(), ctypes.byref(r), ctypes.byref(c)), f"mat {path}")
    out = np.empty((r.value, c.value), dtype=np.float16)
    nz = ctypes.c_int32()
    check(lib().mat_read(path.encode(), r.value, c.value, _ptr(out), ctypes.byref(nz)),
          f"mat {path}")
    return out, nz.value


def mat_write(path: str, A: np.ndarray):
    A = np.ascontiguousarray(A, dtype=np.float16)
    check(lib().mat_write(path.encode(), A.shape[0], A.shape[1], _ptr(A)), f"mat {path}")


# ---- device memory, streams and events (the library's own HIP runtime) ------
def device_count() -> int:
    n = ctypes.c_int32(0)
    rc = lib().bcsr_device_count(ctypes.byref(n))
    return n.value if rc == 0 else 0


def synchronize():
    check(lib().bcsr_device_synchronize(), "hipDeviceSynchronize")


class DeviceArray:
    """A HIP device allocation holding a dense row-major array."""

    def __init__(self, shape, dtype):
        self.shape = tuple(int(x) for x in shape)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape, dtype=np.int64)) * self.dtype.itemsize
        self._ptr = ctypes.c_void_p()
        check(lib().bcsr_dev_malloc(ctypes.byref(self._ptr), self.nbytes), "hipMalloc")

    @classmethod
    def from_numpy(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        check(lib().bcsr_memcpy_htod(d._ptr, _ptr(a), a.nbytes), "hipMemcpy H2D")
        return d

    @property
    def ptr(self):
        return self._ptr

    def offset(self, nbytes):
        return ctypes.c_void_p(self._ptr.value + nbytes)

    def fill_bytes(self, value=0, stream=None):
        check(lib().bcsr_memset_async(self._ptr, value, self.nbytes, stream_handle(stream)),
              "hipMemsetAsync")

    def numpy(self):
        out = np.empty(self.shape, dtype=self.dtype)
        check(lib().bcsr_memcpy_dtoh(_ptr(out), self._ptr, self.nbytes), "hipMemcpy D2H")
        return out

    def free(self):
        if self._ptr:
            check(lib().bcsr_dev_free(self._ptr), "hipFree")
            self._ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Stream:
    def __init__(self):
        self.handle = ctypes.c_void_p()
        check(lib().bcsr_stream_create(ctypes.byref(self.handle)), "hipStreamCreate")

    def synchronize(self):
        check(lib().bcsr_stream_synchronize(self.handle), "hipStreamSynchronize")

    def destroy(self):
        if self.handle:
            check(lib().bcsr_stream_destroy(self.handle), "hipStreamDestroy")
            self.handle = ctypes.c_void_p()


class Event:
    def __init__(self):
        self.handle = ctypes.c_void_p()
        check(lib().bcsr_event_create(ctypes.byref(self.handle)), "hipEventCreate")

    def record(self, stream=None):
        check(lib().bcsr_event_record(self.handle, stream_handle(stream)), "hipEventRecord")

    def elapsed_ms(self, end: "Event") -> float:
        ms = ctypes.c_float()
        check(lib().bcsr_event_elapsed_ms(self.handle, end.handle, ctypes.byref(ms)),
              "hipEventElapsedTime")
        return ms.value

    def destroy(self):
        if self.handle:
            check(lib().bcsr_event_destroy(self.handle), "hipEventDestroy")
            self.handle = ctypes.c_void_p()


class Graph:
    """A HIP graph of this library's launches, captured from a Stream:

        g = Graph.capture(stream, lambda: plan.spmm_ptr(B, C, N, stream))
        g.launch(stream)   # replays the recorded kernels
    """

    def __init__(self, handle):
        self.handle = handle

    @classmethod
    def capture(cls, stream: "Stream", record) -> "Graph":
        check(lib().bcsr_graph_begin(stream_handle(stream)), "bcsr_graph_begin")
        h = ctypes.c_void_p()
        try:
            record()
        except BaseException:
            lib().bcsr_graph_end(stream_handle(stream), ctypes.byref(h))  # stop capturing
            if h:
                lib().bcsr_graph_destroy(h)
            raise
        check(lib().bcsr_graph_end(stream_handle(stream), ctypes.byref(h)), "bcsr_graph_end")
        g = cls(h)
        g._keep = record  # its closure holds the plans and arrays the graph points at
        return g

    def launch(self, stream=None):
        check(lib().bcsr_graph_launch(self.handle, stream_handle(stream)), "bcsr_graph_launch")

    def destroy(self):
        if self.handle:
            check(lib().bcsr_graph_destroy(self.handle), "bcsr_graph_destroy")
            self.handle = ctypes.c_void_p()


def stream_handle(stream):
    """None = the null stream; a Stream; or a raw hipStream_t integer."""
    if stream is None:
        return ctypes.c_void_p()
    if isinstance(stream, Stream):
        return stream.handle
    if isinstance(stream, ctypes.c_void_p):
        return stream
    return ctypes.c_void_p(int(stream))


def make_cfg(tile_n=0, xcd_map=-1, variant=-1, waves_per_block=0):
    """Launch configuration; 0 / -1 fields are chosen automatically."""
    cfg = bcsr_launch_cfg()
    cfg.tile_n, cfg.xcd_map, cfg.variant = tile_n, xcd_map, variant
    cfg.waves_per_block = waves_per_block
    return cfg


class Plan:
    """A BCSR matrix resident on the GPU in the engine's device layout."""

    def __init__(self, h: HostBCSR, precision="f16"):
        self._lib = lib()
        self._handle = ctypes.c_void_p()
        prec = PRECISIONS[precision] if isinstance(precision, str) else int(precision)
        hdr = np.ascontiguousarray(h.hdr, dtype=np.int32)
        idx = np.ascontiguousarray(h.idx, dtype=np.int32) if h.nnzb else np.zeros(1, np.int32)
        if h.data.dtype == np.float16 and prec == PREC_F16:
            data = np.ascontiguousarray(h.data, dtype=np.float16)
            rc = self._lib.bcsr_plan_create(_ptr(hdr), _ptr(idx), _ptr(data) if h.nnzb else None,
                                            h.block_rows, h.block_cols, h.bh, h.bw,
                                            ctypes.byref(self._handle))
        else:
            data = np.ascontiguousarray(h.data, dtype=np.float32)
            rc = self._lib.bcsr_plan_create_f32(_ptr(hdr), _ptr(idx),
                                                _ptr(data) if h.nnzb else None, h.block_rows,
                                                h.block_cols, h.bh, h.bw, prec,
                                                ctypes.byref(self._handle))
        check(rc, "bcsr_plan_create")
        self.rows, self.cols = h.shape
        self.bh, self.bw = h.bh, h.bw
        self.nnzb = h.nnzb

    @property
    def info(self) -> bcsr_plan_info:
        inf = bcsr_plan_info()
        check(self._lib.bcsr_plan_get_info(self._handle, ctypes.byref(inf)), "plan info")
        return inf

    def kernel_name(self, N, cfg=None) -> str:
        buf = ctypes.create_string_buffer(256)
        check(self._lib.bcsr_spmm_kernel_name(self._handle, N,
                                              ctypes.byref(cfg) if cfg is not None else None,
                                              buf, 256), "kernel name")
        return buf.value.decode()

    def spmm_ptr(self, B_ptr, C_ptr, N, stream=None, cfg=None):
        st = stream_handle(stream)
        if cfg is None:
            rc = self._lib.bcsr_spmm(self._handle, B_ptr, C_ptr, N, st)
        else:
            rc = self._lib.bcsr_spmm_cfg(self._handle, B_ptr, C_ptr, N, ctypes.byref(cfg), st)
        check(rc, "bcsr_spmm")

    def spmm_repeat_ptr(self, B_ptr, C_ptr, N, reps, stream=None, cfg=None):
        check(self._lib.bcsr_spmm_repeat(self._handle, B_ptr, C_ptr, N,
                                         ctypes.byref(cfg) if cfg is not None else None, reps,
                                         stream_handle(stream)), "bcsr_spmm_repeat")

    def time_ptr(self, B_ptr, C_ptr, N, stream=None, cfg=None, warmup=2, reps=5) -> float:
        """Average milliseconds of one launch (HIP events, synchronises)."""
        ms = ctypes.c_float(0.0)
        check(self._lib.bcsr_spmm_time(self._handle, B_ptr, C_ptr, N,
                                       ctypes.byref(cfg) if cfg is not None else None, warmup,
                                       reps, stream_handle(stream), ctypes.byref(ms)),
              "bcsr_spmm_time")
        return ms.value

    def supports(self, N, cfg) -> bool:
        """Whether a launch configuration (None: the automatic one) is
        available for this plan and N."""
        buf = ctypes.create_string_buffer(256)
        return self._lib.bcsr_spmm_kernel_name(self._handle, N, ctypes.byref(cfg) if cfg is not None else None,
                                               buf, 256) == 0

    def spmm(self, B: DeviceArray, C: DeviceArray = None, stream=None, cfg=None) -> DeviceArray:
        """C = A . B on the GPU (B fp16 K x N, C fp32 M x N DeviceArrays)."""
        if B.dtype != np.float16 or len(B.shape) != 2 or B.shape[0] != self.cols:
            raise ValueError("B must be a K x N float16 DeviceArray")
        N = B.shape[1]
        if C is None:
            C = DeviceArray((self.rows, N), np.float32)
        elif C.dtype != np.float32 or C.shape != (self.rows, N):
            raise ValueError("C must be an M x N float32 DeviceArray")
        self.spmm_ptr(B.ptr, C.ptr, N, stream, cfg)
        return C

    def matmul(self, B16: np.ndarray, cfg=None) -> np.ndarray:
        """Host convenience: upload B, multiply, download C (synchronous)."""
        B = DeviceArray.from_numpy(np.ascontiguousarray(B16, dtype=np.float16))
        C = self.spmm(B, cfg=cfg)
        synchronize()
        out = C.numpy()
        B.free()
        C.free()
        return out

    def close(self):
        if self._handle:
            check(self._lib.bcsr_plan_destroy(self._handle), "bcsr_plan_destroy")
            self._handle = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Many products C_i = plans[i] . B_i in one persistent launch.

    plans: Plan objects of one precision; Bs / Cs: DeviceArrays (K_i x N_i
    fp16, M_i x N_i fp32); tiles: per-problem tile widths (None/0 = auto)."""

    def __init__(self, plans, Bs, Cs, tiles=None, waves_per_cu=0, bin_work=0, queues=0):
        self._lib = lib()
        n = len(plans)
        if len(Bs) != n or len(Cs) != n:
            raise ValueError("plans, Bs and Cs must have the same length")
        self._keep = (list(plans), list(Bs), list(Cs))
        hp = (_P * max(n, 1))(*[p._handle.value for p in plans])
        hb = (_P * max(n, 1))(*[b.ptr for b in Bs])
        hc = (_P * max(n, 1))(*[c.ptr for c in Cs])
        nn = np.array([b.shape[1] for b in Bs] or [0], dtype=np.int32)
        tt = np.array(list(tiles) if tiles is not None else [0] * n, dtype=np.int32)
        if tt.size == 0:
            tt = np.zeros(1, np.int32)
        cfg = bcsr_batch_cfg(waves_per_cu, bin_work, queues)
        self._handle = ctypes.c_void_p()
        check(self._lib.bcsr_batch_create(hp, hb, hc, _ptr(nn), _ptr(tt), n, ctypes.byref(cfg),
                                          ctypes.byref(self._handle)), "bcsr_batch_create")
        self.count = n

    @property
    def info(self) -> bcsr_batch_info:
        inf = bcsr_batch_info()
        check(self._lib.bcsr_batch_get_info(self._handle, ctypes.byref(inf)), "batch info")
        return inf

    @property
    def tiles(self):
        out = np.zeros(max(self.count, 1), np.int32)
        check(self._lib.bcsr_batch_tiles(self._handle, _ptr(out)), "batch tiles")
        return out[:self.count]

    def run(self, stream=None, graph=False):
        if graph:
            check(self._lib.bcsr_batch_run_graph(self._handle, stream_handle(stream)),
                  "bcsr_batch_run_graph")
        else:
            check(self._lib.bcsr_batch_run(self._handle, stream_handle(stream)), "bcsr_batch_run")

    def time(self, stream=None, graph=False, warmup=2, reps=5) -> float:
        """Average milliseconds of one batch run (HIP events, synchronises)."""
        ms = ctypes.c_float(0.0)
        check(self._lib.bcsr_batch_time(self._handle, int(graph), warmup, reps,
                                        stream_handle(stream), ctypes.byref(ms)), "bcsr_batch_time")
        return ms.value

    def close(self):
        if self._handle:
            check(self._lib.bcsr_batch_destroy(self._handle), "bcsr_batch_destroy")
            self._handle = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
