"""Host code under AddressSanitizer + UndefinedBehaviorSanitizer (CPU only):
tests/sanitize_driver.cpp drives the host library's generator, .mat I/O and
format constructors and the oracle's C restatement, built with the ROCm
clang++ (host code, no GPU code) and run with leak detection on."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLANG = "/opt/rocm/llvm/bin/clang++"
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=undefined", "-fno-omit-frame-pointer"]


@pytest.mark.skipif(not os.path.exists(CLANG), reason="ROCm clang++ not present")
def test_host_library_and_oracle_under_asan_ubsan(tmp_path):
    csrc = os.path.join(ROOT, "matrixmultiplicationcuda_amd", "csrc")
    oracle_o = str(tmp_path / "oracle.o")
    subprocess.run([CLANG, "-x", "c", "-std=gnu11", "-O1", "-g", *SAN, "-c",
                    os.path.join(ROOT, "oracle", "oracle.c"), "-o", oracle_o], check=True, timeout=300)
    exe = str(tmp_path / "sanitize_driver")
    srcs = [os.path.join(ROOT, "tests", "sanitize_driver.cpp")] + [
        os.path.join(csrc, f) for f in ("matrix.cpp", "matrix_generator.cpp", "formats.cpp", "capi_host.cpp")]
    subprocess.run([CLANG, "-std=c++17", "-O1", "-g", *SAN, "-I" + os.path.join(ROOT, "include"),
                    "-I" + os.path.join(ROOT, "oracle"), *srcs, oracle_o, "-lpthread", "-o", exe],
                   check=True, timeout=600)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([exe, str(tmp_path)], capture_output=True, text=True, timeout=600, env=env)
    assert r.returncode == 0, r.stderr[-4000:]
    assert r.stdout.startswith("ok "), r.stdout
    assert "runtime error" not in r.stderr, r.stderr[-4000:]
