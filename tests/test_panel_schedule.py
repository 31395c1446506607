"""The panel kernel's chunk schedule (matrixmultiplicationcuda_amd/csrc/
panel_schedule.cpp) replayed on the CPU by tests/panel_emulator.cpp: every
staged load, record field and MFMA operand index of kernels_panel.hip is
checked against its buffer, every block is consumed exactly once per column
tile, and the emulated product equals a direct BCSR x B product.  Built with
AddressSanitizer and UndefinedBehaviorSanitizer (host code only)."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_panel_schedule_emulation(tmp_path):
    csrc = os.path.join(ROOT, "matrixmultiplicationcuda_amd", "csrc")
    exe = str(tmp_path / "panel_emulator")
    subprocess.run(["g++", "-O1", "-g", "-std=c++17", "-Wall", "-Werror", "-fsanitize=address,undefined",
                    "-fno-sanitize-recover=undefined", "-I" + csrc,
                    os.path.join(ROOT, "tests", "panel_emulator.cpp"), os.path.join(csrc, "panel_schedule.cpp"),
                    "-o", exe], check=True, timeout=300)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=600, env=env)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.startswith("ok "), r.stdout
