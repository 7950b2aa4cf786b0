"""The host-side code under AddressSanitizer and UndefinedBehaviorSanitizer,
on the CPU: the oracle, the library's host immersed-boundary set-up and host
helpers, and the batch arena allocator (tests/native/sanitize_check.cpp).
Built with ROCm's clang: the HIP sources with the sanitizers on their host
code only (-Xarch_host), the oracle as plain C++, linked against the HIP
runtime; run without a GPU.  Leak reports from inside the HIP runtime's own
start-up are suppressed; the program's are not."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
ROCM = "/opt/rocm"
CLANG = os.path.join(ROCM, "llvm", "bin", "clang++")
HIPCC = os.path.join(ROCM, "bin", "hipcc")

pytestmark = pytest.mark.skipif(not (os.path.exists(CLANG) and os.path.exists(HIPCC)), reason="ROCm clang not found")

FLAGS = ["-O1", "-g", "-fno-omit-frame-pointer", "-std=c++17", "-ffp-contract=off"]
HOST_SAN = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined", "-Xarch_host",
            "-fno-sanitize-recover=undefined"]
HIP = ["-x", "hip", "--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    assert r.returncode == 0, " ".join(cmd) + "\n" + r.stdout


def test_host_code_under_asan_and_ubsan(tmp_path):
    d = str(tmp_path)
    _run([CLANG] + FLAGS + ["-fsanitize=address,undefined", "-fno-sanitize-recover=undefined", "-pthread", "-c",
                            os.path.join(ROOT, "oracle", "cfs_oracle.cpp"), "-o", d + "/oracle.o"])
    _run([HIPCC] + HIP + FLAGS + HOST_SAN + ["-c", os.path.join(ROOT, "2d_compressible_flow_solver_amd", "csrc",
                                                                 "ib_setup.hip"), "-o", d + "/ib_setup.o"])
    _run([HIPCC] + HIP + FLAGS + HOST_SAN + ["-c", os.path.join(HERE, "native", "sanitize_check.cpp"),
                                             "-o", d + "/check.o"])
    exe = d + "/sanitize_check"
    _run([CLANG, "-fsanitize=address,undefined", d + "/check.o", d + "/ib_setup.o", d + "/oracle.o",
          "-L" + ROCM + "/lib", "-lamdhip64", "-Wl,-rpath," + ROCM + "/lib", "-pthread", "-o", exe])
    supp = d + "/lsan.supp"
    with open(supp, "w") as f:
        f.write("leak:libamdhip64\nleak:libhsa-runtime64\n")
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:halt_on_error=1",
               LSAN_OPTIONS="suppressions=" + supp, UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([exe], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, env=env, timeout=600)
    print(r.stdout)
    assert r.returncode == 0, r.stdout
    lines = r.stdout.strip().splitlines()
    assert lines[-1] == "0" and {"oracle ok", "host ok", "arena ok"} <= set(lines)
