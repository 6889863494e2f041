"""libbricks.so's host code under AddressSanitizer and UndefinedBehaviorSanitizer,
on the CPU: the library is compiled with the sanitizers on its host side only
(each -fsanitize preceded by -Xarch_host; the GPU code is built as usual) and
tests/cpp/host_api_check.cpp drives every host-only C-ABI call -- grid
numbering, lists, ghost pairs and images, plans of every shape and length,
BrickDecomp layouts, argument errors.  Any out-of-bounds access, leak or
undefined behaviour aborts the run."""
import concurrent.futures as cf
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = "/opt/rocm/bin/hipcc"
CLANG = "/opt/rocm/llvm/bin/clang++"

pytestmark = pytest.mark.skipif(not (os.path.exists(HIPCC) and os.path.exists(CLANG)),
                                reason="ROCm compilers not present")


def test_host_api_under_asan_and_ubsan(tmp_path):
    csrc = os.path.join(ROOT, "bricklib_amd", "csrc")
    san = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
           "-Xarch_host", "-fno-omit-frame-pointer"]
    flags = ["-O1", "-g", "-std=c++17", "--offload-arch=gfx950", "-fPIC", "-Xclang",
             "-target-feature", "-Xclang", "-packed-fp32-ops", "-ffp-contract=off",
             "-I", os.path.join(ROOT, "include"), "-I", csrc]
    srcs = sorted(f for f in os.listdir(csrc) if f.endswith(".hip"))

    def compile_one(f):
        obj = str(tmp_path / (f[:-4] + ".o"))
        subprocess.run([HIPCC, *flags, *san, "-c", os.path.join(csrc, f), "-o", obj],
                       check=True, capture_output=True)
        return obj

    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(compile_one, srcs))
    lib = str(tmp_path / "libbricks.so")
    subprocess.run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", lib,
                    "-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
                    "-shared-libasan", "-Wl,-rpath,/opt/rocm/lib"], check=True,
                   capture_output=True)
    exe = str(tmp_path / "host_api_check")
    subprocess.run([CLANG, "-O1", "-g", "-std=c++17", "-fsanitize=address,undefined",
                    "-shared-libasan", "-fno-omit-frame-pointer", "-I",
                    os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "cpp", "host_api_check.cpp"), "-L",
                    str(tmp_path), "-lbricks", "-Wl,-rpath," + str(tmp_path),
                    "-Wl,-rpath,/opt/rocm/lib", "-o", exe], check=True, capture_output=True)
    rt = subprocess.run([CLANG, "-print-file-name=libclang_rt.asan-x86_64.so"], check=True,
                        capture_output=True, text=True).stdout.strip()
    env = dict(os.environ,
               ASAN_OPTIONS="detect_leaks=1:halt_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1",
               LD_LIBRARY_PATH=os.path.dirname(rt) + ":" + os.environ.get("LD_LIBRARY_PATH", ""))
    r = subprocess.run([exe], env=env, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0 and "host api ok" in r.stdout, (r.stdout[-2000:], r.stderr[-4000:])
