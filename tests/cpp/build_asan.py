"""Host-AddressSanitizer build of the library and of the two C++ drivers (drop_in_example.cpp,
fused_driver.cpp), into tests/cpp/asan/ (not in git).  Device code is unchanged: -fsanitize goes to
the host pass only (-Xarch_host).  Run on a GPU machine as
    ASAN_OPTIONS=detect_leaks=0 tests/cpp/asan/drop_in_asan
    ASAN_OPTIONS=detect_leaks=0 tests/cpp/asan/fused_asan
(leak detection off: the HIP runtime keeps allocations until exit)."""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
from bachelors_amd import build as B  # noqa: E402

OUT = os.path.join(HERE, "asan")
OBJ = os.path.join(OUT, "obj")
ASAN = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fno-omit-frame-pointer", "-Xarch_host", "-g"]
CXX = "/opt/rocm/llvm/bin/clang++"   # the sanitizer runtime of the compiler that instrumented the library


def _compile(src):
    out = os.path.join(OBJ, src.replace(".hip", ".o"))
    r = subprocess.run([B.HIPCC] + B.FLAGS + ASAN + ["-c", os.path.join(B.CSRC, src), "-o", out],
                       capture_output=True, text=True)
    if r.returncode:
        raise RuntimeError(r.stderr[-3000:])
    return out


def main():
    os.makedirs(OBJ, exist_ok=True)
    with cf.ThreadPoolExecutor(8) as ex:
        objs = list(ex.map(_compile, B.SOURCES))
    lib = os.path.join(OUT, "libbachelors_fmm.so")
    r = subprocess.run([B.HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-pthread", "-o", lib] + objs,
                       capture_output=True, text=True)
    if r.returncode:
        raise RuntimeError(r.stderr[-3000:])
    common = ["-fsanitize=address", "-fno-omit-frame-pointer", "-g", "-O1", "-std=c++17",
              "-I" + os.path.join(ROOT, "include"), "-L" + OUT, "-lbachelors_fmm", "-Wl,-rpath," + OUT]
    hip = ["-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", "-L/opt/rocm/lib", "-lamdhip64"]
    for src, exe, extra in (("drop_in_example.cpp", "drop_in_asan", []), ("fused_driver.cpp", "fused_asan", hip)):
        r = subprocess.run([CXX, os.path.join(HERE, src), "-o", os.path.join(OUT, exe)] + common + extra,
                           capture_output=True, text=True)
        if r.returncode:
            raise RuntimeError(r.stderr[-3000:])
    shutil.rmtree(OBJ)
    print("built", OUT)


if __name__ == "__main__":
    main()
