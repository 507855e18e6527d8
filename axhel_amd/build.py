"""Builds the gfx950 libraries in-tree.

  axhel_amd/lib/libaxhel_hip.so  HIP kernels + the C ABI (include/axhel_hip.h)
  axhel_amd/lib/libaxhel.so      the unipi::axhel C++ drop-in API over the C ABI
  axhel_amd/lib/axhel_cxx_check  a C++ program exercising the drop-in API
  axhel_amd/lib/axhel_serve_bench  the serving-loop load driver (Poisson arrivals)
  axhel_amd/lib/libaxhel_arith_hooks.so  test hooks: each device arithmetic
                                 primitive applied element-wise (tests only)
  axhel_amd/lib/cmake/AXHEL/     the CMake package (find_package(AXHEL), AXHEL::axhel)
  axhel_amd/lib/axhel_cxx_check_cmake  the same check program built by CMake
                                 through that package (axhel_amd/cmake/example)

Usage: python -m axhel_amd.build [-v] [-j N]
"""
import concurrent.futures
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
OBJDIR = os.path.join(HERE, "_build")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

HIP_SOURCES = ["ntt_cfg_16a.hip", "ntt_cfg_16b.hip", "ntt_cfg_14_15.hip", "ntt_cfg_17.hip", "ntt_cfg_18_19.hip", "ntt_cfg_20.hip",
               "ntt_cfg_1_13.hip",
               "common.hip", "eltwise.hip", "ntt.hip", "batch.hip", "serve.hip", "ubench.hip", "polymul.hip", "rns.hip"]
HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # packed FP32 instructions stay off in every HIP translation unit
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-command-line-argument",
    "-I", os.path.join(ROOT, "include"),
]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("build step failed: " + " ".join(cmd[:3]) + " ... " + cmd[-1])
    return r


def _compile_hip(src, verbose):
    obj = os.path.join(OBJDIR, os.path.splitext(src)[0] + ".o")
    _run([HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj], verbose)
    return obj


def build(verbose=False, jobs=None):
    os.makedirs(LIBDIR, exist_ok=True)
    os.makedirs(OBJDIR, exist_ok=True)
    jobs = jobs or min(8, os.cpu_count() or 4)
    sources = [s for s in HIP_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with concurrent.futures.ThreadPoolExecutor(max_workers=jobs) as ex:
        hooks_obj = ex.submit(_compile_hip, "arith_hooks.hip", verbose)
        objs = list(ex.map(lambda s: _compile_hip(s, verbose), sources))
        hooks_obj = hooks_obj.result()
    hip_lib = os.path.join(LIBDIR, "libaxhel_hip.so")
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-Wl,-soname,libaxhel_hip.so", "-o", hip_lib + ".tmp"]
         + objs + ["-lpthread", "-ldl"], verbose)
    os.replace(hip_lib + ".tmp", hip_lib)
    hooks_lib = os.path.join(LIBDIR, "libaxhel_arith_hooks.so")
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", hooks_lib + ".tmp", hooks_obj], verbose)
    os.replace(hooks_lib + ".tmp", hooks_lib)

    # C++ drop-in API: plain C++ over the C ABI (no HIP headers).
    cxx_lib = os.path.join(LIBDIR, "libaxhel.so")
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", "-Wextra",
          "-I", os.path.join(ROOT, "include"), os.path.join(CSRC, "cxx_api.cpp"),
          "-L", LIBDIR, "-laxhel_hip", "-Wl,-rpath,$ORIGIN", "-Wl,-soname,libaxhel.so", "-o", cxx_lib + ".tmp"],
         verbose)
    os.replace(cxx_lib + ".tmp", cxx_lib)

    check = os.path.join(LIBDIR, "axhel_cxx_check")
    _run(["g++", "-O2", "-std=c++17", "-Wall", "-Wextra", "-I", os.path.join(ROOT, "include"),
          os.path.join(CSRC, "cxx_check.cpp"), "-L", LIBDIR, "-laxhel", "-laxhel_hip",
          "-Wl,-rpath,$ORIGIN", "-o", check + ".tmp"], verbose)
    os.replace(check + ".tmp", check)

    # serving benchmark driver (host program over the C ABI)
    sb = os.path.join(LIBDIR, "axhel_serve_bench")
    _run([HIPCC, "-O2", "-std=c++17", f"--offload-arch={ARCH}", "-Wno-unused-command-line-argument",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-I", os.path.join(ROOT, "include"), os.path.join(CSRC, "serve_bench.hip"),
          "-L", LIBDIR, "-laxhel_hip", "-lpthread", "-Wl,-rpath,$ORIGIN", "-o", sb + ".tmp"], verbose)
    os.replace(sb + ".tmp", sb)

    # CMake package beside the libraries, and a consumer built through it
    pkg = os.path.join(LIBDIR, "cmake", "AXHEL")
    os.makedirs(pkg, exist_ok=True)
    for f in ("AXHELConfig.cmake", "AXHELConfigVersion.cmake"):
        shutil.copyfile(os.path.join(HERE, "cmake", f + ".in"), os.path.join(pkg, f))
    bdir = os.path.join(OBJDIR, "cmake_example")
    shutil.rmtree(bdir, ignore_errors=True)
    _run(["cmake", "-S", os.path.join(HERE, "cmake", "example"), "-B", bdir, "-DCMAKE_BUILD_TYPE=Release",
          f"-DCMAKE_PREFIX_PATH={HERE}", "-DCMAKE_CXX_COMPILER=g++"], verbose)
    _run(["cmake", "--build", bdir, "-j", "4"], verbose)
    ck = os.path.join(LIBDIR, "axhel_cxx_check_cmake")
    shutil.copyfile(os.path.join(bdir, "axhel_cxx_check_cmake"), ck + ".tmp")
    os.chmod(ck + ".tmp", 0o755)
    os.replace(ck + ".tmp", ck)
    return hip_lib


if __name__ == "__main__":
    j = None
    if "-j" in sys.argv:
        j = int(sys.argv[sys.argv.index("-j") + 1])
    build(verbose="-v" in sys.argv, jobs=j)
    print("built", LIBDIR)
