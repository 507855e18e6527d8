"""Host sanitizers on the CPU: the serving scheduler (NULL backend) and the
C++ drop-in's error paths, built with AddressSanitizer, and the scheduler
with ThreadSanitizer, then stressed by tests/sanitize/host_check.cpp.

The recipe: every HIP translation unit of the library compiled host-only
(--cuda-host-only), each with -Xarch_host before the -fsanitize= option so
only host code is instrumented, embedding the device code the library build
already made for that unit (its .hip_fatbin section, passed back with
-fcuda-include-gpubinary) instead of compiling the kernels again; the C++
drop-in and the driver built with the same clang; everything linked into
one executable so the sanitizer runtime is the executable's.  The reference builds its Debug configuration
with AddressSanitizer (cmake/axhel/axhel-util.cmake:66-74)."""
import concurrent.futures
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "axhel_amd", "csrc")
OBJDIR = os.path.join(ROOT, "axhel_amd", "_build")  # the library build's objects
HIPCC = "/opt/rocm/bin/hipcc"
CLANG = "/opt/rocm/llvm/bin/clang++"

pytestmark = [
    pytest.mark.skipif(not (os.path.exists(HIPCC) and os.path.exists(CLANG)), reason="ROCm clang missing"),
    pytest.mark.skipif(not os.path.exists(os.path.join(ROOT, "axhel_amd", "_build", "serve.o")),
                       reason="library objects not built (python -m axhel_amd.build)"),
]


def hip_sources():
    from axhel_amd import build
    return build.HIP_SOURCES


def build_checker(kind, out_dir):
    """kind: 'address' or 'thread'; returns the executable's path."""
    os.makedirs(out_dir, exist_ok=True)
    common = ["-O1", "-g", "-fno-omit-frame-pointer", "-std=c++17", "-I", os.path.join(ROOT, "include")]

    def hip_obj(src):
        stem = os.path.splitext(src)[0]
        obj = os.path.join(out_dir, stem + ".o")
        fatbin = os.path.join(out_dir, stem + ".fatbin")
        r = subprocess.run(["objcopy", "-O", "binary", "--only-section=.hip_fatbin", os.path.join(OBJDIR, stem + ".o"),
                            fatbin], capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        # a unit without kernels has no device code to embed
        embed = ["-Xclang", "-fcuda-include-gpubinary", "-Xclang", fatbin] if os.path.getsize(fatbin) else []
        cmd = [HIPCC, "--offload-arch=gfx950", "--cuda-host-only"] + embed + ["-Xarch_host", f"-fsanitize={kind}",
               "-Wno-unused-command-line-argument"] + common + ["-c", os.path.join(CSRC, src), "-o", obj]
        r = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
        assert r.returncode == 0, r.stderr[-3000:]
        return obj

    def cxx_obj(path):
        obj = os.path.join(out_dir, os.path.basename(path) + ".o")
        r = subprocess.run([CLANG, f"-fsanitize={kind}"] + common + ["-c", path, "-o", obj],
                           capture_output=True, text=True, timeout=600)
        assert r.returncode == 0, r.stderr[-3000:]
        return obj

    with concurrent.futures.ThreadPoolExecutor(max_workers=8) as ex:
        objs = list(ex.map(hip_obj, hip_sources()))
    objs += [cxx_obj(os.path.join(CSRC, "cxx_api.cpp")),
             cxx_obj(os.path.join(ROOT, "tests", "sanitize", "host_check.cpp"))]
    exe = os.path.join(out_dir, f"host_check_{kind}")
    r = subprocess.run([CLANG, f"-fsanitize={kind}", "-o", exe] + objs +
                       ["-L/opt/rocm/lib", "-lamdhip64", "-lpthread", "-ldl", "-Wl,-rpath,/opt/rocm/lib"],
                       capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-3000:]
    return exe


def run(exe, args, env_extra):
    env = dict(os.environ)
    env.update(env_extra)
    env["HIP_VISIBLE_DEVICES"] = ""  # the NULL backend needs no device
    return subprocess.run([exe] + args, capture_output=True, text=True, timeout=600, env=env)


def test_address_sanitizer(tmp_path):
    exe = build_checker("address", str(tmp_path / "asan"))
    r = run(exe, [], {"ASAN_OPTIONS": "detect_leaks=1:abort_on_error=0:halt_on_error=1"})
    assert r.returncode == 0 and "host_check ok" in r.stdout, r.stdout[-2000:] + r.stderr[-6000:]
    assert "ERROR: AddressSanitizer" not in r.stderr and "ERROR: LeakSanitizer" not in r.stderr


def test_thread_sanitizer(tmp_path):
    exe = build_checker("thread", str(tmp_path / "tsan"))
    r = run(exe, ["--scheduler-only"], {"TSAN_OPTIONS": "halt_on_error=1"})
    assert r.returncode == 0 and "host_check ok" in r.stdout, r.stdout[-2000:] + r.stderr[-6000:]
    assert "WARNING: ThreadSanitizer" not in r.stderr
