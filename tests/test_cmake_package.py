"""The CMake package: a caller written as the reference's callers are
(integration/seal/SEAL-4.4.5-AXHEL.patch:19-45: find_package(AXHEL CONFIG
REQUIRED), then target_link_libraries(... AXHEL::axhel)) configures, builds
and links against this build.  CPU only: configure, compile, link and check
the loader resolves both libraries; tests/test_gpu_cxx.py runs the result."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG_PREFIX = os.path.join(ROOT, "axhel_amd")
LIBDIR = os.path.join(ROOT, "axhel_amd", "lib")

pytestmark = pytest.mark.skipif(shutil.which("cmake") is None, reason="cmake not installed")


def configure(src, build, *extra):
    return subprocess.run(["cmake", "-S", src, "-B", build, f"-DCMAKE_PREFIX_PATH={PKG_PREFIX}",
                           "-DCMAKE_CXX_COMPILER=g++", *extra], capture_output=True, text=True, timeout=300)


def test_package_files_sit_beside_the_libraries():
    for f in ("AXHELConfig.cmake", "AXHELConfigVersion.cmake"):
        assert os.path.exists(os.path.join(LIBDIR, "cmake", "AXHEL", f)), f


def test_consumer_configures_builds_and_links(tmp_path):
    build = str(tmp_path / "build")
    r = configure(os.path.join(ROOT, "axhel_amd", "cmake", "example"), build)
    assert r.returncode == 0, r.stdout + r.stderr
    r = subprocess.run(["cmake", "--build", build, "-j", "4"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    exe = os.path.join(build, "axhel_cxx_check_cmake")
    assert os.access(exe, os.X_OK)
    ldd = subprocess.run(["ldd", exe], capture_output=True, text=True, timeout=60).stdout
    for lib in ("libaxhel.so", "libaxhel_hip.so"):
        line = next(x for x in ldd.splitlines() if lib + " =>" in x)
        assert os.path.join(LIBDIR, lib) in line, line


def test_target_properties_and_version_rule(tmp_path):
    """AXHEL::axhel carries the public headers and links AXHEL::axhel_hip;
    the version rule is the reference's ExactVersion (1.0.0)."""
    src = tmp_path / "src"
    src.mkdir()
    (src / "CMakeLists.txt").write_text("\n".join([
        "cmake_minimum_required(VERSION 3.13)",
        "project(probe LANGUAGES CXX)",
        "find_package(AXHEL 1.0.0 CONFIG REQUIRED)",
        "get_target_property(inc AXHEL::axhel INTERFACE_INCLUDE_DIRECTORIES)",
        "get_target_property(dep AXHEL::axhel INTERFACE_LINK_LIBRARIES)",
        'message(STATUS "AXHEL_INC=${inc}")',
        'message(STATUS "AXHEL_DEP=${dep}")',
        'message(STATUS "AXHEL_VER=${AXHEL_VERSION}")',
    ]) + "\n")
    r = configure(str(src), str(tmp_path / "b1"))
    assert r.returncode == 0, r.stdout + r.stderr
    assert f"AXHEL_INC={os.path.join(ROOT, 'include')}" in r.stdout
    assert os.path.exists(os.path.join(ROOT, "include", "axhel", "axhel.hpp"))
    assert "AXHEL_DEP=AXHEL::axhel_hip" in r.stdout and "AXHEL_VER=1.0.0" in r.stdout
    (src / "CMakeLists.txt").write_text("\n".join([
        "cmake_minimum_required(VERSION 3.13)",
        "project(probe LANGUAGES CXX)",
        "find_package(AXHEL 2.0 CONFIG REQUIRED)",
    ]) + "\n")
    r = configure(str(src), str(tmp_path / "b2"))
    assert r.returncode != 0 and "1.0.0" in r.stdout + r.stderr
