"""The C++20 drop-in headers (include/blocksp_include.hpp) used the way the
reference's users write code (tests/cpp/dropin_test.cpp)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpp", "dropin_test.cpp")
BIN = os.path.join(ROOT, "tests", "cpp", "dropin_test")


@pytest.mark.parametrize("src", [SRC, os.path.join(ROOT, "tests", "cpp", "capi_test.cpp")],
                         ids=["dropin_test", "capi_test"])
def test_dropin_headers_compile(src):
    res = subprocess.run(["g++", "-std=c++20", "-fsyntax-only", "-Wall", "-Wextra", "-Werror",
                          "-I" + os.path.join(ROOT, "include"), src], capture_output=True, text=True)
    assert res.returncode == 0, res.stderr


def test_c_header_is_plain_c():
    hdr = os.path.join(ROOT, "include", "blocksp_c.h")
    res = subprocess.run(["gcc", "-std=c99", "-fsyntax-only", "-Wall", "-Werror", "-x", "c", hdr],
                         capture_output=True, text=True)
    assert res.returncode == 0, res.stderr


@pytest.mark.gpu
def test_dropin_program_on_gpu():
    assert os.path.exists(BIN), "build tests/cpp/dropin_test first (make -C tests/cpp)"
    print("running the C++ drop-in test", flush=True)
    res = subprocess.run([BIN], capture_output=True, text=True, timeout=300)
    print(res.stdout)
    assert res.returncode == 0, res.stdout + res.stderr
    assert "DROPIN PASSED" in res.stdout


SAN_BINS = [("dropin_test_asan", "DROPIN PASSED"), ("capi_test_asan", "CAPI PASSED")]
# allocations the HIP runtime keeps until the process ends are not this library's
LSAN_SUPPRESSIONS = "leak:libamdhip64\nleak:libhsa-runtime64\n"


@pytest.mark.gpu
@pytest.mark.parametrize("name,marker", SAN_BINS)
def test_programs_with_host_sanitizers(name, marker, tmp_path):
    """The drop-in program and the C-ABI program (tests/cpp/capi_test.cpp) linked with the
    library's host code built under AddressSanitizer + UndefinedBehaviorSanitizer
    (make -C tests/cpp asan): no invalid access, leak or undefined behaviour on the host side."""
    exe = os.path.join(ROOT, "tests", "cpp", name)
    assert os.path.exists(exe), "build the sanitized programs first (make -C tests/cpp asan)"
    supp = tmp_path / "lsan.supp"
    supp.write_text(LSAN_SUPPRESSIONS)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:halt_on_error=1",
               LSAN_OPTIONS=f"suppressions={supp}", UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    print("running", name, flush=True)
    res = subprocess.run([exe], capture_output=True, text=True, timeout=300, env=env)
    print(res.stdout[-4000:])
    assert res.returncode == 0, res.stdout[-3000:] + res.stderr[-6000:]
    assert marker in res.stdout
    assert "ERROR: AddressSanitizer" not in res.stderr and "runtime error:" not in res.stderr, res.stderr[-6000:]
