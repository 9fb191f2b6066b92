"""The C++20 drop-in headers (include/blocksp_include.hpp) used the way the
reference's users write code (tests/cpp/dropin_test.cpp)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpp", "dropin_test.cpp")
BIN = os.path.join(ROOT, "tests", "cpp", "dropin_test")


def test_dropin_headers_compile():
    res = subprocess.run(["g++", "-std=c++20", "-fsyntax-only", "-Wall", "-Wextra", "-Werror",
                          "-I" + os.path.join(ROOT, "include"), SRC], capture_output=True, text=True)
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
