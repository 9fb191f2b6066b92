"""Small dense routines of the solvers (host code), on the CPU: Francis double-shift
eigenvalues against closed forms and the inverse-iteration invariant subspace of the deflated
restart against the complex-Schur route (bsp_dense.hpp, tests/cpp/dense_test.cpp), the complex
solve, least squares and deflated restart of the complex pGMRES (bsp_zdense.hpp,
tests/cpp/zdense_test.cpp), both also built with AddressSanitizer and UndefinedBehaviorSanitizer,
and the device least-squares team routines run by one host thread (restart_test.cpp)."""
import os
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

SANITIZE = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-g", "-O1"]


@pytest.mark.parametrize("name,marker", [("dense_test", "DENSE PASSED"), ("zdense_test", "ZDENSE PASSED")])
@pytest.mark.parametrize("sanitize", [False, True], ids=["O2", "asan_ubsan"])
def test_dense_routines(name, marker, sanitize):
    src = os.path.join(ROOT, "tests", "cpp", name + ".cpp")
    flags = SANITIZE if sanitize else ["-O2"]
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, name)
        res = subprocess.run(["g++", "-std=c++20", *flags, "-Wall", "-Wextra", "-Werror", src, "-o", exe],
                             capture_output=True, text=True)
        assert res.returncode == 0, res.stderr
        env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
        run = subprocess.run([exe], capture_output=True, text=True, timeout=300, env=env)
        print(run.stdout[-3000:])
        assert run.returncode == 0, run.stdout[-3000:] + run.stderr[-3000:]
        assert marker in run.stdout


def test_restart_team_routines():
    """The device least squares and deflated restart (bsp_restart.hpp) run by one host thread
    against the host routines (tests/cpp/restart_test.cpp); compiled host-only by hipcc."""
    src = os.path.join(ROOT, "tests", "cpp", "restart_test.cpp")
    hipcc = os.environ.get("HIPCC") or "/opt/rocm/bin/hipcc"
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "restart_test")
        res = subprocess.run([hipcc, "-x", "hip", "--offload-host-only", "-std=c++20", "-O2", "-Wall", src, "-o", exe],
                             capture_output=True, text=True, cwd=d)
        assert res.returncode == 0, res.stderr
        run = subprocess.run([exe], capture_output=True, text=True, timeout=300)
        print(run.stdout[-2000:])
        assert run.returncode == 0, run.stdout[-4000:] + run.stderr
        assert "RESTART PASSED" in run.stdout
