"""Small dense routines of the solvers (blocksp_amd/csrc/bsp_dense.hpp, host code): Francis
double-shift eigenvalues against closed forms and the inverse-iteration invariant subspace of
the deflated restart against the complex-Schur route (tests/cpp/dense_test.cpp), on the CPU."""
import os
import subprocess
import tempfile

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpp", "dense_test.cpp")


def test_dense_routines():
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "dense_test")
        res = subprocess.run(["g++", "-std=c++20", "-O2", "-Wall", "-Wextra", "-Werror", SRC, "-o", exe],
                             capture_output=True, text=True)
        assert res.returncode == 0, res.stderr
        run = subprocess.run([exe], capture_output=True, text=True, timeout=120)
        print(run.stdout)
        assert run.returncode == 0, run.stdout + run.stderr
        assert "DENSE PASSED" in run.stdout


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
