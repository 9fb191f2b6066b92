"""Convert the reference's LDG_2D_vecLap Matrix-Market fixture into a compact npz.

Reads test_mats/LDG_2D_vecLap{,_b,_sol}.mtx from the BlockSp checkout (data files
the reference's own test holds: blocksp_tests.hpp:312-342) and stores the entries
in file order (the order matters: it is the reference's block insertion order,
blocksp_tests.hpp:617-642).  Known answer: ||x - x_sol||_inf ~ 0.00029
(blocksp_tests.hpp:318).

Usage: python tests/golden/make_ldg_fixture.py /root/reference/test_mats
"""
import os
import sys

import numpy as np


def read_matrix(path):
    with open(path) as f:
        M, N, nnz = (int(t) for t in f.readline().split())
        data = np.loadtxt(f, dtype=np.float64, max_rows=nnz)
    i = data[:, 0].astype(np.int32) - 1
    j = data[:, 1].astype(np.int32) - 1
    return M, N, i, j, data[:, 2].copy()


def read_vector(path):
    with open(path) as f:
        M = int(f.readline().split()[0])
        v = np.loadtxt(f, dtype=np.float64, max_rows=M)
    return v


def main(src):
    M, N, i, j, v = read_matrix(os.path.join(src, "LDG_2D_vecLap.mtx"))
    b = read_vector(os.path.join(src, "LDG_2D_vecLap_b.mtx"))
    x = read_vector(os.path.join(src, "LDG_2D_vecLap_sol.mtx"))
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ldg_2d_veclap.npz")
    np.savez_compressed(out, M=M, N=N, i=i, j=j, v=v, b=b, x_sol=x)
    print(out, os.path.getsize(out))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/test_mats")
