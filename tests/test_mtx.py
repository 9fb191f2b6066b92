"""Matrix-Market converters: the C++ drop-in's tests::mtxMatrix_to_BlockSp /
mtxArray_to_solArray (tests/cpp/mtx_test.cpp) and the Python reader (blocksp_amd.mtx), on the
CPU.  Pins: the reference's own LDG_2D_vecLap fixture (entries in file order,
blocksp_tests.hpp:312-342), and a brute-force restatement of setFromGlobal's insertion
(blocksp_utility.hpp:614-624 into blocksp_containers.hpp:147-164: a block is appended to its
row on first touch, an element keeps the last value written) on small files that have
repeated entries, unsorted columns, a symmetric mirror and a size line after comment lines."""
import os
import subprocess

import numpy as np
import pytest

from blocksp_amd import mtx

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpp", "mtx_test.cpp")


@pytest.fixture(scope="module")
def exe(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("mtx") / "mtx_test")
    res = subprocess.run(["g++", "-std=c++20", "-O1", "-Wall", "-Wextra", "-Werror",
                          "-I" + os.path.join(ROOT, "include"), SRC, "-o", out], capture_output=True, text=True)
    assert res.returncode == 0, res.stderr
    return out


@pytest.fixture(scope="module")
def ldg():
    return np.load(os.path.join(ROOT, "tests", "golden", "ldg_2d_veclap.npz"))


def write_matrix(path, M, N, i, j, v, head_lines=()):
    with open(path, "w") as f:
        for h in head_lines:
            f.write(h + "\n")
        f.write(f"{M} {N} {len(v)}\n")
        for a, b, x in zip(i, j, v):
            f.write(f"{int(a) + 1} {int(b) + 1} {float(x)!r}\n")


def write_vector(path, v):
    with open(path, "w") as f:
        f.write(f"{len(v)} 1\n")
        for x in v:
            f.write(f"{float(x)!r}\n")


def run(exe, *args):
    res = subprocess.run([exe, *map(str, args)], capture_output=True, text=True, timeout=120)
    return res.returncode, res.stdout


def parse_matrix(out, bs, cplx=False):
    lines = out.splitlines()
    m, n = (int(t) for t in lines[0].split())
    cols = [[int(t) for t in lines[1 + r].split()[1:]] for r in range(m)]
    vals = [complex(*map(float, ln.split())) if cplx else float(ln) for ln in lines[1 + m:]]
    nb = sum(len(c) for c in cols)
    arr = np.array(vals).reshape(nb, bs, bs)
    return m, n, cols, arr


def brute_force(M, N, i, j, v, bs, symmetric, conj=False):
    """setFromGlobal entry by entry: per block row a first-touch list of columns and a dict
    of blocks; the mirrored entry follows its original."""
    rows = [[] for _ in range(M // bs)]
    blocks = {}

    def put(a, b, x):
        r, c = a // bs, b // bs
        if (r, c) not in blocks:
            rows[r].append(c)
            blocks[(r, c)] = np.zeros((bs, bs), dtype=complex if conj else float)
        blocks[(r, c)][a % bs, b % bs] = x

    for a, b, x in zip(i, j, v):
        put(a, b, x)
        if symmetric and a != b:
            put(b, a, np.conj(x) if conj else x)
    vals = [blocks[(r, c)] for r in range(len(rows)) for c in rows[r]]
    return rows, np.array(vals).reshape(-1, bs, bs)


SMALL = dict(M=8, N=8,
             i=[5, 0, 7, 0, 3, 2, 5, 6, 0],
             j=[1, 0, 7, 6, 3, 4, 1, 2, 6],
             v=[1.5, 2.0, -3.25, 0.125, 7.0, -1.0, 9.5, 4.0, 0.375])   # (5,1) and (0,6) written twice


@pytest.mark.parametrize("bs", [1, 2, 4, 8])
@pytest.mark.parametrize("symmetric", [False, True])
def test_small_file_matches_brute_force(exe, tmp_path, bs, symmetric):
    path = tmp_path / "small.mtx"
    s = SMALL
    write_matrix(path, s["M"], s["N"], s["i"], s["j"], s["v"])
    rc, out = run(exe, "mat", path, 1, int(symmetric), bs)
    assert rc == 0, out
    m, n, cols, vals = parse_matrix(out, bs)
    rows, bvals = brute_force(s["M"], s["N"], s["i"], s["j"], s["v"], bs, symmetric)
    assert (m, n) == (8 // bs, 8 // bs)
    assert cols == rows
    np.testing.assert_array_equal(vals, bvals)
    # the Python reader and assembly agree with both
    M, N, i, j, v = mtx.read_mtx_entries(path)
    rp, ci, pv = mtx.entries_to_bsr(M, N, i, j, v, bs, symmetric=symmetric)
    assert [list(ci[rp[r]:rp[r + 1]]) for r in range(m)] == rows
    np.testing.assert_array_equal(pv, bvals)


def test_size_line_after_comments(exe, tmp_path):
    """start_line names the size line (GotoLine skips start_line - 1 lines); 0 and 1 both
    mean the first line."""
    s = SMALL
    path = tmp_path / "hdr.mtx"
    write_matrix(path, s["M"], s["N"], s["i"], s["j"], s["v"],
                 head_lines=("%%MatrixMarket matrix coordinate real general", "% a comment"))
    rc, out = run(exe, "mat", path, 3, 0, 2)
    assert rc == 0, out
    _, _, cols, vals = parse_matrix(out, 2)
    rows, bvals = brute_force(s["M"], s["N"], s["i"], s["j"], s["v"], 2, False)
    assert cols == rows
    np.testing.assert_array_equal(vals, bvals)
    M, N, i, j, v = mtx.read_mtx_entries(path, start_line=3)
    np.testing.assert_array_equal(i, s["i"])
    np.testing.assert_array_equal(v, s["v"])
    plain = tmp_path / "plain.mtx"
    write_matrix(plain, s["M"], s["N"], s["i"], s["j"], s["v"])
    assert run(exe, "mat", plain, 0, 0, 2) == run(exe, "mat", plain, 1, 0, 2)


def test_errors(exe, tmp_path):
    s = SMALL
    path = tmp_path / "small.mtx"
    write_matrix(path, s["M"], s["N"], s["i"], s["j"], s["v"])
    rc, out = run(exe, "mat", path, 1, 0, 1)
    assert rc == 0
    odd = tmp_path / "odd.mtx"
    write_matrix(odd, 6, 6, [0], [0], [1.0])
    rc, out = run(exe, "mat", odd, 1, 0, 4)                 # 4 does not divide 6
    assert rc == 3 and out.startswith("ERROR") and "does not divide" in out
    rc, out = run(exe, "mat", tmp_path / "missing.mtx", 1, 0, 1)
    assert rc == 3 and "cannot open" in out
    short = tmp_path / "short.mtx"
    short.write_text("8 8 5\n1 1 2.0\n2 2 3.0\n")
    rc, out = run(exe, "mat", short, 1, 0, 1)
    assert rc == 3 and "ends early" in out
    with pytest.raises(ValueError):
        mtx.read_mtx_entries(short)
    bad = tmp_path / "bad.mtx"
    bad.write_text("8 8 1\n9 1 2.0\n")                       # row 9 of 8
    rc, out = run(exe, "mat", bad, 1, 0, 1)
    assert rc == 3 and "bad entry" in out
    vec = tmp_path / "v.mtx"
    write_vector(vec, [1.0, 2.0, 3.0, 4.0, 5.0, 6.0])           # B*D = 4 does not divide 6
    rc, out = run(exe, "vec", vec, 1)
    assert rc == 3 and "dividing" in out
    with pytest.raises(ValueError):
        mtx.entries_to_bsr(6, 6, [0], [0], [1.0], 4)


def test_ldg_fixture_through_both_readers(exe, tmp_path, ldg):
    """The reference's LDG_2D_vecLap system (8192 x 8192, 188193 entries of the symmetric
    half) read as BlockSp<double, 4> with symmetric = true, the call of
    blocksp_tests.hpp:327 at B * D = 4: the C++ reader, the Python reader + assembly and the
    fixture's own entries agree exactly."""
    M, N = int(ldg["M"]), int(ldg["N"])
    path = tmp_path / "ldg.mtx"
    write_matrix(path, M, N, ldg["i"], ldg["j"], ldg["v"])
    Mr, Nr, i, j, v = mtx.read_mtx_entries(path, start_line=1)
    assert (Mr, Nr) == (M, N)
    np.testing.assert_array_equal(i, ldg["i"])
    np.testing.assert_array_equal(j, ldg["j"])
    np.testing.assert_array_equal(v, ldg["v"])
    rc, out = run(exe, "mat", path, 1, 1, 4)
    assert rc == 0, out[:200]
    m, n, cols, vals = parse_matrix(out, 4)
    rp, ci, pv = mtx.entries_to_bsr(M, N, i, j, v, 4, symmetric=True)
    assert (m, n) == (M // 4, N // 4)
    assert [len(c) for c in cols] == list(np.diff(rp))
    np.testing.assert_array_equal(np.concatenate([np.asarray(c) for c in cols]), ci)
    np.testing.assert_array_equal(vals, pv)
    # the full matrix is symmetric
    dense = np.zeros((M, N))
    for r in range(m):
        for q in range(rp[r], rp[r + 1]):
            dense[4 * r:4 * r + 4, 4 * ci[q]:4 * ci[q] + 4] = pv[q]
    np.testing.assert_array_equal(dense, dense.T)
    # right-hand side: the flat solArray layout (ie*B*D + dim*B + ib) is the file order
    bpath = tmp_path / "b.mtx"
    write_vector(bpath, ldg["b"])
    rc, out = run(exe, "vec", bpath, 1)
    assert rc == 0
    xb = np.array([float(t) for t in out.split()])
    np.testing.assert_array_equal(xb, ldg["b"])
    np.testing.assert_array_equal(mtx.vector_to_solarray(mtx.read_mtx_vector(bpath, 1), 2, 2), ldg["b"])


def test_vector_layout_with_several_dims(exe, tmp_path):
    """solArray<double, 2, 2>: value i lands at element i // 4, dim (i % 4) // 2, entry i % 2
    (setFromGlobal's layout, blocksp_utility.hpp:630-639) -- the flat index i."""
    v = np.arange(1, 13, dtype=float) * 0.5
    path = tmp_path / "v.mtx"
    write_vector(path, v)
    rc, out = run(exe, "vec", path, 1)
    assert rc == 0
    got = np.array([float(t) for t in out.split()])
    el, rem = np.arange(12) // 4, np.arange(12) % 4
    want = np.zeros((3, 2, 2))
    want[el, rem // 2, rem % 2] = v
    np.testing.assert_array_equal(got, want.reshape(-1))


def test_complex_matrix_and_vector(exe, tmp_path):
    """Complex T: "i j re im" entries, the symmetric mirror conjugated
    (blocksp_tests.hpp:628-635), and "re im" vector entries."""
    i, j = [0, 3, 2, 1], [0, 1, 3, 1]
    z = [1 + 2j, -0.5 + 0.25j, 3 - 1j, 2 + 0j]
    path = tmp_path / "c.mtx"
    with open(path, "w") as f:
        f.write(f"4 4 {len(z)}\n")
        for a, b, x in zip(i, j, z):
            f.write(f"{a + 1} {b + 1} {x.real!r} {x.imag!r}\n")
    rc, out = run(exe, "cmat", path, 1, 1)
    assert rc == 0, out
    _, _, cols, vals = parse_matrix(out, 2, cplx=True)
    rows, bvals = brute_force(4, 4, i, j, z, 2, True, conj=True)
    assert cols == rows
    np.testing.assert_array_equal(vals, bvals)
    vpath = tmp_path / "cv.mtx"
    with open(vpath, "w") as f:
        f.write("4 1\n")
        for x in z:
            f.write(f"{x.real!r} {x.imag!r}\n")
    rc, out = run(exe, "cvec", vpath, 1)
    assert rc == 0
    got = [complex(*map(float, ln.split())) for ln in out.splitlines()]
    assert got == z
