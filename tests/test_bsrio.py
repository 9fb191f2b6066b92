"""BSR cache files: numpy reader/writer round trip (CPU) and the C ABI (GPU)."""
import numpy as np
import pytest

from blocksp_amd import bsrio
from blocksp_amd import problems as pr


def test_numpy_round_trip_keeps_order_and_bits(tmp_path):
    p = pr.stencil2d_b2(6, 5, seed=0)
    ci = p.colidx.copy()
    # scramble each row's column order: the file must keep the stored order
    for i in range(p.n):
        ci[p.rowptr[i]:p.rowptr[i + 1]] = ci[p.rowptr[i]:p.rowptr[i + 1]][::-1]
    f = tmp_path / "a.bsr"
    bsrio.save(f, p.n, p.n, p.bs, p.rowptr, ci, p.vals)
    for mm in (True, False):
        n, m, bs, rp, c2, v = bsrio.load(f, mmap=mm)
        assert (n, m, bs) == (p.n, p.n, p.bs)
        np.testing.assert_array_equal(rp, p.rowptr)
        np.testing.assert_array_equal(c2, ci)
        np.testing.assert_array_equal(v, p.vals)
    assert f.stat().st_size == 32 + 4 * (p.n + 1) + 4 * p.nnzb + 8 * p.nnzb * p.bs * p.bs


def test_bad_files_are_refused(tmp_path):
    f = tmp_path / "x.bsr"
    f.write_bytes(b"NOTABSR\0" + bytes(24))
    with pytest.raises(ValueError):
        bsrio.load(f)


@pytest.mark.gpu
def test_c_abi_save_load_matches_numpy(ctx, tmp_path):
    import blocksp_amd as bsp
    p = pr.dg2d_veclap(6, 5, seed=2)
    A = bsp.BSR.from_problem(ctx, p)
    f1 = tmp_path / "c.bsr"
    A.save(f1)
    n, m, bs, rp, ci, v = bsrio.load(f1)
    np.testing.assert_array_equal(rp, p.rowptr)
    np.testing.assert_array_equal(ci, p.colidx)
    np.testing.assert_array_equal(v, p.vals)
    f2 = tmp_path / "n.bsr"
    bsrio.save(f2, p.n, p.n, p.bs, p.rowptr, p.colidx, 2.0 * p.vals)
    B = bsp.BSR.load(ctx, f2)
    _, _, vb = B.download()
    np.testing.assert_array_equal(vb, 2.0 * p.vals)
