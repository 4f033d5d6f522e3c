"""Matrix exchange formats on the host side of libffhip (no device work):
FreeFEM's Morse text writer (`f << A` of a CSR-state HashMatrix,
HashMatrix.hpp:485-508) and its reader (HashMatrix(istream),
HashMatrix.cpp:7-30, 136-214) for the Morse text and the COO dump.

Bars: against the committed fixtures written by FreeFEM's own compiled
HashMatrix (tests/golden/make_io_fixtures.py) the writer reproduces every byte
after the first comment line (which carries a memory address), and the reader
returns the reference's sorted CSR bit for bit (equal to the oracle's own
assembly of the same matrix); where oracle/_ref is built, the same against the
reference called directly, on larger matrices, both file kinds, duplicates,
comment lines and the half flag.
"""
import os

import numpy as np
import pytest

import ffo

ffhip = pytest.importorskip("ffhip")
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
needs_ref = pytest.mark.skipif(not ffo.ref_available(), reason="oracle/_ref not built")


def _cube_matrix(n, tgv_labels=(1, 2, 3, 4, 5, 6)):
    m, rp, ci, va, onbc, b = ffo.poisson_p1_3d(n, labels=tgv_labels)
    return rp, ci, va


def _lines(path):
    with open(path, "rb") as f:
        return f.read().split(b"\n")


def test_writer_matches_reference_fixture(tmp_path):
    rp, ci, va = _cube_matrix(2)
    out = tmp_path / "a.txt"
    ffhip.csr_write_morse(str(out), rp, ci, va)
    mine, ref = _lines(out), _lines(os.path.join(GOLDEN, "p1_cube2_morse.txt"))
    assert mine[0].startswith(b"# Sparse Matrix (Morse)  0x")
    assert mine[1:] == ref[1:]


@pytest.mark.parametrize("name", ["p1_cube2_morse.txt", "p1_cube2_coo.txt"])
def test_reader_on_reference_fixtures(name):
    rp, ci, va, m, half = ffhip.csr_read_text(os.path.join(GOLDEN, name))
    ro, co, vo = _cube_matrix(2)
    assert m == 27 and half == 0
    assert np.array_equal(rp, ro) and np.array_equal(ci, co) and np.array_equal(va, vo)


def test_round_trip_is_exact(tmp_path):
    rng = np.random.default_rng(7)
    rp, ci, va = _cube_matrix(4)
    va = va * rng.uniform(-3, 3, va.size)  # arbitrary doubles, signs
    va[::17] = 0.0
    va[5] = -0.0
    va[7] = 1e-310  # below removeeps' threshold: written as 0
    p = tmp_path / "r.txt"
    ffhip.csr_write_morse(str(p), rp, ci, va)
    r2, c2, v2, m, half = ffhip.csr_read_text(str(p))
    expect = np.where(np.abs(va) < 1e-305, 0.0, va)  # removeeps: -0 and 1e-310 are written as 0
    assert np.array_equal(r2, rp) and np.array_equal(c2, ci)
    assert np.array_equal(v2.view(np.int64), expect.view(np.int64))


def test_reader_semantics(tmp_path):
    # duplicates (last assignment wins), comment lines, unsorted entries, half flag, rectangular
    txt = ("# Sparse Matrix (Morse)  0x0\n# a comment\n#another\n"
           "3 4 1  5\n 3 1 2.5\n 1 1 1.0\n 3 1 -7\n 2 4 0\n 1 3 1e-3\n")
    p = tmp_path / "s.txt"
    p.write_text(txt)
    rp, ci, va, m, half = ffhip.csr_read_text(str(p))
    assert m == 4 and half == 1
    assert rp.tolist() == [0, 2, 3, 4]
    assert ci.tolist() == [0, 2, 3, 0]
    assert va.tolist() == [1.0, 1e-3, 0.0, -7.0]
    coo = ("#  HashMatrix Matrix (COO) 0x0\n#    n       m        nnz     half     fortran   state  \n"
           "2 2 3 0 1 0 0 \n 2 2 4.0\n 1 2 3.0\n 2 2 5.0\n")  # fortran (1-based) indices
    p2 = tmp_path / "c.txt"
    p2.write_text(coo)
    rp, ci, va, m, half = ffhip.csr_read_text(str(p2))
    assert rp.tolist() == [0, 1, 2] and ci.tolist() == [1, 1] and va.tolist() == [3.0, 5.0]
    p3 = tmp_path / "bad.txt"
    p3.write_text("# something else\n1 1 0 1\n1 1 1\n")
    with pytest.raises(ffhip.FFHipError):
        ffhip.csr_read_text(str(p3))
    p4 = tmp_path / "oob.txt"
    p4.write_text("# Sparse Matrix (Morse)\n2 2 0  1\n3 1 1.0\n")
    with pytest.raises(ffhip.FFHipError):
        ffhip.csr_read_text(str(p4))


@needs_ref
@pytest.mark.parametrize("n", [3, 6])
def test_writer_and_reader_vs_reference(tmp_path, n):
    R = ffo.ref()
    m = ffo.cube(n)
    for stage in ("coo", "morse"):
        h = R.ref_hm_new(m["nv"])
        R.ref_hm_add_p1_3d(h, m["nv"], m["nt"], m["xyz"], m["tets"], ffo.COEF_SIN3, 1.0, None)
        onbc = ffo.mark_bc(m["nv"], m["btri"], m["blab"], 3, (1, 3, 5))
        R.ref_hm_setbc(h, onbc, 1e30)
        ref_file = tmp_path / f"ref_{stage}.txt"
        assert R.ref_hm_write(h, str(ref_file).encode(), int(stage == "morse")) == 0
        nnz = R.ref_hm_nnz(h)
        rr, cr, vr = np.zeros(m["nv"] + 1, np.int64), np.zeros(nnz, np.int32), np.zeros(nnz)
        R.ref_hm_csr(h, rr, cr, vr)
        R.ref_hm_free(h)
        # the reference reads its own file ...
        nmh = np.zeros(3, np.int64)
        rr2, cr2, vr2 = np.zeros(m["nv"] + 1, np.int64), np.zeros(nnz, np.int32), np.zeros(nnz)
        got = R.ref_hm_read(str(ref_file).encode(), nnz, nmh, rr2.ctypes.data, cr2.ctypes.data,
                            vr2.ctypes.data)
        assert got == nnz
        # ... and so does libffhip: the same CSR, bit for bit
        rp, ci, va, mm, half = ffhip.csr_read_text(str(ref_file))
        assert np.array_equal(rp, rr2) and np.array_equal(ci, cr2)
        assert np.array_equal(va.view(np.int64), vr2.view(np.int64))
        assert np.array_equal(rp, rr) and np.array_equal(ci, cr)
        if stage == "morse":
            mine = tmp_path / "mine.txt"
            ffhip.csr_write_morse(str(mine), rr, cr, vr)
            assert _lines(mine)[1:] == _lines(ref_file)[1:]


@needs_ref
def test_reference_coo_order_is_row_major():
    # [I, J, C] = A copies the arrays of HashMatrix::COO(): rows then columns
    R = ffo.ref()
    m = ffo.cube(3)
    h = R.ref_hm_new(m["nv"])
    R.ref_hm_add_p1_3d(h, m["nv"], m["nt"], m["xyz"], m["tets"], ffo.COEF_SIN3, 1.0, None)
    nnz = R.ref_hm_nnz(h)
    I, J, C = np.zeros(nnz, np.int64), np.zeros(nnz, np.int64), np.zeros(nnz)
    assert R.ref_hm_coo(h, I, J, C) == nnz
    R.ref_hm_free(h)
    rp, ci, va = _cube_matrix(3, tgv_labels=())
    assert np.array_equal(I, np.repeat(np.arange(rp.size - 1), np.diff(rp)))
    assert np.array_equal(J, ci) and np.array_equal(C, va)
