"""GPU side of the matrix exchange formats: [I, J, C] = A from a device
matrix (ffhip_mat_coo; lgmat.cpp:2670-2710 over HashMatrix::COO,
HashMatrix.cpp:837-843), the Morse text of a device matrix
(ffhip_mat_write_morse) and a text file read into a device matrix
(ffhip_mat_read_text).

Bars: bit-exact (pure data movement): the triplets equal the oracle's row-major
expansion of the same CSR and the downloaded values; FreeFEM's extra
(n, n, 0) entry when (n-1, n-1) is absent; the device matrix's text equals the
host writer's on the downloaded CSR byte for byte after the first line; the
reference-written fixtures read into device matrices equal the oracle's CSR,
a half (lower-triangle) file gives the full symmetric matrix, and a matrix
read from text solves like the oracle's.
"""
import os

import numpy as np
import pytest

import ffo

pytestmark = pytest.mark.gpu

ffhip = pytest.importorskip("ffhip")
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_coo_export_host_and_device():
    import torch
    m = ffhip.cube(5)
    A = ffhip.Matrix.p1(m)
    A.assemble_p1(ffhip.COEF_SIN3)
    A.set_bc((1, 2, 3, 4, 5, 6))
    rp, ci, va = A.download()
    mo, ro, co, vo, onbc, b = ffo.poisson_p1_3d(5)
    assert np.array_equal(rp, ro) and np.array_equal(ci, co)
    I, J, C = A.coo()
    assert I.size == ci.size  # (n-1, n-1) is present: no extra entry
    assert np.array_equal(I, np.repeat(np.arange(rp.size - 1), np.diff(rp)))
    assert np.array_equal(J, ci.astype(np.int64))
    assert np.array_equal(C.view(np.int64), va.view(np.int64))
    Id, Jd, Cd = A.coo(device=True)
    assert torch.equal(Id.cpu(), torch.from_numpy(I)) and torch.equal(Jd.cpu(), torch.from_numpy(J))
    assert np.array_equal(Cd.cpu().numpy().view(np.int64), va.view(np.int64))


def test_coo_extra_entry_when_last_diagonal_missing():
    rp = np.array([0, 2, 3, 4], np.int64)
    ci = np.array([0, 2, 1, 0], np.int32)
    va = np.array([4.0, -1.0, 3.0, -1.0])
    A = ffhip.Matrix.upload(rp, ci, va)
    I, J, C = A.coo()
    assert I.tolist() == [0, 0, 1, 2, 3] and J.tolist() == [0, 2, 1, 0, 3]
    assert C.tolist() == [4.0, -1.0, 3.0, -1.0, 0.0]
    import torch
    Id, Jd, Cd = A.coo(device=True)
    assert Id.cpu().tolist() == I.tolist() and Jd.cpu().tolist() == J.tolist()
    assert Cd.cpu().tolist() == C.tolist()


def test_device_morse_text_equals_host_writer(tmp_path):
    m = ffhip.cube(4, 3, 5)
    A = ffhip.Matrix.p1(m)
    A.assemble_p1(ffhip.COEF_SIN3)
    A.set_bc((1, 2))
    rp, ci, va = A.download()
    A.write_morse(str(tmp_path / "dev.txt"))
    ffhip.csr_write_morse(str(tmp_path / "host.txt"), rp, ci, va)
    a = (tmp_path / "dev.txt").read_bytes().split(b"\n")
    b = (tmp_path / "host.txt").read_bytes().split(b"\n")
    assert a[1:] == b[1:] and len(a) == ci.size + 5


@pytest.mark.parametrize("name", ["p1_cube2_morse.txt", "p1_cube2_coo.txt"])
def test_read_reference_fixture_into_device_matrix_and_solve(name):
    A = ffhip.Matrix.read_text(os.path.join(GOLDEN, name))
    rp, ci, va = A.download()
    mo, ro, co, vo, onbc, b = ffo.poisson_p1_3d(2)
    assert np.array_equal(rp, ro) and np.array_equal(ci, co)
    assert np.array_equal(va.view(np.int64), vo.view(np.int64))
    xo, so = ffo.pcg(ro, co, vo, b, eps=1e-10)
    x, st = ffhip.CG(A).solve(b, eps=1e-10)
    assert abs(st["iters"] - so["iters"]) <= 1
    assert np.linalg.norm(x - xo) <= 1e-8 * np.linalg.norm(xo)


def test_half_file_gives_full_symmetric_matrix(tmp_path):
    mo, ro, co, vo, onbc, b = ffo.poisson_p1_3d(3, labels=())
    n = ro.size - 1
    rows = np.repeat(np.arange(n), np.diff(ro))
    low = co <= rows  # lower triangle, FreeFEM's half storage
    rl = np.zeros(n + 1, np.int64)
    np.add.at(rl, rows[low] + 1, 1)
    rl = np.cumsum(rl)
    p = tmp_path / "half.txt"
    ffhip.csr_write_morse(str(p), rl, co[low], vo[low], half=1)
    A = ffhip.Matrix.read_text(str(p))
    rp, ci, va = A.download()
    assert np.array_equal(rp, ro) and np.array_equal(ci, co)
    assert np.array_equal(va, vo)
