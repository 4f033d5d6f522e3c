"""Pins the oracle's half-storage product (sym=1 matrices, HashMatrix::half:
the half branch of addMatMul, HashMatrix.cpp:1128-1136) against FreeFEM's own
compiled HashMatrix holding the same lower triangle with half = 1 (bit for
bit), and against the full symmetric matrix's product (the half product adds
y_i's terms in the sorted-column order of the full row, so bit for bit too)."""
import numpy as np
import pytest

import ffo

needs_ref = pytest.mark.skipif(not ffo.ref_available(), reason="oracle/_ref not built")


def _lower(rp, ci, va):
    n = rp.size - 1
    rows = np.repeat(np.arange(n), np.diff(rp))
    keep = ci <= rows
    rl = np.zeros(n + 1, np.int64)
    np.add.at(rl, rows[keep] + 1, 1)
    return np.cumsum(rl), np.ascontiguousarray(ci[keep]), np.ascontiguousarray(va[keep])


@pytest.mark.parametrize("n", [3, 7])
def test_half_product_equals_full_product(n):
    m, rp, ci, va, onbc, b = ffo.poisson_p1_3d(n, labels=(1, 4))
    x = np.random.default_rng(n).standard_normal(rp.size - 1)
    rl, cl, vl = _lower(rp, ci, va)
    yh = ffo.spmv_half(rl, cl, vl, x)
    yf = ffo.spmv(rp, ci, va, x)
    assert np.array_equal(yh, yf)
    assert cl.size < ci.size


@needs_ref
@pytest.mark.parametrize("n", [2, 5])
def test_half_product_vs_reference_hashmatrix(n):
    m, rp, ci, va, onbc, b = ffo.poisson_p1_3d(n)
    rl, cl, vl = _lower(rp, ci, va)
    x = np.random.default_rng(3).standard_normal(rp.size - 1)
    R = ffo.ref()
    h = R.ref_hm_from_csr_half(rp.size - 1, rl, cl, vl)
    y = np.zeros(rp.size - 1)
    R.ref_hm_matmul(h, x, y, 0)
    R.ref_hm_free(h)
    assert np.array_equal(y, ffo.spmv_half(rl, cl, vl, x))


def test_half_product_brute_force():
    rng = np.random.default_rng(5)
    n = 9
    S = rng.standard_normal((n, n))
    S = S + S.T
    S[np.abs(S) < 0.8] = 0.0
    np.fill_diagonal(S, 4.0)
    L = np.tril(S)
    rl = np.zeros(n + 1, np.int64)
    ci, va = [], []
    for i in range(n):
        nz = np.nonzero(L[i])[0]
        ci.extend(nz)
        va.extend(L[i, nz])
        rl[i + 1] = len(ci)
    x = rng.standard_normal(n)
    y = ffo.spmv_half(rl, np.array(ci, np.int32), np.array(va), x)
    assert np.abs(y - S @ x).max() <= 1e-13 * np.abs(S @ x).max()


def test_transposed_product_equals_product_of_explicit_transpose():
    # ffo_spmv_t (addMatMul with the transpose flag, HashMatrix.cpp:1094): y_j adds
    # a_ij x_i over the rows in order, as the transpose's row j does
    rng = np.random.default_rng(9)
    n = 40
    D = rng.standard_normal((n, n)) * (rng.uniform(size=(n, n)) < 0.2)
    rp = np.zeros(n + 1, np.int64)
    ci, va = [], []
    for i in range(n):
        nz = np.nonzero(D[i])[0]
        ci.extend(nz)
        va.extend(D[i, nz])
        rp[i + 1] = len(ci)
    ci, va = np.array(ci, np.int32), np.array(va)
    DT = D.T.copy()
    rt = np.zeros(n + 1, np.int64)
    ct, vt = [], []
    for i in range(n):
        nz = np.nonzero(DT[i])[0]
        ct.extend(nz)
        vt.extend(DT[i, nz])
        rt[i + 1] = len(ct)
    x = rng.standard_normal(n)
    yt = ffo.spmv(rp, ci, va, x, trans=True)
    assert np.array_equal(yt, ffo.spmv(rt, np.array(ct, np.int32), np.array(vt), x))
    assert np.abs(yt - D.T @ x).max() <= 1e-13 * max(1.0, np.abs(D.T @ x).max())
