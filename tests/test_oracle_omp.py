"""The all-core (OpenMP) variant of the oracle's 3D P1 path, timed as the
all-core CPU baseline, against the serial restatement: the pattern and the
element-order values (HashMatrix::operator+=, HashMatrix.cpp:1294-1332) bit
for bit, the Dirichlet rows (SetBC, HashMatrix.cpp:1194-1238) bit for bit, and
the Jacobi-CG (CG.cpp:195-265) to the same criterion within one iteration of
the serial one and 1e-9 of its solution (the dot products are summed in
another order).  A planted swap of two element contributions to one entry
must change that entry's bits, showing the check can see summation order."""
import numpy as np
import pytest

import ffo

LABELS = (1, 2, 3, 4, 5, 6)


@pytest.mark.parametrize("n,kind", [(6, ffo.COEF_SIN3), (11, ffo.COEF_P0), (9, ffo.COEF_CONST)])
def test_omp_matches_serial_oracle(n, kind):
    m = ffo.cube(n)
    ke = np.exp(np.random.default_rng(n).standard_normal(m["nt"])) if kind == ffo.COEF_P0 else None
    S = ffo.OmpP1(m, kind, 1.0, ke)
    rp, ci, va = S.arrays()
    rpo, cio = ffo.pattern(m["nv"], m["tets"], 4)
    vo = ffo.assemble_p1_3d(m, rpo, cio, kind, 1.0, ke)
    assert np.array_equal(rp, rpo) and np.array_equal(ci, cio)
    assert np.array_equal(va.view(np.uint64), vo.view(np.uint64))
    onbc = ffo.mark_bc(m["nv"], m["btri"], m["blab"], 3, LABELS)
    S.set_bc(onbc)
    ffo.set_bc(rpo, cio, vo, onbc)
    rp, ci, va = S.arrays()
    assert np.array_equal(va.view(np.uint64), vo.view(np.uint64))
    b = ffo.rhs_p1_3d(m, 1.0)
    ffo.rhs_bc(b, onbc)
    x, st = S.pcg(b, eps=1e-10)
    xo, so = ffo.pcg(rpo, cio, vo, b, eps=1e-10)
    assert st["nret"] == 1 and abs(st["iters"] - so["iters"]) <= 1
    assert np.linalg.norm(x - xo) <= 1e-9 * np.linalg.norm(xo)
    S.free()


def test_summation_order_is_visible():
    """the diagonal entry of an interior vertex, summed over its elements in
    element order, is the oracle's value; adding two neighbouring
    contributions the other way round changes its bits for some pair"""
    m = ffo.cube(5)
    rp, ci = ffo.pattern(m["nv"], m["tets"], 4)
    va = ffo.assemble_p1_3d(m, rp, ci, ffo.COEF_SIN3)
    tets = m["tets"].reshape(-1, 4)
    v = int(np.argmax(np.bincount(m["tets"])))
    parts = []
    for e in np.flatnonzero((tets == v).any(axis=1)):
        Ke = np.zeros(16)
        ffo.lib().ffo_p1_element_3d(m["xyz"], np.ascontiguousarray(tets[e]), ffo.COEF_SIN3, 1.0, 0.0, Ke)
        lv = int(np.flatnonzero(tets[e] == v)[0])
        parts.append(Ke[5 * lv])
    slot = rp[v] + int(np.flatnonzero(ci[rp[v]:rp[v + 1]] == v)[0])
    fwd = 0.
    for p in parts:
        fwd += p
    assert fwd == va[slot]
    diff = False
    for i in range(len(parts) - 1):
        acc = 0.
        for p in parts[:i] + [parts[i + 1], parts[i]] + parts[i + 2:]:
            acc += p
        diff |= acc != fwd
    assert diff
