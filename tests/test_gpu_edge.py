"""Edge cases of the P1 path on the GPU against the oracle: a vertex that no
element uses (its HashMatrix row stays empty: no diagonal, Jacobi 1, a zero
right-hand side entry), in 3D and 2D, through pattern, values, tgv rows, right-
hand side and both solvers; and a one-element mesh (a single dense 4 x 4
block). Bars as in the P1 tests: patterns bit-exact, values within 1e-12 of the
largest entry, solutions within 1e-8."""
import numpy as np
import pytest

import ffo
from test_gpu_p1 import _vals_close

pytestmark = pytest.mark.gpu

ffhip = pytest.importorskip("ffhip")


def _with_isolated_vertex_3d(n):
    m = ffo.cube(n)
    xyz = np.concatenate([m["xyz"], [2.0, 2.0, 2.0]])
    return dict(m, nv=m["nv"] + 1, xyz=np.ascontiguousarray(xyz))


def test_isolated_vertex_3d_pattern_values_and_solves():
    mo = _with_isolated_vertex_3d(4)
    iso = mo["nv"] - 1
    rp, ci = ffo.pattern(mo["nv"], mo["tets"], 4)
    assert rp[iso + 1] == rp[iso]  # the empty row
    vo = ffo.assemble_p1_3d(mo, rp, ci, ffo.COEF_SIN3)
    onbc = ffo.mark_bc(mo["nv"], mo["btri"], mo["blab"], 3, (1, 2, 3))
    ffo.set_bc(rp, ci, vo, onbc)
    bo = ffo.rhs_p1_3d(mo, 1.0)
    ffo.rhs_bc(bo, onbc, 1e30, 0.0)
    assert bo[iso] == 0.0
    m = ffhip.mesh3_upload(mo["xyz"], mo["tets"], mo["btri"], mo["blab"])
    A = ffhip.Matrix.p1(m)
    A.assemble_p1(ffhip.COEF_SIN3)
    A.set_bc((1, 2, 3))
    rpa, cia, va = A.download()
    assert np.array_equal(rpa, rp) and np.array_equal(cia, ci)
    assert _vals_close(va, vo)
    b = ffhip.rhs_p1(m, 1.0, (1, 2, 3))
    assert b[iso] == 0.0
    assert np.abs(b - bo)[bo < 1e29].max() <= 1e-12 * np.abs(bo[bo < 1e29]).max()
    xo, so = ffo.pcg(rp, ci, vo, bo, eps=1e-10)
    for mode in (ffhip.CG_FP64, ffhip.CG_MIXED):
        x, st = ffhip.CG(A, mode).solve(b, eps=1e-10)
        assert st["converged"] == 1 and x[iso] == 0.0
        assert np.linalg.norm(x - xo) <= 1e-8 * np.linalg.norm(xo)
    x, st = ffhip.CG(A, ffhip.CG_FP64).solve(b, eps=1e-10)
    assert abs(st["iters"] - so["iters"]) <= 1


def test_isolated_vertex_2d():
    s = ffo.square(5, 4)
    xy = np.ascontiguousarray(np.concatenate([s["xy"], [3.0, 3.0]]))
    mo = dict(s, nv=s["nv"] + 1, xy=xy)
    rp, ci = ffo.pattern(mo["nv"], mo["tris"], 3)
    vo = ffo.assemble_p1_2d(mo, rp, ci, ffo.COEF_SIN3)
    m = ffhip.mesh2_upload(mo["xy"], mo["tris"], mo["bedges"], mo["blab"])
    A = ffhip.Matrix.p1(m)
    A.assemble_p1(ffhip.COEF_SIN3)
    rpa, cia, va = A.download()
    assert np.array_equal(rpa, rp) and np.array_equal(cia, ci)
    assert np.abs(va - vo).max() <= 1e-12 * np.abs(vo).max()


def test_single_element():
    xyz = np.array([0., 0., 0., 1., 0., 0., 0., 1., 0., 0., 0., 1.])
    tets = np.array([0, 1, 2, 3], np.int32)
    btri = np.array([3, 2, 1, 0, 2, 3, 3, 1, 0, 0, 1, 2], np.int32)
    blab = np.array([1, 2, 3, 4], np.int32)
    mo = dict(nv=4, nt=1, nbe=4, xyz=xyz, tets=tets, btri=btri, blab=blab)
    rp, ci = ffo.pattern(4, tets, 4)
    assert rp.tolist() == [0, 4, 8, 12, 16]
    vo = ffo.assemble_p1_3d(mo, rp, ci, ffo.COEF_CONST, 2.0)
    m = ffhip.mesh3_upload(xyz, tets, btri, blab)
    A = ffhip.Matrix.p1(m)
    A.assemble_p1(ffhip.COEF_CONST, 2.0)
    rpa, cia, va = A.download()
    assert np.array_equal(rpa, rp) and np.array_equal(cia, ci)
    assert np.abs(va - vo).max() <= 1e-14 * np.abs(vo).max()
    assert np.abs(va.reshape(4, 4).sum(1)).max() <= 1e-14  # rows of a Laplacian sum to 0
