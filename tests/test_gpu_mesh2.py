"""2D meshes handed over from the host (ffhip_mesh2_upload, the path a FreeFEM
2D Mesh takes through the plugin): the uploaded square(12,9) gives the same
mesh arrays, the oracle's P1 pattern bit for bit and values within 1e-12 of the
oracle's (stiffness with the analytic coefficient, a reaction term and an
int1d Robin term), the same matrix as the device-generated square, and bad
vertex ids are refused before anything reaches the device."""
import numpy as np
import pytest

import ffo
from test_gpu_p1 import _vals_close

pytestmark = pytest.mark.gpu

ffhip = pytest.importorskip("ffhip")


def test_uploaded_square_matches_oracle_and_generated():
    mo = ffo.square(12, 9)
    m = ffhip.mesh2_upload(mo["xy"], mo["tris"], mo["bedges"], mo["blab"])
    d = m.download()
    assert np.array_equal(d["xyz"], mo["xy"]) and np.array_equal(d["elems"], mo["tris"])
    assert np.array_equal(d["belems"], mo["bedges"]) and np.array_equal(d["blab"], mo["blab"])
    rp, ci = ffo.pattern(mo["nv"], mo["tris"], 3)
    vo = ffo.assemble_p1_2d(mo, rp, ci, ffo.COEF_SIN3, mkind=ffo.COEF_CONST, m0=0.5)
    ffo.p1_boundary(mo, (2, 3), ffo.COEF_CONST, 4.0, rowptr=rp, colidx=ci, vals=vo)
    ffo.set_bc(rp, ci, vo, ffo.mark_bc(mo["nv"], mo["bedges"], mo["blab"], 2, (1, 4)))
    A = ffhip.Matrix.p1(m)
    A.assemble_p1(ffhip.COEF_SIN3, 1.0, None, ffhip.COEF_CONST, 0.5, None)
    A.assemble_p1_boundary((2, 3), ffhip.COEF_CONST, 4.0)
    A.set_bc((1, 4))
    rpa, cia, va = A.download()
    assert np.array_equal(rpa, rp) and np.array_equal(cia, ci)
    assert _vals_close(va, vo)
    g = ffhip.square(12, 9)
    B = ffhip.Matrix.p1(g)
    B.assemble_p1(ffhip.COEF_SIN3, 1.0, None, ffhip.COEF_CONST, 0.5, None)
    B.assemble_p1_boundary((2, 3), ffhip.COEF_CONST, 4.0)
    B.set_bc((1, 4))
    assert np.array_equal(B.download()[2], va)


def test_upload_rejects_bad_ids():
    mo = ffo.square(3, 3)
    tris = mo["tris"].copy()
    tris[4] = mo["nv"]
    with pytest.raises(ffhip.FFHipError):
        ffhip.mesh2_upload(mo["xy"], tris, mo["bedges"], mo["blab"])
    bed = mo["bedges"].copy()
    bed[1] = -1
    with pytest.raises(ffhip.FFHipError):
        ffhip.mesh2_upload(mo["xy"], mo["tris"], bed, mo["blab"])
