"""GPU parity of scalar P2 (Vh(Th,P2)) Poisson through libffhip's C-ABI against
the oracle: pattern and Dirichlet rows bit-exact (cube and shuffled meshes),
values within 1e-12 of the largest entry (constant, analytic, P0 and
quadrature-point kappa), right-hand side within 1e-14, CG solutions within
1e-8 of the oracle's at eps = 1e-10 (FP64 and MIXED, sliced ELL)."""
import numpy as np
import pytest

import ffo
from test_gpu_p2 import _gpu_mesh, _shuffled_cube

pytestmark = pytest.mark.gpu

ffhip = pytest.importorskip("ffhip")
LABELS = (1, 2, 3, 4, 5, 6)


@pytest.mark.parametrize("case", ["cube3", "shuffled3", "cube5_p0"])
def test_p2_scalar_system_vs_oracle(case):
    m = _shuffled_cube(3, 11) if case == "shuffled3" else ffo.cube(5 if case == "cube5_p0" else 3)
    kind = ffo.COEF_P0 if case == "cube5_p0" else ffo.COEF_SIN3
    ke = np.exp(0.5 * np.random.default_rng(2).standard_normal(m["nt"])) if kind == ffo.COEF_P0 else None
    o = ffo.poisson_p2(0, kind=kind, kappa_e=ke, mesh=m)
    sp = ffhip.P2VSpace(_gpu_mesh(m))
    A = ffhip.Matrix.p2(sp)
    assert A.n == o["nnodes"] and A.nnz == o["colidx"].size
    A.assemble_p2(kind, 1.0, ke)
    onbc = A.set_bc_p2(LABELS, want_onbc=True)
    rp, ci, va = A.download()
    assert np.array_equal(rp, o["rowptr"]) and np.array_equal(ci, o["colidx"])
    assert np.array_equal(onbc, o["onbc"])
    free = va < 1e29
    assert np.array_equal(free, o["vals"] < 1e29)
    assert np.abs(va[free] - o["vals"][free]).max() <= 1e-12 * np.abs(o["vals"][free]).max()
    b = ffhip.rhs_p2(sp, 1.0, LABELS)
    assert np.abs(b - o["b"]).max() <= 1e-14 * np.abs(o["b"]).max()
    xo, so = ffo.pcg(o["rowptr"], o["colidx"], o["vals"], o["b"], eps=1e-10)
    for mode in (ffhip.CG_FP64, ffhip.CG_MIXED):
        cg = ffhip.CG(A, mode)
        assert cg.layout()[0] in ("sell16", "sell", "csr")
        x, st = cg.solve(b, eps=1e-10)
        assert st["converged"] == 1
        assert np.linalg.norm(x - xo) <= 1e-8 * np.linalg.norm(xo)


def test_p2_scalar_qp_coefficient_equals_analytic():
    m = ffo.cube(4)
    sp = ffhip.P2VSpace(_gpu_mesh(m))
    A = ffhip.Matrix.p2(sp)
    A.assemble_p2(ffhip.COEF_SIN3)
    _, _, v_an = A.download()
    pts, w = np.zeros(42), np.zeros(14)
    ffo.lib().ffo_qf_tet5(pts, w)
    X = m["xyz"].reshape(-1, 3)[m["tets"].reshape(-1, 4)]
    P = pts.reshape(14, 3)
    lam = np.c_[1 - (P[:, 0] + P[:, 1] + P[:, 2]), P]
    Q = np.einsum("qi,eik->eqk", lam, X)
    kq = 1 + 0.5 * np.sin(2 * np.pi * Q[..., 0]) * np.sin(2 * np.pi * Q[..., 1]) * np.sin(2 * np.pi * Q[..., 2])
    B = ffhip.Matrix.p2(sp)
    B.assemble_p2(ffhip.COEF_QP, 0.0, kq.ravel())
    _, _, v_qp = B.download()
    assert np.abs(v_qp - v_an).max() <= 1e-13 * np.abs(v_an).max()
