"""GPU parity of LinearCG / AffineCG with a matrix operator
(ConjuguedGradient2, MatriceCreuse.hpp:180-240; ffhip_cg_solve_affine) against
the oracle's restatement.

Bars: the reference's own unit/base/LinearCG.edp checks (return 1,
max |A u - b| < 1e-12); iterates after k iterations within 1e-9 of the
oracle's (relative to the size of the update x_k - x_0: only the summation
order of the dot products and mat-vec differs); converged solutions at
eps = 1e-10 within 1e-8 relative and iteration counts within 2; the return
codes (2, 0 after nbiter + 1 iterations, NaN error for a zero operator) and
the threshold written back into veps exactly as the oracle's.
"""
import numpy as np
import pytest

import ffo

pytestmark = pytest.mark.gpu

ffhip = pytest.importorskip("ffhip")


def _diag10():
    n = 10
    rp = np.arange(n + 1, dtype=np.int64)
    ci = np.arange(n, dtype=np.int32)
    va = np.arange(1.0, n + 1)
    return rp, ci, va


def _shifted_neumann(n):
    """cube(n) P1 stiffness without Dirichlet rows plus 0.5 on the diagonal:
    SPD with no tgv rows, the system LinearCG's identity preconditioner
    handles (on tgv rows an identity-preconditioned CG is at the mercy of
    rounding, in the reference as here)."""
    m, rp, ci, va, onbc, b = ffo.poisson_p1_3d(n, labels=())
    rows = np.repeat(np.arange(rp.size - 1), np.diff(rp))
    va = va.copy()
    va[ci == rows] += 0.5
    return rp, ci, va, b


def _start(n, seed, onbc, b, tgv=1e30):
    """random start; on tgv rows b/tgv, the value SetInitWithBC gives them.
    LinearCG leaves such rows to the iteration: an O(1) start there makes
    g = tgv x carry the rounding of tgv * x0 (which FMA contraction changes)
    into every dot product, in the reference as here."""
    x0 = np.random.default_rng(seed).standard_normal(n)
    if onbc is not None:
        x0[onbc != 0] = b[onbc != 0] / tgv
    return x0


def test_linearcg_edp_cases():
    rp, ci, va = _diag10()
    A = ffhip.Matrix.upload(rp, ci, va)
    cg = ffhip.CG(A)
    b = np.ones(10)
    u = np.zeros(10)
    ret, st = cg.linear_cg(u, b)  # LinearCG(DJ, u)
    assert ret == 1 and np.abs(va * u - b).max() < 1e-12
    u = np.zeros(10)
    veps = np.array([-1e-6])  # nbiter=20, eps=Eps, veps=Veps, precon=matId
    ret, st = cg.linear_cg(u, b, eps=1e-6, nbiter=20, precon="identity", veps=veps)
    assert ret == 1 and np.abs(va * u - b).max() < 1e-12
    assert veps[0] == -1e-6
    uo, so = ffo.cg2(rp, ci, va, b, eps=-1e-6, itmax=20)
    assert st["iters"] == so["iters"]


@pytest.mark.parametrize("layout", ["sell", "sell32", "csr"])
@pytest.mark.parametrize("precon", [None, "jacobi"])
def test_iterates_vs_oracle(layout, precon):
    if precon:
        m, rp, ci, va, onbc, b = ffo.poisson_p1_3d(6)
    else:
        rp, ci, va, b = _shifted_neumann(6)
    A = ffhip.Matrix.upload(rp, ci, va)
    cg = ffhip.CG(A, layout=layout)
    x0 = _start(rp.size - 1, 5, None if not precon else onbc, b)
    for k in (1, 4, 9):
        x = x0.copy()
        ret, st = cg.linear_cg(x, b, eps=1e-300, nbiter=k - 1, precon=precon)
        xo, so = ffo.cg2(rp, ci, va, b, x0=x0, eps=1e-300, itmax=k - 1,
                         precon=1 if precon else 0)
        assert ret == 0 and so["nret"] == 0 and st["iters"] == so["iters"] == k
        scale = np.abs(xo - x0).max()
        assert np.abs(x - xo).max() <= 1e-9 * scale


@pytest.mark.parametrize("mode", [0, 1])
def test_converged_vs_oracle_on_gpu_assembled_system(mode):
    n = 8
    mesh = ffhip.cube(n)
    A = ffhip.Matrix.p1(mesh)
    A.assemble_p1(ffhip.COEF_SIN3, 1.0)
    A.set_bc((1, 2, 3, 4, 5, 6))
    rp, ci, va = A.download()
    b = ffhip.rhs_p1(mesh, 1.0, (1, 2, 3, 4, 5, 6))
    cg = ffhip.CG(A, mode)  # a MIXED handle still runs LinearCG in FP64
    x = np.zeros(A.n)
    veps = np.array([1e-10])
    ret, st = cg.linear_cg(x, b, nbiter=2000, precon="jacobi", veps=veps)
    xo, so = ffo.cg2(rp, ci, va, b, eps=1e-10, itmax=2000, precon=1)
    assert ret == so["nret"] == 1
    assert abs(st["iters"] - so["iters"]) <= 2
    assert np.linalg.norm(x - xo) <= 1e-8 * np.linalg.norm(xo)
    assert veps[0] == pytest.approx(-so["reps2"], rel=1e-12)
    assert st["gcg0"] == pytest.approx(so["g20"], rel=1e-12)


def test_device_pointers_and_affine_zero_rhs():
    import torch
    m, rp, ci, va, onbc, b = ffo.poisson_p1_3d(5)
    A = ffhip.Matrix.upload(rp, ci, va)
    cg = ffhip.CG(A)
    x0 = _start(rp.size - 1, 1, onbc, b)
    xd = torch.tensor(x0, device="cuda")
    bd = torch.tensor(b, device="cuda")
    ret, st = cg.linear_cg(xd, bd, eps=1e-300, nbiter=6, precon="jacobi")
    xo, so = ffo.cg2(rp, ci, va, b, x0=x0, eps=1e-300, itmax=6, precon=1)
    assert np.abs(xd.cpu().numpy() - xo).max() <= 1e-9 * np.abs(xo - x0).max()
    # AffineCG(A*x, x): zero right-hand side (b NULL), no preconditioner
    rp, ci, va, b = _shifted_neumann(5)
    cg = ffhip.CG(ffhip.Matrix.upload(rp, ci, va))
    x = x0.copy()
    ret, st = cg.linear_cg(x, None, eps=1e-300, nbiter=4)
    xo, so = ffo.cg2(rp, ci, va, None, x0=x0, eps=1e-300, itmax=4)
    assert np.abs(x - xo).max() <= 1e-9 * np.abs(xo - x0).max()


def test_return_codes():
    rp, ci, va = _diag10()
    A = ffhip.Matrix.upload(rp, ci, va)
    cg = ffhip.CG(A)
    # already the solution: 2, nothing done, eps written back unchanged
    x = np.ones(10)
    veps = np.array([1e-6])
    ret, st = cg.linear_cg(x, va.copy(), veps=veps)
    assert ret == 2 and st["iters"] == 0 and veps[0] == -1e-6 and np.array_equal(x, np.ones(10))
    # not converged within nbiter: LinearCG returns 0, no error
    m, rp2, ci2, va2, onbc, b2 = ffo.poisson_p1_3d(5)
    cg2 = ffhip.CG(ffhip.Matrix.upload(rp2, ci2, va2))
    ret, st = cg2.linear_cg(np.zeros(rp2.size - 1), b2, eps=1e-12, nbiter=3, precon="jacobi")
    assert ret == 0 and st["iters"] == 4
    # zero operator: the reference's ExecError "Matrix is not defined (/0)"
    Z = ffhip.Matrix.upload(rp, ci, np.zeros(10))
    with pytest.raises(ffhip.FFHipError) as e:
        ffhip.CG(Z).linear_cg(np.zeros(10), np.ones(10))
    assert e.value.code == -4


def test_vector_p2_block_layout():
    m = ffo.cube(3)
    o = ffo.elasticity_p2(3, mesh=m)
    sp = ffhip.P2VSpace(ffhip.mesh3_upload(m["xyz"], m["tets"], m["btri"], m["blab"]))
    A = ffhip.Matrix.p2v(sp)
    A.assemble_p2v_elasticity(1.0, 0.5)
    A.set_bc_p2v([4])
    rp, ci, va = A.download()
    b = ffhip.rhs_p2v(sp, (0.0, 0.0, -1.0), [4])
    cg = ffhip.CG(A, layout="bsr3")
    assert cg.layout()[0] == "bsr3"
    x = np.zeros(A.n)
    ret, st = cg.linear_cg(x, b, eps=1e-300, nbiter=7, precon="jacobi")
    xo, so = ffo.cg2(rp, ci, va, b, eps=1e-300, itmax=7, precon=1)
    assert np.abs(x - xo).max() <= 1e-9 * np.abs(xo).max()
    x = np.zeros(A.n)
    ret, st = cg.linear_cg(x, b, eps=1e-10, nbiter=5000, precon="jacobi")
    xo, so = ffo.cg2(rp, ci, va, b, eps=1e-10, itmax=5000, precon=1)
    assert ret == so["nret"] == 1 and abs(st["iters"] - so["iters"]) <= 2
    assert np.linalg.norm(x - xo) <= 1e-8 * np.linalg.norm(xo)
