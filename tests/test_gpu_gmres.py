"""GPU parity of solver=GMRES (ffhip_gmres_solve: SolverGMRES::dosolver over
fgmres, VirtualSolverCG.hpp:236-255, CG.cpp:347-517) against the oracle's
restatement, itself bit-exact against FreeFEM's compiled fgmres.

Bars: iteration count within 1 of the oracle's (only the summation order of
the dot products and mat-vec differs) and the solution within 1e-8 relative
at eps = 1e-10; restarts, an absolute tolerance, the transposed system and the
iteration cap behave as the oracle's; non-convergence raises NOCONV.
"""
import numpy as np
import pytest

import ffo

pytestmark = pytest.mark.gpu

ffhip = pytest.importorskip("ffhip")


def nonsymmetric(n, up=0.5, seed=0):
    """cube(n) P1 Poisson with tgv rows whose strictly-upper off-diagonal
    entries are scaled by `up` (10 % random perturbation): non-symmetric."""
    m, rp, ci, va, onbc, b = ffo.poisson_p1_3d(n)
    rows = np.repeat(np.arange(rp.size - 1), np.diff(rp))
    rng = np.random.default_rng(seed)
    va = va.copy()
    upper = (ci > rows) & (va < 1e29)
    va[upper] *= up * (1 + 0.1 * rng.standard_normal(upper.sum()))
    return rp, ci, va, b


@pytest.mark.parametrize("layout", ["sell", "csr"])
@pytest.mark.parametrize("restart", [1000, 8])
def test_gmres_vs_oracle(layout, restart):
    rp, ci, va, b = nonsymmetric(6)
    A = ffhip.Matrix.upload(rp, ci, va)
    cg = ffhip.CG(A, layout=layout)
    x, st = cg.gmres(b, eps=1e-10, restart=restart)
    xo, so = ffo.fgmres(rp, ci, va, b, eps=1e-10, restart=restart)
    assert st["converged"] == so["converged"] == 1
    assert abs(st["iters"] - so["iters"]) <= 1
    assert st["refinements"] == so["restarts"] or abs(st["iters"] - so["iters"]) == 1
    assert np.linalg.norm(x - xo) <= 1e-8 * np.linalg.norm(xo)
    assert st["gcg0"] == pytest.approx(so["normb"], rel=1e-12)
    assert st["true_relres"] < 1e-8
    if restart == 8:
        assert st["refinements"] > 0


def test_gmres_absolute_eps_transposed_and_device_pointers():
    import torch
    rp, ci, va, b = nonsymmetric(5, seed=2)
    A = ffhip.Matrix.upload(rp, ci, va)
    cg = ffhip.CG(A)
    x, st = cg.gmres(b, eps=-1e-9, restart=30)
    xo, so = ffo.fgmres(rp, ci, va, b, eps=-1e-9, restart=30)
    assert st["converged"] == 1 and abs(st["iters"] - so["iters"]) <= 1 and st["gcg0"] == 1.0
    assert np.linalg.norm(x - xo) <= 1e-7 * np.linalg.norm(xo)
    # A^T x = b: the oracle on the explicit transpose
    n = rp.size - 1
    rows = np.repeat(np.arange(n), np.diff(rp))
    order = np.lexsort((rows, ci))
    rpt = np.zeros(n + 1, np.int64)
    np.add.at(rpt, ci + 1, 1)
    rpt = np.cumsum(rpt)
    cit, vat = rows[order].astype(np.int32), va[order]
    bd = torch.tensor(b, device="cuda")
    xd, st = cg.gmres(bd, eps=1e-10, trans=True)
    xo, so = ffo.fgmres(rpt, cit, vat, b, eps=1e-10)
    assert st["converged"] == 1 and abs(st["iters"] - so["iters"]) <= 1
    assert np.linalg.norm(xd.cpu().numpy() - xo) <= 1e-8 * np.linalg.norm(xo)


def test_gmres_iteration_cap_raises():
    rp, ci, va, b = nonsymmetric(5)
    cg = ffhip.CG(ffhip.Matrix.upload(rp, ci, va))
    with pytest.raises(ffhip.FFHipError) as e:
        cg.gmres(b, eps=1e-12, itmax=5, restart=4)
    assert e.value.code == -3
    x, st = cg.gmres(b, eps=1e-12, itmax=5, restart=4, raise_on_error=False)
    xo, so = ffo.fgmres(rp, ci, va, b, eps=1e-12, itmax=5, restart=4)
    assert st["converged"] == so["converged"] == 0 and st["iters"] == so["iters"]
    assert st["refinements"] == so["restarts"]
