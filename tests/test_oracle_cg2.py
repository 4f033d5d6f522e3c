"""Pins the oracle's ConjuguedGradient2 restatement (LinearCG / AffineCG with a
matrix operator, MatriceCreuse.hpp:180-240 via lgfem.cpp:705-758).

The reference's ConjuguedGradient2 is a template in MatriceCreuse.hpp, whose
includes pull the script interpreter in, so it is not compiled into
oracle/_ref; the restatement is pinned instead by

* the reference's own test unit/base/LinearCG.edp (10 x 10 diagonal system,
  LinearCG must return 1 and leave max |A u - b| < HWERROR = 1e-12), in both
  of its forms (defaults; nbiter=20 with veps=-1e-6 and an identity precon);
* exact-arithmetic equivalence with ConjugueGradient (CG.cpp:195-265, whose
  restatement is pinned bit-exact against the compiled reference): started
  from the same x0 with the same Jacobi preconditioner, both produce the same
  Krylov iterates, so after k iterations the two x agree to rounding — a
  dropped term, a wrong sign or a wrong step length in either fails this;
* brute force on small SPD systems (solution of numpy.linalg.solve) and the
  finite-termination property (n distinct eigenvalues: at most n iterations);
* the return codes: 2 when the start is already the solution (eps untouched),
  -1 (the ExecError "Matrix is not defined") for a zero operator, and the
  threshold written back into eps (relative eps^2 g2_0, absolute -eps).
"""
import numpy as np
import pytest

import ffo


def _csr(Ad):
    Ad = np.asarray(Ad, np.float64)
    n = Ad.shape[0]
    rp = np.zeros(n + 1, np.int64)
    ci, va = [], []
    for i in range(n):
        nz = np.nonzero(Ad[i])[0]
        ci.extend(nz)
        va.extend(Ad[i, nz])
        rp[i + 1] = len(ci)
    return rp, np.array(ci, np.int32), np.array(va, np.float64)


def test_linearcg_edp_defaults():
    # Conv = LinearCG(DJ, u) with DJ(u) = (i+1) u - b, b = 1, u = 0
    rp, ci, va = _csr(np.diag(np.arange(1.0, 11.0)))
    b = np.ones(10)
    u, st = ffo.cg2(rp, ci, va, b, eps=1e-6, itmax=100)
    assert st["nret"] == 1
    assert np.abs(np.arange(1.0, 11.0) * u - b).max() < 1e-12
    assert st["reps2"] == pytest.approx(1e-12 * st["g20"], rel=1e-15)


def test_linearcg_edp_named_parameters():
    # LinearCG(DJ, u, nbiter=20, eps=Eps, veps=Veps, precon=matId, ...):
    # veps = -1e-6 overrides eps (absolute threshold), veps = -reps2 after
    rp, ci, va = _csr(np.diag(np.arange(1.0, 11.0)))
    b = np.ones(10)
    u, st = ffo.cg2(rp, ci, va, b, eps=-1e-6, itmax=20, precon=0)
    assert st["nret"] == 1
    assert np.abs(np.arange(1.0, 11.0) * u - b).max() < 1e-12
    assert -st["reps2"] == -1e-6


@pytest.mark.parametrize("k", [1, 2, 5, 9])
def test_iterates_equal_conjugue_gradient(k):
    # cube(4) Poisson with tgv rows; both solvers start from the same x0
    _, rp, ci, va, onbc, b = ffo.poisson_p1_3d(4)
    n = rp.size - 1
    x0 = np.random.default_rng(k).standard_normal(n)
    x0[onbc != 0] = b[onbc != 0] / 1e30  # what SetInitWithBC sets before ConjugueGradient
    x_pcg, st_pcg = ffo.pcg(rp, ci, va, b, x0=x0, eps=1e-300, itmax=k)
    x_cg2, st_cg2 = ffo.cg2(rp, ci, va, b, x0=x0, eps=1e-300, itmax=k - 1, precon=1)
    assert st_pcg["iters"] == k and st_cg2["iters"] == k
    scale = np.abs(x_pcg - x0).max()
    assert np.abs(x_cg2 - x_pcg).max() <= 1e-9 * scale


def test_converged_solution_matches_conjugue_gradient():
    _, rp, ci, va, onbc, b = ffo.poisson_p1_3d(5)
    x_pcg, st_pcg = ffo.pcg(rp, ci, va, b, eps=1e-10)
    n = rp.size - 1
    x0 = np.zeros(n)
    x0[onbc != 0] = b[onbc != 0] / 1e30  # the SetInitWithBC start of ConjugueGradient's caller
    x_cg2, st_cg2 = ffo.cg2(rp, ci, va, b, x0=x0, eps=1e-10, itmax=1000, precon=1)
    assert st_cg2["nret"] == 1
    assert abs(st_cg2["iters"] - st_pcg["iters"]) <= 1
    assert np.abs(x_cg2 - x_pcg).max() <= 1e-8 * np.abs(x_pcg).max()


@pytest.mark.parametrize("seed", range(4))
def test_small_spd_brute_force(seed):
    rng = np.random.default_rng(seed)
    n = 8
    Q = rng.standard_normal((n, n))
    Ad = Q @ Q.T + n * np.eye(n)
    Ad[np.abs(Ad) < 0.3] = 0.0  # some sparsity, still SPD (diagonally dominant)
    Ad = 0.5 * (Ad + Ad.T)
    if np.linalg.eigvalsh(Ad).min() <= 0:
        pytest.skip("not SPD after sparsification")
    rp, ci, va = _csr(Ad)
    b = rng.standard_normal(n)
    for precon in (0, 1):
        x, st = ffo.cg2(rp, ci, va, b, eps=1e-14, itmax=50, precon=precon)
        assert st["nret"] == 1 and st["iters"] <= n + 2
        xe = np.linalg.solve(Ad, b)
        assert np.abs(x - xe).max() <= 1e-10 * np.abs(xe).max()


def test_affine_zero_rhs_and_return_codes():
    rp, ci, va = _csr(np.diag(np.arange(1.0, 11.0)))
    # AffineCG(A, x): zero right-hand side, converges to 0
    x0 = np.linspace(-1, 1, 10)
    x, st = ffo.cg2(rp, ci, va, None, x0=x0, eps=1e-12, itmax=100)
    assert st["nret"] == 1 and np.abs(x).max() < 1e-10
    # start at the solution: 2, "Nothing to do", eps untouched
    b = np.arange(1.0, 11.0)
    x, st = ffo.cg2(rp, ci, va, b, x0=np.ones(10), eps=1e-6)
    assert st["nret"] == 2 and st["iters"] == 0 and st["reps2"] == 1e-6
    # zero operator: (h, Ah) = 0 -> ExecError
    x, st = ffo.cg2(rp, ci, np.zeros_like(va), b, eps=1e-6)
    assert st["nret"] == -1
    # not converged within nbiter: 0 after nbiter + 1 iterations
    _, rp, ci, va, onbc, b = ffo.poisson_p1_3d(4)
    x, st = ffo.cg2(rp, ci, va, b, eps=1e-12, itmax=3, precon=1)
    assert st["nret"] == 0 and st["iters"] == 4
