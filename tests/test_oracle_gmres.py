"""Pins the oracle's restarted GMRES (SolverGMRES::dosolver + fgmres,
VirtualSolverCG.hpp:236-255, CG.cpp:347-517) bit-exact against FreeFEM's own
fgmres compiled in oracle/_ref, on symmetric and non-symmetric systems with
tgv rows, with restarts, relative and absolute tolerances and an iteration
cap; and by the solution of numpy.linalg.solve on a small dense system."""
import numpy as np
import pytest

import ffo

needs_ref = pytest.mark.skipif(not ffo.ref_available(), reason="oracle/_ref not built")


def nonsymmetric(n=4, up=0.5, seed=0):
    """cube(n) P1 Poisson with tgv rows whose strictly-upper off-diagonal
    entries are scaled by `up` (and a random 10 % perturbation): the same
    pattern, a non-symmetric M-matrix."""
    m, rp, ci, va, onbc, b = ffo.poisson_p1_3d(n)
    rows = np.repeat(np.arange(rp.size - 1), np.diff(rp))
    rng = np.random.default_rng(seed)
    va = va.copy()
    upper = (ci > rows) & (va < 1e29)
    va[upper] *= up * (1 + 0.1 * rng.standard_normal(upper.sum()))
    return rp, ci, va, b, onbc


def ref_gmres(rp, ci, va, b, eps, itmax, restart, x0=None):
    n = rp.size - 1
    h = ffo.ref().ref_hm_from_csr(n, rp, ci, va)
    x = np.zeros(n) if x0 is None else np.array(x0, np.float64)
    st = np.zeros(4)
    rc = ffo.ref().ref_hm_gmres(h, np.ascontiguousarray(b, np.float64), x, eps, itmax, restart, 1e30, st)
    ffo.ref().ref_hm_free(h)
    return x, int(st[0]), rc


@needs_ref
@pytest.mark.parametrize("case", ["sym", "nonsym", "nonsym_restart5", "abs_eps", "capped"])
def test_fgmres_bitexact_vs_reference(case):
    if case == "sym":
        _, rp, ci, va, onbc, b = ffo.poisson_p1_3d(4)
        eps, itmax, m = 1e-10, 0, 50
    else:
        rp, ci, va, b, onbc = nonsymmetric(4)
        eps, itmax, m = 1e-10, 0, 50
        if case == "nonsym_restart5":
            m = 5
        elif case == "abs_eps":
            eps = -1e-9
        elif case == "capped":
            itmax, m = 7, 5
    x0 = np.random.default_rng(3).standard_normal(rp.size - 1) * 0.1
    xo, so = ffo.fgmres(rp, ci, va, b, x0=x0, eps=eps, itmax=itmax, restart=m)
    xr, itr, rcr = ref_gmres(rp, ci, va, b, eps, itmax, m, x0=x0)
    assert np.array_equal(xo, xr)
    assert so["iters"] == itr and so["converged"] == rcr
    assert so["converged"] == (0 if case == "capped" else 1)
    if case == "nonsym_restart5":
        assert so["restarts"] > 0


def test_fgmres_small_dense_vs_numpy():
    rng = np.random.default_rng(1)
    n = 12
    Ad = rng.standard_normal((n, n)) + n * np.eye(n)  # non-symmetric, well conditioned
    rp = np.arange(0, n * n + 1, n, dtype=np.int64)
    ci = np.tile(np.arange(n, dtype=np.int32), n)
    b = rng.standard_normal(n)
    x, st = ffo.fgmres(rp, ci, Ad.ravel().copy(), b, eps=1e-14, restart=n)
    assert st["converged"] == 1 and st["iters"] <= n
    xe = np.linalg.solve(Ad, b)
    assert np.abs(x - xe).max() <= 1e-11 * np.abs(xe).max()
    # with restarts it still converges, in more iterations
    x2, st2 = ffo.fgmres(rp, ci, Ad.ravel().copy(), b, eps=1e-14, itmax=300, restart=3)
    assert st2["converged"] == 1 and st2["restarts"] > 0
    assert np.abs(x2 - xe).max() <= 1e-10 * np.abs(xe).max()
