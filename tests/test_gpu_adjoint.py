"""GPU checks of the adjoint gradient dJ/dkappa_e (J = 1/2 int (u-ud)^2) behind
the torch.autograd.Function, against the oracle's composition of FreeFEM's
varfs and against central finite differences.

Bars: J within 1e-9 relative of the oracle; the gradient within 1e-7 (norm,
relative) of the oracle's; central differences (h = 1e-6) on 20 seeded random
elements within 1e-3 relative (+1e-4 max|g| absolute); at the cube(48)
configuration a directional difference along a random direction within 1e-5,
and central differences (h = 1e-6) on 20 seeded random elements within 1e-3
relative.

At cube(48) one element moves J by about 1e-16 (|g_e| ~ 1e-10, J ~ 3e-6), so
J(kappa + h e) - J(kappa - h e) is formed without cancellation: since the
misfit is the quadratic form of the mass matrix M of the same Tet_5 rule,
J(u+) - J(u-) = 1/2 (u+ - u-)^T M (u+ + u- - 2 ud) exactly, and the two
misfit vectors M (u -/+ ud) give it as one dot product.  Subtracting two
rounded J values instead leaves the rounding of their 663,552-term sums
(about 3e-19) in a difference of about 2e-16.
"""
import numpy as np
import pytest

import ffo

pytestmark = pytest.mark.gpu

ffhip = pytest.importorskip("ffhip")


def _kappa(nt, seed=20261202):
    # kappa_e = exp(g_e), g_e ~ N(0, 0.5^2) i.i.d.
    return np.exp(0.5 * np.random.default_rng(seed).standard_normal(nt))


def _problem(n, eps):
    import ffhip.adjoint as adj
    return adj, adj.P1Diffusion(n, eps=eps)


def test_adjoint_matches_oracle():
    import torch
    n = 6
    adj, P = _problem(n, 1e-12)
    k = _kappa(P.mesh.nt)
    kt = torch.tensor(k, device="cuda", requires_grad=True)
    J = adj.objective(kt, P)
    J.backward()
    o = ffo.adjoint_p1_3d(n, k, eps=1e-12)
    assert abs(J.item() - o["J"]) <= 1e-9 * abs(o["J"])
    g = kt.grad.cpu().numpy()
    assert np.linalg.norm(g - o["g"]) <= 1e-7 * np.linalg.norm(o["g"])


def test_adjoint_central_differences_on_elements():
    import torch
    n = 5
    adj, P = _problem(n, 1e-14)
    k = _kappa(P.mesh.nt, seed=7)
    kt = torch.tensor(k, device="cuda", requires_grad=True)
    J = adj.objective(kt, P)
    J.backward()
    g = kt.grad.cpu().numpy()
    h = 1e-6
    els = np.random.default_rng(11).choice(P.mesh.nt, 20, replace=False)
    for e in els:
        kp, km = k.copy(), k.copy()
        kp[e] += h
        km[e] -= h
        Jp = adj.objective(torch.tensor(kp, device="cuda"), P).item()
        Jm = adj.objective(torch.tensor(km, device="cuda"), P).item()
        fd = (Jp - Jm) / (2 * h)
        assert abs(fd - g[e]) <= 1e-3 * abs(g[e]) + 1e-4 * np.abs(g).max(), (e, fd, g[e])


def test_adjoint_directional_derivative_cube48():
    """configuration: cube(48), 663,552 elements, lognormal kappa_e"""
    import torch
    n = 48
    adj, P = _problem(n, 1e-12)
    assert P.mesh.nt == 663552
    k = _kappa(P.mesh.nt)
    kt = torch.tensor(k, device="cuda", requires_grad=True)
    J = adj.objective(kt, P)
    J.backward()
    g = kt.grad
    d = torch.tensor(np.random.default_rng(5).standard_normal(P.mesh.nt), device="cuda")
    h = 1e-6
    Jp = adj.objective(kt.detach() + h * d, P).item()
    Jm = adj.objective(kt.detach() - h * d, P).item()
    fd = (Jp - Jm) / (2 * h)
    an = torch.dot(g, d).item()
    assert abs(fd - an) <= 1e-5 * abs(an), (fd, an)
    # both solves met FreeFEM's criterion
    for key in ("forward", "adjoint"):
        assert P.last[key]["converged"] == 1


def test_adjoint_central_differences_cube48():
    """configuration: cube(48), 663,552 elements, lognormal kappa_e (seed
    20261202), central differences with h = 1e-6 on 20 seeded elements"""
    import torch
    n = 48
    adj, P = _problem(n, 1e-12)
    k = _kappa(P.mesh.nt)
    kt = torch.tensor(k, device="cuda", requires_grad=True)
    J = adj.objective(kt, P)
    J.backward()
    g = kt.grad.cpu().numpy()
    h = 1e-6
    els = np.random.default_rng(11).choice(P.mesh.nt, 20, replace=False)
    worst = 0.0
    for e in els:
        us = []
        for s in (1.0, -1.0):
            kk = k.copy()
            kk[e] += s * h
            us.append(P.solve(torch.tensor(kk, device="cuda")).clone())
            assert P.last["forward"]["converged"] == 1
        _, rp = ffhip.p1_misfit(P.mesh, us[0], P.ud)
        _, rm = ffhip.p1_misfit(P.mesh, us[1], P.ud)
        dJ = 0.5 * torch.dot(us[0] - us[1], rp + rm).item()
        fd = dJ / (2 * h)
        rel = abs(fd - g[e]) / abs(g[e])
        worst = max(worst, rel)
        assert rel <= 1e-3, (e, fd, g[e])
    print(f"cube(48) central differences on 20 elements: worst relative deviation {worst:.2e}",
          flush=True)
