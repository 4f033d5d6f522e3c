"""Pins the oracle's P2 / vector-P2 elasticity restatement.

* node numbering and [P2,P2,P2] dofs: bit-exact against FreeFEM's own
  FESpace3 (BuildDFNumbering, GTypeOfFESum) compiled in oracle/_ref;
* P2 basis values and gradients: bit-exact against FreeFEM's
  TypeOfFE_P2Lagrange3d::FB;
* element matrix: symmetric, the six rigid motions in its kernel, and the
  exact strain energy of linear displacement fields (P2 contains them);
* pattern size: FreeFEM's measured nnz for cube(32) (69,108,489);
* right-hand side: the exact integrals of the P2 basis functions.
"""
import numpy as np
import pytest

import ffo

needs_ref = pytest.mark.skipif(not ffo.ref_available(), reason="oracle/_ref not built")


@needs_ref
@pytest.mark.parametrize("n", [1, 2, 4])
def test_p2_numbering_bitexact_vs_freefem_fespace(n):
    m = ffo.cube(n)
    nn, en, ed = ffo.p2v_dofs(m)
    ndf2, ed2 = ffo.ref_fespace_dofs(m, 2)
    ndf3, ed3 = ffo.ref_fespace_dofs(m, 3)
    assert nn == ndf2 == (2 * n + 1) ** 3
    assert np.array_equal(en, ed2)
    assert ndf3 == 3 * nn and np.array_equal(ed, ed3)


@needs_ref
def test_p2_basis_bitexact_vs_freefem_fb():
    rng = np.random.default_rng(2)
    for _ in range(50):
        v = rng.standard_normal(12)
        mes = ffo.lib().ffo_tet_mesure(v[0:3].copy(), v[3:6].copy(), v[6:9].copy(), v[9:12].copy())
        if mes < 0:
            v[3:6], v[6:9] = v[6:9].copy(), v[3:6].copy()
            mes = -mes
        G = np.zeros(12)
        ffo.lib().ffo_tet_gradlambda(v, mes, G)
        ph = rng.dirichlet(np.ones(4))[1:].copy()
        f, d, fr, dr = np.zeros(10), np.zeros(30), np.zeros(10), np.zeros(30)
        ffo.lib().ffo_p2_basis(ph, G, f, d)
        ffo.ref().ref_p2_fb(v, ph, fr, dr)
        assert np.array_equal(f, fr) and np.array_equal(d, dr)


def _element(seed, lam=1.0, mu=0.5):
    rng = np.random.default_rng(seed)
    X = rng.standard_normal(12)
    t = np.arange(4, dtype=np.int32)
    if ffo.lib().ffo_tet_mesure(X[0:3].copy(), X[3:6].copy(), X[6:9].copy(), X[9:12].copy()) < 0:
        t = np.array([0, 2, 1, 3], np.int32)
    Ke = np.zeros(900)
    ffo.lib().ffo_p2_elast_element(X, t, lam, mu, Ke)
    V = X.reshape(4, 3)[t]
    mes = ffo.lib().ffo_tet_mesure(*(V[i].copy() for i in range(4)))
    edges = [(0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3)]
    nodes = np.vstack([V] + [0.5 * (V[a] + V[b]) for a, b in edges])  # 10 P2 nodes
    return Ke.reshape(30, 30), nodes, mes


def _nodal(field, nodes):
    # comp-blocked local dofs: index 10*c + i
    U = np.array([field(p) for p in nodes])  # 10 x 3
    return U.T.ravel()


def test_p2_elast_element_kernel_and_energy():
    lam, mu = 1.0, 0.5
    for seed in range(5):
        K, nodes, mes = _element(seed, lam, mu)
        assert np.allclose(K, K.T, atol=1e-12 * np.abs(K).max())
        scale = np.abs(K).max()
        rigid = [lambda p, k=k: np.eye(3)[k] for k in range(3)]
        rigid += [lambda p, a=a: np.cross(np.eye(3)[a], p) for a in range(3)]
        for r in rigid:
            u = _nodal(r, nodes)
            assert np.abs(K @ u).max() <= 1e-11 * scale * np.abs(u).max()
        rng = np.random.default_rng(100 + seed)
        for _ in range(3):
            E = rng.standard_normal((3, 3))
            u = _nodal(lambda p: E @ p, nodes)
            Es = 0.5 * (E + E.T)
            exact = mes * (lam * np.trace(E) ** 2 + 2 * mu * np.sum(Es * Es))
            assert abs(u @ K @ u - exact) <= 1e-11 * abs(exact)


def test_p2_elast_element_exact_energy_of_quadratic_fields():
    """u_c(x) = x^T Q_c x + p_c . x + r_c (each component quadratic): P2
    interpolates u exactly, and the energy density lambda (div u)^2 +
    2 mu eps(u):eps(u) is a quadratic, integrated exactly by the P2 nodal rule
    int_K f = |K| (-1/20 sum_vertices f + 1/5 sum_edge-midpoints f).  The
    quadratic modes exercise the edge basis functions' gradients
    4 (l_a grad l_b + l_b grad l_a), which affine fields leave at their
    averages; a wrong edge gradient, Lame term or coupling index fails here."""
    lam, mu = 1.0, 0.5
    for seed in range(5):
        K, nodes, mes = _element(seed, lam, mu)
        rng = np.random.default_rng(300 + seed)
        for _ in range(3):
            Q = rng.standard_normal((3, 3, 3))
            Q = Q + Q.transpose(0, 2, 1)  # Q[c] symmetric
            P = rng.standard_normal((3, 3))
            r0 = rng.standard_normal(3)
            field = lambda x: np.array([x @ Q[c] @ x for c in range(3)]) + P @ x + r0
            grad = lambda x: 2 * Q @ x + P  # grad[c, a] = d u_c / d x_a

            def density(x):
                G = grad(x)
                E = 0.5 * (G + G.T)
                return lam * np.trace(G) ** 2 + 2 * mu * np.sum(E * E)

            exact = mes * (-sum(density(x) for x in nodes[:4]) / 20 +
                           sum(density(x) for x in nodes[4:]) / 5)
            u = _nodal(field, nodes)
            assert abs(u @ K @ u - exact) <= 1e-11 * abs(exact)


def test_p2_elast_system_small():
    r = ffo.elasticity_p2(3)
    n = 3
    assert r["ndof"] == 3 * (2 * n + 1) ** 3
    # clamped face x = 0 (label 4): all (2n+1)^2 nodes, 3 components
    assert r["onbc"].sum() == 3 * (2 * n + 1) ** 2
    b = r["b"]
    free = r["onbc"] == 0
    # total force = -volume on the free z-dofs plus the part on clamped rows (zeroed)
    nn, en, ed = r["nnodes"], r["elem_nodes"], r["elem_dofs"]
    assert b[2::3][free[2::3]].sum() < 0
    x, st = ffo.pcg(r["rowptr"], r["colidx"], r["vals"], b, eps=1e-10)
    assert st["nret"] == 1
    # a downward load on a clamped cube moves it down
    assert x[2::3][free[2::3]].mean() < 0


def test_p2_rhs_exact_basis_integrals():
    # int_K l_i (2 l_i - 1) = -|K|/20 (vertices), int_K 4 l_a l_b = |K|/5 (edges);
    # Tet_5 integrates both exactly
    m = ffo.cube(2)
    nn, en, ed = ffo.p2v_dofs(m)
    b = np.zeros(3 * nn)
    ffo.lib().ffo_rhs_p2_elast(m["nt"], m["xyz"], m["tets"], ed, np.array([0.0, 0.0, -1.0]), 3 * nn, b)
    X = m["xyz"].reshape(-1, 3)
    ref = np.zeros(nn)
    for k, t in enumerate(m["tets"].reshape(-1, 4)):
        mes = ffo.lib().ffo_tet_mesure(*(X[i].copy() for i in t))
        nodes = en[10 * k:10 * k + 10]
        ref[nodes[:4]] += -mes / 20
        ref[nodes[4:]] += mes / 5
    assert np.allclose(b[2::3], -ref, rtol=1e-13, atol=1e-16)
    assert np.abs(b[0::3]).max() == 0 and np.abs(b[1::3]).max() == 0
    assert abs(b[2::3].sum() + 1.0) < 1e-13


def test_p2_pattern_size_cube32_matches_freefem():
    m = ffo.cube(32)
    nn, en = ffo.p2_nodes(m)
    assert nn == 274625
    rp, ci = ffo.pattern(nn, en, 10)
    assert 9 * int(rp[-1]) == 69108489


# ---- scalar P2: int3d(kappa grad u . grad v) on Vh(Th,P2)
def _p2_lap(seed, kind=ffo.COEF_CONST, c0=1.0):
    rng = np.random.default_rng(seed)
    X = rng.standard_normal(12)
    t = np.arange(4, dtype=np.int32)
    if ffo.lib().ffo_tet_mesure(X[0:3].copy(), X[3:6].copy(), X[6:9].copy(), X[9:12].copy()) < 0:
        t = np.array([0, 2, 1, 3], np.int32)
    Ke = np.zeros(100)
    ffo.lib().ffo_p2_lap_element(X, t, kind, c0, 0.0, Ke)
    V = X.reshape(4, 3)[t]
    mes = ffo.lib().ffo_tet_mesure(*(V[i].copy() for i in range(4)))
    edges = [(0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3)]
    nodes = np.vstack([V] + [0.5 * (V[a] + V[b]) for a, b in edges])
    return Ke.reshape(10, 10), nodes, mes


def test_p2_lap_element_kernel_and_exact_energy():
    for seed in range(6):
        K, nodes, mes = _p2_lap(seed, c0=1.7)
        scale = np.abs(K).max()
        assert np.allclose(K, K.T, atol=1e-13 * scale)
        assert np.abs(K.sum(axis=1)).max() <= 1e-12 * scale  # constants in the kernel
        rng = np.random.default_rng(50 + seed)
        for _ in range(3):
            Q = rng.standard_normal((3, 3)); Q = Q + Q.T
            p = rng.standard_normal(3)
            u = np.array([x @ Q @ x + p @ x for x in nodes])  # P2 interpolates quadratics exactly
            g2 = lambda x: np.sum((2 * Q @ x + p) ** 2)  # |grad u|^2, a quadratic
            # int_K of a quadratic: |K| (-1/20 sum_vertices + 1/5 sum_edge midpoints)
            exact = 1.7 * mes * (-sum(g2(x) for x in nodes[:4]) / 20 + sum(g2(x) for x in nodes[4:]) / 5)
            assert abs(u @ K @ u - exact) <= 1e-11 * abs(exact)


def test_p2_lap_system_small():
    r = ffo.poisson_p2(3)
    nn = r["nnodes"]
    assert nn == 7 ** 3 and r["onbc"].sum() == 7 ** 3 - 5 ** 3  # boundary nodes of the 7^3 lattice
    m, en = r["mesh"], r["elem_nodes"]
    b = np.zeros(nn)
    ffo.lib().ffo_rhs_p2(m["nt"], m["xyz"], m["tets"], en, 1.0, nn, b)
    assert abs(b.sum() - 1.0) < 1e-13  # the P2 basis sums to 1: int(f) over the unit cube
    assert np.array_equal(r["b"][r["onbc"] == 0], b[r["onbc"] == 0])
    x, st = ffo.pcg(r["rowptr"], r["colidx"], r["vals"], r["b"], eps=1e-10)
    assert st["nret"] == 1 and x[r["onbc"] == 0].min() > 0  # -lap u = 1, u = 0 on the boundary
    # the pattern: the scalar node rows of the [P2,P2,P2] pattern
    assert 9 * r["colidx"].size == ffo.elasticity_p2(3)["colidx"].size
