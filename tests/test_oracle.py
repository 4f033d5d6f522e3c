"""Pins the oracle (CPU restatement) before the GPU results are compared with it.

Each check is chosen so that a plausible slip in the restatement (a dropped
quadrature point, a wrong face-label map, a transposed element matrix, a wrong
sign in the CG recurrences) fails one of them:
  * the compiled reference subset (oracle/_ref) where it is built: Tet_5/T_5
    tables, Tet geometry, HashMatrix pattern + values + SetBC, ConjugueGradient;
  * the committed golden fixtures made from that subset (tests/golden);
  * closed forms: quadrature exactness, cube/square nnz = nv + 2E, volumes,
    P1 element matrix = |K| kappa grad.grad, integral of the basis.
"""
import math

import numpy as np
import pytest

import ffo

GOLDEN = __file__.rsplit("/", 1)[0] + "/golden/"
needs_ref = pytest.mark.skipif(not ffo.ref_available(), reason="oracle/_ref not built")


def _monomial_tet(a, b, c):
    # exact integral of x^a y^b z^c over the unit reference tet, normalised to volume 1
    return 6.0 * math.factorial(a) * math.factorial(b) * math.factorial(c) / math.factorial(a + b + c + 3)


def test_qf_tet5_exact_to_degree_5_and_not_6():
    p = np.zeros(42); w = np.zeros(14)
    assert ffo.lib().ffo_qf_tet5(p, w) == 14
    P = p.reshape(14, 3)
    assert abs(w.sum() - 1.0) < 1e-15
    for deg in range(6):
        for a in range(deg + 1):
            for b in range(deg + 1 - a):
                c = deg - a - b
                q = np.sum(w * P[:, 0] ** a * P[:, 1] ** b * P[:, 2] ** c)
                assert abs(q - _monomial_tet(a, b, c)) < 1e-14, (a, b, c)
    q6 = np.sum(w * P[:, 0] ** 6)
    assert abs(q6 - _monomial_tet(6, 0, 0)) > 1e-6


def test_qf_tri5_exact_to_degree_5():
    p = np.zeros(14); w = np.zeros(7)
    ffo.lib().ffo_qf_tri5(p, w)
    P = p.reshape(7, 2)
    for deg in range(6):
        for a in range(deg + 1):
            b = deg - a
            exact = 2.0 * math.factorial(a) * math.factorial(b) / math.factorial(a + b + 2)
            assert abs(np.sum(w * P[:, 0] ** a * P[:, 1] ** b) - exact) < 1e-14


@needs_ref
def test_qf_tables_bitexact_vs_reference():
    for n, fo, fr in ((14, "ffo_qf_tet5", "ref_qf_tet5"), (7, "ffo_qf_tri5", "ref_qf_tri5")):
        d = 3 if n == 14 else 2
        p, w, pr, wr = np.zeros(d * n), np.zeros(n), np.zeros(d * n), np.zeros(n)
        getattr(ffo.lib(), fo)(p, w)
        getattr(ffo.ref(), fr)(pr, wr)
        assert np.array_equal(p, pr) and np.array_equal(w, wr)


@needs_ref
def test_tet_geometry_bitexact_vs_reference():
    rng = np.random.default_rng(7)
    for _ in range(200):
        v = rng.standard_normal(12)
        mes = ffo.lib().ffo_tet_mesure(v[0:3].copy(), v[3:6].copy(), v[6:9].copy(), v[9:12].copy())
        if mes <= 0:
            v[3:6], v[6:9] = v[6:9].copy(), v[3:6].copy()
            mes = -mes
        mes = ffo.lib().ffo_tet_mesure(v[0:3].copy(), v[3:6].copy(), v[6:9].copy(), v[9:12].copy())
        G = np.zeros(12)
        ffo.lib().ffo_tet_gradlambda(v, mes, G)
        mr = np.zeros(1); Gr = np.zeros(12)
        ffo.ref().ref_tet_geom(v, mr, Gr)
        assert mes == mr[0] and np.array_equal(G, Gr)


def _faces(tets):
    T = tets.reshape(-1, 4)
    F = np.concatenate([T[:, [1, 2, 3]], T[:, [0, 2, 3]], T[:, [0, 1, 3]], T[:, [0, 1, 2]]])
    F.sort(axis=1)
    _, cnt = np.unique(F, axis=0, return_counts=True)
    return cnt


@pytest.mark.parametrize("n", [1, 2, 3, 5])
def test_cube_mesh_is_conforming_unit_volume(n):
    m = ffo.cube(n)
    assert (m["nv"], m["nt"], m["nbe"]) == ((n + 1) ** 3, 6 * n ** 3, 12 * n * n)
    X = m["xyz"].reshape(-1, 3)
    T = m["tets"].reshape(-1, 4)
    vols = np.array([ffo.lib().ffo_tet_mesure(*(X[t[k]].copy() for k in range(4))) for t in T])
    assert np.all(vols > 0) and abs(vols.sum() - 1.0) < 1e-13
    cnt = _faces(m["tets"])
    # boundary faces appear once, interior faces twice (conforming split)
    assert set(np.unique(cnt)) <= {1, 2}
    assert np.sum(cnt == 1) == m["nbe"]
    # each label owns one face of the cube: x=0 -> 4, x=1 -> 2, y=0 -> 1, y=1 -> 3, z=0 -> 5, z=1 -> 6
    B = m["btri"].reshape(-1, 3)
    for lab, (ax, val) in {4: (0, 0.), 2: (0, 1.), 1: (1, 0.), 3: (1, 1.), 5: (2, 0.), 6: (2, 1.)}.items():
        sel = B[m["blab"] == lab]
        assert sel.shape[0] == 2 * n * n
        assert np.all(X[sel.ravel(), ax] == val)
    # every tet holds the hex body diagonal (split around vertex 0 -> vertex 7)
    for t in T[:6]:
        assert 0 in t and (1 + (n + 1) + (n + 1) ** 2) in t


def _cube_nnz(n):
    E = 3 * n * (n + 1) ** 2 + 3 * n * n * (n + 1) + n ** 3
    return (n + 1) ** 3 + 2 * E


@pytest.mark.parametrize("n", [1, 2, 4, 7, 12])
def test_pattern_closed_form_and_bruteforce(n):
    m = ffo.cube(n)
    rp, ci = ffo.pattern(m["nv"], m["tets"], 4)
    assert rp[-1] == _cube_nnz(n)
    rb, cb = ffo.pattern_bruteforce(m["nv"], m["tets"], 4)
    assert np.array_equal(rp, rb) and np.array_equal(ci, cb)


def test_pattern_closed_form_large():
    # figures quoted for the reference runs: cube(20) 128,581; cube(48) 1,707,697
    assert _cube_nnz(20) == 128581 and _cube_nnz(48) == 1707697
    assert _cube_nnz(64) == 4018753 and _cube_nnz(160) == 61979041
    m = ffo.cube(20)
    rp, _ = ffo.pattern(m["nv"], m["tets"], 4)
    assert rp[-1] == 128581


@pytest.mark.parametrize("n", [3, 10])
def test_square_pattern_closed_form(n):
    s = ffo.square(n)
    assert (s["nv"], s["nt"], s["nbe"]) == ((n + 1) ** 2, 2 * n * n, 4 * n)
    rp, ci = ffo.pattern(s["nv"], s["tris"], 3)
    E = 2 * n * (n + 1) + n * n
    assert rp[-1] == s["nv"] + 2 * E
    assert ffo.square(200)["nv"] == 40401 and s["nv"] + 2 * E > 0
    assert (40401 + 2 * (2 * 200 * 201 + 200 * 200)) == 281201


@needs_ref
@pytest.mark.parametrize("n,kind", [(2, ffo.COEF_SIN3), (5, ffo.COEF_SIN3), (4, ffo.COEF_CONST), (3, ffo.COEF_P0)])
def test_pattern_and_values_bitexact_vs_reference_hashmatrix(n, kind):
    m = ffo.cube(n)
    ke = np.exp(np.random.default_rng(n).standard_normal(m["nt"]) * 0.5) if kind == ffo.COEF_P0 else None
    rp, ci = ffo.pattern(m["nv"], m["tets"], 4)
    v = ffo.assemble_p1_3d(m, rp, ci, kind, 1.0, ke)
    R = ffo.ref()
    h = R.ref_hm_new(m["nv"])
    R.ref_hm_add_p1_3d(h, m["nv"], m["nt"], m["xyz"], m["tets"], kind, 1.0,
                       None if ke is None else ke.ctypes.data_as(ffo.ctypes.c_void_p))
    nnz = R.ref_hm_nnz(h)
    rr, cr, vr = np.zeros(m["nv"] + 1, np.int64), np.zeros(nnz, np.int32), np.zeros(nnz)
    R.ref_hm_csr(h, rr, cr, vr)
    R.ref_hm_free(h)
    assert np.array_equal(rp, rr) and np.array_equal(ci, cr)
    assert np.array_equal(v, vr)


@pytest.mark.parametrize("name", ["p1_cube4_sin3.npz", "p1_cube6_const.npz", "p1_cube5_p0.npz"])
def test_oracle_vs_golden_fixtures(name):
    g = np.load(GOLDEN + name)
    n, kind = int(g["n"]), int(g["kind"])
    ke = g["kappa_e"] if "kappa_e" in g.files else None
    m, rp, ci, v, onbc, b = ffo.poisson_p1_3d(n, kind, 1.0, ke)
    assert np.array_equal(rp, g["rowptr"]) and np.array_equal(ci, g["colidx"])
    assert np.array_equal(v, g["vals"])
    assert np.array_equal(onbc, g["onbc"]) and np.array_equal(b, g["b"])
    x, st = ffo.pcg(rp, ci, v, b, eps=1e-10)
    # the reference CG runs in hash order, the oracle in CSR order: same
    # iterations up to summation-order rounding
    assert abs(st["iters"] - int(g["iters"])) <= 1
    assert np.max(np.abs(x - g["x"])) <= 1e-9 * np.max(np.abs(g["x"]))


def test_p1_element_closed_form():
    rng = np.random.default_rng(3)
    X = rng.standard_normal(12)
    t = np.arange(4, dtype=np.int32)
    mes = ffo.lib().ffo_tet_mesure(X[0:3].copy(), X[3:6].copy(), X[6:9].copy(), X[9:12].copy())
    if mes < 0:
        t = np.array([0, 2, 1, 3], np.int32)
        mes = -mes
    G = np.zeros(12)
    ffo.lib().ffo_tet_gradlambda(X[np.repeat(t, 3) * 3 + np.tile(np.arange(3), 4)].copy(), mes, G)
    G = G.reshape(4, 3)
    Ke = np.zeros(16)
    ffo.lib().ffo_p1_element_3d(X, t, ffo.COEF_CONST, 2.5, 0.0, Ke)
    Ke = Ke.reshape(4, 4)
    ref = 2.5 * mes * G @ G.T
    assert np.allclose(Ke, ref, rtol=1e-13, atol=1e-13 * np.abs(ref).max())
    assert np.allclose(Ke.sum(axis=1), 0, atol=1e-13 * np.abs(Ke).max())
    # gradients of barycentrics: sum to zero, and grad(l_i) . (v_j - v_0) = delta_ij - delta_i0
    V = X[np.repeat(t, 3) * 3 + np.tile(np.arange(3), 4)].reshape(4, 3)
    assert np.allclose(G.sum(axis=0), 0, atol=1e-12)
    for j in range(1, 4):
        assert np.allclose(G @ (V[j] - V[0]), np.eye(4)[:, j] - np.eye(4)[:, 0], atol=1e-12)


def test_p1_sin3_element_matches_exact_integral_on_small_tet():
    # kappa integral via Tet_5 vs a fine composite rule: agree to the rule's accuracy
    m = ffo.cube(1)
    Ke = np.zeros(16)
    ffo.lib().ffo_p1_element_3d(m["xyz"], m["tets"][:4].copy(), ffo.COEF_SIN3, 1.0, 0.0, Ke)
    Kc = np.zeros(16)
    ffo.lib().ffo_p1_element_3d(m["xyz"], m["tets"][:4].copy(), ffo.COEF_CONST, 1.0, 0.0, Kc)
    ratio = Ke.reshape(4, 4)[0, 0] / Kc.reshape(4, 4)[0, 0]
    # mean of kappa over the tet (4,0,6,7) of the unit cube by Monte Carlo
    rng = np.random.default_rng(0)
    L = rng.dirichlet(np.ones(4), 400000)
    X = m["xyz"].reshape(-1, 3)[m["tets"][:4]]
    P = L @ X
    kap = 1 + 0.5 * np.sin(2 * np.pi * P[:, 0]) * np.sin(2 * np.pi * P[:, 1]) * np.sin(2 * np.pi * P[:, 2])
    assert abs(ratio - kap.mean()) < 5e-3


def test_rhs_is_integral_of_basis():
    m = ffo.cube(3)
    b = ffo.rhs_p1_3d(m, 1.0)
    assert abs(b.sum() - 1.0) < 1e-14
    X = m["xyz"].reshape(-1, 3)
    ref = np.zeros(m["nv"])
    for t in m["tets"].reshape(-1, 4):
        ref[t] += ffo.lib().ffo_tet_mesure(*(X[k].copy() for k in t)) / 4
    assert np.allclose(b, ref, rtol=1e-13)


def test_bc_marks_every_boundary_vertex():
    n = 4
    m = ffo.cube(n)
    onbc = ffo.mark_bc(m["nv"], m["btri"], m["blab"], 3, (1, 2, 3, 4, 5, 6))
    assert onbc.sum() == (n + 1) ** 3 - (n - 1) ** 3
    only4 = ffo.mark_bc(m["nv"], m["btri"], m["blab"], 3, (4,))
    assert only4.sum() == (n + 1) ** 2
    assert np.all(m["xyz"].reshape(-1, 3)[only4 == 1, 0] == 0.0)


@needs_ref
def test_cg_bitexact_vs_reference_in_csr_order():
    m, rp, ci, v, onbc, b = ffo.poisson_p1_3d(7)
    x, st = ffo.pcg(rp, ci, v, b, eps=1e-10)
    R = ffo.ref()
    h = R.ref_hm_new(m["nv"])
    R.ref_hm_add_p1_3d(h, m["nv"], m["nt"], m["xyz"], m["tets"], ffo.COEF_SIN3, 1.0, None)
    R.ref_hm_setbc(h, onbc, 1e30)
    nnz = R.ref_hm_nnz(h)
    rr, cr, vr = np.zeros(m["nv"] + 1, np.int64), np.zeros(nnz, np.int32), np.zeros(nnz)
    R.ref_hm_csr(h, rr, cr, vr)  # sorts the HashMatrix: its COO product now runs in CSR order
    xr, s = np.zeros(m["nv"]), np.zeros(4)
    R.ref_hm_cg(h, b, xr, 1e-10, 0, 1e30, 0, s)
    R.ref_hm_free(h)
    assert int(s[0]) == st["iters"] and np.array_equal(x, xr)


def test_cg_converges_and_true_residual_small():
    m, rp, ci, v, onbc, b = ffo.poisson_p1_3d(10)
    x, st = ffo.pcg(rp, ci, v, b, eps=1e-10)
    assert st["nret"] == 1 and st["gcg"] < 1e-20 * st["gcg0"]
    r = b - ffo.spmv(rp, ci, v, x)
    inner = onbc == 0
    assert np.linalg.norm(r[inner]) <= 1e-8 * np.linalg.norm(b[inner])
    assert np.all(x[onbc == 1] == 0.0) or np.max(np.abs(x[onbc == 1])) < 1e-25
    # -div(kappa grad u) = 1 has a positive solution
    assert x[inner].min() > 0


def test_cg_planted_sign_error_is_caught():
    # a plausible slip (wrong sign of rho) diverges: the convergence check would catch it
    m, rp, ci, v, onbc, b = ffo.poisson_p1_3d(4)
    x, st = ffo.pcg(rp, ci, -v, b, eps=1e-10, itmax=50)
    r = b - ffo.spmv(rp, ci, v, x)
    assert np.linalg.norm(r[onbc == 0]) > 1e-3 * np.linalg.norm(b[onbc == 0])


def test_2d_p1_stencil():
    n = 6
    s = ffo.square(n)
    rp, ci = ffo.pattern(s["nv"], s["tris"], 3)
    v = ffo.assemble_p1_2d(s, rp, ci)
    # interior row of the 5-point Laplacian: 4 on the diagonal, -1 on axis
    # neighbours, 0 on the split diagonal
    i = 3 + 3 * (n + 1)
    row = dict(zip(ci[rp[i]:rp[i + 1]], v[rp[i]:rp[i + 1]]))
    assert abs(row[i] - 4) < 1e-13
    for j in (i - 1, i + 1, i - (n + 1), i + (n + 1)):
        assert abs(row[j] + 1) < 1e-13
    for j in (i - n - 2, i + n + 2):
        assert abs(row[j]) < 1e-13
    b = ffo.rhs_p1_2d(s, 1.0)
    assert abs(b.sum() - 1.0) < 1e-14
