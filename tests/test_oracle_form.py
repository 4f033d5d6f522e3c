"""Pins the oracle's P1 form with a reaction (mass) term,
int3d(Th)(kappa*grad(u)'*grad(v) + m*u*v) (Element_Op with an op_id x op_id
term, problem.cpp:6384-6437; P1 values P012_3d.cpp:123-175), 2D likewise:

* element mass matrix with constant m equals the closed form
  m |K| (1 + delta_ij) / 20 (tet), m |T| (1 + delta_ij) / 12 (triangle):
  Tet_5 and T_5 integrate the degree-2 products exactly;
* linearity: the combined form equals stiffness + mass assembled apart, and
  the mass-only form equals the adjoint's mass varf restatement;
* P0 mass coefficients scale each element's block;
* kappa = 0, m = 1 gives a matrix whose entries sum to the mesh volume.
"""
import numpy as np

import ffo


def test_element_mass_closed_form_3d_and_2d():
    rng = np.random.default_rng(0)
    for _ in range(20):
        X = rng.standard_normal(12)
        t = np.arange(4, dtype=np.int32)
        mes = ffo.lib().ffo_tet_mesure(X[0:3].copy(), X[3:6].copy(), X[6:9].copy(), X[9:12].copy())
        if mes < 0:
            t = np.array([0, 2, 1, 3], np.int32)
            mes = -mes
        m = 0.3 + rng.random()
        Ke = np.zeros(16)
        ffo.lib().ffo_p1_form_element_3d(X, t, ffo.COEF_CONST, 0.0, 0.0, ffo.COEF_CONST, m, 0.0, Ke)
        exact = m * mes / 20 * (np.ones((4, 4)) + np.eye(4))
        assert np.abs(Ke.reshape(4, 4) - exact).max() <= 1e-14 * exact.max()
    sq = ffo.square(3, 4)
    xy, tris = sq["xy"], sq["tris"].reshape(-1, 3)
    for k in range(len(tris)):
        Ke = np.zeros(9)
        ffo.lib().ffo_p1_form_element_2d(xy, tris[k].copy(), ffo.COEF_CONST, 0.0, 0.0,
                                         ffo.COEF_CONST, 2.5, 0.0, Ke)
        P = xy.reshape(-1, 2)[tris[k]]
        area = 0.5 * abs((P[1, 0] - P[0, 0]) * (P[2, 1] - P[0, 1]) - (P[1, 1] - P[0, 1]) * (P[2, 0] - P[0, 0]))
        exact = 2.5 * area / 12 * (np.ones((3, 3)) + np.eye(3))
        assert np.abs(Ke.reshape(3, 3) - exact).max() <= 1e-14 * exact.max()


def test_assembled_form_linearity_and_volume():
    m = ffo.cube(4)
    rp, ci = ffo.pattern(m["nv"], m["tets"], 4)
    stiff = ffo.assemble_p1_3d(m, rp, ci, ffo.COEF_SIN3)
    mass = ffo.assemble_p1_3d(m, rp, ci, ffo.COEF_CONST, 0.0, mkind=ffo.COEF_CONST, m0=0.7)
    both = ffo.assemble_p1_3d(m, rp, ci, ffo.COEF_SIN3, mkind=ffo.COEF_CONST, m0=0.7)
    assert np.abs(both - (stiff + mass)).max() <= 1e-14 * np.abs(both).max()
    unit = ffo.assemble_p1_3d(m, rp, ci, ffo.COEF_CONST, 0.0, mkind=ffo.COEF_CONST, m0=1.0)
    assert abs(unit.sum() - 1.0) < 1e-13  # sum_ij M_ij = volume of the unit cube
    assert np.array_equal(unit, ffo.mass_p1_3d(m, rp, ci))
    # P0 mass coefficient: linear in m_e
    me = np.random.default_rng(1).random(m["nt"])
    a = ffo.assemble_p1_3d(m, rp, ci, ffo.COEF_CONST, 0.0, mkind=ffo.COEF_P0, mass_e=me)
    b = ffo.assemble_p1_3d(m, rp, ci, ffo.COEF_CONST, 0.0, mkind=ffo.COEF_P0, mass_e=2 * me)
    assert np.abs(b - 2 * a).max() <= 1e-15 * np.abs(b).max()


def test_assembled_form_2d():
    sq = ffo.square(6, 6)
    rp, ci = ffo.pattern(sq["nv"], sq["tris"], 3)
    unit = ffo.assemble_p1_2d(sq, rp, ci, ffo.COEF_CONST, 0.0, mkind=ffo.COEF_CONST, m0=1.0)
    assert abs(unit.sum() - 1.0) < 1e-13
    stiff = ffo.assemble_p1_2d(sq, rp, ci, ffo.COEF_CONST, 1.0)
    both = ffo.assemble_p1_2d(sq, rp, ci, ffo.COEF_CONST, 1.0, mkind=ffo.COEF_CONST, m0=3.0)
    assert np.abs(both - (stiff + 3.0 * unit)).max() <= 1e-14 * np.abs(both).max()
