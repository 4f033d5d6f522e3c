"""Pins the oracle's boundary integrals on P1 (int2d(Th3, labels)(alpha*u*v),
int2d(Th3, labels)(g*v) and their int1d forms in 2D: AssembleBilinearForm's
boundary branch, problem.cpp:1273-1327, Element_Op's face branch
problem.cpp:6515-6561 and 2D edges 6223-6240, Element_rhs on a face
problem.cpp:8517-8590).

Element_Op is interpreter code and does not compile on its own, so the
restatement is pinned by what the mathematics fixes:

* constant alpha: each listed face adds alpha |F| (1 + delta_ij) / 12 (edges:
  alpha |E| (1 + delta_ij) / 6) at its vertex pairs and nothing elsewhere — the
  face rule T_5 (degree 5) and Gauss-Legendre 3 are exact for it;
* alpha = x (given at the quadrature points, QP kind): exact moments
  int_F l_i l_j l_k = 2|F| a!b!c!/(a+b+c+2)! (edges |E| a!b!/(a+b+1)!), so a
  wrong point mapping, a wrong face table or a dropped 1/2 fails;
* the analytic SIN3 coefficient evaluated through the element mapping equals
  the QP kind fed sin values at the boundary element's own points;
* the linear form: int_F g l_i for constant and linear g (same moments);
* label selection: unlisted labels add nothing; all six cube faces add 6 alpha.
"""
import math

import numpy as np
import pytest

import ffo


def _pattern(m, dim):
    el = m["tets"] if dim == 3 else m["tris"]
    rp = np.zeros(m["nv"] + 1, np.int64)
    ffo.lib().ffo_pattern_rowptr(m["nv"], m["nt"], dim + 1, el, rp)
    ci = np.zeros(rp[-1], np.int32)
    ffo.lib().ffo_pattern_colidx(m["nv"], m["nt"], dim + 1, el, rp, ci)
    return rp, ci


def _dense(rp, ci, va, n):
    A = np.zeros((n, n))
    for i in range(n):
        A[i, ci[rp[i]:rp[i + 1]]] = va[rp[i]:rp[i + 1]]
    return A


def _bnd(m, dim):
    be = m["btri"] if dim == 3 else m["bedges"]
    X = m["xyz"] if dim == 3 else m["xy"]
    return be.reshape(-1, dim), X.reshape(-1, dim)


def _measure(P):
    if P.shape[0] == 3:
        return 0.5 * np.linalg.norm(np.cross(P[1] - P[0], P[2] - P[0]))
    return np.linalg.norm(P[1] - P[0])


def _moment(dim, meas, exps):
    """int over a face (dim 3) / edge (dim 2) of prod l_i^a_i"""
    num = np.prod([math.factorial(a) for a in exps])
    s = sum(exps)
    if dim == 3:
        return 2 * meas * num / math.factorial(s + 2)
    return meas * num / math.factorial(s + 1)


def _exact(m, dim, labels, alpha_lin=None, alpha=1.0):
    """dense boundary matrix and vector of int alpha l_i l_j and int alpha l_i,
    alpha = constant or the linear function x (alpha_lin = 'x')"""
    be, X = _bnd(m, dim)
    n = m["nv"]
    M = np.zeros((n, n))
    b = np.zeros(n)
    for f in range(be.shape[0]):
        if m["blab"][f] not in labels:
            continue
        vid = be[f]
        P = X[vid]
        meas = _measure(P)
        nb = dim
        for i in range(nb):
            e = [0] * nb
            e[i] += 1
            if alpha_lin is None:
                b[vid[i]] += alpha * _moment(dim, meas, e)
            else:
                for k in range(nb):
                    ek = list(e)
                    ek[k] += 1
                    b[vid[i]] += P[k, 0] * _moment(dim, meas, ek)
            for j in range(nb):
                e2 = [0] * nb
                e2[i] += 1
                e2[j] += 1
                if alpha_lin is None:
                    M[vid[i], vid[j]] += alpha * _moment(dim, meas, e2)
                else:
                    for k in range(nb):
                        e3 = list(e2)
                        e3[k] += 1
                        M[vid[i], vid[j]] += P[k, 0] * _moment(dim, meas, e3)
    return M, b


def _own_points(m, dim):
    """physical points of every boundary element's own rule (mesh order)"""
    be, X = _bnd(m, dim)
    if dim == 3:
        qp, qw = np.zeros(14), np.zeros(7)
        ffo.lib().ffo_qf_tri5(qp, qw)
        s, t = qp[0::2], qp[1::2]
        lam = np.stack([1 - s - t, s, t], 1)  # (7, 3)
    else:
        x, w = np.zeros(3), np.zeros(3)
        ffo.lib().ffo_qf_gl3(x, w)
        lam = np.stack([1 - x, x], 1)
    return np.einsum("qk,fkc->fqc", lam, X[be])  # (nbe, nq, dim)


MESHES = {
    "cube4": (3, lambda: ffo.cube(4)),
    "cube_3x2x5": (3, lambda: ffo.cube(3, 2, 5)),
    "square6x5": (2, lambda: ffo.square(6, 5)),
}


@pytest.mark.parametrize("name", list(MESHES))
@pytest.mark.parametrize("labels", [(1,), (2, 5), (1, 2, 3, 4, 5, 6)])
def test_constant_alpha_closed_form(name, labels):
    dim, mk = MESHES[name]
    m = mk()
    labels = tuple(l for l in labels if dim == 3 or l <= 4)
    rp, ci = _pattern(m, dim)
    va = np.zeros(rp[-1])
    b = np.zeros(m["nv"])
    ffo.p1_boundary(m, labels, ffo.COEF_CONST, 2.5, rowptr=rp, colidx=ci, vals=va, b=b)
    M, bx = _exact(m, dim, labels, alpha=2.5)
    A = _dense(rp, ci, va, m["nv"])
    assert np.abs(A - M).max() <= 1e-15 * max(1.0, np.abs(M).max()) * 10
    assert np.abs(b - bx).max() <= 1e-14 * np.abs(bx).max()


def test_all_cube_faces_total():
    m = ffo.cube(5)
    rp, ci = _pattern(m, 3)
    va = np.zeros(rp[-1])
    ffo.p1_boundary(m, (1, 2, 3, 4, 5, 6), ffo.COEF_CONST, 1.0, rowptr=rp, colidx=ci, vals=va)
    assert va.sum() == pytest.approx(6.0, rel=1e-13)  # the cube's surface
    va2 = np.zeros(rp[-1])
    ffo.p1_boundary(m, (77,), ffo.COEF_CONST, 1.0, rowptr=rp, colidx=ci, vals=va2)
    assert not va2.any()


@pytest.mark.parametrize("name", list(MESHES))
def test_linear_alpha_exact_moments(name):
    dim, mk = MESHES[name]
    m = mk()
    labels = (1, 2, 3, 4) if dim == 2 else (1, 2, 3, 4, 5, 6)
    rp, ci = _pattern(m, dim)
    pts = _own_points(m, dim)
    qpv = np.ascontiguousarray(pts[:, :, 0])  # alpha = x at the rule's points
    va = np.zeros(rp[-1])
    b = np.zeros(m["nv"])
    ffo.p1_boundary(m, labels, ffo.COEF_QP, 0.0, coef=qpv, rowptr=rp, colidx=ci, vals=va, b=b)
    M, bx = _exact(m, dim, labels, alpha_lin="x")
    A = _dense(rp, ci, va, m["nv"])
    assert np.abs(A - M).max() <= 1e-14 * np.abs(M).max()
    assert np.abs(b - bx).max() <= 1e-14 * np.abs(bx).max()


@pytest.mark.parametrize("name", list(MESHES))
def test_sin3_through_element_mapping_equals_qp(name):
    dim, mk = MESHES[name]
    m = mk()
    labels = (1, 3) if dim == 2 else (2, 4, 6)
    rp, ci = _pattern(m, dim)
    pts = _own_points(m, dim)
    tp = 2 * np.pi
    s = 0.5 * np.sin(tp * pts[:, :, 0]) * np.sin(tp * pts[:, :, 1])
    if dim == 3:
        s = s * np.sin(tp * pts[:, :, 2])
    qpv = np.ascontiguousarray(1 + s)
    va1, va2 = np.zeros(rp[-1]), np.zeros(rp[-1])
    ffo.p1_boundary(m, labels, ffo.COEF_SIN3, 0.0, rowptr=rp, colidx=ci, vals=va1)
    ffo.p1_boundary(m, labels, ffo.COEF_QP, 0.0, coef=qpv, rowptr=rp, colidx=ci, vals=va2)
    assert np.abs(va1 - va2).max() <= 1e-14 * np.abs(va2).max()
    assert np.abs(va1).max() > 0


def test_p0_coefficient_is_linear():
    m = ffo.cube(3)
    rp, ci = _pattern(m, 3)
    ce = np.random.default_rng(3).uniform(0.5, 2.0, m["nbe"])
    va = np.zeros(rp[-1])
    ffo.p1_boundary(m, (1, 2), ffo.COEF_P0, 0.0, coef=ce, rowptr=rp, colidx=ci, vals=va)
    ref = np.zeros(rp[-1])
    for f in range(m["nbe"]):  # face by face, constant coefficient on each
        if m["blab"][f] in (1, 2):
            mf = dict(m, nbe=1, btri=np.ascontiguousarray(m["btri"][3 * f:3 * f + 3]),
                      blab=np.ascontiguousarray(m["blab"][f:f + 1]))
            ffo.p1_boundary(mf, (m["blab"][f],), ffo.COEF_CONST, ce[f], rowptr=rp, colidx=ci,
                            vals=ref)
    assert np.abs(va - ref).max() <= 1e-15 * np.abs(ref).max() * 10


@pytest.mark.skipif(not ffo.ref_available(), reason="oracle/_ref not built")
def test_gauss_legendre3_table_bitexact_vs_reference():
    x, w, xr, wr = np.zeros(3), np.zeros(3), np.zeros(3), np.zeros(3)
    assert ffo.lib().ffo_qf_gl3(x, w) == 3 and ffo.ref().ref_qf_gl3(xr, wr) == 3
    assert np.array_equal(x, xr) and np.array_equal(w, wr)
