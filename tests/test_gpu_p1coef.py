"""GPU parity of the P1 finite-element-function coefficient kind
(FFHIP_COEF_P1: vertex values, interpolated on the device at each quadrature
point as FreeFEM evaluates a P1 function there) against the QP kind fed the
oracle's interpolation of the same function (ffo.p1_at_qp, FreeFEM's
P012_3d.cpp:98-106 / FESpace.cpp:1078-1085 order) — the values the plugin
handed over before, evaluated by the interpreter at every point.

Bars: matrices, right-hand sides within 1e-13 (max-norm, relative); the
device-pointer path bit-equal to the host-pointer path.
"""
import numpy as np
import pytest

import ffo

pytestmark = pytest.mark.gpu
ffhip = pytest.importorskip("ffhip")


def _close(a, b, tol=1e-13):
    return np.max(np.abs(a - b)) <= tol * np.max(np.abs(b))


def _u(nv, seed):
    return np.exp(0.4 * np.random.default_rng(seed).standard_normal(nv))


@pytest.mark.parametrize("with_mass", [False, True])
def test_p1_kind_3d_matrix(with_mass):
    n = 6
    mo = ffo.cube(n)
    k, mm = _u(mo["nv"], 1), _u(mo["nv"], 2)
    kq, mq = ffo.p1_at_qp(mo, k), ffo.p1_at_qp(mo, mm)
    M = ffhip.cube(n)
    A1, A2 = ffhip.Matrix.p1(M), ffhip.Matrix.p1(M)
    if with_mass:
        A1.assemble_p1(ffhip.COEF_P1, 0.0, k, ffhip.COEF_P1, 0.0, mm)
        A2.assemble_p1(ffhip.COEF_QP, 0.0, kq, ffhip.COEF_QP, 0.0, mq)
    else:
        A1.assemble_p1(ffhip.COEF_P1, 0.0, k)
        A2.assemble_p1(ffhip.COEF_QP, 0.0, kq)
    assert _close(A1.download()[2], A2.download()[2])


@pytest.mark.parametrize("n", [6, 13])
def test_p1_kind_one_call_assembly(n):
    """the one-call pattern+values path the plugin takes, with kappa's host
    coefficient array, against the QP kind fed the oracle's interpolation"""
    mo = ffo.cube(n)
    k = _u(mo["nv"], 3)
    kq = ffo.p1_at_qp(mo, k)
    M = ffhip.cube(n)
    A1 = ffhip.Matrix.p1_assembled(M, ffhip.COEF_P1, 0.0, k)
    A2 = ffhip.Matrix.p1(M)
    A2.assemble_p1(ffhip.COEF_QP, 0.0, kq)
    r1, c1, v1 = A1.download()
    r2, c2, v2 = A2.download()
    assert np.array_equal(r1, r2) and np.array_equal(c1, c2)
    assert _close(v1, v2)


def test_p1_kind_device_pointer_equals_host_pointer():
    import torch
    n = 5
    M = ffhip.cube(n)
    k = _u(M.nv, 3)
    A1, A2 = ffhip.Matrix.p1(M), ffhip.Matrix.p1(M)
    A1.assemble_p1(ffhip.COEF_P1, 0.0, k)
    A2.assemble_p1(ffhip.COEF_P1, 0.0, torch.tensor(k, device="cuda"))
    assert np.array_equal(A1.download()[2], A2.download()[2])


def test_p1_kind_2d_matrix_with_mass():
    mo = ffo.square(9, 7)
    k, mm = _u(mo["nv"], 4), _u(mo["nv"], 5)
    kq, mq = ffo.p1_at_qp(mo, k), ffo.p1_at_qp(mo, mm)
    M = ffhip.square(9, 7)
    A1, A2 = ffhip.Matrix.p1(M), ffhip.Matrix.p1(M)
    A1.assemble_p1(ffhip.COEF_P1, 0.0, k, ffhip.COEF_P1, 0.0, mm)
    A2.assemble_p1(ffhip.COEF_QP, 0.0, kq, ffhip.COEF_QP, 0.0, mq)
    assert _close(A1.download()[2], A2.download()[2])


@pytest.mark.parametrize("dim", [3, 2])
def test_p1_kind_rhs(dim):
    mo = ffo.cube(5) if dim == 3 else ffo.square(8)
    M = ffhip.cube(5) if dim == 3 else ffhip.square(8)
    f = _u(mo["nv"], 6) - 0.5
    fq = ffo.p1_at_qp(mo, f)
    labels = (1, 3)
    b1 = ffhip.rhs_p1_form(M, ffhip.COEF_P1, 0.0, f, labels, 1e30, 0.25)
    b2 = ffhip.rhs_p1_form(M, ffhip.COEF_QP, 0.0, fq, labels, 1e30, 0.25)
    off = b2 < 1e29
    assert np.array_equal(b1[~off], b2[~off])
    assert _close(b1[off], b2[off])


def test_p1_kind_scalar_p2_matrix():
    n = 4
    mo = ffo.cube(n)
    k = _u(mo["nv"], 7)
    kq = ffo.p1_at_qp(mo, k)
    M = ffhip.cube(n)
    sp = ffhip.P2VSpace(M)
    A1, A2 = ffhip.Matrix.p2(sp), ffhip.Matrix.p2(sp)
    A1.assemble_p2(ffhip.COEF_P1, 0.0, k)
    A2.assemble_p2(ffhip.COEF_QP, 0.0, kq)
    assert _close(A1.download()[2], A2.download()[2])
