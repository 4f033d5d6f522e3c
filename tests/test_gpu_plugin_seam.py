"""The plugin's solver glue (freefem-sources_amd/plugin/ffhip_glue.hpp:
SolverCore, the Resident hand-over, values_into_hashmatrix) on the GPU, with
FreeFEM's compiled HashMatrix and ConjugueGradient on the other side
(oracle/seam_driver.cpp in oracle/_ref/libffref_ref.so).

Bars:
* a HashMatrix solved through the glue (uploaded by HashMatrix::CSR; a half
  HashMatrix, sym=1, expanded to L + L^T - diag L on the device) against the
  compiled ConjugueGradient on the same HashMatrix (SolverCG's Jacobi, tgv
  rows): FP64 iterations within 2, x within 1e-8 (relative 2-norm), the eps
  written back for veps within 1e-12 of the reference's (sqrt(eps^2 gCg0),
  CG.cpp:226); the mixed solver within 1e-8;
* the assembly-to-solver hand-over on cube(8): the solver adopts the device
  matrix without an upload; a pattern-keeping re-assembly (kappa = 2) writes
  values equal to the oracle's assembly (1e-12) into the HashMatrix and is
  again solved without an upload; a script change through HashMatrix::pij is
  noticed (one upload) and the solution then matches the compiled
  ConjugueGradient on the changed HashMatrix.
"""
import numpy as np
import pytest

import ffo

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not ffo.ref_available(), reason="oracle/_ref not built")]
ffhip = pytest.importorskip("ffhip")


def _lower(rp, ci, va):
    n = rp.size - 1
    rows = np.repeat(np.arange(n), np.diff(rp))
    keep = ci <= rows
    rl = np.concatenate([[0], np.cumsum(np.bincount(rows[keep], minlength=n))]).astype(np.int64)
    return rl, ci[keep].copy(), va[keep].copy()


def _ref_cg(R, h, b, eps):
    x = np.zeros(b.size)
    st = np.zeros(4)
    res = R.ref_hm_cg(h, b, x, eps, 0, 1e30, 0, st)
    assert res == 1
    return x, int(st[0]), st[1]


@pytest.mark.parametrize("half", [0, 1])
def test_core_solve_vs_reference_cg(half):
    R = ffo.ref()
    _, rp, ci, va, _, b = ffo.poisson_p1_3d(7)
    n = rp.size - 1
    if half:
        h = R.seam_hm_from_csr(n, 1, *_lower(rp, ci, va))
    else:
        h = R.seam_hm_from_csr(n, 0, rp, ci, va)
    xr, itr, epsr = _ref_cg(R, h, b, 1e-10)
    for mode in (ffhip.CG_FP64, ffhip.CG_MIXED):
        x = np.zeros(n)
        st = np.zeros(5)
        assert R.seam_core_solve(h, mode, b, x, 1e-10, 0, 1e30, 0, st) == 0
        assert st[2] == 1 and st[4] == 1  # converged; one upload of the HashMatrix
        assert np.linalg.norm(x - xr) <= 1e-8 * np.linalg.norm(xr)
        if mode == ffhip.CG_FP64:
            assert abs(int(st[0]) - itr) <= 2
            assert abs(st[1] - epsr) <= 1e-12 * epsr
    R.ref_hm_free(h)


def test_resident_handover_reuse_and_script_change():
    R = ffo.ref()
    n, c1 = 8, 2.0
    irow = 4 + 9 * (4 + 9 * 4)  # the centre vertex: an interior row
    nv = (n + 1) ** 3
    b = np.zeros(nv)
    x1, x2, x3 = np.zeros(nv), np.zeros(nv), np.zeros(nv)
    s1, s2, s3 = np.zeros(5), np.zeros(5), np.zeros(5)
    m, rp, ci, vo, onbc, bo = ffo.poisson_p1_3d(n)
    hv2 = np.zeros(ci.size)
    rc = ffo.ctypes.c_int(0)
    h = R.seam_resident_flow(n, irow, c1, 1e-10, b, x1, s1, hv2, x2, s2, x3, s3, ffo.ctypes.byref(rc))
    assert rc.value == 0 and h
    assert np.max(np.abs(b - bo)) <= 1e-13 * np.max(np.abs(bo))
    # 1. the hand-over: no upload, the oracle's solution
    xo1, _ = ffo.pcg(rp, ci, vo, bo, eps=1e-10)
    assert s1[2] == 1 and s1[4] == 0
    assert np.linalg.norm(x1 - xo1) <= 1e-8 * np.linalg.norm(xo1)
    # 2. pattern kept, values of kappa = c1 in the HashMatrix, no upload
    _, _, _, vo2, _, _ = ffo.poisson_p1_3d(n, ffo.COEF_CONST, c1)
    big = vo2 >= 1e29
    assert np.array_equal(hv2[big], vo2[big])
    assert np.max(np.abs(hv2[~big] - vo2[~big])) <= 1e-12 * np.max(np.abs(vo2[~big]))
    xo2, _ = ffo.pcg(rp, ci, vo2, bo, eps=1e-10)
    assert s2[2] == 1 and s2[4] == 0
    assert np.linalg.norm(x2 - xo2) <= 1e-8 * np.linalg.norm(xo2)
    # 3. A(irow, irow) *= 2 by the script: one upload, the compiled CG's answer
    xr, itr, epsr = _ref_cg(R, h, b, 1e-10)
    assert s3[2] == 1 and s3[4] == 1
    assert abs(int(s3[0]) - itr) <= 2 and abs(s3[1] - epsr) <= 1e-12 * epsr
    assert np.linalg.norm(x3 - xr) <= 1e-8 * np.linalg.norm(xr)
    assert np.linalg.norm(x3 - x2) > 1e-6 * np.linalg.norm(x2)  # the change mattered
    R.ref_hm_free(h)
