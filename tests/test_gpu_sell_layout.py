"""The solver's sliced-ELL layout with 16-bit column deltas on rows of every
width: FP64 values stored in entry pairs (plus one plane of single values for
an odd slice width), rows longer than the 16-entry unrolled round (the tail
loop), slices narrower than 16, and a last partial slice.  A symmetric,
strictly diagonally dominant matrix with row lengths 1 to about 40 and columns within
3000 of their row: the FP64 and FP32 mat-vecs of the sell16 layout against the
oracle's CSR product (HashMatrix::addMatMul order, HashMatrix.cpp:1086-1154),
and FP64 / MIXED CG solves (plain and transposed) against the oracle's PCG
(CG.cpp:195-265)."""
import numpy as np
import pytest

import ffo

pytestmark = pytest.mark.gpu
ffhip = pytest.importorskip("ffhip")


def _banded_spd(n, seed):
    rng = np.random.default_rng(seed)
    rows = [set([i]) for i in range(n)]
    narrow = 3 * 64  # three slices of short rows (slice widths below 16)
    for i in range(n):
        if i < narrow:
            lo, hi, want = 0, narrow - 1, int(rng.integers(0, 3))
        else:  # about 1..41 entries per row once mirrored
            lo, hi, want = max(narrow, i - 3000), min(n - 1, i + 3000), int(rng.integers(0, 21))
        for _ in range(want):
            j = int(rng.integers(lo, hi + 1))
            rows[i].add(j)
            rows[j].add(i)
    rp = np.zeros(n + 1, np.int64)
    for i in range(n):
        rp[i + 1] = rp[i] + len(rows[i])
    ci = np.concatenate([np.array(sorted(r), np.int32) for r in rows])
    va = np.zeros(ci.size)
    off = {}
    for i in range(n):
        for k in range(rp[i], rp[i + 1]):
            j = int(ci[k])
            if j != i:
                key = (min(i, j), max(i, j))
                if key not in off:
                    off[key] = -rng.uniform(0.1, 1.0)
                va[k] = off[key]
    for i in range(n):
        seg = slice(rp[i], rp[i + 1])
        d = np.flatnonzero(ci[seg] == i)[0] + rp[i]
        va[d] = np.abs(va[seg]).sum() + 1.0 + rng.uniform(0, 1)
    return rp, ci, va


@pytest.fixture(scope="module")
def system():
    n = 64 * 47 + 29  # a last partial slice
    rp, ci, va = _banded_spd(n, 11)
    widths = np.diff(rp)
    slice_w = [widths[s:s + 64].max() for s in range(0, n, 64)]
    assert widths.min() == 1 and max(slice_w[:3]) < 16
    assert {w % 2 for w in slice_w} == {0, 1}  # odd and even slice widths
    assert widths.max() > 32  # two tail rounds past the 16-entry round
    return rp, ci, va


def test_sell16_matvec_fp64_fp32_every_width(system):
    import torch
    rp, ci, va = system
    A = ffhip.Matrix.upload(rp, ci, va)
    cg = ffhip.CG(A, ffhip.CG_MIXED, layout="sell")
    assert cg.layout()[0] == "sell16"
    x = np.random.default_rng(2).standard_normal(rp.size - 1)
    y = ffo.spmv(rp, ci, va, x)
    xd = torch.tensor(x, device="cuda")
    yd = torch.empty_like(xd)
    cg.bench_spmv(ffhip.CG_FP64, 2, xd, yd)
    assert np.allclose(yd.cpu().numpy(), y, rtol=1e-13, atol=1e-13 * np.abs(y).max())
    # FP32: D^-1/2 A D^-1/2 in single precision
    d = va[[np.flatnonzero(ci[rp[i]:rp[i + 1]] == i)[0] + rp[i] for i in range(rp.size - 1)]]
    s = 1.0 / np.sqrt(d)
    y32 = s * ffo.spmv(rp, ci, va, s * x)
    x32 = torch.tensor(x, device="cuda", dtype=torch.float32)
    y32d = torch.empty_like(x32)
    cg.bench_spmv(ffhip.CG_MIXED, 2, x32, y32d)
    assert np.allclose(y32d.cpu().numpy(), y32, rtol=2e-5, atol=2e-5 * np.abs(y32).max())
    cg.free()


@pytest.mark.parametrize("trans", [False, True])
@pytest.mark.parametrize("mode", [0, 1])
def test_sell16_cg_every_width(system, mode, trans):
    rp, ci, va = system
    b = np.random.default_rng(3).standard_normal(rp.size - 1)
    A = ffhip.Matrix.upload(rp, ci, va)
    cg = ffhip.CG(A, mode, layout="sell")
    x, st = cg.solve(b, eps=1e-10, trans=trans)
    xo, so = ffo.pcg(rp, ci, va, b, eps=1e-10, trans=trans)
    assert st["converged"] == 1
    if mode == ffhip.CG_FP64:
        assert abs(st["iters"] - so["iters"]) <= 1
    assert np.linalg.norm(x - xo) <= 1e-8 * np.linalg.norm(xo)
    cg.free()
