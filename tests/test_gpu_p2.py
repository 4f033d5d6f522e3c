"""GPU parity of the vector-P2 Lame elasticity path against the oracle.

Bars: node numbering, pattern, Dirichlet dofs: bit-exact; element-order
assembled values within 1e-12 of the largest entry (the element matrices
are summed in a different order on the matrix cores); right-hand side
within 1e-14 relative; CG solutions within 1e-8 (relative 2-norm) of the
oracle's ConjugueGradient.  At the cube(32) configuration (823,875 dofs,
69,108,489 nonzeros): FreeFEM's pattern size, the CSR layout, a converged
mixed-precision solve that agrees with the FP64 solve, and symmetry of A.
"""
import numpy as np
import pytest

import ffo

pytestmark = pytest.mark.gpu

ffhip = pytest.importorskip("ffhip")


def _shuffled_cube(n, seed):
    """cube(n) with vertices relabelled, elements reordered and each element's
    vertices rotated by an even permutation (orientation kept)."""
    m = ffo.cube(n)
    rng = np.random.default_rng(seed)
    nv, nt = m["nv"], m["nt"]
    perm = rng.permutation(nv).astype(np.int32)  # old -> new
    xyz = np.zeros_like(m["xyz"]).reshape(-1, 3)
    xyz[perm] = m["xyz"].reshape(-1, 3)
    tets = perm[m["tets"].reshape(-1, 4)][rng.permutation(nt)]
    even = np.array([[0, 1, 2, 3], [1, 2, 0, 3], [2, 0, 1, 3], [0, 2, 3, 1], [1, 0, 3, 2],
                     [3, 1, 0, 2]], np.int64)
    # only permutations with positive sign: check them once
    sign = [np.linalg.det(np.eye(4)[p]) for p in even]
    even = even[np.array(sign) > 0]
    pick = even[rng.integers(0, len(even), nt)]
    tets = np.take_along_axis(tets, pick, axis=1).astype(np.int32)
    btri = perm[m["btri"].reshape(-1, 3)].astype(np.int32)
    return dict(nv=nv, nt=nt, nbe=m["nbe"], xyz=xyz.ravel(), tets=tets.ravel(), btri=btri.ravel(),
                blab=m["blab"].copy())


def _gpu_mesh(m):
    return ffhip.mesh3_upload(m["xyz"], m["tets"], m["btri"], m["blab"])


@pytest.mark.parametrize("case", ["cube1", "cube2", "cube5", "shuffled4"])
def test_p2_numbering_bitexact(case):
    m = _shuffled_cube(4, 3) if case == "shuffled4" else ffo.cube(int(case[4:]))
    nn, en = ffo.p2_nodes(m)
    sp = ffhip.P2VSpace(_gpu_mesh(m))
    assert sp.nnodes == nn and sp.ndof == 3 * nn
    assert np.array_equal(sp.elem_nodes(), en)


@pytest.mark.parametrize("case", ["cube3", "shuffled3"])
def test_p2_system_vs_oracle(case):
    m = _shuffled_cube(3, 9) if case == "shuffled3" else ffo.cube(3)
    o = ffo.elasticity_p2(3, mesh=m)
    sp = ffhip.P2VSpace(_gpu_mesh(m))
    A = ffhip.Matrix.p2v(sp)
    assert A.n == o["ndof"] and A.nnz == o["colidx"].size
    A.assemble_p2v_elasticity(1.0, 0.5)
    onbc = A.set_bc_p2v([4], want_onbc=True)
    rp, ci, va = A.download()
    assert np.array_equal(rp, o["rowptr"]) and np.array_equal(ci, o["colidx"])
    assert np.array_equal(onbc, o["onbc"])
    free = va < 1e29
    assert np.array_equal(free, o["vals"] < 1e29)
    scale = np.abs(o["vals"][free]).max()
    assert np.abs(va[free] - o["vals"][free]).max() <= 1e-12 * scale
    b = ffhip.rhs_p2v(sp, (0.0, 0.0, -1.0), [4])
    ob = o["b"]
    assert np.abs(b - ob).max() <= 1e-14 * np.abs(ob).max()
    xo, so = ffo.pcg(o["rowptr"], o["colidx"], o["vals"], ob, eps=1e-10)
    for layout in ("bsr3", "csr"):
        for mode in (ffhip.CG_FP64, ffhip.CG_MIXED):
            cg = ffhip.CG(A, mode, layout=layout)
            assert cg.layout()[0] == layout
            x, st = cg.solve(b, eps=1e-10)
            assert st["converged"] == 1
            assert np.linalg.norm(x - xo) <= 1e-8 * np.linalg.norm(xo)
            if mode == ffhip.CG_FP64:
                assert abs(st["iters"] - so["iters"]) <= 2
        # transposed solve (A is symmetric: the same solution)
        xt, stt = ffhip.CG(A, ffhip.CG_FP64, layout=layout).solve(b, eps=1e-10, trans=True)
        assert np.linalg.norm(xt - xo) <= 1e-8 * np.linalg.norm(xo)


def test_p2_layout_is_blocks_for_vector_p2():
    import torch
    m = ffo.cube(4)
    sp = ffhip.P2VSpace(_gpu_mesh(m))
    A = ffhip.Matrix.p2v(sp)
    A.assemble_p2v_elasticity()
    A.set_bc_p2v([4])
    lay, ents = ffhip.CG(A).layout()
    assert lay == "bsr3" and A.nnz <= ents <= 1.5 * A.nnz  # node rows sorted by length
    # the block mat-vec kernels (FP64, and FP32 of D^-1/2 A D^-1/2) against the CSR product
    rp, ci, va = A.download()
    x = np.random.default_rng(3).standard_normal(A.n)
    y = ffo.spmv(rp, ci, va, x)
    cg = ffhip.CG(A, ffhip.CG_MIXED, layout="bsr3")
    xd = torch.tensor(x, device="cuda")
    yd = torch.empty_like(xd)
    cg.bench_spmv(ffhip.CG_FP64, 2, xd, yd)
    assert np.allclose(yd.cpu().numpy(), y, rtol=1e-13, atol=1e-13 * np.abs(y).max())


def test_p2_graph_replay_bitwise_equals_eager():
    m = ffo.cube(4)
    sp = ffhip.P2VSpace(_gpu_mesh(m))
    A = ffhip.Matrix.p2v(sp)
    A.assemble_p2v_elasticity()
    A.set_bc_p2v([4])
    b = ffhip.rhs_p2v(sp, (0.0, 0.0, -1.0), [4])
    for mode in (ffhip.CG_FP64, ffhip.CG_MIXED):
        xg, sg = ffhip.CG(A, mode, layout="bsr3").solve(b, eps=1e-10)
        xe, se = ffhip.CG(A, mode, layout="bsr3", graph=False).solve(b, eps=1e-10)
        assert sg["converged"] == 1 and np.array_equal(xg, xe) and sg["iters"] == se["iters"]


def test_p2_cube32_configuration():
    """configuration: cube(32), [P2,P2,P2], lambda=1, mu=0.5, f=(0,0,-1),
    clamped label 4, FP32 inner CG with FP64 refinement"""
    import torch
    M = ffhip.cube(32)
    sp = ffhip.P2VSpace(M)
    assert sp.ndof == 823875
    A = ffhip.Matrix.p2v(sp)
    assert A.nnz == 69108489  # FreeFEM's measured nnz for this matrix
    A.assemble_p2v_elasticity(1.0, 0.5)
    # symmetry of the assembled matrix (before the tgv rows): y.Ax == x.Ay
    x = torch.randn(sp.ndof, dtype=torch.float64, device="cuda")
    y = torch.randn(sp.ndof, dtype=torch.float64, device="cuda")
    Ax = torch.empty_like(x)
    Ay = torch.empty_like(x)
    A.spmv(x, Ax)
    A.spmv(y, Ay)
    s1 = torch.dot(y, Ax).item()
    s2 = torch.dot(x, Ay).item()
    assert abs(s1 - s2) <= 1e-12 * (abs(s1) + abs(s2))
    A.set_bc_p2v([4])
    b = torch.empty(sp.ndof, dtype=torch.float64, device="cuda")
    ffhip.rhs_p2v(sp, (0.0, 0.0, -1.0), [4], out=b)
    cgm = ffhip.CG(A, ffhip.CG_MIXED)
    lay, ents = cgm.layout()
    assert lay == "bsr3" and ents <= 1.1 * A.nnz  # 3x3 blocks, node rows sorted by length
    xm, sm = cgm.solve(b, eps=1e-10)
    x64, s64 = ffhip.CG(A, ffhip.CG_FP64).solve(b, eps=1e-10)
    assert sm["converged"] == 1 and s64["converged"] == 1
    assert torch.linalg.norm(xm - x64).item() <= 1e-7 * torch.linalg.norm(x64).item()
    uz = x64[2::3]
    assert uz.mean().item() < 0  # the load pulls the clamped cube down
    print("cube32 P2: fp64 iters", s64["iters"], "ms", s64["ms_solve"], "mixed inner iters",
          sm["iters"], "ms", sm["ms_solve"], "refinements", sm["refinements"])
