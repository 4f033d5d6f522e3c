"""GPU parity of the P1 path through libffhip's C-ABI against the oracle.

Bars: integer mesh / pattern data bit-exact; matrix values within
max|A_gpu - A_ref| <= 1e-12 * max|A_ref| (off the tgv diagonal, which must be
exactly tgv); right-hand side within 1e-13 relative; CG solutions within 1e-8
relative of the oracle's FP64 solution at eps = 1e-10, iteration counts within
2 of the oracle's (summation order differs), and the FP64 criterion
(g, D^-1 g) < eps^2 (g0, D^-1 g0) met by both GPU modes.
"""
import numpy as np
import pytest

import ffo

pytestmark = pytest.mark.gpu

ffhip = pytest.importorskip("ffhip")
LABELS = (1, 2, 3, 4, 5, 6)


def _vals_close(va, vo, tol=1e-12):
    big = vo >= 1e29
    assert np.array_equal(va[big], vo[big])
    scale = np.max(np.abs(vo[~big]))
    return np.max(np.abs(va[~big] - vo[~big])) <= tol * scale


@pytest.mark.parametrize("dims", [(1, 1, 1), (2, 2, 2), (5, 5, 5), (3, 4, 5), (8, 8, 8)])
def test_cube_mesh_bitexact(dims):
    nx, ny, nz = dims
    m = ffhip.cube(nx, ny, nz, labels=(11, 12, 13, 14, 15, 16))
    d = m.download()
    o = ffo.cube(nx, ny, nz, labels=(11, 12, 13, 14, 15, 16))
    assert (d["nv"], d["nt"], d["nbe"]) == (o["nv"], o["nt"], o["nbe"])
    assert np.array_equal(d["xyz"], o["xyz"])
    assert np.array_equal(d["elems"], o["tets"])
    assert np.array_equal(d["belems"], o["btri"]) and np.array_equal(d["blab"], o["blab"])


@pytest.mark.parametrize("n", [7, 40])
def test_square_mesh_bitexact(n):
    m = ffhip.square(n, n + 3)
    d = m.download()
    o = ffo.square(n, n + 3)
    assert np.array_equal(d["xyz"], o["xy"]) and np.array_equal(d["elems"], o["tris"])
    assert np.array_equal(d["belems"], o["bedges"]) and np.array_equal(d["blab"], o["blab"])


@pytest.mark.parametrize("n", [1, 2, 6, 13])
def test_pattern_bitexact(n):
    m = ffhip.cube(n)
    A = ffhip.Matrix.p1(m)
    rp, ci, _ = A.download()
    o = ffo.cube(n)
    rpo, cio = ffo.pattern(o["nv"], o["tets"], 4)
    assert np.array_equal(rp, rpo) and np.array_equal(ci, cio)


@pytest.mark.parametrize("name", ["p1_cube4_sin3.npz", "p1_cube6_const.npz", "p1_cube5_p0.npz"])
def test_matrix_and_rhs_vs_golden(name):
    g = np.load(__file__.rsplit("/", 1)[0] + "/golden/" + name)
    n, kind = int(g["n"]), int(g["kind"])
    m = ffhip.cube(n)
    A = ffhip.Matrix.p1(m)
    A.assemble_p1(kind, 1.0, g["kappa_e"] if "kappa_e" in g.files else None)
    onbc = A.set_bc(LABELS, want_onbc=True)
    rp, ci, va = A.download()
    assert np.array_equal(rp, g["rowptr"]) and np.array_equal(ci, g["colidx"])
    assert np.array_equal(onbc, g["onbc"])
    assert _vals_close(va, g["vals"])
    b = ffhip.rhs_p1(m, 1.0, LABELS)
    assert np.max(np.abs(b - g["b"])) <= 1e-13 * np.max(np.abs(g["b"]))
    x, st = ffhip.CG(A).solve(b, eps=1e-10)
    assert abs(st["iters"] - int(g["iters"])) <= 2
    assert np.linalg.norm(x - g["x"]) <= 1e-8 * np.linalg.norm(g["x"])


@pytest.mark.parametrize("n,kind", [(10, ffo.COEF_SIN3), (9, ffo.COEF_CONST), (7, ffo.COEF_P0)])
def test_assembly_vs_oracle(n, kind):
    mo = ffo.cube(n)
    ke = np.exp(0.5 * np.random.default_rng(n).standard_normal(mo["nt"])) if kind == ffo.COEF_P0 else None
    m, rpo, cio, vo, onbco, bo = ffo.poisson_p1_3d(n, kind, 1.0, ke)
    M = ffhip.cube(n)
    A = ffhip.Matrix.p1(M)
    A.assemble_p1(kind, 1.0, ke)
    onbc = A.set_bc(LABELS, want_onbc=True)
    rp, ci, va = A.download()
    assert np.array_equal(rp, rpo) and np.array_equal(ci, cio) and np.array_equal(onbc, onbco)
    assert _vals_close(va, vo)
    b = ffhip.rhs_p1(M, 1.0, LABELS)
    assert np.max(np.abs(b - bo)) <= 1e-13 * np.max(np.abs(bo))


def test_kappa_on_device_matches_host():
    import torch
    n = 6
    M = ffhip.cube(n)
    ke = np.exp(0.5 * np.random.default_rng(1).standard_normal(M.nt))
    A1 = ffhip.Matrix.p1(M)
    A1.assemble_p1(ffhip.COEF_P0, 1.0, ke)
    A2 = ffhip.Matrix.p1(M)
    A2.assemble_p1(ffhip.COEF_P0, 1.0, torch.tensor(ke, device="cuda"))
    assert np.array_equal(A1.download()[2], A2.download()[2])


@pytest.mark.parametrize("layout", ["sell", "sell32", "csr"])
@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("n", [6, 20])
def test_cg_vs_oracle(mode, n, layout):
    m, rpo, cio, vo, onbco, bo = ffo.poisson_p1_3d(n)
    xo, so = ffo.pcg(rpo, cio, vo, bo, eps=1e-10)
    M = ffhip.cube(n)
    A = ffhip.Matrix.p1(M)
    A.assemble_p1(ffhip.COEF_SIN3)
    A.set_bc(LABELS)
    b = ffhip.rhs_p1(M, 1.0, LABELS)
    cg = ffhip.CG(A, mode, layout=layout)
    assert cg.layout()[0] == {"sell": "sell16", "sell32": "sell", "csr": "csr"}[layout]
    x, st = cg.solve(b, eps=1e-10)
    assert st["converged"] == 1
    assert st["true_gcg"] < 1e-20 * st["gcg0"] * (1.0 + 1e-6)
    assert abs(st["gcg0"] - so["gcg0"]) <= 1e-12 * so["gcg0"]
    if mode == 0:
        assert abs(st["iters"] - so["iters"]) <= 2
    else:
        assert st["iters"] <= 1.25 * so["iters"] + 5
    assert np.linalg.norm(x - xo) <= 1e-8 * np.linalg.norm(xo)


@pytest.mark.parametrize("layout", ["sell", "sell32", "csr"])
def test_cg_deterministic_and_warm_start(layout):
    n = 12
    M = ffhip.cube(n)
    A = ffhip.Matrix.p1(M)
    A.assemble_p1(ffhip.COEF_SIN3)
    A.set_bc(LABELS)
    b = ffhip.rhs_p1(M, 1.0, LABELS)
    cg = ffhip.CG(A, layout=layout)
    x1, s1 = cg.solve(b, eps=1e-10)
    x2, s2 = cg.solve(b, eps=1e-10)
    assert np.array_equal(x1, x2) and s1["iters"] == s2["iters"]
    # warm start (x is not reset, RNM.hpp:568-569) with an absolute criterion
    # (eps < 0, CG.cpp:226): from the solution the first iteration already
    # satisfies it, as in the oracle
    x3, s3 = cg.solve(b, x=x1, eps=-1e-9)
    rp, ci, va = A.download()
    xo, so = ffo.pcg(rp, ci, va, b, x0=x1, eps=-1e-9)
    assert s3["iters"] == so["iters"] == 1


def test_cg_nonconvergence_raises():
    n = 8
    M = ffhip.cube(n)
    A = ffhip.Matrix.p1(M)
    A.assemble_p1(ffhip.COEF_SIN3)
    A.set_bc(LABELS)
    b = ffhip.rhs_p1(M, 1.0, LABELS)
    with pytest.raises(ffhip.FFHipError) as e:
        ffhip.CG(A).solve(b, eps=1e-10, itmax=3)
    assert e.value.code == -3


def test_upload_path_and_spmv():
    import torch
    m, rpo, cio, vo, onbco, bo = ffo.poisson_p1_3d(7)
    A = ffhip.Matrix.upload(rpo, cio, vo)
    x = np.random.default_rng(0).standard_normal(rpo.size - 1)
    xd = torch.tensor(x, device="cuda")
    yd = torch.empty_like(xd)
    A.spmv(xd, yd)
    y = ffo.spmv(rpo, cio, vo, x)
    assert np.allclose(yd.cpu().numpy(), y, rtol=1e-13, atol=1e-13 * np.abs(y).max())
    A.spmv(xd, yd, trans=True)
    yt = ffo.spmv(rpo, cio, vo, x, trans=True)
    assert np.allclose(yd.cpu().numpy(), yt, rtol=1e-13, atol=1e-13 * np.abs(yt).max())
    x2, st = ffhip.CG(A).solve(bo, eps=1e-10)
    xo, so = ffo.pcg(rpo, cio, vo, bo, eps=1e-10)
    assert np.linalg.norm(x2 - xo) <= 1e-8 * np.linalg.norm(xo)


@pytest.mark.parametrize("layout", ["sell", "sell32", "csr"])
@pytest.mark.parametrize("mode", [0, 1])
def test_transposed_solve(mode, layout):
    mo = ffo.cube(6)
    ke = np.exp(0.5 * np.random.default_rng(3).standard_normal(mo["nt"]))
    m, rpo, cio, vo, onbco, bo = ffo.poisson_p1_3d(6, ffo.COEF_P0, 1.0, ke)
    A = ffhip.Matrix.upload(rpo, cio, vo)
    x, st = ffhip.CG(A, mode, layout=layout).solve(bo, eps=1e-10, trans=True)
    xo, so = ffo.pcg(rpo, cio, vo, bo, eps=1e-10, trans=True)
    assert np.linalg.norm(x - xo) <= 1e-8 * np.linalg.norm(xo)


def test_2d_square_poisson_vs_oracle():
    n = 30
    s = ffo.square(n)
    rpo, cio = ffo.pattern(s["nv"], s["tris"], 3)
    vo = ffo.assemble_p1_2d(s, rpo, cio)
    onbc = ffo.mark_bc(s["nv"], s["bedges"], s["blab"], 2, (1, 2, 3, 4))
    ffo.set_bc(rpo, cio, vo, onbc)
    bo = ffo.rhs_p1_2d(s, 1.0)
    ffo.rhs_bc(bo, onbc)
    M = ffhip.square(n)
    A = ffhip.Matrix.p1(M)
    A.assemble_p1(ffhip.COEF_CONST, 1.0)
    on2 = A.set_bc((1, 2, 3, 4), want_onbc=True)
    rp, ci, va = A.download()
    assert np.array_equal(rp, rpo) and np.array_equal(ci, cio) and np.array_equal(on2, onbc)
    assert _vals_close(va, vo)
    b = ffhip.rhs_p1(M, 1.0, (1, 2, 3, 4))
    assert np.max(np.abs(b - bo)) <= 1e-13 * np.max(np.abs(bo))
    xo, so = ffo.pcg(rpo, cio, vo, bo, eps=1e-10)
    x, st = ffhip.CG(A).solve(b, eps=1e-10)
    assert abs(st["iters"] - so["iters"]) <= 2
    assert np.linalg.norm(x - xo) <= 1e-8 * np.linalg.norm(xo)


def _cube_nnz(n):
    E = 3 * n * (n + 1) ** 2 + 3 * n * n * (n + 1) + n ** 3
    return (n + 1) ** 3 + 2 * E


@pytest.mark.parametrize("n", [64])
def test_full_size_properties(n):
    """cube(64), the 3D Poisson configuration: nnz closed form, zero row sums
    before Dirichlet (constants are in the kernel), symmetry, both CG modes
    reach the FP64 criterion."""
    import torch
    M = ffhip.cube(n)
    A = ffhip.Matrix.p1(M)
    assert A.nnz == _cube_nnz(n)
    A.assemble_p1(ffhip.COEF_SIN3)
    one = torch.ones(A.n, dtype=torch.float64, device="cuda")
    y = torch.empty_like(one)
    A.spmv(one, y)
    rp, ci, va = A.download()
    assert y.abs().max().item() <= 1e-12 * np.abs(va).max()
    yt = torch.empty_like(one)
    xr = torch.rand(A.n, dtype=torch.float64, device="cuda")
    A.spmv(xr, y)
    A.spmv(xr, yt, trans=True)
    assert (y - yt).abs().max().item() <= 1e-13 * y.abs().max().item()
    A.set_bc(LABELS)
    b = torch.empty(A.n, dtype=torch.float64, device="cuda")
    ffhip.rhs_p1(M, 1.0, LABELS, out=b)
    sols = []
    for mode in (0, 1):
        x, st = ffhip.CG(A, mode).solve(b, eps=1e-10)
        assert st["converged"] == 1 and st["true_gcg"] < 1.0001e-20 * st["gcg0"]
        sols.append(x)
    assert (sols[0] - sols[1]).norm().item() <= 1e-8 * sols[0].norm().item()


def test_square200_configuration():
    """square(200), the 2D Poisson configuration: FreeFEM's nnz (281,201),
    zero row sums before Dirichlet, the CG solve against the oracle's, and
    the discrete solution against the exact series value at the centre."""
    import torch
    n = 200
    M = ffhip.square(n)
    A = ffhip.Matrix.p1(M)
    assert A.nnz == 281201
    A.assemble_p1(ffhip.COEF_CONST, 1.0)
    one = torch.ones(A.n, dtype=torch.float64, device="cuda")
    y = torch.empty_like(one)
    A.spmv(one, y)
    rp, ci, va = A.download()
    assert y.abs().max().item() <= 1e-12 * np.abs(va).max()
    A.set_bc([1, 2, 3, 4])
    b = ffhip.rhs_p1(M, 1.0, [1, 2, 3, 4])
    x, st = ffhip.CG(A).solve(b, eps=1e-10)
    assert st["converged"] == 1
    s = ffo.square(n)
    rpo, cio = ffo.pattern(s["nv"], s["tris"], 3)
    vo = ffo.assemble_p1_2d(s, rpo, cio, ffo.COEF_CONST, 1.0)
    onb = ffo.mark_bc(s["nv"], s["bedges"], s["blab"], 2, (1, 2, 3, 4))
    ffo.set_bc(rpo, cio, vo, onb, 1e30)
    bo = ffo.rhs_p1_2d(s, 1.0)
    ffo.rhs_bc(bo, onb, 1e30, 0.0)
    xo, so = ffo.pcg(rpo, cio, vo, bo, eps=1e-10)
    assert abs(st["iters"] - so["iters"]) <= 2
    assert np.linalg.norm(x - xo) <= 1e-8 * np.linalg.norm(xo)
    # -Lap u = 1 on the unit square, u = 0 on the boundary: u(1/2,1/2) = 0.0736713532...
    centre = (n // 2) * (n + 1) + n // 2
    assert abs(x[centre] - 0.07367135328) < 2e-5


def test_qp_coefficient_matches_analytic():
    """kappa handed over at the 14 Tet_5 points (the plugin's path for any
    interpreted coefficient) gives the SIN3 matrix to rounding."""
    n = 7
    o = ffo.cube(n)
    p = np.zeros(42); w = np.zeros(14)
    ffo.lib().ffo_qf_tet5(p, w)
    P = p.reshape(14, 3)
    X = o["xyz"].reshape(-1, 3)[o["tets"].reshape(-1, 4)]            # nt x 4 x 3
    lam = np.concatenate([1 - P.sum(1, keepdims=True), P], axis=1)    # 14 x 4
    Q = np.einsum("qi,eik->eqk", lam, X)                              # nt x 14 x 3
    kq = 1 + 0.5 * np.sin(2 * np.pi * Q[..., 0]) * np.sin(2 * np.pi * Q[..., 1]) * np.sin(2 * np.pi * Q[..., 2])
    M = ffhip.cube(n)
    A1 = ffhip.Matrix.p1(M)
    A1.assemble_p1(ffhip.COEF_SIN3)
    A2 = ffhip.Matrix.p1(M)
    A2.assemble_p1(ffhip.COEF_QP, 0.0, np.ascontiguousarray(kq.ravel()))
    v1, v2 = A1.download()[2], A2.download()[2]
    assert np.max(np.abs(v1 - v2)) <= 1e-13 * np.max(np.abs(v1))


def test_set_values_and_update():
    m, rpo, cio, vo, onbco, bo = ffo.poisson_p1_3d(6)
    A = ffhip.Matrix.upload(rpo, cio, vo)
    cg = ffhip.CG(A)
    x1, _ = cg.solve(bo, eps=1e-10)
    A.set_values(2.0 * vo)   # tgv rows become 2e30, still the tgv of the matrix
    cg.update_values()
    x2, _ = cg.solve(bo, eps=1e-10)
    assert np.linalg.norm(2.0 * x2 - x1) <= 1e-8 * np.linalg.norm(x1)


def test_layout_choice_and_spmv_bench_both_layouts():
    """P1 rows have nearly equal length: AUTO streams sliced ELL with little
    padding; both layouts' plain mat-vec kernels give A x."""
    import torch
    n = 10
    M = ffhip.cube(n)
    A = ffhip.Matrix.p1(M)
    A.assemble_p1(ffhip.COEF_SIN3)
    rp, ci, va = A.download()
    nnz = int(rp[-1])
    cg = ffhip.CG(A, ffhip.CG_MIXED)
    lay, ents = cg.layout()
    assert lay == "sell16" and nnz <= ents <= 1.25 * nnz
    x = np.random.default_rng(1).standard_normal(rp.size - 1)
    y = ffo.spmv(rp, ci, va, x)
    for layout in ("sell", "sell32", "csr"):
        c = ffhip.CG(A, ffhip.CG_MIXED, layout=layout)
        xd = torch.tensor(x, device="cuda")
        yd = torch.empty_like(xd)
        c.bench_spmv(ffhip.CG_FP64, 2, xd, yd)
        assert np.allclose(yd.cpu().numpy(), y, rtol=1e-13, atol=1e-13 * np.abs(y).max())


def test_far_columns_use_32bit_layout():
    """Vertex ids relabelled at random (cube(32): 35,937 vertices): columns lie
    farther than 2^15 from their rows, so the solver may not stream 16-bit
    deltas; the solution is the relabelled solution of the cube mesh."""
    m = ffo.cube(32)
    rng = np.random.default_rng(4)
    perm = rng.permutation(m["nv"]).astype(np.int32)  # old -> new
    xyz = np.zeros_like(m["xyz"]).reshape(-1, 3)
    xyz[perm] = m["xyz"].reshape(-1, 3)
    tets = perm[m["tets"]].astype(np.int32)
    btri = perm[m["btri"]].astype(np.int32)
    outs = []
    for (X, T, B) in ((m["xyz"], m["tets"], m["btri"]), (xyz.ravel(), tets, btri)):
        M = ffhip.mesh3_upload(X, T, B, m["blab"])
        A = ffhip.Matrix.p1(M)
        A.assemble_p1(ffhip.COEF_SIN3)
        A.set_bc(LABELS)
        b = ffhip.rhs_p1(M, 1.0, LABELS)
        for mode in (ffhip.CG_MIXED, ffhip.CG_FP64):
            cg = ffhip.CG(A, mode)
            outs.append((cg.layout()[0], cg.solve(b, eps=1e-10)))
    # relabelled: int32 columns or CSR (a 16-bit layout tried first is filled
    # again with int32 columns once the fill finds a far column)
    assert outs[0][0] == "sell16" and outs[1][0] == "sell16"
    assert outs[2][0] != "sell16" and outs[3][0] != "sell16"
    for k in (0, 1):
        x0, x1 = outs[k][1][0], outs[2 + k][1][0]
        assert outs[2 + k][1][1]["converged"] == 1
        assert np.linalg.norm(x1[perm] - x0) <= 1e-8 * np.linalg.norm(x0)
    # FP64: the same iteration count on the relabelled mesh (the same system)
    assert abs(outs[3][1][1]["iters"] - outs[1][1][1]["iters"]) <= 1


def _fan_mesh(m=20):
    """A vertex with 2m incident tets (more than the 32 the half-wave kernels
    take), ring vertices with 4, apexes with m."""
    ang = 2 * np.pi * np.arange(m) / m
    xyz = np.vstack([[0.0, 0.0, 0.0], np.c_[np.cos(ang), np.sin(ang), np.zeros(m)],
                     [0.0, 0.0, 1.0], [0.0, 0.0, -1.0]])
    top, bot = m + 1, m + 2
    tets, tris = [], []
    for i in range(m):
        a, b = 1 + i, 1 + (i + 1) % m
        tets += [[0, a, b, top], [0, b, a, bot]]
        tris += [[a, b, top], [b, a, bot]]
    tets = np.array(tets, np.int32)
    X = xyz
    for t in tets:
        if np.linalg.det(np.c_[X[t[1]] - X[t[0]], X[t[2]] - X[t[0]], X[t[3]] - X[t[0]]]) < 0:
            t[1], t[2] = t[2], t[1]
    return dict(nv=len(xyz), nt=len(tets), nbe=len(tris), xyz=xyz.ravel().copy(),
                tets=tets.ravel().copy(), btri=np.array(tris, np.int32).ravel().copy(),
                blab=np.ones(len(tris), np.int32))


@pytest.mark.parametrize("ring", [20, 100, 253])
def test_high_degree_vertex_pattern_and_values(ring):
    """centre rows of 23 (lane-per-row merge, 64 LDS columns), 103 and 256
    columns (the wide merge, 256 columns: the byte offsets' limit)"""
    m = _fan_mesh(ring)
    rp, ci = ffo.pattern(m["nv"], m["tets"], 4)
    vo = ffo.assemble_p1_3d(m, rp, ci, ffo.COEF_CONST, 1.0, None)
    M = ffhip.mesh3_upload(m["xyz"], m["tets"], m["btri"], m["blab"])
    A = ffhip.Matrix.p1(M)
    A.assemble_p1(ffhip.COEF_CONST, 1.0)
    r2, c2, v2 = A.download()
    assert np.array_equal(r2, rp) and np.array_equal(c2, ci)
    assert np.abs(v2 - vo).max() <= 1e-13 * np.abs(vo).max()
    assert rp[1] - rp[0] == m["nv"] == ring + 3  # the centre couples to every vertex


def test_row_wider_than_256_columns_is_refused_without_leak():
    """a centre row of 303 columns: FFHIP_ERR_UNSUPPORTED (the element-to-row
    map stores byte offsets), and the failed call frees what it allocated:
    repeating it does not lower the device's free memory"""
    import torch
    m = _fan_mesh(300)
    M = ffhip.mesh3_upload(m["xyz"], m["tets"], m["btri"], m["blab"])
    free = []
    for _ in range(4):
        with pytest.raises(ffhip.FFHipError) as ei:
            ffhip.Matrix.p1(M)
        assert ei.value.code == ffhip.ERR_UNSUPPORTED
        torch.cuda.synchronize()
        free.append(torch.cuda.mem_get_info()[0])
    assert free[-1] >= free[1]


@pytest.mark.parametrize("layout", ["sell", "csr"])
@pytest.mark.parametrize("mode", [0, 1])
def test_graph_replay_bitwise_equals_eager(mode, layout):
    # the solve loops replay one captured chunk of 16 iterations, with the
    # host's polls (and MIXED's restarts, which alternate the chunk parity)
    # between replays; later solves replay it again.  The same kernels in the
    # same order with order-fixed reductions: the bits must equal the
    # kernel-by-kernel launches'.
    m, rpo, cio, vo, onbco, bo = ffo.poisson_p1_3d(8)
    A = ffhip.Matrix.upload(rpo, cio, vo)
    cg_g = ffhip.CG(A, mode, layout=layout)
    cg_e = ffhip.CG(A, mode, layout=layout, graph=False)
    rng = np.random.default_rng(7)
    for k in range(3):
        b = bo * (1 + k) + (k > 0) * rng.standard_normal(bo.size) * (onbco == 0)
        for trans in (False, True):
            xg, sg = cg_g.solve(b, eps=1e-10, trans=trans)
            xe, se = cg_e.solve(b, eps=1e-10, trans=trans)
            assert sg["converged"] == 1 and sg["iters"] > 16
            assert np.array_equal(xg, xe)
            assert sg["iters"] == se["iters"] and sg["refinements"] == se["refinements"]
    # profiling launches kernel by kernel with timing events: same bits
    cg_g.set_profiling(True)
    xp, sp = cg_g.solve(bo, eps=1e-10)
    xe, se = cg_e.solve(bo, eps=1e-10)
    assert np.array_equal(xp, xe)
    assert abs(sp["timed"] - sp["iters"]) <= 1 and sp["ms_matvec_avg"] > 0 and sp["ms_update_avg"] > 0


@pytest.mark.parametrize("kind", ["sin3", "const", "p0", "qp", "p1"])
def test_pattern_assemble_one_call_equals_two_steps(kind):
    """ffhip_pattern_assemble_p1 (pattern and values in one call, the values
    formed in the pattern pass) gives the matrix of ffhip_pattern_p1 +
    ffhip_assemble_p1 bit for bit, and the oracle's within 1e-12"""
    n = 9
    M = ffhip.cube(n)
    rng = np.random.default_rng(3)
    arrs = {"sin3": (ffhip.COEF_SIN3, 1.0, None), "const": (ffhip.COEF_CONST, 2.5, None),
            "p0": (ffhip.COEF_P0, 0.0, np.exp(rng.standard_normal(M.nt))),
            "qp": (ffhip.COEF_QP, 0.0, np.exp(rng.standard_normal(14 * M.nt))),
            "p1": (ffhip.COEF_P1, 0.0, np.exp(rng.standard_normal(M.nv)))}
    k, c0, arr = arrs[kind]
    A1 = ffhip.Matrix.p1_assembled(M, k, c0, arr)
    A2 = ffhip.Matrix.p1(M)
    A2.assemble_p1(k, c0, arr)
    r1, c1, v1 = A1.download()
    r2, c2, v2 = A2.download()
    assert np.array_equal(r1, r2) and np.array_equal(c1, c2)
    assert np.array_equal(v1.view(np.int64), v2.view(np.int64))
    if kind in ("sin3", "const", "p0"):
        mo = ffo.cube(n)
        rpo, cio = ffo.pattern(mo["nv"], mo["tets"], 4)
        vo = ffo.assemble_p1_3d(mo, rpo, cio, {"sin3": ffo.COEF_SIN3, "const": ffo.COEF_CONST,
                                               "p0": ffo.COEF_P0}[kind], c0, arr)
        assert np.array_equal(r1, rpo) and np.array_equal(c1, cio)
        assert np.max(np.abs(v1 - vo)) <= 1e-12 * np.max(np.abs(vo))


@pytest.mark.parametrize("case", ["star", "star12_centre_last", "no_boundary_triangles"])
def test_pattern_assemble_one_call_fallback_sorts_incidences(case):
    """one-call assembly on meshes whose incidence lists all fit the 32-entry
    buckets (the pattern pass is to sort them into element order) but whose
    direct pass is refused: a vertex of 30 tets that share only it (a row of
    91 columns), a vertex of 12 such tets numbered last (37 columns, the
    diagonal at offset 36: the row's values past the first 32 slots), or a
    cube without its boundary triangles (row lengths unlike the vertex-link
    counts).  The matrix equals the two-step one, and a right-hand side
    formed afterwards from the incidence lists equals that of a fresh mesh,
    bit for bit (jittered coordinates and a P1 coefficient make the element
    order of every sum visible)."""
    rng = np.random.default_rng(5)
    if case.startswith("star"):
        k = 30 if case == "star" else 12
        pts, tets = [np.zeros(3)], []
        for _ in range(k):
            d = rng.standard_normal(3)
            q = 2.0 * d / np.linalg.norm(d) + 0.3 * rng.standard_normal((3, 3))
            i0 = len(pts)
            pts += list(q)
            tets.append([0, i0, i0 + 1, i0 + 2])
        X = np.array(pts)
        if case != "star":  # the centre numbered last
            perm = np.r_[1:len(X), 0]          # new vertex i is old vertex perm[i]
            inv = np.empty_like(perm)
            inv[perm] = np.arange(len(X))
            X = X[perm]
            tets = [[int(inv[a]) for a in t] for t in tets]
        for t in tets:
            if np.linalg.det(np.c_[X[t[1]] - X[t[0]], X[t[2]] - X[t[0]], X[t[3]] - X[t[0]]]) < 0:
                t[1], t[2] = t[2], t[1]
        xyz = X.ravel()
        tets = np.array(tets, np.int32).ravel()
    else:
        mo = ffo.cube(6)
        X = mo["xyz"].reshape(-1, 3).copy()
        inner = np.all((X > 1e-12) & (X < 1 - 1e-12), axis=1)
        X[inner] += 0.02 * rng.standard_normal((int(inner.sum()), 3))
        xyz, tets = X.ravel(), mo["tets"]
    none3, none1 = np.zeros(0, np.int32), np.zeros(0, np.int32)
    M1 = ffhip.mesh3_upload(xyz, tets, none3, none1)
    M2 = ffhip.mesh3_upload(xyz, tets, none3, none1)
    A1 = ffhip.Matrix.p1_assembled(M1, ffhip.COEF_SIN3)
    A2 = ffhip.Matrix.p1(M2)
    A2.assemble_p1(ffhip.COEF_SIN3)
    d1, d2 = A1.download(), A2.download()
    assert np.array_equal(d1[0], d2[0]) and np.array_equal(d1[1], d2[1])
    assert np.array_equal(d1[2].view(np.int64), d2[2].view(np.int64))
    f = np.exp(rng.standard_normal(M1.nv))
    b1 = ffhip.rhs_p1_form(M1, ffhip.COEF_P1, 0.0, f)
    b2 = ffhip.rhs_p1_form(M2, ffhip.COEF_P1, 0.0, f)
    assert np.array_equal(b1.view(np.int64), b2.view(np.int64))


@pytest.mark.parametrize("kind", ["sin3", "p0"])
def test_pattern_assemble_one_call_twice_on_one_mesh(kind):
    """the first one-call assembly on a mesh builds its incidence lists and
    the element scales in one pass; a second one on the same mesh (lists
    kept) computes the scales on their own: both matrices bit for bit equal,
    and equal to the two-step matrix"""
    n = 7
    M = ffhip.cube(n)
    rng = np.random.default_rng(11)
    k, c0, arr = ((ffhip.COEF_SIN3, 1.0, None) if kind == "sin3"
                  else (ffhip.COEF_P0, 0.0, np.exp(rng.standard_normal(M.nt))))
    A1 = ffhip.Matrix.p1_assembled(M, k, c0, arr)
    A2 = ffhip.Matrix.p1_assembled(M, k, c0, arr)
    A3 = ffhip.Matrix.p1(M)
    A3.assemble_p1(k, c0, arr)
    d1, d2, d3 = A1.download(), A2.download(), A3.download()
    for a, b in ((d1, d2), (d1, d3)):
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
        assert np.array_equal(a[2].view(np.int64), b[2].view(np.int64))


def test_pattern_assemble_one_call_with_lane_per_row_rows():
    """a vertex of 103 columns: the one-call path falls back to the element-row
    values for every row; the matrix equals the two-step one"""
    m = _fan_mesh(100)
    M = ffhip.mesh3_upload(m["xyz"], m["tets"], m["btri"], m["blab"])
    A1 = ffhip.Matrix.p1_assembled(M, ffhip.COEF_SIN3)
    A2 = ffhip.Matrix.p1(M)
    A2.assemble_p1(ffhip.COEF_SIN3)
    d1, d2 = A1.download(), A2.download()
    assert all(np.array_equal(a, b) for a, b in zip(d1, d2))
