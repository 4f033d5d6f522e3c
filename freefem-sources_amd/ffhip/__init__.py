"""Python host mirror of the ffhip FreeFEM plugin, over libffhip's C-ABI.

The production entry point is the FreeFEM plugin (plugin/ffhip.cpp,
`load "ffhip"`); this module mirrors the same operations for tests and
benchmarks:

    cube(n)                     -> mesh3 cube(n,n,n)        (msh3.cpp:7879)
    Matrix.p1(mesh)             -> HashMatrix pattern         (HashMatrix.cpp:858)
    Matrix.assemble_p1(...)     -> matrix A = va(Vh,Vh)       (problem.cpp:1116)
    Matrix.set_bc(labels)       -> + on(labels, u=0)          (problem.cpp:10039)
    rhs_p1(mesh, f, labels)     -> b = vb(0,Vh)               (problem.cpp:10878)
    CG(A).solve(b, x, eps)      -> u[] = A^-1*b, solver=CG    (VirtualSolverCG.hpp:169)

Errors raise FFHipError with the library's message (the plugin maps the same
codes to ExecError).  There is no CPU fallback: if libffhip.so is missing or
no GPU is present, every call raises.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.environ.get("FFHIP_LIB") or os.path.join(os.path.dirname(HERE), "libffhip.so")

COEF_NONE, COEF_CONST, COEF_SIN3, COEF_P0, COEF_QP, COEF_P1 = -1, 0, 1, 2, 3, 4
CG_FP64, CG_MIXED = 0, 1
CG_LAYOUT_AUTO, CG_LAYOUT_SELL, CG_LAYOUT_CSR, CG_LAYOUT_BSR3 = 0, 0x10, 0x20, 0x100
CG_INDEX16, CG_INDEX32 = 0x40, 0x80
CG_EAGER = 0x200
CG2_IDENTITY, CG2_JACOBI = 0, 1
ERR_ARG, ERR_HIP, ERR_NOCONV, ERR_NAN, ERR_UNSUPPORTED, ERR_INTERNAL = -1, -2, -3, -4, -5, -6
ERRORS = {-1: "bad argument", -2: "HIP error", -3: "CG did not converge", -4: "NaN in CG",
          -5: "unsupported", -6: "internal error"}


class FFHipError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"ffhip error {code} ({ERRORS.get(code, '?')}): {msg}")
        self.code = code


class CGStats(ctypes.Structure):
    _fields_ = [("iters", ctypes.c_int), ("converged", ctypes.c_int),
                ("refinements", ctypes.c_int), ("reserved", ctypes.c_int),
                ("gcg0", ctypes.c_double), ("gcg", ctypes.c_double),
                ("true_gcg", ctypes.c_double), ("true_relres", ctypes.c_double),
                ("ms_solve", ctypes.c_double), ("ms_matvec_avg", ctypes.c_double),
                ("ms_update_avg", ctypes.c_double), ("timed", ctypes.c_int),
                ("reserved2", ctypes.c_int), ("ms_direction_avg", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if not k.startswith("reserved")}


_vp = ctypes.c_void_p
_i = ctypes.c_int
_l = ctypes.c_int64
_d = ctypes.c_double
_pvp = ctypes.POINTER(ctypes.c_void_p)
_pl = ctypes.POINTER(ctypes.c_int64)

_SIGS = {
    "ffhip_last_error": (ctypes.c_char_p, []),
    "ffhip_version": (ctypes.c_char_p, []),
    "ffhip_init": (_i, [_i, _pvp]),
    "ffhip_ctx_free": (_i, [_vp]),
    "ffhip_sync": (_i, [_vp]),
    "ffhip_mesh3_cube": (_i, [_vp, _i, _i, _i, _vp, _pvp]),
    "ffhip_mesh2_square": (_i, [_vp, _i, _i, _vp, _pvp]),
    "ffhip_mesh3_upload": (_i, [_vp, _l, _l, _l, _vp, _vp, _vp, _vp, _pvp]),
    "ffhip_mesh2_upload": (_i, [_vp, _l, _l, _l, _vp, _vp, _vp, _vp, _pvp]),
    "ffhip_mesh_info": (_i, [_vp, _pl, _pl, _pl]),
    "ffhip_mesh_download": (_i, [_vp, _vp, _vp, _vp, _vp]),
    "ffhip_mesh_free": (_i, [_vp]),
    "ffhip_pattern_p1": (_i, [_vp, _vp, _pvp]),
    "ffhip_mat_info": (_i, [_vp, _pl, _pl]),
    "ffhip_assemble_p1": (_i, [_vp, _vp, _vp, _i, _d, _vp, _i]),
    "ffhip_assemble_p1_form": (_i, [_vp, _vp, _vp, _i, _d, _vp, _i, _d, _vp, _i]),
    "ffhip_set_bc": (_i, [_vp, _vp, _vp, _vp, _i, _d, _vp]),
    "ffhip_rhs_p1": (_i, [_vp, _vp, _d, _vp, _i, _d, _d, _vp, _i]),
    "ffhip_rhs_p1_form": (_i, [_vp, _vp, _i, _d, _vp, _vp, _i, _d, _d, _vp, _vp, _i]),
    "ffhip_assemble_p1_boundary": (_i, [_vp, _vp, _vp, _vp, _i, _i, _d, _vp, _i]),
    "ffhip_mat_coo": (_i, [_vp, _pl, _vp, _vp, _vp, _i]),
    "ffhip_mat_lower": (_i, [_vp, _pvp]),
    "ffhip_mat_set_half": (_i, [_vp, _i]),
    "ffhip_mat_full": (_i, [_vp, _pvp]),
    "ffhip_mat_half": (_i, [_vp, ctypes.POINTER(_i)]),
    "ffhip_csr_write_morse": (_i, [_l, _l, _i, _vp, _vp, _vp, ctypes.c_char_p]),
    "ffhip_mat_write_morse": (_i, [_vp, ctypes.c_char_p]),
    "ffhip_csr_read_text": (_i, [ctypes.c_char_p, _l, _pl, _pl, ctypes.POINTER(_i), _pl, _vp, _vp,
                                 _vp]),
    "ffhip_mat_read_text": (_i, [_vp, ctypes.c_char_p, _pvp]),
    "ffhip_rhs_p1_boundary": (_i, [_vp, _vp, _vp, _i, _i, _d, _vp, _vp, _i, _vp, _i]),
    "ffhip_mat_download": (_i, [_vp, _vp, _vp, _vp]),
    "ffhip_mat_upload": (_i, [_vp, _l, _l, _vp, _vp, _vp, _pvp]),
    "ffhip_mat_device_ptrs": (_i, [_vp, _pvp, _pvp, _pvp]),
    "ffhip_mat_set_values": (_i, [_vp, _vp]),
    "ffhip_mat_free": (_i, [_vp]),
    "ffhip_spmv": (_i, [_vp, _vp, _vp, _vp, _i]),
    "ffhip_cg_create": (_i, [_vp, _vp, _i, _pvp]),
    "ffhip_cg_update_values": (_i, [_vp]),
    "ffhip_cg_layout": (_i, [_vp, ctypes.POINTER(_i), _pl]),
    "ffhip_cg_solve": (_i, [_vp, _vp, _vp, _i, _d, _i, _d, _i, ctypes.POINTER(CGStats)]),
    "ffhip_cg_solve_affine": (_i, [_vp, _vp, _vp, _i, _d, _i, _i, ctypes.POINTER(_d),
                                   ctypes.POINTER(CGStats)]),
    "ffhip_gmres_solve": (_i, [_vp, _vp, _vp, _i, _d, _i, _i, _d, _i, ctypes.POINTER(CGStats)]),
    "ffhip_cg_free": (_i, [_vp]),
    "ffhip_cg_set_profiling": (_i, [_vp, _i]),
    "ffhip_bench_spmv": (_i, [_vp, _i, _i, _vp, _vp, ctypes.POINTER(_d)]),
    "ffhip_last_ms": (_d, [_vp]),
    "ffhip_bench_copy": (_i, [_vp, _l, _i, ctypes.POINTER(_d)]),
    "ffhip_p1_misfit": (_i, [_vp, _vp, _vp, _vp, _vp, _i, ctypes.POINTER(_d), _vp]),
    "ffhip_pattern_assemble_p1": (_i, [_vp, _vp, _i, _d, _vp, _i, _pvp]),
    "ffhip_p1_misfit_noalloc": (_i, [_vp, _vp, _vp, _vp, _vp, _i, ctypes.POINTER(_d), _vp]),
    "ffhip_p1_grad_kappa": (_i, [_vp, _vp, _vp, _vp, _vp]),
    "ffhip_p2v_create": (_i, [_vp, _vp, _pvp]),
    "ffhip_p2v_info": (_i, [_vp, _pl, _pl]),
    "ffhip_p2v_download": (_i, [_vp, _vp]),
    "ffhip_p2v_free": (_i, [_vp]),
    "ffhip_pattern_p2v": (_i, [_vp, _vp, _pvp]),
    "ffhip_assemble_p2v_elasticity": (_i, [_vp, _vp, _vp, _d, _d]),
    "ffhip_set_bc_p2v": (_i, [_vp, _vp, _vp, _vp, _i, _d, _vp]),
    "ffhip_rhs_p2v": (_i, [_vp, _vp, _vp, _vp, _i, _d, _vp, _vp, _i]),
    "ffhip_pattern_p2": (_i, [_vp, _vp, _pvp]),
    "ffhip_assemble_p2": (_i, [_vp, _vp, _vp, _i, _d, _vp, _i]),
    "ffhip_set_bc_p2": (_i, [_vp, _vp, _vp, _vp, _i, _d, _vp]),
    "ffhip_rhs_p2": (_i, [_vp, _vp, _d, _vp, _i, _d, _d, _vp, _i]),
}

_LIB = None


def lib():
    """Load libffhip.so (fails loudly when it was not built).

    torch, when importable, is imported first: its wheel ships its own HIP
    runtime (SONAME libamdhip64.so.7), and libffhip must bind to that same
    runtime rather than load a second one from /opt/rocm, so that device
    pointers and the device itself are shared within the process."""
    global _LIB
    if _LIB is None:
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        if not os.path.exists(LIBPATH):
            raise FFHipError(-6, f"{LIBPATH} not built (run __graft_entry__.build())")
        L = ctypes.CDLL(LIBPATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def check(rc):
    if rc != 0:
        raise FFHipError(rc, lib().ffhip_last_error().decode())


def _hp(a):
    """host pointer of a C-contiguous numpy array (or None)"""
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.c_void_p)


def _dp(t):
    """device pointer of a torch tensor (or an int address)"""
    if t is None:
        return None
    if isinstance(t, int):
        return ctypes.c_void_p(t)
    assert t.is_contiguous() and t.is_cuda
    # the library stream is ordered with the legacy default stream; a caller
    # on another torch stream is drained first
    import torch
    torch.cuda.current_stream().synchronize()
    return ctypes.c_void_p(t.data_ptr())


class Context:
    def __init__(self, device=0):
        h = ctypes.c_void_p()
        check(lib().ffhip_init(device, ctypes.byref(h)))
        self.h = h

    def sync(self):
        check(lib().ffhip_sync(self.h))

    def last_ms(self):
        return lib().ffhip_last_ms(self.h)

    def bench_copy(self, nbytes, reps=20):
        """average ms of a plain device copy of nbytes (read + write)"""
        ms = _d()
        check(lib().ffhip_bench_copy(self.h, int(nbytes), reps, ctypes.byref(ms)))
        return ms.value

    def close(self):
        if self.h:
            lib().ffhip_ctx_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_CTX = None


def context():
    global _CTX
    if _CTX is None:
        _CTX = Context(0)
    return _CTX


class Mesh:
    def __init__(self, h, ctx, dim):
        self.h, self.ctx, self.dim = h, ctx, dim
        nv, nt, nbe = _l(), _l(), _l()
        check(lib().ffhip_mesh_info(h, ctypes.byref(nv), ctypes.byref(nt), ctypes.byref(nbe)))
        self.nv, self.nt, self.nbe = nv.value, nt.value, nbe.value

    def download(self):
        d = self.dim
        nloc = d + 1
        xyz = np.zeros(d * self.nv)
        el = np.zeros(nloc * self.nt, np.int32)
        be = np.zeros(d * self.nbe, np.int32)
        bl = np.zeros(self.nbe, np.int32)
        check(lib().ffhip_mesh_download(self.h, _hp(xyz), _hp(el), _hp(be), _hp(bl)))
        return dict(nv=self.nv, nt=self.nt, nbe=self.nbe, xyz=xyz, elems=el, belems=be, blab=bl)

    def free(self):
        if self.h:
            lib().ffhip_mesh_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def cube(nx, ny=None, nz=None, labels=(1, 2, 3, 4, 5, 6), ctx=None):
    ctx = ctx or context()
    ny = nx if ny is None else ny
    nz = nx if nz is None else nz
    lab = np.asarray(labels, np.int32)
    h = ctypes.c_void_p()
    check(lib().ffhip_mesh3_cube(ctx.h, nx, ny, nz, _hp(lab), ctypes.byref(h)))
    return Mesh(h, ctx, 3)


def square(nx, ny=None, labels=(1, 2, 3, 4), ctx=None):
    ctx = ctx or context()
    ny = nx if ny is None else ny
    lab = np.asarray(labels, np.int32)
    h = ctypes.c_void_p()
    check(lib().ffhip_mesh2_square(ctx.h, nx, ny, _hp(lab), ctypes.byref(h)))
    return Mesh(h, ctx, 2)


def mesh2_upload(xy, tris, bedges, blab, ctx=None):
    """a 2D mesh from host arrays (FreeFEM Mesh: vertices, triangles, boundary edges)"""
    ctx = ctx or context()
    xy = np.ascontiguousarray(xy, np.float64)
    tris = np.ascontiguousarray(tris, np.int32)
    bedges = np.ascontiguousarray(bedges, np.int32)
    blab = np.ascontiguousarray(blab, np.int32)
    h = ctypes.c_void_p()
    check(lib().ffhip_mesh2_upload(ctx.h, xy.size // 2, tris.size // 3, blab.size, _hp(xy), _hp(tris),
                                   _hp(bedges), _hp(blab), ctypes.byref(h)))
    return Mesh(h, ctx, 2)


def mesh3_upload(xyz, tets, btri, blab, ctx=None):
    ctx = ctx or context()
    xyz = np.ascontiguousarray(xyz, np.float64)
    tets = np.ascontiguousarray(tets, np.int32)
    btri = np.ascontiguousarray(btri, np.int32)
    blab = np.ascontiguousarray(blab, np.int32)
    h = ctypes.c_void_p()
    check(lib().ffhip_mesh3_upload(ctx.h, xyz.size // 3, tets.size // 4, blab.size, _hp(xyz),
                                   _hp(tets), _hp(btri), _hp(blab), ctypes.byref(h)))
    return Mesh(h, ctx, 3)


class Matrix:
    def __init__(self, h, ctx, mesh=None):
        self.h, self.ctx, self.mesh = h, ctx, mesh
        n, nnz = _l(), _l()
        check(lib().ffhip_mat_info(h, ctypes.byref(n), ctypes.byref(nnz)))
        self.n, self.nnz = n.value, nnz.value

    @staticmethod
    def p1(mesh):
        h = ctypes.c_void_p()
        check(lib().ffhip_pattern_p1(mesh.ctx.h, mesh.h, ctypes.byref(h)))
        return Matrix(h, mesh.ctx, mesh)

    @staticmethod
    def p1_assembled(mesh, kind=COEF_SIN3, c0=1.0, kappa_e=None):
        """`matrix A = va(Vh,Vh)` of int3d(kappa grad u . grad v) on a 3D P1
        space in one call (pattern and values; the same matrix as p1() then
        assemble_p1()).  kappa_e: numpy (host) or torch cuda tensor."""
        on_dev, ptr = 0, None
        if kappa_e is not None:
            if isinstance(kappa_e, np.ndarray):
                kappa_e = np.ascontiguousarray(kappa_e, np.float64)
                ptr = _hp(kappa_e)
            else:
                ptr, on_dev = _dp(kappa_e), 1
        h = ctypes.c_void_p()
        check(lib().ffhip_pattern_assemble_p1(mesh.ctx.h, mesh.h, kind, c0, ptr, on_dev,
                                              ctypes.byref(h)))
        return Matrix(h, mesh.ctx, mesh)

    @staticmethod
    def upload(rowptr, colidx, vals, ctx=None):
        ctx = ctx or context()
        rowptr = np.ascontiguousarray(rowptr, np.int64)
        colidx = np.ascontiguousarray(colidx, np.int32)
        vals = np.ascontiguousarray(vals, np.float64)
        h = ctypes.c_void_p()
        check(lib().ffhip_mat_upload(ctx.h, rowptr.size - 1, colidx.size, _hp(rowptr), _hp(colidx),
                                     _hp(vals), ctypes.byref(h)))
        return Matrix(h, ctx)

    @staticmethod
    def p2v(space):
        """pattern of `matrix A = va(Vh,Vh)` on Vh(Th,[P2,P2,P2])"""
        h = ctypes.c_void_p()
        check(lib().ffhip_pattern_p2v(space.ctx.h, space.h, ctypes.byref(h)))
        A = Matrix(h, space.ctx, space.mesh)
        A.space = space
        return A

    @staticmethod
    def p2(space):
        """pattern of `matrix A = va(Vh,Vh)` on the scalar Vh(Th,P2) (dof = node
        of `space`)"""
        h = ctypes.c_void_p()
        check(lib().ffhip_pattern_p2(space.ctx.h, space.h, ctypes.byref(h)))
        A = Matrix(h, space.ctx, space.mesh)
        A.space = space
        return A

    def assemble_p2(self, kind=COEF_SIN3, c0=1.0, kappa_e=None):
        """int3d(kappa grad u . grad v) values on the scalar P2 space"""
        on_dev, ptr = 0, None
        if kappa_e is not None:
            if isinstance(kappa_e, np.ndarray):
                kappa_e = np.ascontiguousarray(kappa_e, np.float64)
                ptr = _hp(kappa_e)
            else:
                ptr, on_dev = _dp(kappa_e), 1
        check(lib().ffhip_assemble_p2(self.ctx.h, self.space.h, self.h, kind, c0, ptr, on_dev))
        return self.ctx.last_ms()

    def set_bc_p2(self, labels, tgv=1e30, want_onbc=False):
        lab = np.asarray(labels, np.int32)
        onbc = np.zeros(self.n, np.uint8) if want_onbc else None
        check(lib().ffhip_set_bc_p2(self.ctx.h, self.space.h, self.h, _hp(lab), lab.size, tgv,
                                    _hp(onbc)))
        return onbc if want_onbc else self.ctx.last_ms()

    def assemble_p2v_elasticity(self, lam=1.0, mu=0.5):
        """int3d(lambda div(u) div(v) + 2 mu epsilon(u):epsilon(v)) values"""
        check(lib().ffhip_assemble_p2v_elasticity(self.ctx.h, self.space.h, self.h, lam, mu))
        return self.ctx.last_ms()

    def set_bc_p2v(self, labels, tgv=1e30, want_onbc=False):
        lab = np.asarray(labels, np.int32)
        onbc = np.zeros(self.n, np.uint8) if want_onbc else None
        check(lib().ffhip_set_bc_p2v(self.ctx.h, self.space.h, self.h, _hp(lab), lab.size, tgv,
                                     _hp(onbc)))
        return onbc if want_onbc else self.ctx.last_ms()

    def assemble_p1(self, kind=COEF_SIN3, c0=1.0, kappa_e=None, mkind=COEF_NONE, m0=0.0,
                    mass_e=None):
        """int3d(kappa grad u . grad v) [+ int3d(m u v) when mkind != COEF_NONE].
        kappa_e / mass_e: numpy (host) or torch cuda tensors (device), nt
        values (P0), nt x nq (QP) or nv (P1, vertex values of a P1
        function); both on the same side."""
        arrs = [a for a in (kappa_e, mass_e) if a is not None]
        on_dev = int(any(not isinstance(a, np.ndarray) for a in arrs))
        assert all((not isinstance(a, np.ndarray)) == bool(on_dev) for a in arrs), \
            "kappa_e and mass_e must both be host or both device arrays"

        def ptr(a):
            if a is None:
                return None
            if isinstance(a, np.ndarray):
                return _hp(a)
            return _dp(a)
        if kappa_e is not None and isinstance(kappa_e, np.ndarray):
            kappa_e = np.ascontiguousarray(kappa_e, np.float64)
        if mass_e is not None and isinstance(mass_e, np.ndarray):
            mass_e = np.ascontiguousarray(mass_e, np.float64)
        if mkind == COEF_NONE:
            check(lib().ffhip_assemble_p1(self.ctx.h, self.mesh.h, self.h, kind, c0, ptr(kappa_e),
                                          on_dev))
        else:
            check(lib().ffhip_assemble_p1_form(self.ctx.h, self.mesh.h, self.h, kind, c0,
                                               ptr(kappa_e), mkind, m0, ptr(mass_e), on_dev))
        return self.ctx.last_ms()

    def assemble_p1_boundary(self, labels, kind=COEF_CONST, c0=1.0, coef=None):
        """add int2d(Th, labels)(alpha*u*v) (int1d on a 2D mesh); alpha as
        kind/c0/coef: coef numpy (host) or torch cuda tensor (device), nbe
        values (P0) or nbe x nq (QP, nq = 7 on faces, 3 on edges)"""
        lab = np.asarray(labels, np.int32)
        on_dev = int(coef is not None and not isinstance(coef, np.ndarray))
        if coef is not None and not on_dev:
            coef = np.ascontiguousarray(coef, np.float64)
        cp = None if coef is None else (_dp(coef) if on_dev else _hp(coef))
        check(lib().ffhip_assemble_p1_boundary(self.ctx.h, self.mesh.h, self.h, _hp(lab), lab.size,
                                               kind, c0, cp, on_dev))
        return self.ctx.last_ms()

    def coo(self, device=False):
        """[I, J, C] = A: int64 rows and columns (0-based), values, in row
        order, with FreeFEM's extra (n, n, 0) when A has no (n-1, n-1) entry.
        device=True returns torch cuda tensors."""
        cnt = _l()
        check(lib().ffhip_mat_coo(self.h, ctypes.byref(cnt), None, None, None, 0))
        k = cnt.value
        if device:
            import torch
            I = torch.empty(k, dtype=torch.int64, device="cuda")
            J = torch.empty(k, dtype=torch.int64, device="cuda")
            C = torch.empty(k, dtype=torch.float64, device="cuda")
            check(lib().ffhip_mat_coo(self.h, ctypes.byref(cnt), _dp(I), _dp(J), _dp(C), 1))
        else:
            I, J, C = np.zeros(k, np.int64), np.zeros(k, np.int64), np.zeros(k)
            check(lib().ffhip_mat_coo(self.h, ctypes.byref(cnt), _hp(I), _hp(J), _hp(C), 0))
        return I, J, C

    def lower(self):
        """the half (sym=1) matrix: lower triangle of this symmetric matrix"""
        h = ctypes.c_void_p()
        check(lib().ffhip_mat_lower(self.h, ctypes.byref(h)))
        return Matrix(h, self.ctx)

    def full(self):
        """the full symmetric matrix of a half one"""
        h = ctypes.c_void_p()
        check(lib().ffhip_mat_full(self.h, ctypes.byref(h)))
        return Matrix(h, self.ctx)

    def set_half(self, half=1):
        check(lib().ffhip_mat_set_half(self.h, int(half)))

    @property
    def half(self):
        v = _i()
        check(lib().ffhip_mat_half(self.h, ctypes.byref(v)))
        return v.value

    def write_morse(self, path):
        """`ofstream f(path); f << A;` (FreeFEM's Morse text)"""
        check(lib().ffhip_mat_write_morse(self.h, os.fsencode(path)))

    @staticmethod
    def read_text(path, ctx=None):
        """`ifstream f(path); f >> A;` into a device matrix (Morse or COO text)"""
        ctx = ctx or context()
        h = ctypes.c_void_p()
        check(lib().ffhip_mat_read_text(ctx.h, os.fsencode(path), ctypes.byref(h)))
        return Matrix(h, ctx)

    def set_bc(self, labels, tgv=1e30, want_onbc=False):
        lab = np.asarray(labels, np.int32)
        onbc = np.zeros(self.n, np.uint8) if want_onbc else None
        check(lib().ffhip_set_bc(self.ctx.h, self.mesh.h, self.h, _hp(lab), lab.size, tgv, _hp(onbc)))
        return onbc if want_onbc else self.ctx.last_ms()

    def download(self):
        rp = np.zeros(self.n + 1, np.int64)
        ci = np.zeros(self.nnz, np.int32)
        va = np.zeros(self.nnz)
        check(lib().ffhip_mat_download(self.h, _hp(rp), _hp(ci), _hp(va)))
        return rp, ci, va

    def set_values(self, vals):
        vals = np.ascontiguousarray(vals, np.float64)
        assert vals.size == self.nnz
        check(lib().ffhip_mat_set_values(self.h, _hp(vals)))

    def device_ptrs(self):
        a, b, c = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
        check(lib().ffhip_mat_device_ptrs(self.h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)))
        return a.value, b.value, c.value

    def spmv(self, x, y, trans=False):
        """device vectors (torch cuda tensors)"""
        check(lib().ffhip_spmv(self.ctx.h, self.h, _dp(x), _dp(y), int(trans)))

    def free(self):
        if self.h:
            lib().ffhip_mat_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def rhs_p1(mesh, f=1.0, labels=(), tgv=1e30, g=0.0, out=None):
    """b = vb(0,Vh) for int(f*v) + on(labels, u=g).  out: torch cuda tensor
    (device) or None (returns a numpy array)."""
    lab = np.asarray(labels, np.int32)
    if out is None:
        b = np.zeros(mesh.nv)
        check(lib().ffhip_rhs_p1(mesh.ctx.h, mesh.h, f, _hp(lab), lab.size, tgv, g, _hp(b), 0))
        return b
    check(lib().ffhip_rhs_p1(mesh.ctx.h, mesh.h, f, _hp(lab), lab.size, tgv, g, _dp(out), 1))
    return out


def rhs_p1_form(mesh, kind=COEF_CONST, f0=1.0, fcoef=None, labels=(), tgv=1e30, g0=0.0, gvals=None,
                out=None):
    """b = vb(0,Vh) for int(f*v) + on(labels, u=g): f as kind/f0/fcoef (P0: nt
    values, QP: nt x nq, P1: nv vertex values), g per dof (gvals) or constant g0.  out: torch cuda
    tensor (then fcoef/gvals are device tensors too) or None (numpy)."""
    lab = np.asarray(labels, np.int32)
    if out is None:
        fc = None if fcoef is None else np.ascontiguousarray(fcoef, np.float64)
        gv = None if gvals is None else np.ascontiguousarray(gvals, np.float64)
        b = np.zeros(mesh.nv)
        check(lib().ffhip_rhs_p1_form(mesh.ctx.h, mesh.h, kind, f0, _hp(fc), _hp(lab), lab.size, tgv,
                                      g0, _hp(gv), _hp(b), 0))
        return b
    check(lib().ffhip_rhs_p1_form(mesh.ctx.h, mesh.h, kind, f0, _dp(fcoef), _hp(lab), lab.size, tgv,
                                  g0, _dp(gvals), _dp(out), 1))
    return out


def rhs_p1_boundary(mesh, b, labels, kind=COEF_CONST, g0=1.0, coef=None, dirichlet=()):
    """b += int2d(Th, labels)(g*v) (int1d on a 2D mesh), leaving the rows of
    the Dirichlet labels unchanged.  b: numpy array (updated in place) or
    torch cuda tensor; coef like Matrix.assemble_p1_boundary's."""
    lab = np.asarray(labels, np.int32)
    dl = np.asarray(dirichlet, np.int32)
    on_dev = int(not isinstance(b, np.ndarray))
    if coef is not None:
        assert isinstance(coef, np.ndarray) != bool(on_dev), "b and coef on the same side"
        if not on_dev:
            coef = np.ascontiguousarray(coef, np.float64)
    if not on_dev:
        assert b.dtype == np.float64 and b.flags.c_contiguous
    cp = None if coef is None else (_dp(coef) if on_dev else _hp(coef))
    check(lib().ffhip_rhs_p1_boundary(mesh.ctx.h, mesh.h, _hp(lab), lab.size, kind, g0, cp,
                                      _hp(dl) if dl.size else None, dl.size,
                                      _dp(b) if on_dev else _hp(b), on_dev))
    return b


def csr_write_morse(path, rowptr, colidx, vals, m=None, half=0):
    """FreeFEM's Morse text of a host CSR (host-only: no device work)"""
    rowptr = np.ascontiguousarray(rowptr, np.int64)
    colidx = np.ascontiguousarray(colidx, np.int32)
    vals = np.ascontiguousarray(vals, np.float64)
    n = rowptr.size - 1
    check(lib().ffhip_csr_write_morse(n, n if m is None else m, half, _hp(rowptr), _hp(colidx),
                                      _hp(vals), os.fsencode(path)))


def csr_read_text(path):
    """HashMatrix(istream) of a Morse / COO text file as a sorted host CSR:
    returns (rowptr, colidx, vals, m, half) (host-only: no device work)"""
    n, m, nnz = _l(), _l(), _l()
    half = _i()
    p = os.fsencode(path)
    check(lib().ffhip_csr_read_text(p, 0, ctypes.byref(n), ctypes.byref(m), ctypes.byref(half),
                                    ctypes.byref(nnz), None, None, None))
    rp = np.zeros(n.value + 1, np.int64)
    ci = np.zeros(nnz.value, np.int32)
    va = np.zeros(nnz.value)
    check(lib().ffhip_csr_read_text(p, nnz.value, ctypes.byref(n), ctypes.byref(m),
                                    ctypes.byref(half), ctypes.byref(nnz), _hp(rp), _hp(ci), _hp(va)))
    return rp, ci, va, m.value, half.value


class P2VSpace:
    """Vh(Th,[P2,P2,P2]) on a device tetrahedral mesh"""

    def __init__(self, mesh):
        h = ctypes.c_void_p()
        check(lib().ffhip_p2v_create(mesh.ctx.h, mesh.h, ctypes.byref(h)))
        self.h, self.mesh, self.ctx = h, mesh, mesh.ctx
        nn, nd = _l(), _l()
        check(lib().ffhip_p2v_info(h, ctypes.byref(nn), ctypes.byref(nd)))
        self.nnodes, self.ndof = nn.value, nd.value

    def elem_nodes(self):
        en = np.zeros(10 * self.mesh.nt, np.int32)
        check(lib().ffhip_p2v_download(self.h, _hp(en)))
        return en

    def free(self):
        if self.h:
            lib().ffhip_p2v_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def rhs_p2(space, f=1.0, labels=(), tgv=1e30, g=0.0, out=None):
    """b of int3d(f*v) + on(labels, u=g) on the scalar Vh(Th,P2)"""
    lab = np.asarray(labels, np.int32)
    if out is None:
        b = np.zeros(space.nnodes)
        check(lib().ffhip_rhs_p2(space.ctx.h, space.h, f, _hp(lab), lab.size, tgv, g, _hp(b), 0))
        return b
    check(lib().ffhip_rhs_p2(space.ctx.h, space.h, f, _hp(lab), lab.size, tgv, g, _dp(out), 1))
    return out


def rhs_p2v(space, f=(0.0, 0.0, -1.0), labels=(), tgv=1e30, g=None, out=None):
    """b of int3d(f'*v) + on(labels, u=g) on [P2,P2,P2].  out: torch cuda
    tensor (device) or None (returns a numpy array)."""
    lab = np.asarray(labels, np.int32)
    fv = np.ascontiguousarray(f, np.float64)
    gv = None if g is None else np.ascontiguousarray(g, np.float64)
    if out is None:
        b = np.zeros(space.ndof)
        check(lib().ffhip_rhs_p2v(space.ctx.h, space.h, _hp(fv), _hp(lab), lab.size, tgv, _hp(gv),
                                  _hp(b), 0))
        return b
    check(lib().ffhip_rhs_p2v(space.ctx.h, space.h, _hp(fv), _hp(lab), lab.size, tgv, _hp(gv),
                              _dp(out), 1))
    return out


class CG:
    """Jacobi-CG solver of a device matrix (SolverCG).  graph=False launches
    the iteration kernels one by one instead of replaying a captured chunk of
    16 iterations.  layout: "auto", "sell"
    (sliced ELL, 16-bit column deltas where they fit), "sell32" (sliced ELL,
    32-bit columns), "csr" (16 lanes per row) or "bsr3" (3x3 blocks of a
    [P2,P2,P2] matrix); auto takes bsr3 for [P2,P2,P2] matrices, else sliced
    ELL unless it pads by more than 25 %."""

    def __init__(self, A, mode=CG_FP64, layout="auto", graph=True):
        h = ctypes.c_void_p()
        flag = {"auto": CG_LAYOUT_AUTO, "sell": CG_LAYOUT_SELL,
                "sell32": CG_LAYOUT_SELL | CG_INDEX32, "csr": CG_LAYOUT_CSR,
                "bsr3": CG_LAYOUT_BSR3}[layout]
        if not graph:
            flag |= CG_EAGER
        check(lib().ffhip_cg_create(A.ctx.h, A.h, mode | flag, ctypes.byref(h)))
        self.h, self.A, self.mode = h, A, mode

    def layout(self):
        """(layout name "sell16" | "sell" | "csr" | "bsr3", entries streamed per mat-vec)"""
        k, e = ctypes.c_int(), ctypes.c_int64()
        check(lib().ffhip_cg_layout(self.h, ctypes.byref(k), ctypes.byref(e)))
        if k.value & CG_LAYOUT_SELL:
            return ("sell16" if k.value & CG_INDEX16 else "sell"), e.value
        return ("bsr3" if k.value & CG_LAYOUT_BSR3 else "csr"), e.value

    def set_profiling(self, on=True):
        check(lib().ffhip_cg_set_profiling(self.h, int(on)))

    def update_values(self):
        check(lib().ffhip_cg_update_values(self.h))

    def solve(self, b, x=None, eps=1e-10, itmax=0, tgv=1e30, trans=False, raise_on_error=True):
        """numpy b/x (host) or torch cuda tensors (device).  Returns (x, stats)."""
        st = CGStats()
        if isinstance(b, np.ndarray):
            b = np.ascontiguousarray(b, np.float64)
            x = np.zeros(b.size) if x is None else np.array(x, np.float64)
            rc = lib().ffhip_cg_solve(self.h, _hp(b), _hp(x), 0, eps, itmax, tgv, int(trans),
                                      ctypes.byref(st))
        else:
            import torch
            x = torch.zeros_like(b) if x is None else x
            rc = lib().ffhip_cg_solve(self.h, _dp(b), _dp(x), 1, eps, itmax, tgv, int(trans),
                                      ctypes.byref(st))
        if raise_on_error:
            check(rc)
        return x, st.as_dict()

    def gmres(self, b, x=None, eps=1e-6, itmax=0, restart=1000, tgv=1e30, trans=False,
              raise_on_error=True):
        """solver=GMRES (SolverGMRES::dosolver over fgmres) with this solver's
        matrix, Jacobi preconditioner and layout, in FP64.  restart is the
        Krylov dimension (FreeFEM's dimKrylov / NbSpace, 1000 by default).
        Returns (x, stats): iters, converged, refinements = restarts,
        gcg0 = ||b||, gcg = final relative GMRES residual."""
        st = CGStats()
        if isinstance(b, np.ndarray):
            b = np.ascontiguousarray(b, np.float64)
            x = np.zeros(b.size) if x is None else np.array(x, np.float64)
            rc = lib().ffhip_gmres_solve(self.h, _hp(b), _hp(x), 0, eps, itmax, restart, tgv,
                                         int(trans), ctypes.byref(st))
        else:
            import torch
            x = torch.zeros_like(b) if x is None else x
            rc = lib().ffhip_gmres_solve(self.h, _dp(b), _dp(x), 1, eps, itmax, restart, tgv,
                                         int(trans), ctypes.byref(st))
        if raise_on_error:
            check(rc)
        return x, st.as_dict()

    def linear_cg(self, x, b=None, eps=1e-6, nbiter=100, precon=None, veps=None,
                  raise_on_error=True):
        """LinearCG(A*x, x, b, eps=, nbiter=, precon=, veps=) / AffineCG(A*x, x)
        (lgfem.cpp:705-758): ConjuguedGradient2 with this solver's matrix as
        the operator, in FP64.  precon: None (identity) or "jacobi".  x (the
        initial guess, updated in place) and b: numpy arrays (host) or torch
        cuda tensors (device); b None is the zero right-hand side.  veps, a
        one-element numpy array, replaces eps when given and receives -reps2
        afterwards, as the script's veps does.  Returns (ret, stats): ret is
        LinearCG's value, 1 converged, 2 nothing to do, 0 not converged."""
        pc = {None: CG2_IDENTITY, "identity": CG2_IDENTITY, "jacobi": CG2_JACOBI}[precon]
        if veps is not None:
            eps = float(veps[0])
        st = CGStats()
        eo = _d()
        if isinstance(x, np.ndarray):
            assert x.dtype == np.float64 and x.flags["C_CONTIGUOUS"]
            bb = None if b is None else np.ascontiguousarray(b, np.float64)
            rc = lib().ffhip_cg_solve_affine(self.h, _hp(bb), _hp(x), 0, eps, nbiter, pc,
                                             ctypes.byref(eo), ctypes.byref(st))
        else:
            rc = lib().ffhip_cg_solve_affine(self.h, _dp(b), _dp(x), 1, eps, nbiter, pc,
                                             ctypes.byref(eo), ctypes.byref(st))
        if raise_on_error:
            check(rc)
        if veps is not None:
            veps[0] = -eo.value
        d = st.as_dict()
        d["reps2"] = eo.value
        return d["converged"], d

    def bench_spmv(self, mode, reps, x, y):
        ms = _d()
        check(lib().ffhip_bench_spmv(self.h, mode, reps, _dp(x), _dp(y), ctypes.byref(ms)))
        return ms.value

    def free(self):
        if self.h:
            lib().ffhip_cg_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def p1_misfit(mesh, u, ud, labels=(), want_rhs=True):
    """J = 1/2 int (u-ud)^2 and r = M (u-ud) with Dirichlet rows zeroed
    (u, ud torch cuda tensors).  Returns (J, r)."""
    import torch
    lab = np.asarray(labels, np.int32)
    J = _d()
    r = torch.empty_like(u) if want_rhs else None
    check(lib().ffhip_p1_misfit(mesh.ctx.h, mesh.h, _dp(u), _dp(ud), _hp(lab), lab.size,
                                ctypes.byref(J), _dp(r)))
    return J.value, r


def p1_grad_kappa(mesh, u, lam, out=None):
    """g_e = -|K| grad u_K . grad lambda_K (torch cuda tensors)"""
    import torch
    g = torch.empty(mesh.nt, dtype=torch.float64, device=u.device) if out is None else out
    check(lib().ffhip_p1_grad_kappa(mesh.ctx.h, mesh.h, _dp(u), _dp(lam), _dp(g)))
    return g
