"""ctypes bindings for the oracle (libffo.so) and the compiled reference subset
(oracle/_ref/libffref_ref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as what results are compared against.  The product never imports it.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_REF = None

_i32p = np.ctypeslib.ndpointer(np.int32, flags="C")
_i64p = np.ctypeslib.ndpointer(np.int64, flags="C")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C")
_u8p = np.ctypeslib.ndpointer(np.uint8, flags="C")
_vp = ctypes.c_void_p
_i = ctypes.c_int
_l = ctypes.c_int64
_d = ctypes.c_double

COEF_CONST, COEF_SIN3, COEF_P0, COEF_QP = 0, 1, 2, 3


def lib():
    """Load libffo.so (build it with `make -C oracle`); FFO_LIB names another
    build of the same sources (the sanitized oracle/libffo_asan.so)."""
    global _LIB
    if _LIB is None:
        path = os.environ.get("FFO_LIB") or os.path.join(HERE, "libffo.so")
        L = ctypes.CDLL(path)
        sig = {
            "ffo_cube_counts": (None, [_i, _i, _i, _i64p]),
            "ffo_cube": (_i, [_i, _i, _i, _i32p, _f64p, _i32p, _i32p, _i32p]),
            "ffo_square_counts": (None, [_i, _i, _i64p]),
            "ffo_square": (_i, [_i, _i, _i32p, _f64p, _i32p, _i32p, _i32p]),
            "ffo_qf_tet5": (_i, [_f64p, _f64p]),
            "ffo_qf_tri5": (_i, [_f64p, _f64p]),
            "ffo_tet_mesure": (_d, [_f64p, _f64p, _f64p, _f64p]),
            "ffo_tet_gradlambda": (None, [_f64p, _d, _f64p]),
            "ffo_pattern_rowptr": (_l, [_i, _l, _i, _i32p, _i64p]),
            "ffo_pattern_colidx": (_i, [_i, _l, _i, _i32p, _i64p, _i32p]),
            "ffo_pattern_bruteforce": (_l, [_i, _l, _i, _i32p, _i64p, _i32p]),
            "ffo_p1_element_3d": (_i, [_f64p, _i32p, _i, _d, _d, _f64p]),
            "ffo_p1_element_2d": (_i, [_f64p, _i32p, _i, _d, _d, _f64p]),
            "ffo_assemble_p1_3d": (_i, [_i, _l, _f64p, _i32p, _i, _d, _vp, _i64p, _i32p, _f64p]),
            "ffo_assemble_p1_2d": (_i, [_i, _l, _f64p, _i32p, _i, _d, _vp, _i64p, _i32p, _f64p]),
            "ffo_p1_form_element_3d": (_i, [_f64p, _i32p, _i, _d, _d, _i, _d, _d, _f64p]),
            "ffo_p2_lap_element": (_i, [_f64p, _i32p, _i, _d, _d, _f64p]),
            "ffo_assemble_p2_lap": (_i, [_l, _f64p, _i32p, _i32p, _i, _d, _vp, _l, _i64p, _i32p, _f64p]),
            "ffo_rhs_p2": (_i, [_l, _f64p, _i32p, _i32p, _d, _l, _f64p]),
            "ffo_p2_mark_bc": (_i, [_i, _l, _i32p, _i32p, _l, _i32p, _i32p, _i, _i32p, _i, _l, _u8p]),
            "ffo_p1_form_element_2d": (_i, [_f64p, _i32p, _i, _d, _d, _i, _d, _d, _f64p]),
            "ffo_assemble_p1_form_3d": (_i, [_i, _l, _f64p, _i32p, _i, _d, _vp, _i, _d, _vp, _i64p, _i32p, _f64p]),
            "ffo_assemble_p1_form_2d": (_i, [_i, _l, _f64p, _i32p, _i, _d, _vp, _i, _d, _vp, _i64p, _i32p, _f64p]),
            "ffo_rhs_p1_3d": (_i, [_i, _l, _f64p, _i32p, _d, _f64p]),
            "ffo_rhs_p1_2d": (_i, [_i, _l, _f64p, _i32p, _d, _f64p]),
            "ffo_mark_bc": (_i, [_i, _l, _i, _i32p, _i32p, _i, _i32p, _u8p]),
            "ffo_assemble_p1_mass_3d": (_i, [_i, _l, _f64p, _i32p, _i64p, _i32p, _f64p]),
            "ffo_p1_misfit_3d": (_d, [_l, _f64p, _i32p, _f64p, _f64p]),
            "ffo_p0_grad_3d": (_i, [_l, _f64p, _i32p, _f64p, _f64p, _f64p]),
            "ffo_p2_nodes": (_l, [_i, _l, _i32p, _i32p]),
            "ffo_p2_basis": (None, [_f64p, _f64p, _f64p, _f64p]),
            "ffo_p2_elast_element": (_i, [_f64p, _i32p, _d, _d, _f64p]),
            "ffo_p2v_elem_dofs": (_i, [_l, _i32p, _i32p]),
            "ffo_assemble_p2_elast": (_i, [_l, _f64p, _i32p, _i32p, _d, _d, _l, _i64p, _i32p, _f64p]),
            "ffo_rhs_p2_elast": (_i, [_l, _f64p, _i32p, _i32p, _f64p, _l, _f64p]),
            "ffo_p2v_mark_bc": (_i, [_i, _l, _i32p, _i32p, _l, _i32p, _i32p, _i, _i32p, _l, _u8p]),
            "ffo_set_bc": (_i, [_i, _i64p, _i32p, _f64p, _u8p, _d]),
            "ffo_rhs_bc": (_i, [_i, _u8p, _d, _d, _f64p]),
            "ffo_spmv": (None, [_i, _i64p, _i32p, _f64p, _f64p, _f64p]),
            "ffo_spmv_t": (None, [_i, _i64p, _i32p, _f64p, _f64p, _f64p]),
            "ffo_spmv_half": (None, [_i, _i64p, _i32p, _f64p, _f64p, _f64p]),
            "ffo_gettgv": (_d, [_i, _i64p, _i32p, _f64p, ctypes.POINTER(_i)]),
            "ffo_pcg": (_i, [_i, _i64p, _i32p, _f64p, _f64p, _f64p, _d, _i, _d, _i, _f64p]),
            "ffo_cg2": (_i, [_i, _i64p, _i32p, _f64p, _f64p, _f64p, _f64p, _i, _i, _f64p]),
            "ffo_fgmres": (_i, [_i, _i64p, _i32p, _f64p, _f64p, _f64p, _d, _i, _i, _d, _f64p]),
            "ffo_qf_gl3": (_i, [_f64p, _f64p]),
            "ffo_rhs_p1_coef": (_i, [_i, _i, _l, _f64p, _i32p, _i, _d, _vp, _f64p]),
            "ffo_p1_at_qp": (_i, [_i, _l, _i32p, _f64p, _f64p]),
            "ffo_p1_boundary": (_i, [_i, _i, _l, _f64p, _i32p, _l, _i32p, _i32p, _i, _i32p, _i, _d,
                                     _vp, _vp, _vp, _vp, _vp]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


_OMP = None


def omp_lib():
    """Load libffo_omp.so, the all-core (OpenMP) variant of the 3D P1 path."""
    global _OMP
    if _OMP is None:
        L = ctypes.CDLL(os.path.join(HERE, "libffo_omp.so"))
        sig = {
            "ffo_omp_max_threads": (_i, []),
            "ffo_omp_assemble_p1_3d": (_vp, [_i, _l, _f64p, _i32p, _i, _d, _vp]),
            "ffo_omp_nnz": (_l, [_vp]),
            "ffo_omp_arrays": (None, [_vp, _i64p, _i32p, _f64p]),
            "ffo_omp_setbc": (None, [_vp, _u8p, _d]),
            "ffo_omp_pcg": (_i, [_vp, _f64p, _f64p, _d, _i, _d, _f64p]),
            "ffo_omp_free": (None, [_vp]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _OMP = L
    return _OMP


class OmpP1:
    """3D P1 system assembled on all host cores (pattern, element-order values
    bit-identical to pattern() + assemble_p1_3d())."""

    def __init__(self, m, kind=COEF_SIN3, c0=1.0, kappa_e=None):
        L = omp_lib()
        ke = None if kappa_e is None else np.ascontiguousarray(kappa_e, np.float64)
        self.h = L.ffo_omp_assemble_p1_3d(m["nv"], m["nt"], m["xyz"], m["tets"], kind, c0,
                                          None if ke is None else ke.ctypes.data)
        if not self.h:
            raise RuntimeError("ffo_omp_assemble_p1_3d failed")
        self.n = m["nv"]
        self.nnz = L.ffo_omp_nnz(self.h)

    def arrays(self):
        rp = np.zeros(self.n + 1, np.int64)
        ci = np.zeros(self.nnz, np.int32)
        va = np.zeros(self.nnz)
        omp_lib().ffo_omp_arrays(self.h, rp, ci, va)
        return rp, ci, va

    def set_bc(self, onbc, tgv=1e30):
        omp_lib().ffo_omp_setbc(self.h, np.ascontiguousarray(onbc, np.uint8), tgv)

    def pcg(self, b, eps=1e-10, itmax=0, tgv=1e30):
        x = np.zeros(self.n)
        st = np.zeros(4)
        omp_lib().ffo_omp_pcg(self.h, np.ascontiguousarray(b, np.float64), x, eps, itmax, tgv, st)
        return x, {"iters": int(st[0]), "gcg0": st[1], "gcg": st[2], "nret": int(st[3])}

    def free(self):
        if self.h:
            omp_lib().ffo_omp_free(self.h)
            self.h = None

    def __del__(self):
        self.free()


def ref_available():
    return os.path.exists(os.path.join(HERE, "_ref", "libffref_ref.so"))


def ref():
    """Load the compiled reference subset (built by oracle/build_ref.sh).

    The library also holds the plugin-glue driver, which links libffhip: as
    for the ffhip package, torch (when importable) is imported first so the
    process has one HIP runtime, the one torch's wheel ships."""
    global _REF
    if _REF is None:
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = ctypes.CDLL(os.path.join(HERE, "_ref", "libffref_ref.so"))
        sig = {
            "ref_qf_tet5": (_i, [_f64p, _f64p]),
            "ref_qf_tri5": (_i, [_f64p, _f64p]),
            "ref_qf_gl3": (_i, [_f64p, _f64p]),
            "ref_tet_geom": (_i, [_f64p, _f64p, _f64p]),
            "ref_hm_new": (_vp, [_i]),
            "ref_hm_free": (None, [_vp]),
            "ref_hm_add_p1_3d": (_i, [_vp, _i, _l, _f64p, _i32p, _i, _d, _vp]),
            "ref_hm_setbc": (_i, [_vp, _u8p, _d]),
            "ref_hm_nnz": (_l, [_vp]),
            "ref_hm_gettgv": (_d, [_vp, ctypes.POINTER(_i)]),
            "ref_hm_matmul": (_i, [_vp, _f64p, _f64p, _i]),
            "ref_hm_csr": (_i, [_vp, _i64p, _i32p, _f64p]),
            "ref_hm_cg": (_i, [_vp, _f64p, _f64p, _d, _i, _d, _i, _f64p]),
            "ref_hm_from_csr": (_vp, [_i, _i64p, _i32p, _f64p]),
            "ref_hm_gmres": (_i, [_vp, _f64p, _f64p, _d, _i, _i, _d, _f64p]),
            "ref_fespace_dofs": (_l, [_i, _f64p, _l, _i32p, _i, _vp, ctypes.POINTER(_i)]),
            "ref_p2_fb": (_i, [_f64p, _f64p, _f64p, _f64p]),
            "ref_hm_write": (_i, [_vp, ctypes.c_char_p, _i]),
            "ref_hm_read": (_l, [ctypes.c_char_p, _l, _i64p, _vp, _vp, _vp]),
            "ref_hm_coo": (_l, [_vp, _i64p, _i64p, _f64p]),
            "ref_hm_from_csr_half": (_vp, [_i, _i64p, _i32p, _f64p]),
            # the plugin's glue run against the compiled objects (seam_driver.cpp)
            "seam_mesh3_roundtrip": (_i, [_i, _f64p, _l, _i32p, _l, _i32p, _i32p, _f64p, _i32p,
                                          _i32p, _i32p]),
            "seam_hm_from_csr": (_vp, [_i, _i, _i64p, _i32p, _f64p]),
            "seam_hm_state": (_i, [_vp, _i32p]),
            "seam_core_solve": (_i, [_vp, _i, _f64p, _f64p, _d, _i, _d, _i, _f64p]),
            "seam_resident_flow": (_vp, [_i, _i, _d, _d, _f64p, _f64p, _f64p, _f64p, _f64p, _f64p,
                                         _f64p, _f64p, ctypes.POINTER(_i)]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _REF = L
    return _REF


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


# ---------------------------------------------------------------- meshes
def cube(n, ny=None, nz=None, labels=(1, 2, 3, 4, 5, 6)):
    ny = n if ny is None else ny
    nz = n if nz is None else nz
    c = np.zeros(3, np.int64)
    lib().ffo_cube_counts(n, ny, nz, c)
    nv, nt, nbe = (int(v) for v in c)
    xyz = np.zeros(3 * nv)
    tets = np.zeros(4 * nt, np.int32)
    btri = np.zeros(3 * nbe, np.int32)
    blab = np.zeros(nbe, np.int32)
    rc = lib().ffo_cube(n, ny, nz, np.asarray(labels, np.int32), xyz, tets, btri, blab)
    assert rc == 0, rc
    return dict(nv=nv, nt=nt, nbe=nbe, xyz=xyz, tets=tets, btri=btri, blab=blab)


def square(n, ny=None, labels=(1, 2, 3, 4)):
    ny = n if ny is None else ny
    c = np.zeros(3, np.int64)
    lib().ffo_square_counts(n, ny, c)
    nv, nt, nbe = (int(v) for v in c)
    xy = np.zeros(2 * nv)
    tris = np.zeros(3 * nt, np.int32)
    bed = np.zeros(2 * nbe, np.int32)
    blab = np.zeros(nbe, np.int32)
    rc = lib().ffo_square(n, ny, np.asarray(labels, np.int32), xy, tris, bed, blab)
    assert rc == 0, rc
    return dict(nv=nv, nt=nt, nbe=nbe, xy=xy, tris=tris, bedges=bed, blab=blab)


# ---------------------------------------------------------------- pattern
def pattern(n, elems, nloc):
    nel = elems.size // nloc
    rowptr = np.zeros(n + 1, np.int64)
    nnz = lib().ffo_pattern_rowptr(n, nel, nloc, elems, rowptr)
    colidx = np.zeros(nnz, np.int32)
    rc = lib().ffo_pattern_colidx(n, nel, nloc, elems, rowptr, colidx)
    assert rc == 0, rc
    return rowptr, colidx


def pattern_bruteforce(n, elems, nloc):
    nel = elems.size // nloc
    rowptr = np.zeros(n + 1, np.int64)
    colidx = np.zeros(nel * nloc * nloc, np.int32)
    nnz = lib().ffo_pattern_bruteforce(n, nel, nloc, elems, rowptr, colidx)
    return rowptr, colidx[:nnz].copy()


def assemble_p1_3d(m, rowptr, colidx, kind=COEF_SIN3, c0=1.0, kappa_e=None, mkind=-1, m0=0.0,
                   mass_e=None):
    """int3d(kappa grad u . grad v [+ m u v]); mkind < 0: no mass term"""
    vals = np.zeros(colidx.size)
    ke = None if kappa_e is None else np.ascontiguousarray(kappa_e, np.float64)
    me = None if mass_e is None else np.ascontiguousarray(mass_e, np.float64)
    if mkind < 0:
        rc = lib().ffo_assemble_p1_3d(m["nv"], m["nt"], m["xyz"], m["tets"], kind, c0,
                                      _ptr(ke), rowptr, colidx, vals)
    else:
        rc = lib().ffo_assemble_p1_form_3d(m["nv"], m["nt"], m["xyz"], m["tets"], kind, c0, _ptr(ke),
                                           mkind, m0, _ptr(me), rowptr, colidx, vals)
    assert rc == 0, rc
    return vals


def assemble_p1_2d(m, rowptr, colidx, kind=COEF_CONST, c0=1.0, kappa_e=None, mkind=-1, m0=0.0,
                   mass_e=None):
    vals = np.zeros(colidx.size)
    ke = None if kappa_e is None else np.ascontiguousarray(kappa_e, np.float64)
    me = None if mass_e is None else np.ascontiguousarray(mass_e, np.float64)
    if mkind < 0:
        rc = lib().ffo_assemble_p1_2d(m["nv"], m["nt"], m["xy"], m["tris"], kind, c0,
                                      _ptr(ke), rowptr, colidx, vals)
    else:
        rc = lib().ffo_assemble_p1_form_2d(m["nv"], m["nt"], m["xy"], m["tris"], kind, c0, _ptr(ke),
                                           mkind, m0, _ptr(me), rowptr, colidx, vals)
    assert rc == 0, rc
    return vals


def rhs_p1_3d(m, f=1.0):
    b = np.zeros(m["nv"])
    lib().ffo_rhs_p1_3d(m["nv"], m["nt"], m["xyz"], m["tets"], f, b)
    return b


def p1_at_qp(m, u):
    """a P1 function (vertex values u) at the Tet_5 / T_5 points of every
    element, FreeFEM's evaluation: (nt * nq) values, element-major"""
    dim = 3 if "tets" in m else 2
    nq = 14 if dim == 3 else 7
    out = np.zeros(m["nt"] * nq)
    rc = lib().ffo_p1_at_qp(dim, m["nt"], m["tets"] if dim == 3 else m["tris"],
                            np.ascontiguousarray(u, np.float64), out)
    assert rc == nq, rc
    return out


def rhs_p1_coef(m, kind, f0=0.0, fcoef=None):
    """int(f*v) with a coefficient (CONST / SIN3 / P0 / QP), 3D or 2D mesh"""
    dim = 3 if "tets" in m else 2
    b = np.zeros(m["nv"])
    fc = None if fcoef is None else np.ascontiguousarray(fcoef, np.float64)
    rc = lib().ffo_rhs_p1_coef(dim, m["nv"], m["nt"], m["xyz"] if dim == 3 else m["xy"],
                               m["tets"] if dim == 3 else m["tris"], kind, f0, _ptr(fc), b)
    assert rc == 0, rc
    return b


def rhs_p1_2d(m, f=1.0):
    b = np.zeros(m["nv"])
    lib().ffo_rhs_p1_2d(m["nv"], m["nt"], m["xy"], m["tris"], f, b)
    return b


def p1_boundary(m, labels, kind=COEF_CONST, c0=1.0, coef=None, rowptr=None, colidx=None,
                vals=None, b=None):
    """FreeFEM's int2d(Th, labels)(alpha*u*v) added to vals (CSR of the P1
    pattern) and/or int2d(Th, labels)(g*v) added to b; int1d on a 2D mesh."""
    dim = 3 if "tets" in m else 2
    X = m["xyz"] if dim == 3 else m["xy"]
    el = m["tets"] if dim == 3 else m["tris"]
    be = m["btri"] if dim == 3 else m["bedges"]
    labs = np.asarray(labels, np.int32)
    cf = None if coef is None else np.ascontiguousarray(coef, np.float64)
    for a in (vals, b):
        assert a is None or (a.dtype == np.float64 and a.flags.c_contiguous)
    rc = lib().ffo_p1_boundary(dim, m["nv"], m["nt"], X, el, m["nbe"], be, m["blab"], labs.size,
                               labs, kind, c0, _ptr(cf), _ptr(rowptr), _ptr(colidx), _ptr(vals),
                               _ptr(b))
    assert rc == 0, rc


def mark_bc(n, bfaces, blab, nvf, labels):
    onbc = np.zeros(n, np.uint8)
    labs = np.asarray(labels, np.int32)
    lib().ffo_mark_bc(n, blab.size, nvf, bfaces, blab, labs.size, labs, onbc)
    return onbc


def set_bc(rowptr, colidx, vals, onbc, tgv=1e30):
    rc = lib().ffo_set_bc(onbc.size, rowptr, colidx, vals, onbc, tgv)
    assert rc == 0, rc


def rhs_bc(b, onbc, tgv=1e30, g=0.0):
    lib().ffo_rhs_bc(onbc.size, onbc, tgv, g, b)


def spmv_half(rowptr, colidx, vals, x):
    """y = A x of a half (sym=1, lower triangle) matrix"""
    y = np.zeros(rowptr.size - 1)
    lib().ffo_spmv_half(rowptr.size - 1, rowptr, colidx, vals, np.ascontiguousarray(x, np.float64), y)
    return y


def spmv(rowptr, colidx, vals, x, trans=False):
    y = np.zeros(rowptr.size - 1)
    f = lib().ffo_spmv_t if trans else lib().ffo_spmv
    f(rowptr.size - 1, rowptr, colidx, vals, np.ascontiguousarray(x, np.float64), y)
    return y


def pcg(rowptr, colidx, vals, b, x0=None, eps=1e-10, itmax=0, tgv=1e30, trans=False):
    n = rowptr.size - 1
    x = np.zeros(n) if x0 is None else np.array(x0, np.float64)
    st = np.zeros(4)
    lib().ffo_pcg(n, rowptr, colidx, vals, np.ascontiguousarray(b, np.float64), x, eps,
                  itmax, tgv, int(trans), st)
    return x, dict(iters=int(st[0]), gcg0=st[1], gcg=st[2], nret=int(st[3]))


def cg2(rowptr, colidx, vals, b=None, x0=None, eps=1e-6, itmax=100, precon=0):
    """LinearCG(A*x, x, b) / AffineCG: ConjuguedGradient2 with the matrix
    product as the operator (b None: zero right-hand side).  Returns x and
    {iters, g20, g2, nret, reps2}."""
    n = rowptr.size - 1
    x = np.zeros(n) if x0 is None else np.array(x0, np.float64)
    bb = np.zeros(n) if b is None else np.ascontiguousarray(b, np.float64)
    e = np.array([eps], np.float64)
    st = np.zeros(4)
    lib().ffo_cg2(n, rowptr, colidx, vals, bb, x, e, itmax, precon, st)
    return x, dict(iters=int(st[0]), g20=st[1], g2=st[2], nret=int(st[3]), reps2=float(e[0]))


def fgmres(rowptr, colidx, vals, b, x0=None, eps=1e-6, itmax=0, restart=50, tgv=1e30):
    """solver=GMRES: SolverGMRES::dosolver over fgmres.  Returns x and
    {iters, relres, normb, restarts, converged}."""
    n = rowptr.size - 1
    x = np.zeros(n) if x0 is None else np.array(x0, np.float64)
    st = np.zeros(4)
    rc = lib().ffo_fgmres(n, rowptr, colidx, vals, np.ascontiguousarray(b, np.float64), x, eps,
                          itmax, restart, tgv, st)
    return x, dict(iters=int(st[0]), relres=st[1], normb=st[2], restarts=int(st[3]), converged=rc)


def poisson_p1_3d(n, kind=COEF_SIN3, c0=1.0, kappa_e=None, labels=(1, 2, 3, 4, 5, 6),
                  tgv=1e30, f=1.0):
    """cube(n) P1 Poisson system exactly as `matrix A = va(Vh,Vh)` and
    `b = vb(0,Vh)` build it: pattern, values, tgv rows, right-hand side."""
    m = cube(n)
    rowptr, colidx = pattern(m["nv"], m["tets"], 4)
    vals = assemble_p1_3d(m, rowptr, colidx, kind, c0, kappa_e)
    onbc = mark_bc(m["nv"], m["btri"], m["blab"], 3, labels)
    set_bc(rowptr, colidx, vals, onbc, tgv)
    b = rhs_p1_3d(m, f)
    rhs_bc(b, onbc, tgv, 0.0)
    return m, rowptr, colidx, vals, onbc, b


# ---------------------------------------------------------------- adjoint
def mass_p1_3d(m, rowptr, colidx):
    vals = np.zeros(colidx.size)
    rc = lib().ffo_assemble_p1_mass_3d(m["nv"], m["nt"], m["xyz"], m["tets"], rowptr, colidx, vals)
    assert rc == 0, rc
    return vals


def misfit_p1_3d(m, u, ud):
    return lib().ffo_p1_misfit_3d(m["nt"], m["xyz"], m["tets"], np.ascontiguousarray(u, np.float64),
                                  np.ascontiguousarray(ud, np.float64))


def p0_grad_3d(m, u, lam):
    g = np.zeros(m["nt"])
    lib().ffo_p0_grad_3d(m["nt"], m["xyz"], m["tets"], np.ascontiguousarray(u, np.float64),
                         np.ascontiguousarray(lam, np.float64), g)
    return g


def adjoint_p1_3d(n, kappa_e, eps=1e-12, labels=(1, 2, 3, 4, 5, 6)):
    """J(kappa) = 1/2 int (u - ud)^2 and dJ/dkappa_e composed from FreeFEM's
    varfs as a script would: forward solve with P0 kappa, ud from kappa = 1,
    mass varf applied to u - ud with Dirichlet rows zeroed, transposed solve
    (A'^-1, lgmat.cpp:3215-3220), and the P0 varf -int grad u . grad l."""
    m = cube(n)
    rp, ci = pattern(m["nv"], m["tets"], 4)
    onbc = mark_bc(m["nv"], m["btri"], m["blab"], 3, labels)
    b = rhs_p1_3d(m, 1.0)
    rhs_bc(b, onbc)
    v1 = assemble_p1_3d(m, rp, ci, COEF_CONST, 1.0)
    set_bc(rp, ci, v1, onbc)
    ud, _ = pcg(rp, ci, v1, b, eps=eps)
    vk = assemble_p1_3d(m, rp, ci, COEF_P0, 1.0, kappa_e)
    set_bc(rp, ci, vk, onbc)
    u, _ = pcg(rp, ci, vk, b, eps=eps)
    J = misfit_p1_3d(m, u, ud)
    mv = mass_p1_3d(m, rp, ci)
    r = spmv(rp, ci, mv, u - ud)
    r[onbc == 1] = 0.0
    lam, _ = pcg(rp, ci, vk, r, eps=eps, trans=True)
    g = p0_grad_3d(m, u, lam)
    return dict(mesh=m, u=u, ud=ud, J=J, lam=lam, g=g, rowptr=rp, colidx=ci, vals=vk, b=b, onbc=onbc)


# ---------------------------------------------------------------- P2 elasticity
def p2_nodes(m):
    en = np.zeros(10 * m["nt"], np.int32)
    nn = lib().ffo_p2_nodes(m["nv"], m["nt"], m["tets"], en)
    return int(nn), en


def p2v_dofs(m):
    nn, en = p2_nodes(m)
    ed = np.zeros(30 * m["nt"], np.int32)
    lib().ffo_p2v_elem_dofs(m["nt"], en, ed)
    return nn, en, ed


def elasticity_p2(n, lam=1.0, mu=0.5, f=(0.0, 0.0, -1.0), labels=(4,), tgv=1e30, mesh=None):
    """cube(n) (or `mesh`, a dict like cube's) [P2,P2,P2] Lame system as
    `matrix A = va(Vh,Vh)` builds it: pattern, values, tgv rows (clamped
    labels), body-force right-hand side."""
    m = cube(n) if mesh is None else mesh
    nn, en, ed = p2v_dofs(m)
    ndof = 3 * nn
    rp, ci = pattern(ndof, ed, 30)
    vals = np.zeros(ci.size)
    rc = lib().ffo_assemble_p2_elast(m["nt"], m["xyz"], m["tets"], ed, lam, mu, ndof, rp, ci, vals)
    assert rc == 0, rc
    onbc = np.zeros(ndof, np.uint8)
    labs = np.asarray(labels, np.int32)
    rc = lib().ffo_p2v_mark_bc(m["nv"], m["nt"], m["tets"], en, m["nbe"], m["btri"], m["blab"],
                              labs.size, labs, ndof, onbc)
    assert rc == 0, rc
    set_bc(rp, ci, vals, onbc, tgv)
    b = np.zeros(ndof)
    lib().ffo_rhs_p2_elast(m["nt"], m["xyz"], m["tets"], ed, np.asarray(f, np.float64), ndof, b)
    rhs_bc(b, onbc, tgv, 0.0)
    return dict(mesh=m, nnodes=nn, elem_nodes=en, elem_dofs=ed, ndof=ndof, rowptr=rp, colidx=ci,
                vals=vals, onbc=onbc, b=b)


def poisson_p2(n, kind=COEF_SIN3, c0=1.0, kappa_e=None, labels=(1, 2, 3, 4, 5, 6), tgv=1e30,
               f=1.0, mesh=None):
    """cube(n) (or `mesh`) scalar P2 system of int3d(kappa grad u . grad v) +
    int3d(f v) + on(labels, u=0) on Vh(Th,P2): pattern, values, tgv rows, rhs."""
    m = cube(n) if mesh is None else mesh
    nn, en = p2_nodes(m)
    rp, ci = pattern(nn, en, 10)
    vals = np.zeros(ci.size)
    ke = None if kappa_e is None else np.ascontiguousarray(kappa_e, np.float64)
    rc = lib().ffo_assemble_p2_lap(m["nt"], m["xyz"], m["tets"], en, kind, c0, _ptr(ke), nn, rp, ci, vals)
    assert rc == 0, rc
    onbc = np.zeros(nn, np.uint8)
    labs = np.asarray(labels, np.int32)
    rc = lib().ffo_p2_mark_bc(m["nv"], m["nt"], m["tets"], en, m["nbe"], m["btri"], m["blab"],
                             labs.size, labs, 1, nn, onbc)
    assert rc == 0, rc
    set_bc(rp, ci, vals, onbc, tgv)
    b = np.zeros(nn)
    lib().ffo_rhs_p2(m["nt"], m["xyz"], m["tets"], en, f, nn, b)
    rhs_bc(b, onbc, tgv, 0.0)
    return dict(mesh=m, nnodes=nn, elem_nodes=en, rowptr=rp, colidx=ci, vals=vals, onbc=onbc, b=b)


def ref_fespace_dofs(m, kind):
    """element dofs of FreeFEM's own FESpace3 (kind 1 P1, 2 P2, 3 [P2,P2,P2])"""
    R = ref()
    nd = ctypes.c_int()
    ndf = R.ref_fespace_dofs(m["nv"], m["xyz"], m["nt"], m["tets"], kind, None, ctypes.byref(nd))
    ed = np.zeros(m["nt"] * nd.value, np.int32)
    R.ref_fespace_dofs(m["nv"], m["xyz"], m["nt"], m["tets"], kind, ed.ctypes.data_as(ctypes.c_void_p),
                       ctypes.byref(nd))
    return int(ndf), ed
