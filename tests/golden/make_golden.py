"""Writes the golden fixtures under tests/golden/ from FreeFEM's own compiled
femlib code (oracle/_ref/libffref_ref.so, built by oracle/build_ref.sh from the
reference tree): HashMatrix accumulation + SetBC + CSR sort, and the reference
ConjugueGradient.  The element loop feeding HashMatrix is the driver's
restatement of Element_Op (problem.cpp:6384-6437); meshes come from the
oracle's BuildCube restatement.

Run in a container that has the reference tree:  python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))
import ffo  # noqa: E402


def p1_case(n, kind, name, kappa_e=None):
    R = ffo.ref()
    m = ffo.cube(n)
    h = R.ref_hm_new(m["nv"])
    ke = None if kappa_e is None else kappa_e.ctypes.data_as(ffo.ctypes.c_void_p)
    R.ref_hm_add_p1_3d(h, m["nv"], m["nt"], m["xyz"], m["tets"], kind, 1.0, ke)
    onbc = ffo.mark_bc(m["nv"], m["btri"], m["blab"], 3, (1, 2, 3, 4, 5, 6))
    R.ref_hm_setbc(h, onbc, 1e30)
    b = ffo.rhs_p1_3d(m, 1.0)
    ffo.rhs_bc(b, onbc, 1e30, 0.0)
    # CG in the reference's hash (insertion) order, as `u[] = A^-1*b` runs it
    x = np.zeros(m["nv"])
    st = np.zeros(4)
    R.ref_hm_cg(h, b, x, 1e-10, 0, 1e30, 0, st)
    nnz = R.ref_hm_nnz(h)
    rp = np.zeros(m["nv"] + 1, np.int64)
    ci = np.zeros(nnz, np.int32)
    va = np.zeros(nnz)
    R.ref_hm_csr(h, rp, ci, va)
    R.ref_hm_free(h)
    out = dict(n=n, kind=kind, rowptr=rp, colidx=ci, vals=va, b=b, x=x, iters=int(st[0]),
               onbc=onbc)
    if kappa_e is not None:
        out["kappa_e"] = kappa_e
    np.savez_compressed(os.path.join(HERE, name), **out)
    print(name, "nv", m["nv"], "nnz", nnz, "iters", int(st[0]))


if __name__ == "__main__":
    p1_case(4, ffo.COEF_SIN3, "p1_cube4_sin3.npz")
    p1_case(6, ffo.COEF_CONST, "p1_cube6_const.npz")
    rng = np.random.default_rng(20261202)
    m = ffo.cube(5)
    p1_case(5, ffo.COEF_P0, "p1_cube5_p0.npz", np.exp(0.5 * rng.standard_normal(m["nt"])))
