"""Writes the matrix-file fixtures under tests/golden/ with FreeFEM's own
compiled HashMatrix (oracle/_ref/libffref_ref.so, built by oracle/build_ref.sh):
the cube(2) P1 stiffness (analytic kappa) with tgv rows on labels 1-6,
assembled by HashMatrix::operator+= over the element loop, then

* p1_cube2_coo.txt:   `f << A` while A is still a hash (the COO dump, entries
                      in insertion order, 0-based);
* p1_cube2_morse.txt: `f << A` after HashMatrix::CSR (the Morse text).

The first comment line of each file carries the matrix address FreeFEM prints.

Run in a container that has the reference tree:  python tests/golden/make_io_fixtures.py
"""
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))
import ffo  # noqa: E402


def main():
    R = ffo.ref()
    m = ffo.cube(2)
    h = R.ref_hm_new(m["nv"])
    R.ref_hm_add_p1_3d(h, m["nv"], m["nt"], m["xyz"], m["tets"], ffo.COEF_SIN3, 1.0, None)
    onbc = ffo.mark_bc(m["nv"], m["btri"], m["blab"], 3, (1, 2, 3, 4, 5, 6))
    R.ref_hm_setbc(h, onbc, 1e30)
    assert R.ref_hm_write(h, os.path.join(HERE, "p1_cube2_coo.txt").encode(), 0) == 0
    assert R.ref_hm_write(h, os.path.join(HERE, "p1_cube2_morse.txt").encode(), 1) == 0
    R.ref_hm_free(h)


if __name__ == "__main__":
    main()
