"""The cube split of the oracle (ffo_mesh.c, kCubeSplit6) pinned by re-running
FreeFEM's own selection on FreeFEM's own tables: the 58 candidate simplexes
and 74 hex splits (msh3.cpp:7682-7742) are read from the reference source,
the search of BuildCube (msh3.cpp:7891-8014) — the split whose six face
diagonals all get code 0 — is replayed here, and the mesh BuildCube then
builds (vertex numbering p = i + nj j + nk k, tets in hex order with the
split's simplexes in order, boundary triangles from the tet faces
Tet::nvface with labels label[nff[k]], msh3.cpp:8024-8110) is compared entry
by entry, in order, with the oracle's: tets, the vertex order inside each
tet (which fixes P2 first-appearance numbering), boundary triangles and
labels.  A planted swap of two tets or of two vertices inside a tet must
fail the same comparison (checked on a perturbed copy).

Reads the reference tree as text; skipped where it is absent.
"""
import os
import re

import numpy as np
import pytest

import ffo

MSH3 = "/root/reference/src/fflib/msh3.cpp"
MESH3DN = "/root/reference/src/femlib/Mesh3dn.cpp"
pytestmark = pytest.mark.skipif(not os.path.exists(MSH3), reason="reference tree not present")


def _int_table(text, name, rows, cols):
    start = text.index(name)
    body = text[text.index("{", start):]
    depth, end = 0, 0
    for i, ch in enumerate(body):
        depth += ch == "{"
        depth -= ch == "}"
        if depth == 0:
            end = i
            break
    body = re.sub(r"/\*.*?\*/", "", body[:end + 1], flags=re.S)
    body = re.sub(r"//[^\n]*", "", body)
    vals = [int(v) for v in re.findall(r"-?\d+", body)]
    assert len(vals) == rows * cols, (name, len(vals))
    return np.array(vals).reshape(rows, cols)


def _tables():
    src = open(MSH3).read()
    simplexes = _int_table(src, "const int simplexes[nbsimplexes][4]", 58, 4)
    hex_decoupe = _int_table(src, "const int hex_decoupe[74][7]", 74, 7)
    nvface = _int_table(open(MESH3DN).read(), "nvfaceTet[4][3]", 4, 3)
    return simplexes, hex_decoupe, nvface


def _search_code6(simplexes, hex_decoupe, nvface):
    """BuildCube's selection loop (msh3.cpp:7913-8003), restated"""
    l8 = [0] * 8
    for c in range(2):
        for b in range(2):
            for a in range(2):
                l8[a + 2 * b + 4 * c] = (1 * (a == 0) + 2 * (a == 1) + 4 * (b == 0) + 8 * (b == 1) +
                                         16 * (c == 0) + 32 * (c == 1))
    cdiag = [0] * 8
    col = 0
    code6 = -1
    for d in range(74):
        dd = [-1] * 7
        for kt in range(2, hex_decoupe[d][0] + 1):
            nu = simplexes[hex_decoupe[d][kt]]
            for f in range(4):
                nf = [nu[nvface[f][0]], nu[nvface[f][1]], nu[nvface[f][2]]]
                lab = l8[nf[0]] & l8[nf[1]] & l8[nf[2]]
                for k in range(6):
                    if lab == (1 << k):
                        ni, j0 = k // 2, k % 2
                        oi = (1 << ni) * j0
                        j1, j2 = 1 << (ni + 1) % 3, 1 << (ni + 2) % 3
                        if j1 > j2:
                            j1, j2 = j2, j1
                        i4 = [oi, oi + j1, oi + j1 + j2, oi + j2]
                        col += 1
                        c1 = col
                        col += 1
                        c2 = col
                        cdiag[i4[0]] = cdiag[i4[2]] = c1
                        cdiag[i4[1]] = cdiag[i4[3]] = c2
                        for e in range(3):
                            i0, i1 = nf[e], nf[(e + 1) % 3]
                            if cdiag[i0] == cdiag[i1]:
                                di = 0 if cdiag[i0] == c1 else 1
                                if dd[k] < 0:
                                    dd[k] = di
                                else:
                                    assert dd[k] == di
        code = sum(dd[i] * (1 << i) for i in range(6))
        if code == 0:
            code6 = d
    return code6


def _build_cube(nx, ny, nz, labels, simplexes, split, nvface):
    """BuildCube's element and boundary loops (msh3.cpp:8024-8110), restated"""
    nj, nk = nx + 1, (nx + 1) * (ny + 1)
    lab = np.zeros((nz + 1) * (ny + 1) * (nx + 1), np.int64)
    for k in range(nz + 1):
        for j in range(ny + 1):
            for i in range(nx + 1):
                lab[i + nj * j + nk * k] = (1 * (i == 0) + 2 * (i == nx) + 4 * (j == 0) +
                                            8 * (j == ny) + 16 * (k == 0) + 32 * (k == nz))
    nff = [3, 1, 0, 2, 4, 5]
    tets, btri, blab = [], [], []
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                n = [(i + a) + nj * (j + b) + nk * (k + c)
                     for c in range(2) for b in range(2) for a in range(2)]
                for d in range(6):
                    nu = [n[simplexes[split[d + 1]][q]] for q in range(4)]
                    tets.append(nu)
                    for f in range(4):
                        nf = [nu[nvface[f][0]], nu[nvface[f][1]], nu[nvface[f][2]]]
                        lf = lab[nf[0]] & lab[nf[1]] & lab[nf[2]]
                        for kk in range(6):
                            if lf == (1 << kk):
                                btri.append(nf)
                                blab.append(labels[nff[kk]])
    return (np.array(tets, np.int32).ravel(), np.array(btri, np.int32).ravel(),
            np.array(blab, np.int32))


def test_reference_search_selects_split_70():
    simplexes, hex_decoupe, nvface = _tables()
    code6 = _search_code6(simplexes, hex_decoupe, nvface)
    assert code6 == 70
    assert hex_decoupe[70].tolist() == [6, 7, 15, 19, 27, 26, 11]


@pytest.mark.parametrize("dims", [(1, 1, 1), (2, 3, 4), (5, 5, 5)])
def test_oracle_cube_is_buildcube(dims):
    simplexes, hex_decoupe, nvface = _tables()
    split = hex_decoupe[_search_code6(simplexes, hex_decoupe, nvface)]
    labels = (11, 12, 13, 14, 15, 16)
    tets, btri, blab = _build_cube(*dims, labels, simplexes, split, nvface)
    o = ffo.cube(*dims, labels=labels)
    assert np.array_equal(o["tets"], tets)  # tet order and vertex order inside each tet
    assert np.array_equal(o["btri"], btri) and np.array_equal(o["blab"], blab)
    # planted mistakes fail the same comparison
    t2 = tets.reshape(-1, 4).copy()
    t2[[0, 1]] = t2[[1, 0]]
    assert not np.array_equal(o["tets"], t2.ravel())
    t3 = tets.reshape(-1, 4).copy()
    t3[0, [1, 2]] = t3[0, [2, 1]]
    assert not np.array_equal(o["tets"], t3.ravel())
