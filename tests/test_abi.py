"""CPU checks of the boundary: libffhip.so loads and exports every entry
point include/ffhip.h declares (no compute call: there is no GPU here), and
the FreeFEM plugin type-checks against FreeFEM's own headers where that tree
is present."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "freefem-sources_amd", "libffhip.so")


def declared():
    txt = open(os.path.join(ROOT, "include", "ffhip.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(ffhip_[a-z0-9_]+)\s*\(", txt)))


def test_header_declares_the_path():
    names = declared()
    for must in ("ffhip_init", "ffhip_mesh3_cube", "ffhip_pattern_p1", "ffhip_assemble_p1",
                 "ffhip_set_bc", "ffhip_rhs_p1", "ffhip_cg_create", "ffhip_cg_solve", "ffhip_spmv",
                 "ffhip_cg_solve_affine"):
        assert must in names


def test_library_exports_every_declared_symbol():
    assert os.path.exists(LIB), "libffhip.so not built"
    L = ctypes.CDLL(LIB)
    missing = [n for n in declared() if not hasattr(L, n)]
    assert not missing, missing
    # no compute: only the pure-host entry points
    L.ffhip_version.restype = ctypes.c_char_p
    assert L.ffhip_version().startswith(b"ffhip")


@pytest.mark.skipif(not os.path.isdir("/root/reference/src/fflib"), reason="FreeFEM tree absent")
def test_plugin_typechecks_against_freefem_headers():
    r = subprocess.run(["sh", os.path.join(ROOT, "freefem-sources_amd", "plugin", "check_syntax.sh")],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
