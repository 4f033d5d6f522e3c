"""ffhip_p1_misfit_noalloc (the entry point that allocates no device memory:
its temporaries live in the library's fixed per-device scratch array) against
ffhip_p1_misfit, through the C-ABI: J and r = M (u - ud) with zeroed
Dirichlet rows bit for bit equal on two calls in a row, the second on a
smaller mesh with other labels (the scratch keeps the first call's
contents); sizes beyond its maximum are refused before any launch.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
ffhip = pytest.importorskip("ffhip")


def _misfit(fn, mesh, u, ud, labels):
    lab = np.asarray(labels, np.int32)
    J = ctypes.c_double()
    r = u.new_empty(u.shape)
    rc = fn(mesh.ctx.h, mesh.h, ffhip._dp(u), ffhip._dp(ud), ffhip._hp(lab), lab.size,
            ctypes.byref(J), ffhip._dp(r))
    return rc, J.value, r


@pytest.mark.parametrize("sizes", [((9, 8, 7), (1, 2, 3, 4, 5, 6), (4, 4, 3), (1, 3, 5))])
def test_noalloc_misfit_equals_allocating_entry_point(sizes):
    import torch
    L = ffhip.lib()
    n1, lab1, n2, lab2 = sizes
    for dims, labels, seed in ((n1, lab1, 1), (n2, lab2, 2)):
        mesh = ffhip.cube(*dims)
        ffhip.Matrix.p1(mesh)  # builds the mesh's vertex incidence
        g = torch.Generator().manual_seed(seed)
        u = torch.randn(mesh.nv, generator=g, dtype=torch.float64).cuda()
        ud = torch.randn(mesh.nv, generator=g, dtype=torch.float64).cuda()
        rc1, J1, r1 = _misfit(L.ffhip_p1_misfit, mesh, u, ud, labels)
        rc2, J2, r2 = _misfit(L.ffhip_p1_misfit_noalloc, mesh, u, ud, labels)
        assert rc1 == 0 and rc2 == 0
        assert J2 == J1 and J1 > 0
        assert torch.equal(r1, r2)
        assert (r2 == 0).sum().item() > 0  # Dirichlet rows zeroed


def test_noalloc_misfit_refuses_beyond_its_maximum():
    import torch
    L = ffhip.lib()
    mesh = ffhip.cube(3)
    u = torch.zeros(mesh.nv, dtype=torch.float64, device="cuda")
    rc, _, _ = _misfit(L.ffhip_p1_misfit_noalloc, mesh, u, u, tuple(range(1, 66)))
    assert rc == ffhip.ERR_ARG  # 65 labels, more than 64
    fresh = ffhip.cube(3)  # no vertex incidence yet: r cannot be formed without allocating
    rc, _, _ = _misfit(L.ffhip_p1_misfit_noalloc, fresh, u, u, (1,))
    assert rc == ffhip.ERR_ARG
