"""Small fixed workloads of the library, for profiling runs (rocprofv3) and
quick timing:

    python3 freefem-sources_amd/ffhip/cases.py assembly_p1 160 3
    python3 freefem-sources_amd/ffhip/cases.py assembly_p2 32 3
    python3 freefem-sources_amd/ffhip/cases.py solve_p2 32

assembly_p1: cube(n) P1 Poisson pattern and values (one call) + tgv rows +
mixed solver setup, `reps` times (a fresh device mesh each time).
assembly_p2: cube(n) [P2,P2,P2] numbering + pattern + Lame values, `reps` times.
solve_p2: cube(n) Lame system, FP64 and MIXED solves in every solver layout,
with per-kernel HIP-event timing.
pcie_p1: the host-buffer path of the C-ABI on cube(n) P1: mesh upload from
host arrays, device pattern and values, CSR download to host (wall times).
plugin_p1: the sequence the FreeFEM plugin runs for `matrix A = va(Vh,Vh)`
with kappa a P1 finite element function (wall times): host mesh upload, the
assembly call with kappa's host coefficient array (FFHIP_COEF_P1), the tgv
rows, the CSR download into host arrays (FreeFEM's matrix; the solver keeps
the device copy).
copy: the copy kernel (1 GiB read + 1 GiB written per launch) whose PMC byte
counts scale the other kernels'.
"""
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

LABELS = (1, 2, 3, 4, 5, 6)


def assembly_p1(n=160, reps=3, fp64=0):
    # fp64=1: the FP64 solver's set-up (the benchmark's) instead of the MIXED one
    import ffhip
    for _ in range(reps):
        M = ffhip.cube(n)
        A = ffhip.Matrix.p1_assembled(M, ffhip.COEF_SIN3)
        tpv = M.ctx.last_ms()
        tb = A.set_bc(LABELS)
        cg = ffhip.CG(A, ffhip.CG_FP64 if fp64 else ffhip.CG_MIXED)
        print(f"P1 cube({n}) nnz {A.nnz}: pattern and values {tpv:.3f} ms, bc {tb:.3f} ms", flush=True)
        cg.free(); A.free(); M.free()


def assembly_p2(n=32, reps=3):
    import ffhip
    for _ in range(reps):
        M = ffhip.cube(n)
        sp = ffhip.P2VSpace(M)
        tn = M.ctx.last_ms()
        A = ffhip.Matrix.p2v(sp)
        tp = M.ctx.last_ms()
        tv = A.assemble_p2v_elasticity(1.0, 0.5)
        print(f"P2 cube({n}) nnz {A.nnz}: numbering {tn:.3f} ms, pattern {tp:.3f} ms, values {tv:.3f} ms",
              flush=True)
        A.free(); sp.free(); M.free()


def solve_p2(n=32):
    import torch
    import ffhip
    M = ffhip.cube(n)
    sp = ffhip.P2VSpace(M)
    A = ffhip.Matrix.p2v(sp)
    A.assemble_p2v_elasticity(1.0, 0.5)
    A.set_bc_p2v([4])
    b = torch.empty(sp.ndof, dtype=torch.float64, device="cuda")
    ffhip.rhs_p2v(sp, (0.0, 0.0, -1.0), [4], out=b)
    for layout in ("bsr3", "csr"):
        for mode in (ffhip.CG_FP64, ffhip.CG_MIXED):
            cg = ffhip.CG(A, mode, layout=layout)
            cg.set_profiling(True)
            x, st = cg.solve(b, eps=1e-10)
            print(f"{layout} {'FP64' if mode == ffhip.CG_FP64 else 'MIXED'}: {st['iters']} iterations, "
                  f"{st['ms_solve']:.1f} ms, mat-vec {st['ms_matvec_avg'] * 1e3:.1f} us, "
                  f"update {st['ms_update_avg'] * 1e3:.1f} us", flush=True)
            cg.free()


def pcie_p1(n=160, reps=2):
    import time
    import numpy as np
    import ffhip
    G = ffhip.cube(n)
    m = G.download()
    G.free()
    for _ in range(reps):
        t0 = time.perf_counter()
        M = ffhip.mesh3_upload(m["xyz"], m["elems"], m["belems"], m["blab"])
        M.ctx.sync()
        t1 = time.perf_counter()
        A = ffhip.Matrix.p1(M)
        A.assemble_p1(ffhip.COEF_SIN3)
        A.set_bc(LABELS)
        M.ctx.sync()
        t2 = time.perf_counter()
        rp, ci, va = A.download()
        t3 = time.perf_counter()
        mb_up = (m["xyz"].nbytes + m["elems"].nbytes + m["belems"].nbytes + m["blab"].nbytes) / 1e6
        mb_dn = (rp.nbytes + ci.nbytes + va.nbytes) / 1e6
        print(f"P1 cube({n}): mesh upload {mb_up:.0f} MB {1e3 * (t1 - t0):.1f} ms (host checks included), "
              f"device assembly {1e3 * (t2 - t1):.1f} ms, CSR download {mb_dn:.0f} MB "
              f"{1e3 * (t3 - t2):.1f} ms ({mb_dn / 1e3 / (t3 - t2):.1f} GB/s)", flush=True)
        A.free(); M.free()


def plugin_p1(n=160, reps=2, quiet=False):
    """returns the stage times (ms) of the last repetition"""
    import time
    import numpy as np
    import ffhip
    G = ffhip.cube(n)
    m = G.download()
    G.free()
    X = m["xyz"].reshape(-1, 3)
    kappa = 1 + 0.5 * np.sin(2 * np.pi * X[:, 0]) * np.sin(2 * np.pi * X[:, 1]) * np.sin(2 * np.pi * X[:, 2])
    out = {}
    for _ in range(reps):
        t0 = time.perf_counter()
        M = ffhip.mesh3_upload(m["xyz"], m["elems"], m["belems"], m["blab"])
        M.ctx.sync()
        t1 = time.perf_counter()
        A = ffhip.Matrix.p1_assembled(M, ffhip.COEF_P1, kappa_e=kappa)
        M.ctx.sync()
        t2 = time.perf_counter()
        A.set_bc(LABELS)
        M.ctx.sync()
        t3 = time.perf_counter()
        rp, ci, va = A.download()
        t4 = time.perf_counter()
        nnz = A.nnz
        out = {"mesh_upload_ms": 1e3 * (t1 - t0), "assembly_with_coefficients_ms": 1e3 * (t2 - t1),
               "tgv_rows_ms": 1e3 * (t3 - t2), "csr_download_ms": 1e3 * (t4 - t3),
               "total_ms": 1e3 * (t4 - t0), "nnz": nnz, "Mnnz_per_s": nnz / (t4 - t0) / 1e6}
        if not quiet:
            print(f"plugin path cube({n}): " + ", ".join(f"{k} {v:.1f}" for k, v in out.items()), flush=True)
        del rp, ci, va
        A.free(); M.free()
    return out


def copy(mib=1024, reps=5):
    """the copy kernel the PMC byte counters are scaled by (reads and writes
    mib MiB per launch)"""
    import ffhip
    ctx = ffhip.context()
    ms = ctx.bench_copy(mib << 20, reps)
    print(f"copy {mib} MiB: {ms:.3f} ms, {2 * (mib << 20) / ms / 1e6:.0f} GB/s", flush=True)


if __name__ == "__main__":
    case = sys.argv[1]
    args = [int(a) for a in sys.argv[2:]]
    {"assembly_p1": assembly_p1, "assembly_p2": assembly_p2, "solve_p2": solve_p2,
     "pcie_p1": pcie_p1, "plugin_p1": plugin_p1, "copy": copy}[case](*args)
