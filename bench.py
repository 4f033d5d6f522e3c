"""Benchmark of the assembly + Jacobi-CG path on one MI355X.

Workload: 3D P1 Poisson on cube(160,160,160) (4,173,281 vertices,
24,576,000 tets, 61,979,041 nnz), kappa(x) = 1+0.5 sin(2 pi x) sin(2 pi y)
sin(2 pi z) at the 14 Tet_5 points, f = 1, u = 0 on labels 1..6 (tgv = 1e30),
CG to eps = 1e-10 in FreeFEM's criterion (g, D^-1 g) < eps^2 (g0, D^-1 g0).

One step = what `matrix A = va(Vh,Vh,solver=CG); b = vb(0,Vh); u[] = A^-1*b`
does, twice solved, plus one adjoint gradient evaluation:
  assembly   : CSR pattern + element values + Dirichlet rows + solver setup
  rhs        : b
  solve FP64 : plain FP64 Jacobi-PCG
  solve MIXED: FP32 matrix/inner CG with FP64 residual refinement
  adjoint    : dJ/dkappa_e for J = 1/2 int (u-ud)^2 with lognormal kappa_e
               (target ud solved with kappa = 1; forward, transposed, gradient)
The mesh is built on the device before the timed region (inputs resident in
HBM).  value = assembled Mnnz/s (nnz / assembly time); the CG times, the
iteration counts and the accuracies are reported beside it.

roofline: the dominant kernel is the FP64 CG mat-vec (k_pcg_matvec: SpMV of
the direction H with the dot product (H, AH)), timed per launch with HIP
events on the library stream during the profiled solves of the timed steps.
Algorithmic bytes per launch = those of an FP64 CSR SpMV with int32 column
indices and row pointers, x read once and y written once: 12 nnz + 20 n.
The layout the solver streams moves less (FP64 value + 16-bit column delta
per entry in the sliced ELL layout, H read once, A H written once:
10 nnz + 16 n); its rate is reported beside it (`streamed_layout`).

cpu_baseline: the metric's cube(160) system on the host.  One core:
FreeFEM's own compiled femlib code (oracle/_ref: HashMatrix accumulation +
SetBC + the reference ConjugueGradient with its COO mat-vec) driven by the
oracle's element loop — the whole assembly, and 64 of the 640 CG iterations
(the per-iteration cost at full size; the 640-iteration time is 10x it); the
oracle port when the reference subset is not built.  All cores: the oracle's
OpenMP variant, whole assembly and whole solve (`all_cores`).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "freefem-sources_amd"))

METRIC = "assembly Mnnz/s + CG solve s to 1e-10 (FP64 vs mixed), 3D P1 Poisson cube(160)"
LABELS = (1, 2, 3, 4, 5, 6)
HBM_PEAK_GBS = 8000.0


def measured_traffic(kernel="k_pcg_matvec", layout="sell16"):
    """HBM bytes per launch from the committed PMC summary (FETCH_SIZE and
    WRITE_SIZE scaled by a copy kernel of known size) for the layout the
    solver streamed, or None."""
    path = os.path.join(ROOT, "profiles", "k_pcg_matvec_cube160.md")
    try:
        for line in open(path):
            if line.startswith(f"traffic_bytes_per_launch {kernel} {layout}:"):
                return int(line.split(":")[1])
    except OSError:
        pass
    return None


def cpu_baseline(n=160, iters_1core=64):
    """The same cube(n) P1 system on the host: FreeFEM's compiled femlib code
    on one core (the oracle port where the reference subset is not built),
    and the oracle's OpenMP variant on all host cores."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import ffo  # the oracle: the CPU baseline only

    m = ffo.cube(n)
    onbc = ffo.mark_bc(m["nv"], m["btri"], m["blab"], 3, LABELS)
    b = ffo.rhs_p1_3d(m, 1.0)
    ffo.rhs_bc(b, onbc)
    x = np.zeros(m["nv"])
    st = np.zeros(4)
    if ffo.ref_available():
        R = ffo.ref()
        t0 = time.perf_counter()
        h = R.ref_hm_new(m["nv"])
        R.ref_hm_add_p1_3d(h, m["nv"], m["nt"], m["xyz"], m["tets"], ffo.COEF_SIN3, 1.0, None)
        R.ref_hm_setbc(h, onbc, 1e30)
        t_asm = time.perf_counter() - t0
        nnz = R.ref_hm_nnz(h)
        t0 = time.perf_counter()
        R.ref_hm_cg(h, b, x, 1e-30, iters_1core, 1e30, 0, st)
        t_cg = time.perf_counter() - t0
        R.ref_hm_free(h)
        kind = "reference"
        what = "FreeFEM femlib HashMatrix += element matrices, SetBC, ConjugueGradient (COO addMatMul)"
    else:
        t0 = time.perf_counter()
        rp, ci = ffo.pattern(m["nv"], m["tets"], 4)
        v = ffo.assemble_p1_3d(m, rp, ci, ffo.COEF_SIN3)
        ffo.set_bc(rp, ci, v, onbc)
        t_asm = time.perf_counter() - t0
        nnz = ci.size
        t0 = time.perf_counter()
        ffo.pcg(rp, ci, v, b, eps=1e-30, itmax=iters_1core)
        t_cg = time.perf_counter() - t0
        del rp, ci, v
        kind = "port"
        what = "oracle C restatement"
    ms_it = t_cg / iters_1core * 1e3
    # all host cores: the oracle's OpenMP variant, the whole solve to 1e-10
    threads = ffo.omp_lib().ffo_omp_max_threads()
    t0 = time.perf_counter()
    S = ffo.OmpP1(m, ffo.COEF_SIN3)
    S.set_bc(onbc)
    t_asm_all = time.perf_counter() - t0
    t0 = time.perf_counter()
    _, so = S.pcg(b, eps=1e-10)
    t_cg_all = time.perf_counter() - t0
    S.free()
    return {
        "value": round(nnz / t_asm / 1e6, 4), "unit": "Mnnz/s", "cores": 1, "kind": kind,
        "sample": f"cube({n}) P1 kappa(x), the metric's system: assembly of {nnz} nnz in {t_asm:.2f} s "
                  f"+ {iters_1core} of the 640 Jacobi-CG iterations in {t_cg:.2f} s ({what})",
        "assembly_s": round(t_asm, 3),
        "cg_ms_per_iter": round(ms_it, 3),
        "cg_s_640it": round(ms_it * 640 / 1e3, 2),
        "all_cores": {"value": round(nnz / t_asm_all / 1e6, 4), "unit": "Mnnz/s", "cores": threads,
                      "kind": "port", "assembly_s": round(t_asm_all, 3), "cg_s": round(t_cg_all, 3),
                      "cg_iters": so["iters"],
                      "sample": f"cube({n}) P1 kappa(x): pattern, element-order values and tgv rows "
                                f"(oracle OpenMP variant), then Jacobi-CG to 1e-10, {threads} threads"},
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=160)
    ap.add_argument("--no-cpu", action="store_true",
                    help="the timed steps only: no CPU baseline and no smaller configurations "
                         "(the form the profiles are taken with)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: runs on one GPU only; the multi-GPU sharding of the path (contiguous row "
              "slabs, one slab-face exchange per mat-vec) is not launched", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    import ffhip

    n = args.n
    mesh = ffhip.cube(n)
    b = torch.empty(mesh.nv, dtype=torch.float64, device="cuda")

    def step(record):
        # a fresh device mesh per step (generated before the timer starts), so
        # the timed assembly includes the vertex-to-element incidence lists
        nonlocal mesh
        mesh.free()
        mesh = ffhip.cube(n)
        mesh.ctx.sync()
        t0 = time.perf_counter()
        A = ffhip.Matrix.p1_assembled(mesh, ffhip.COEF_SIN3)
        t_pv = mesh.ctx.last_ms()
        t_bc = A.set_bc(LABELS)
        cg64 = ffhip.CG(A, ffhip.CG_FP64)
        t_asm = time.perf_counter() - t0
        t0 = time.perf_counter()
        ffhip.rhs_p1(mesh, 1.0, LABELS, out=b)
        t_rhs = time.perf_counter() - t0
        # the production solves replay captured 16-iteration chunks (the first
        # solve of a solver captures them); a second, profiled solve launches
        # the same kernels one by one between HIP events for the per-kernel
        # durations
        t0 = time.perf_counter()
        x64, s64 = cg64.solve(b, eps=1e-10)
        t_64 = time.perf_counter() - t0
        t0 = time.perf_counter()
        cgm = ffhip.CG(A, ffhip.CG_MIXED)
        t_msetup = time.perf_counter() - t0
        t0 = time.perf_counter()
        xm, sm = cgm.solve(b, eps=1e-10)
        t_m = time.perf_counter() - t0
        rel = ((xm - x64).norm() / x64.norm()).item()
        if record:
            cg64.set_profiling(True)
            _, p64 = cg64.solve(b, eps=1e-10)
            cgm.set_profiling(True)
            _, pm = cgm.solve(b, eps=1e-10)
        else:
            p64, pm = s64, sm
        out = dict(nnz=A.nnz, n=A.n, t_asm=t_asm, t_pv_dev=t_pv, t_bc_dev=t_bc,
                   t_rhs=t_rhs, t_64=t_64, t_m=t_m, t_msetup=t_msetup, s64=s64, sm=sm, rel=rel,
                   p64=p64, pm=pm, layout64=cg64.layout(), layoutm=cgm.layout())
        cg64.free(); cgm.free(); A.free()
        return out

    # one adjoint dJ/dkappa_e evaluation with lognormal kappa_e (same mesh):
    # ud solve (kappa = 1), forward solve, misfit, transposed solve, gradient
    import ffhip.adjoint as adj
    kappa_e = torch.exp(0.5 * torch.randn(mesh.nt, dtype=torch.float64, device="cuda",
                                          generator=torch.Generator(device="cuda").manual_seed(20261202)))

    def adjoint_eval():
        t0 = time.perf_counter()
        P = adj.P1Diffusion(n, eps=1e-10)
        t_ud = time.perf_counter() - t0
        kt = kappa_e.clone().requires_grad_(True)
        t0 = time.perf_counter()
        J = adj.objective(kt, P)
        t_fwd = time.perf_counter() - t0
        t0 = time.perf_counter()
        J.backward()
        torch.cuda.synchronize()
        t_bwd = time.perf_counter() - t0
        out = dict(J=J.item(), t_setup_ud=t_ud, t_forward=t_fwd, t_adjoint_grad=t_bwd,
                   iters_fwd=P.last["forward"]["iters"], iters_adj=P.last["adjoint"]["iters"],
                   gnorm=kt.grad.norm().item())
        P.cg.free(); P.A.free(); P.mesh.free()
        return out

    # the [P2,P2,P2] Lame configuration: cube(32), lambda = 1, mu = 0.5,
    # f = (0,0,-1), clamped label 4, FP32 inner CG with FP64 refinement
    def p2_eval(n2=32):
        M2 = ffhip.cube(n2)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        sp = ffhip.P2VSpace(M2)
        t_num = M2.ctx.last_ms()
        A2 = ffhip.Matrix.p2v(sp)
        t_pat = M2.ctx.last_ms()
        t_val = A2.assemble_p2v_elasticity(1.0, 0.5)
        t_bc = A2.set_bc_p2v([4])
        cgm = ffhip.CG(A2, ffhip.CG_MIXED)
        t_asm = time.perf_counter() - t0
        b2 = torch.empty(sp.ndof, dtype=torch.float64, device="cuda")
        ffhip.rhs_p2v(sp, (0.0, 0.0, -1.0), [4], out=b2)
        t0 = time.perf_counter()
        xm, sm = cgm.solve(b2, eps=1e-10)
        t_m = time.perf_counter() - t0
        cg64 = ffhip.CG(A2, ffhip.CG_FP64)
        t0 = time.perf_counter()
        x64, s64 = cg64.solve(b2, eps=1e-10)
        t_64 = time.perf_counter() - t0
        out = dict(ndof=sp.ndof, nnz=A2.nnz, layout=cgm.layout()[0], t_asm=t_asm,
                   dev=dict(numbering=t_num, pattern=t_pat, values=t_val, bc=t_bc),
                   t_m=t_m, t_64=t_64, sm=sm, s64=s64,
                   rel=((xm - x64).norm() / x64.norm()).item())
        cgm.free(); cg64.free(); A2.free(); sp.free(); M2.free()
        return out

    # the two smaller configurations of the same path, once, after the timed
    # steps: 2D P1 square(200) (kappa = 1) and 3D P1 cube(64) (analytic kappa),
    # assembly (pattern, values, tgv rows, FP64 solver setup) and the first
    # FP64 solve to 1e-10, beside FreeFEM's own times for the same workloads
    # on one CPU core (its assembly, and its ConjugueGradient iterations/time)
    def small_eval():
        out = {}
        cases = (("square200_2d", lambda: ffhip.square(200), ffhip.COEF_CONST, (1, 2, 3, 4),
                  dict(assembly_s=0.153, cg_iters=414, cg_s=0.276)),
                 ("cube64_3d", lambda: ffhip.cube(64), ffhip.COEF_SIN3, LABELS,
                  dict(assembly_s=8.78, cg_iters=254, cg_s=2.72)))
        for name, mk, kind, labels, ff in cases:
            Ms = mk()
            Ms.ctx.sync()
            t0 = time.perf_counter()
            As = ffhip.Matrix.p1(Ms)
            As.assemble_p1(kind, 1.0)
            As.set_bc(labels)
            cgs = ffhip.CG(As, ffhip.CG_FP64)
            ta = time.perf_counter() - t0
            bs = torch.empty(Ms.nv, dtype=torch.float64, device="cuda")
            ffhip.rhs_p1(Ms, 1.0, labels, out=bs)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            _, ss = cgs.solve(bs, eps=1e-10)
            tc = time.perf_counter() - t0
            out[name] = dict(nnz=As.nnz, assembly_s=round(ta, 5), cg_iters=ss["iters"], cg_s=round(tc, 5),
                             freefem_1core=ff)
            cgs.free(); As.free(); Ms.free()
        return out

    for _ in range(args.warmup):  # the whole step: its first calls load code and grow buffers
        step(False)
        adjoint_eval()
        p2_eval()
    torch.cuda.synchronize()
    T0 = time.perf_counter()
    res = []
    for _ in range(args.steps):
        r = step(True)
        r["adj"] = adjoint_eval()
        r["p2"] = p2_eval()
        res.append(r)
    torch.cuda.synchronize()
    T = time.perf_counter() - T0

    nnz, nv = res[0]["nnz"], res[0]["n"]
    mean = lambda k: float(np.mean([r[k] for r in res]))
    t_asm = mean("t_asm")
    mv_ms = float(np.mean([r["p64"]["ms_matvec_avg"] for r in res]))
    mv_ms32 = float(np.mean([r["pm"]["ms_matvec_avg"] for r in res]))
    # algorithmic bytes of the FP64 SpMV: int32 CSR (value, column index, row
    # pointer), x read once, y written once
    bytes_alg = 12 * nnz + 20 * nv
    # bytes a launch moves in the streamed layout: value + column index per
    # nonzero (16-bit column deltas in "sell16"), H read, A H written (the FP32
    # mat-vec of the mixed solve: p read, q written)
    ib64 = 2 if res[-1]["layout64"][0] == "sell16" else 4
    ib32 = 2 if res[-1]["layoutm"][0] == "sell16" else 4
    bytes_mv = (8 + ib64) * nnz + 16 * nv
    bytes_mv32 = (4 + ib32) * nnz + 8 * nv
    achieved = bytes_alg / (mv_ms * 1e-3) / 1e9
    s64, sm = res[-1]["s64"], res[-1]["sm"]
    copy_bytes = 1 << 30
    copy_ms = mesh.ctx.bench_copy(copy_bytes, 20)
    copy_gbs = 2 * copy_bytes / (copy_ms * 1e-3) / 1e9  # read + write
    line = {
        "metric": METRIC,
        "value": round(nnz / t_asm / 1e6, 2),
        "unit": "Mnnz/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(T / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: cube mesh generated on the device (BuildCube kind=6 numbering), "
                "analytic kappa, no files",
        "config": {"workload": f"3D P1 Poisson cube({n},{n},{n}), kappa=1+0.5sin(2pi x)sin(2pi y)sin(2pi z), "
                               f"f=1, Dirichlet tgv=1e30 on labels 1-6, Jacobi-CG eps=1e-10",
                   "nv": nv, "nnz": nnz},
        "assembly_s": round(t_asm, 5),
        "assembly_device_ms": {"pattern_and_values": round(mean("t_pv_dev"), 3),
                               "bc": round(mean("t_bc_dev"), 3)},
        "rhs_s": round(mean("t_rhs"), 5),
        "cg_fp64_s": round(mean("t_64"), 5),
        "cg_mixed_s": round(mean("t_m"), 5),
        "cg_launch": "hipGraph replay of captured 16-iteration chunks (first solve captures)",
        "cg_kernel_us_profiled": {
            "k_pcg_dirvec": round(float(np.mean([r["p64"]["ms_direction_avg"] for r in res])) * 1e3, 1),
            "k_pcg_matvec": round(mv_ms * 1e3, 1),
            "k_pcg_update": round(float(np.mean([r["p64"]["ms_update_avg"] for r in res])) * 1e3, 1),
            "k_mix_dirvec": round(float(np.mean([r["pm"]["ms_direction_avg"] for r in res])) * 1e3, 1),
            "k_mix_matvec": round(mv_ms32 * 1e3, 1),
            "k_mix_update": round(float(np.mean([r["pm"]["ms_update_avg"] for r in res])) * 1e3, 1)},
        "cg_mixed_setup_s": round(mean("t_msetup"), 5),
        "mixed_speedup_vs_fp64": round(mean("t_64") / mean("t_m"), 3),
        "cg_fp64_iters": s64["iters"],
        "cg_mixed_inner_iters": sm["iters"],
        "cg_mixed_fp64_refinements": sm["refinements"],
        "final_gcg_over_gcg0": {"fp64": s64["true_gcg"] / s64["gcg0"], "mixed": sm["true_gcg"] / sm["gcg0"]},
        "true_relres": {"fp64": s64["true_relres"], "mixed": sm["true_relres"]},
        "mixed_relerr_vs_fp64": res[-1]["rel"],
        "adjoint": {"J": res[-1]["adj"]["J"], "grad_norm": res[-1]["adj"]["gnorm"],
                    "ud_solve_s": round(float(np.mean([r["adj"]["t_setup_ud"] for r in res])), 5),
                    "forward_s": round(float(np.mean([r["adj"]["t_forward"] for r in res])), 5),
                    "adjoint_and_gradient_s": round(float(np.mean([r["adj"]["t_adjoint_grad"] for r in res])), 5),
                    "iters_forward": res[-1]["adj"]["iters_fwd"], "iters_adjoint": res[-1]["adj"]["iters_adj"]},
        "p2_elasticity": {
            "workload": "cube(32) [P2,P2,P2] Lame lambda=1 mu=0.5 f=(0,0,-1), clamped label 4",
            "ndof": res[-1]["p2"]["ndof"], "nnz": res[-1]["p2"]["nnz"],
            "assembly_s": round(float(np.mean([r["p2"]["t_asm"] for r in res])), 5),
            "assembly_Mnnz_per_s": round(res[-1]["p2"]["nnz"] / float(np.mean([r["p2"]["t_asm"] for r in res])) / 1e6, 1),
            "assembly_device_ms": {k: round(float(np.mean([r["p2"]["dev"][k] for r in res])), 3)
                                   for k in ("numbering", "pattern", "values", "bc")},
            "solver_layout": res[-1]["p2"]["layout"],
            "cg_mixed_s": round(float(np.mean([r["p2"]["t_m"] for r in res])), 5),
            "cg_fp64_s": round(float(np.mean([r["p2"]["t_64"] for r in res])), 5),
            "cg_mixed_inner_iters": res[-1]["p2"]["sm"]["iters"],
            "cg_mixed_fp64_refinements": res[-1]["p2"]["sm"]["refinements"],
            "cg_fp64_iters": res[-1]["p2"]["s64"]["iters"],
            "true_relres": {"fp64": res[-1]["p2"]["s64"]["true_relres"],
                            "mixed": res[-1]["p2"]["sm"]["true_relres"]},
            "mixed_relerr_vs_fp64": res[-1]["p2"]["rel"]},
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": measured_traffic("k_pcg_matvec", res[-1]["layout64"][0]) if n == 160 else None,
            "kernel": f"k_pcg_matvec (FP64 {res[-1]['layout64'][0]} mat-vec of the CG direction + (H, AH))",
            "bytes_per_launch": bytes_alg,
            "bytes_definition": "FP64 CSR SpMV, int32 indices: 12 nnz + 20 n",
            "avg_launch_ms": round(mv_ms, 5),
            "streamed_layout": {"layout": res[-1]["layout64"][0], "bytes_per_launch": bytes_mv,
                                "bytes_definition": f"{8 + ib64} nnz + 16 n",
                                "achieved_GBs": round(bytes_mv / (mv_ms * 1e-3) / 1e9, 1)},
            "copy_reference_GBs": round(copy_gbs, 1),
            "fp32_matvec": {"kernel": "k_mix_matvec", "bytes_per_launch": bytes_mv32,
                            "avg_launch_ms": round(mv_ms32, 5),
                            "achieved_GBs": round(bytes_mv32 / (mv_ms32 * 1e-3) / 1e9, 1)},
        },
        "cpu_baseline": None,
    }
    # the plugin's sequence on the same workload: host mesh and P1 coefficient
    # array in, device assembly, CSR back to host (not part of `value`)
    from ffhip import cases
    pp = cases.plugin_p1(n, reps=2, quiet=True)
    line["plugin_path"] = {k: (round(v, 2) if isinstance(v, float) else v) for k, v in pp.items()}
    line["plugin_path"]["sequence"] = ("mesh upload from host arrays, assembly with kappa's P1 host "
                                       "coefficient array, tgv rows, CSR download to host")
    if not args.no_cpu:
        line["other_configs"] = small_eval()
        line["cpu_baseline"] = cpu_baseline()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
