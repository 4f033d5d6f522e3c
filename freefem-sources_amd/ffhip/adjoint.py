"""dJ/dkappa_e for J = 1/2 int (u - ud)^2, -div(kappa grad u) = f, u = 0 on the
boundary, P1 on a device-built cube mesh, behind a torch.autograd.Function.

The objective composes the same FreeFEM operations a script would use
(FreeFEM has no adjoint routine of its own):
  forward   matrix A = va(Vh,Vh) with a P0 kappa_e (problem.cpp:1116-1417),
            u[] = A^-1*b (SolverCG, VirtualSolverCG.hpp:169-190)
  adjoint   r = M (u - ud) with the Dirichlet rows zeroed (mass varf),
            lambda[] = A'^-1*r (transposed solve, lgmat.cpp:3215-3220)
  gradient  g_e = int3d(Th)(-(grad u . grad lambda) * w), w in P0
Every step runs through libffhip's C-ABI on the GPU; element gradients are
recomputed from the coordinates in the gradient kernel, not stored.
"""
import torch

from . import (CG, CG_FP64, COEF_CONST, COEF_P0, Matrix, cube, p1_grad_kappa, p1_misfit,
               rhs_p1)

LABELS = (1, 2, 3, 4, 5, 6)


class P1Diffusion:
    """cube(n) P1 diffusion problem with a per-element coefficient."""

    def __init__(self, n, f=1.0, labels=LABELS, eps=1e-10, mode=CG_FP64):
        self.mesh = cube(n)
        self.labels, self.eps = labels, eps
        self.A = Matrix.p1(self.mesh)
        self.b = torch.empty(self.mesh.nv, dtype=torch.float64, device="cuda")
        rhs_p1(self.mesh, f, labels, out=self.b)
        # target: the solution with kappa = 1
        self.A.assemble_p1(COEF_CONST, 1.0)
        self.A.set_bc(labels)
        self.cg = CG(self.A, mode)
        self.ud, self.st_ud = self.cg.solve(self.b, eps=eps)
        self.last = {}

    def solve(self, kappa_e):
        """u for the given per-element kappa (torch cuda float64, nt values)"""
        self.A.assemble_p1(COEF_P0, 1.0, kappa_e.contiguous())
        self.A.set_bc(self.labels)
        self.cg.update_values()
        u, st = self.cg.solve(self.b, eps=self.eps)
        self.last["forward"] = st
        return u

    def objective(self, u):
        J, r = p1_misfit(self.mesh, u, self.ud, self.labels)
        return J, r

    def adjoint(self, r):
        lam, st = self.cg.solve(r, eps=self.eps, trans=True)
        self.last["adjoint"] = st
        return lam


class KappaObjective(torch.autograd.Function):
    """J(kappa_e); backward gives dJ/dkappa_e by one adjoint solve."""

    @staticmethod
    def forward(ctx, kappa_e, problem):
        u = problem.solve(kappa_e)
        J, r = problem.objective(u)
        ctx.problem = problem
        ctx.save_for_backward(u, r)
        return torch.tensor(J, dtype=torch.float64, device=kappa_e.device)

    @staticmethod
    def backward(ctx, gJ):
        u, r = ctx.saved_tensors
        problem = ctx.problem
        # the adjoint needs A(kappa) as assembled by forward; it still is
        lam = problem.adjoint(r)
        g = p1_grad_kappa(problem.mesh, u, lam)
        return gJ * g, None


def objective(kappa_e, problem):
    return KappaObjective.apply(kappa_e, problem)
