"""Differentiable miniFE solve: per-element conductivity kappa -> u(kappa).

The assembled operator is linear in each element's conductivity k (Hex8.hpp:273:
K_e = sum_gp k detJ w grad psi^T grad psi), so with the reference's Dirichlet
treatment (u=0 rows first, then u=1 rows; SparseMatrix_functions.hpp:431-480)
the free unknowns solve

    K_FF(kappa) u_F = f_F - K_FL(kappa) 1_L,   L = bc_rows_1 minus bc_rows_0,

(the x=1 face's edge nodes are in both sets; the first pass zeroes their
columns with value 0, so they do not lift).  For a loss L(u) the adjoint is
A lambda = dL/du with Dirichlet rows zeroed (A symmetric, same CG), and

    dL/dkappa_e = -lambda_e^T K_e(1) utilde_e,   utilde = u with bc_rows_0 set to 0.

Forward and adjoint solves are FP64 CG on the GPU; the element gradient is the
`mfe_element_gradient` kernel.  `MiniFESolve` exposes this to PyTorch.
"""
from __future__ import annotations

import numpy as np

from . import (CSRMatrix, Context, Vector, assemble_FE_data, cg_solve, default_context, dot,
               element_gradient, generate_matrix_structure, impose_dirichlet, make_local_matrix,
               matvec, simple_mesh_description, waxpby)


class KappaProblem:
    """The box problem with per-element conductivities, assembled on the GPU."""

    def __init__(self, nx: int, ny: int, nz: int, ctx: Context | None = None):
        self.dims = (nx, ny, nz)
        self.ctx = ctx or default_context()
        self.mesh = simple_mesh_description([0, nx, 0, ny, 0, nz])
        self.ne = nx * ny * nz
        self.n = (nx + 1) * (ny + 1) * (nz + 1)
        self.bc0 = self.mesh.bc_rows_0
        self.bc_all = np.union1d(self.mesh.bc_rows_0, self.mesh.bc_rows_1)
        self.A = None
        self.b = None

    def assemble(self, kappa: Vector):
        nx, ny, nz = self.dims
        A = CSRMatrix(self.ctx)
        if generate_matrix_structure(self.mesh, A) != 0:
            raise MemoryError("generate_matrix_structure: out of device memory")
        b = Vector(A.nrows, self.ctx)
        assemble_FE_data(self.mesh, A, b, kappa=kappa)
        impose_dirichlet(0.0, A, b, nx + 1, ny + 1, nz + 1, self.mesh.bc_rows_0)
        impose_dirichlet(1.0, A, b, nx + 1, ny + 1, nz + 1, self.mesh.bc_rows_1)
        make_local_matrix(A)
        self.A, self.b = A, b
        return A, b

    def solve(self, rhs: Vector, x: Vector, rel_tol: float, max_iter: int = 50000):
        r = cg_solve(self.A, rhs, x, max_iter=max_iter, tolerance=rel_tol, relative=True)
        if r.breakdown:
            raise RuntimeError("CG breakdown in the kappa solve")
        return r

    def resolve_from(self, x: Vector, reduction: float, max_iter: int = 50000):
        """Warm-started solve of A x = b from the given x, stopping once the
        residual has dropped by `reduction` relative to ||b - A x0|| (for a
        nearby kappa, e.g. a perturbed or updated field)."""
        r = Vector(self.n, self.ctx)
        matvec(self.A, x, r)
        waxpby(1.0, self.b, -1.0, r, r)
        r0 = float(np.sqrt(dot(r, r)))
        if r0 == 0.0:
            return None
        res = cg_solve(self.A, self.b, x, max_iter=max_iter, tolerance=reduction * r0)
        if res.breakdown:
            raise RuntimeError("CG breakdown in the kappa re-solve")
        return res


def loss_and_gradient(nx, ny, nz, kappa: np.ndarray, u_target: np.ndarray, rel_tol=1e-13,
                      ctx: Context | None = None):
    """L = 0.5 ||u(kappa) - u_target||^2 and dL/dkappa (host arrays in and out)."""
    P = KappaProblem(nx, ny, nz, ctx)
    P.assemble(Vector.from_numpy(kappa, P.ctx))
    u = Vector(P.n, P.ctx)
    P.solve(P.b, u, rel_tol)
    uh = u.to_numpy()
    d = uh - u_target
    L = 0.5 * float(np.dot(d, d))
    rhs = d.copy()
    rhs[P.bc_all] = 0.0
    lam = Vector(P.n, P.ctx)
    if np.any(rhs):
        P.solve(Vector.from_numpy(rhs, P.ctx), lam, rel_tol)
    ut = uh.copy()
    ut[P.bc0] = 0.0
    g = Vector(P.ne, P.ctx)
    element_gradient(P.mesh, lam, Vector.from_numpy(ut, P.ctx), g)
    return L, g.to_numpy(), uh


def solve_u(nx, ny, nz, kappa: np.ndarray, rel_tol=1e-13, ctx: Context | None = None):
    P = KappaProblem(nx, ny, nz, ctx)
    P.assemble(Vector.from_numpy(kappa, P.ctx))
    u = Vector(P.n, P.ctx)
    P.solve(P.b, u, rel_tol)
    return u.to_numpy()


try:
    import torch
except ImportError:  # torch is optional for the numpy-level API above
    torch = None

if torch is not None:

    def _view(t: "torch.Tensor", ctx: Context) -> Vector:
        assert t.is_cuda and t.dtype == torch.float64 and t.is_contiguous()
        return Vector.wrap(t.data_ptr(), t.numel(), ctx, owner=t)

    class MiniFESolve(torch.autograd.Function):
        """u = solve(kappa): forward FP64 CG on the GPU-assembled operator,
        backward = adjoint CG + element-gradient kernel, all through the C-ABI
        on the tensors' memory (no copies)."""

        @staticmethod
        def forward(fctx, kappa, dims, rel_tol=1e-13):
            P = KappaProblem(*dims)
            kappa = kappa.detach().to(torch.float64).contiguous()
            u = torch.zeros(P.n, dtype=torch.float64, device=kappa.device)
            torch.cuda.synchronize()  # the library works on its own stream
            P.assemble(_view(kappa, P.ctx))
            P.solve(P.b, _view(u, P.ctx), rel_tol)
            fctx.problem = P
            fctx.rel_tol = rel_tol
            fctx.save_for_backward(kappa, u)
            return u

        @staticmethod
        def backward(fctx, grad_u):
            P = fctx.problem
            kappa, u = fctx.saved_tensors
            dev = u.device
            rhs = grad_u.detach().to(torch.float64).contiguous().clone()
            rhs[torch.as_tensor(P.bc_all, device=dev, dtype=torch.long)] = 0.0
            lam = torch.zeros_like(u)
            torch.cuda.synchronize()
            if bool(torch.any(rhs != 0)):
                P.solve(_view(rhs, P.ctx), _view(lam, P.ctx), fctx.rel_tol)
            ut = u.clone()
            ut[torch.as_tensor(P.bc0, device=dev, dtype=torch.long)] = 0.0
            grad = torch.empty(P.ne, dtype=torch.float64, device=dev)
            torch.cuda.synchronize()
            element_gradient(P.mesh, _view(lam, P.ctx), _view(ut, P.ctx), _view(grad, P.ctx))
            P.ctx.synchronize()
            return grad, None, None

    def solve(kappa: "torch.Tensor", dims, rel_tol=1e-13) -> "torch.Tensor":
        """Differentiable u(kappa) for the (nx, ny, nz) box."""
        return MiniFESolve.apply(kappa, tuple(dims), rel_tol)
