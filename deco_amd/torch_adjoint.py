"""Differentiable CPR-Krylov solve for PyTorch (adjoint through the C ABI).

CPRSolve.apply(values, b, solver) returns x with A(values) x = b. The backward pass solves
A^T lambda = dL/dx with the transposed CPR preconditioner (FSAI factors swapped, transposed AMG
hierarchy, device A^T products) on the decoupled system, maps lambda back through the CPR left
transform (lambda = W^T mu), and returns

    dL/dA_ij = -lambda_i x_j   on A's sparsity pattern (device sampled kernel)
    dL/db    =  lambda

Tensors must be float64 on the GPU; the library works on their device memory directly. The
library runs on its own stream: every tensor it reads is made (copied, cleared) on torch's stream
before a torch.cuda.synchronize(), and the library's calls return after their own work is done.

A forward or transposed solve that stops without reaching the tolerance (SOL::flag != 0,
solver/SOL.hpp:10-13) raises NotConverged: its x or gradients are not returned.
"""
import ctypes as C

import torch

from . import _lib
from ._lib import check, SolInfo
from .core import cpr_params, sol_params, lib, PTree


def _ptr(t):
    return C.c_void_p(t.data_ptr())


class NotConverged(RuntimeError):
    """A solve stopped at itmax before the relative residual reached reltol. `info` holds the
    solve's SolInfo (iterations, residuals, flag)."""

    def __init__(self, what, info):
        super().__init__(f"{what} did not converge: flag {info.flag} after {info.iter} iterations, "
                         f"relative residual {info.res_rel:.3e}")
        self.info = info


class SparsePattern:
    """A fixed CSR pattern (int32 row pointers and sorted columns, host or device tensors)."""

    def __init__(self, n, rowptr, cols):
        self.n = int(n)
        self.rowptr = torch.as_tensor(rowptr, dtype=torch.int32).cuda().contiguous()
        self.cols = torch.as_tensor(cols, dtype=torch.int32).cuda().contiguous()
        self.nnz = int(self.cols.numel())
        self.rows = torch.repeat_interleave(torch.arange(self.n, device=self.rowptr.device),
                                            torch.diff(self.rowptr.long()))


class CPRSolver:
    """Parameters of the differentiable solve: the reference's "prec" and "sol" subtrees."""

    def __init__(self, pattern, prec, sol, precision=_lib.PREC_FP64):
        self.pattern = pattern
        self.prec = prec if isinstance(prec, PTree) else PTree(prec)
        self.sol = sol
        self.precision = precision
        self.last = None     # handles of the last forward solve (A, M, x) for the backward

    def _csr(self, values):
        h = C.c_void_p()
        p = self.pattern
        check(lib().deco_csr_create_dev(p.n, p.n, p.nnz, _ptr(p.rowptr), _ptr(p.cols), _ptr(values), C.byref(h)))
        return h.value


class _Handles:
    """Owns the decoupled matrix and the CPR handle of one forward solve."""

    def __init__(self, A, M):
        self.A, self.M = A, M

    def __del__(self):
        try:
            if self.M:
                lib().deco_prec_destroy(self.M)
            if self.A:
                lib().deco_csr_destroy(self.A)
        except Exception:
            pass


def _check_tensor(t, name, n):
    if t.dtype != torch.float64 or not t.is_cuda:
        raise TypeError(f"{name} must be a float64 tensor on the GPU, got {t.dtype} on {t.device}")
    if t.numel() != n:
        raise ValueError(f"{name} has {t.numel()} entries, the pattern needs {n}")


def _solve(solver, values, b):
    L = lib()
    _check_tensor(values, "values", solver.pattern.nnz)
    _check_tensor(b, "b", solver.pattern.n)
    # copies on torch's stream, finished before the library's stream reads them
    vals = values.detach().contiguous()
    bw = b.detach().clone().contiguous()
    torch.cuda.synchronize()
    A = solver._csr(vals)
    M = C.c_void_p()
    try:
        check(L.deco_cpr_create(C.byref(cpr_params(solver.prec, solver.precision)), A, _ptr(bw), C.byref(M)))
        x = torch.empty_like(bw)
        info = SolInfo()
        sp = sol_params(solver.sol)
        check(L.deco_solve(C.byref(sp), A, M, _ptr(bw), _ptr(x), C.byref(info)))
        if info.flag != 0:
            raise NotConverged("forward CPR solve", info)
    except Exception:
        if M.value:
            L.deco_prec_destroy(M.value)
        L.deco_csr_destroy(A)
        raise
    return A, M.value, x, info


class CPRSolve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, values, b, solver):
        A, M, x, info = _solve(solver, values, b)
        ctx.solver = solver
        ctx.handles = _Handles(A, M)
        ctx.info = info
        ctx.save_for_backward(x)
        solver.last = info
        return x

    @staticmethod
    def backward(ctx, grad_x):
        (x,) = ctx.saved_tensors
        A, M = ctx.handles.A, ctx.handles.M
        L = lib()
        solver = ctx.solver
        p = solver.pattern
        # every tensor the library reads is written on torch's stream before the synchronize
        g = grad_x.contiguous().to(torch.float64)
        zeros = torch.zeros(p.nnz, dtype=torch.float64, device=g.device)
        mu = torch.empty_like(g)
        lam = torch.empty_like(g)
        dA = torch.empty(p.nnz, dtype=torch.float64, device=g.device)
        torch.cuda.synchronize()
        info = SolInfo()
        check(L.deco_solve_transpose(C.byref(sol_params(solver.sol)), A, M, _ptr(g), _ptr(mu), C.byref(info)))
        solver.last_adjoint = info
        if info.flag != 0:
            raise NotConverged("transposed CPR solve", info)
        check(L.deco_cpr_transform_transpose(M, _ptr(mu), _ptr(lam)))
        Pat = C.c_void_p()
        check(L.deco_csr_create_dev(p.n, p.n, p.nnz, _ptr(p.rowptr), _ptr(p.cols), _ptr(zeros), C.byref(Pat)))
        try:
            check(L.deco_grad_A_pattern(Pat.value, _ptr(lam), _ptr(x), _ptr(dA)))
        finally:
            L.deco_csr_destroy(Pat.value)
        ctx.handles = None
        return dA, lam, None


def cpr_solve(values, b, solver):
    return CPRSolve.apply(values, b, solver)
