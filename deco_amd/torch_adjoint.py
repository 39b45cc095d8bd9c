"""Differentiable CPR-Krylov solve for PyTorch (adjoint through the C ABI).

CPRSolve.apply(values, b, solver) returns x with A(values) x = b. The backward pass solves
A^T lambda = dL/dx with the transposed CPR preconditioner (FSAI factors swapped, transposed AMG
hierarchy, device A^T products) on the decoupled system, maps lambda back through the CPR left
transform (lambda = W^T mu), and returns

    dL/dA_ij = -lambda_i x_j   on A's sparsity pattern (device sampled kernel)
    dL/db    =  lambda

Tensors must be float64 on the GPU; the library works on their device memory directly.
"""
import ctypes as C

import torch

from . import _lib
from ._lib import check, SolInfo
from .core import cpr_params, sol_params, lib, PTree


def _ptr(t):
    return C.c_void_p(t.data_ptr())


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
    torch.cuda.synchronize()
    A = solver._csr(values.contiguous())
    M = C.c_void_p()
    try:
        bw = b.detach().clone().contiguous()
        check(L.deco_cpr_create(C.byref(cpr_params(solver.prec, solver.precision)), A, _ptr(bw), C.byref(M)))
        x = torch.empty_like(bw)
        info = SolInfo()
        sp = sol_params(solver.sol)
        check(L.deco_solve(C.byref(sp), A, M, _ptr(bw), _ptr(x), C.byref(info)))
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
        g = grad_x.contiguous().to(torch.float64)
        torch.cuda.synchronize()
        mu = torch.empty_like(g)
        info = SolInfo()
        check(L.deco_solve_transpose(C.byref(sol_params(solver.sol)), A, M, _ptr(g), _ptr(mu), C.byref(info)))
        lam = torch.empty_like(g)
        check(L.deco_cpr_transform_transpose(M, _ptr(mu), _ptr(lam)))
        p = solver.pattern
        Pat = C.c_void_p()
        zeros = torch.zeros(p.nnz, dtype=torch.float64, device=g.device)
        check(L.deco_csr_create_dev(p.n, p.n, p.nnz, _ptr(p.rowptr), _ptr(p.cols), _ptr(zeros), C.byref(Pat)))
        try:
            dA = torch.empty(p.nnz, dtype=torch.float64, device=g.device)
            check(L.deco_grad_A_pattern(Pat.value, _ptr(lam), _ptr(x), _ptr(dA)))
        finally:
            L.deco_csr_destroy(Pat.value)
        ctx.handles = None
        solver.last_adjoint = info
        return dA, lam, None


def cpr_solve(values, b, solver):
    return CPRSolve.apply(values, b, solver)
