"""The variable-coefficient 27-point operator and its adjoint, as a
torch.autograd.Function over the hpcgx C ABI.

Operator (on an nx*ny*nz grid, per-cell diffusivity kappa > 0):
    a_ij = -(k_i + k_j)/2                         for in-grid neighbours j != i
    a_ii = sum_j (k_i + k_j)/2 + (26 - n_i) k_i   n_i in-grid neighbours
It is symmetric positive definite and equals the HPCG matrix at kappa = 1.
The multigrid hierarchy re-discretises the operator on each coarse grid
with kappa taken by injection (the HPCG coarsening, GenerateCoarseProblem.cpp:71-82).

KappaSolve.apply(kappa, b, grid, tol, variant) returns x = A(kappa)^-1 b.
Backward (implicit differentiation, only the converged x is stored):
    solve A lam = dL/dx            (A symmetric, same MG-PCG)
    dL/db     = lam
    dL/dkappa_k = -1/2 sum_{j in N(k)} (lam_k - lam_j)(x_k - x_j) - (26 - n_k) lam_k x_k
computed by one device kernel (hpcgx_kappa_grad).
"""
import time

import numpy as np
import torch

from . import _lib

VARIANTS = {"fp64": (_lib.HPCGX_CG_FP64, None),
            "mixed_fp32": (_lib.HPCGX_CG_MIXED_FP32, _lib.HPCGX_PREC_FP32),
            "mixed_fp16": (_lib.HPCGX_CG_MIXED_FP16, _lib.HPCGX_PREC_FP16)}


class KappaOperator:
    """Device hierarchy of the kappa operator and its preconditioned CG."""

    def __init__(self, kappa, grid, levels=4, variant="mixed_fp32", device=0):
        nx, ny, nz = grid
        k = kappa.detach().to("cpu", torch.float64).contiguous().numpy().reshape(-1) \
            if isinstance(kappa, torch.Tensor) else np.ascontiguousarray(kappa, np.float64).reshape(-1)
        if k.size != nx * ny * nz:
            raise ValueError("kappa must have nx*ny*nz entries (natural order, x fastest)")
        if not np.all(k > 0):
            raise ValueError("kappa must be positive")
        self.grid = (nx, ny, nz)
        self.variant = variant
        self.ctx = _lib.Context(device)
        self.stream = torch.cuda.current_stream(device)
        self.ctx.set_stream(self.stream.cuda_stream)
        t = time.perf_counter()
        self.ctx.generate(nx, ny, nz, levels=levels, kappa=k)
        prec = VARIANTS[variant][1]
        if prec is not None:
            self.ctx.prepare(prec)
        self.ctx.sync()
        self.setup_s = time.perf_counter() - t
        inf = self.ctx.info(0)
        self.n, self.npad = inf["n"], inf["npad"]
        self.device = torch.device("cuda", device)
        self.last = {}

    def _dev(self, v):
        return torch.empty(self.npad, dtype=torch.float64, device=self.device) if v is None else v

    def solve(self, b, tol=1e-10, max_iter=5000, variant=None):
        """x = A^-1 b for a natural-order fp64 device tensor b; returns x (natural order)."""
        variant = variant or self.variant
        b = b.detach().to(self.device, torch.float64).contiguous().reshape(-1)
        bd = self._dev(None)
        xd = torch.zeros(self.npad, dtype=torch.float64, device=self.device)
        self.ctx.from_natural(bd, b, level=0)
        t = time.perf_counter()
        out = self.ctx.cg(bd, xd, max_iter=max_iter, tol=tol, variant=VARIANTS[variant][0])
        self.ctx.sync()
        out["seconds"] = time.perf_counter() - t
        x = torch.empty(self.n, dtype=torch.float64, device=self.device)
        self.ctx.to_natural(xd, x, level=0)
        self.ctx.sync()
        out["relres"] = out["normr"] / out["normr0"] if out["normr0"] > 0 else 0.0
        if not (out["relres"] <= tol):
            raise RuntimeError(f"kappa solve did not reach {tol}: {out}")
        self.last = out
        return x

    def grad_kappa(self, lam, x):
        """dL/dkappa from the adjoint lam and the solution x (natural order)."""
        ld, xd, gd = self._dev(None), self._dev(None), self._dev(None)
        self.ctx.from_natural(ld, lam.detach().contiguous().reshape(-1), level=0)
        self.ctx.from_natural(xd, x.detach().contiguous().reshape(-1), level=0)
        self.ctx.kappa_grad(ld, xd, gd)
        g = torch.empty(self.n, dtype=torch.float64, device=self.device)
        self.ctx.to_natural(gd, g, level=0)
        self.ctx.sync()
        return g

    def close(self):
        self.ctx.close()


class KappaSolve(torch.autograd.Function):
    """x = A(kappa)^-1 b with gradients for kappa and b.

    kappa may be a tensor of any shape with nx*ny*nz entries (natural order),
    on any device, or a numpy array (then it gets no gradient); each gradient
    is returned in its input's shape, dtype and device, and only when autograd
    asks for it."""

    @staticmethod
    def forward(fctx, kappa, b, grid, tol=1e-10, variant="mixed_fp32"):
        t = time.perf_counter()
        op = KappaOperator(kappa, grid, variant=variant, device=b.device.index or 0)
        x = op.solve(b, tol=tol)
        fctx.op, fctx.tol = op, tol
        fctx.kappa_meta = (kappa.shape, kappa.dtype, kappa.device) if isinstance(kappa, torch.Tensor) else None
        fctx.save_for_backward(x)
        fctx.timing = {"forward_s": time.perf_counter() - t, "setup_s": op.setup_s,
                       "forward_solve_s": op.last["seconds"], "forward_iters": op.last["niters"]}
        KappaSolve.last_timing = fctx.timing
        return x.reshape(b.shape)

    @staticmethod
    def backward(fctx, gx):
        (x,) = fctx.saved_tensors
        t = time.perf_counter()
        op = fctx.op
        need_k = fctx.needs_input_grad[0] and fctx.kappa_meta is not None
        need_b = fctx.needs_input_grad[1]
        gk = gb = None
        try:
            if need_k or need_b:
                lam = op.solve(gx.reshape(-1), tol=fctx.tol)
                fctx.timing.update({"adjoint_solve_s": op.last["seconds"], "adjoint_iters": op.last["niters"]})
                if need_b:
                    gb = lam.reshape(gx.shape).to(gx.dtype)
                if need_k:
                    tg = time.perf_counter()
                    shape, dtype, device = fctx.kappa_meta
                    gk = op.grad_kappa(lam, x).reshape(shape).to(device=device, dtype=dtype)
                    fctx.timing["grad_kernel_s"] = time.perf_counter() - tg
            fctx.timing["backward_s"] = time.perf_counter() - t
            KappaSolve.last_timing = fctx.timing
        finally:
            op.close()
        return gk, gb, None, None, None


KappaSolve.last_timing = {}


def kappa_solve(kappa, b, grid, tol=1e-10, variant="mixed_fp32"):
    """Functional form of KappaSolve.apply."""
    return KappaSolve.apply(kappa, b, grid, tol, variant)
