"""torch.autograd binding of the harmonic solve.

`harmonic_solve(xc, lap)` returns x_f with A_ff x_f = -A_fc x_c (the
HarmonicSolve RHS of HarmonicSolve.lean:35-66, solved by the library's
PCG); its backward pass is the library's adjoint (cn_adjoint_solve, or
cn_adjoint_solve_hsc): lambda = A_ff^-1 dL/dx_f with the same preconditioner
(IC(0) factor or HSC hierarchy), dL/dx_c = -A_fc^T lambda.
Composes with any torch code in front (for example a B-spline basis mapping
control points to x_c) and behind (a loss on x_f).
"""
from __future__ import annotations

from . import _lib as L
from .curvenet import HSC, IC0, _p, _stream, _torch


def _function():
    torch = _torch()

    class HarmonicSolveFn(torch.autograd.Function):
        @staticmethod
        def forward(ctx, xc, lap, pre, tol, precision):
            # the gradient goes back in xc's own device, dtype and shape
            ctx.in_device, ctx.in_dtype, ctx.in_shape = xc.device, xc.dtype, xc.shape
            if xc.dim() not in (1, 2) or xc.shape[0] != lap.n_con:
                raise ValueError(f"xc must be n_con ({lap.n_con}) or n_con x k")
            xc = xc.detach().to(device="cuda", dtype=torch.float64).reshape(lap.n_con, -1)
            xc = xc.contiguous()
            k = xc.shape[1]
            b = torch.empty((lap.n_free, k), dtype=torch.float64, device="cuda")
            L.check(L.load().cn_harmonic_rhs(lap.handle, k, _p(xc), _p(b), _stream()),
                    "harmonic_rhs")
            kw = {"hsc": pre} if isinstance(pre, HSC) else {"ic0": pre}
            pc = "hsc" if isinstance(pre, HSC) else "ic0"
            x, info = lap.solve(b, precond=pc, precision=precision, tol=tol, **kw)
            if info.status != L.CN_OK:
                raise L.CurvenetError(info.status, "harmonic_solve forward")
            ctx.lap, ctx.pre, ctx.tol, ctx.precision = lap, pre, tol, precision
            return x.reshape(-1) if len(ctx.in_shape) == 1 else x

        @staticmethod
        def backward(ctx, gx):
            gx = gx.detach().to(device="cuda", dtype=torch.float64).reshape(ctx.lap.n_free, -1)
            gx = gx.contiguous()
            pre = ctx.pre
            kw = {"hsc": pre} if isinstance(pre, HSC) else {"ic0": pre}
            pc = "hsc" if isinstance(pre, HSC) else "ic0"
            _, gc, _, info = ctx.lap.adjoint(gx, precond=pc, precision=ctx.precision,
                                             tol=ctx.tol, **kw)
            if info.status != L.CN_OK:
                raise L.CurvenetError(info.status, "harmonic_solve backward")
            gc = gc.to(device=ctx.in_device, dtype=ctx.in_dtype).reshape(ctx.in_shape)
            return gc, None, None, None, None

    return HarmonicSolveFn


_FN = None


def harmonic_solve(xc, lap, ic0: IC0 | None = None, tol: float = 1e-9,
                   precision: str = "mixed", hsc: HSC | None = None):
    """x_f (n_free x k, FP64, on the GPU) from constrained values xc
    (n_con x k torch tensor), differentiable in xc.  Preconditioned by the
    HSC hierarchy `hsc` when given, else by the IC(0) factor `ic0` (built
    when None)."""
    global _FN
    if _FN is None:
        _FN = _function()
    pre = hsc if hsc is not None else (ic0 if ic0 is not None else IC0(lap))
    return _FN.apply(xc, lap, pre, tol, precision)
