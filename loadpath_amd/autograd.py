"""torch.autograd bindings of the device objectives: compliance and the
p-mean von Mises stress aggregate as functions of the element densities.

The forward sets E = Emin + x^p (1 - Emin) on a GridElasticity and solves
with its MG-PCG on the GPU; for a CUDA x on a full-box grid the moduli,
the state and the sensitivity never leave the device (compliance). The
backward is the analytic sensitivity
(loadpath/simp.py:1897-1904 for compliance, whose adjoint is the state
itself) or, for the stress aggregate, an adjoint solve on the same
hierarchy (constrained.py:607-664). The functions compose with torch
operations on x (filters, projections), so a design pipeline written in
torch differentiates end to end.

    c = compliance(x, grid, f)            # x: (n_elem,) float64 tensor
    s = stress_pnorm(x, agg, f)           # agg: stress.StressPNorm
    (c + w * s).backward()
"""
from __future__ import annotations

import numpy as np
import torch

__all__ = ["compliance", "stress_pnorm", "project", "Compliance", "StressPNormObjective", "Projection"]


def _np(x: torch.Tensor) -> np.ndarray:
    return x.detach().to("cpu", torch.float64).numpy()


def _on_device(x: torch.Tensor, grid) -> bool:
    """The device path: x a CUDA tensor and the grid's cells its elements in
    C order (a full box with even dims), so x, the moduli, the state and
    the sensitivities stay in HBM."""
    return bool(x.is_cuda) and hasattr(grid, "solve_device") and grid.cell_order_is_grid


def _moduli_device(x, penal, emin):
    """e = Emin + x^p (1 - Emin) and de/dx on the device (lp_simp_moduli)."""
    import ctypes as C

    from . import _lib

    xd = x.detach().to(torch.float64).contiguous()
    e, de = torch.empty_like(xd), torch.empty_like(xd)
    torch.cuda.synchronize()
    _lib.check(_lib.lib().lp_simp_moduli(xd.numel(), C.c_void_p(xd.data_ptr()), float(penal), float(emin),
                                         C.c_void_p(e.data_ptr()), C.c_void_p(de.data_ptr()), 1), "simp_moduli")
    return e, de


def _load_grid(grid, f, device):
    """The load vector as a node-grid tensor on the device, made once per
    (grid, load) pair."""
    cache = grid.__dict__.setdefault("_dev_loads", {})
    key = id(f)
    if key not in cache or cache[key][0] is not f:
        cache[key] = (f, torch.as_tensor(np.ascontiguousarray(grid.to_grid(f))[None], device=device))
    return cache[key][1]


def _strain_energy_device(grid, u):
    import ctypes as C

    from . import _lib

    se = torch.empty((u.shape[0],) + tuple(grid.cells), dtype=torch.float64, device=u.device)
    _lib.check(_lib.lib().lp_strain_energy(grid._h, u.shape[0], C.c_void_p(u.data_ptr()), C.c_void_p(se.data_ptr()),
                                           1), "strain_energy")
    return se


def _set(grid, e):
    if getattr(grid, "_e", None) is not e:
        grid.set_moduli(e)


class Compliance(torch.autograd.Function):
    """c = f . u(x); dc/dx = -p x^(p-1) (1 - Emin) u_e^T Ke u_e."""

    @staticmethod
    def forward(ctx, x, grid, f, penal, emin):
        if _on_device(x, grid):
            # x, E, u and dc stay on the device
            e, de = _moduli_device(x, penal, emin)
            grid.set_moduli_device(e)
            fg = _load_grid(grid, f, x.device)
            u, _, ok, b = grid.solve_device(fg)
            if not ok.all():
                raise RuntimeError("linear solve did not converge; the structure may be a mechanism (check supports)")
            ce = _strain_energy_device(grid, u)[0].reshape(-1)
            ctx.dc = -(de.reshape(-1) * ce).reshape(x.shape)
            return torch.dot(b.reshape(-1), u.reshape(-1)).to(x.dtype)
        xd = _np(x)
        e = emin + xd ** penal * (1 - emin)
        grid.set_moduli(e)
        u = grid.solve_many([f])[0]
        ce = grid.strain_energy(u)
        ctx.dc = -penal * (1 - emin) * xd ** (penal - 1) * ce
        return x.new_tensor(float(np.dot(f, u)))

    @staticmethod
    def backward(ctx, gout):
        dc = ctx.dc if torch.is_tensor(ctx.dc) else torch.as_tensor(ctx.dc, device=gout.device)
        return gout * dc.to(dtype=gout.dtype, device=gout.device), None, None, None, None


class StressPNormObjective(torch.autograd.Function):
    """g = kappa pm(u(x), x); the backward solves K lambda = dg/du on the
    same hierarchy and returns explicit - dE/dx lambda_e^T Ke u_e."""

    @staticmethod
    def forward(ctx, x, agg, f, penal, emin, kappa):
        xd = _np(x)
        e = emin + xd ** penal * (1 - emin)
        agg.grid.set_moduli(e)
        u = agg.grid.solve_many([f])[0]
        ev = agg.evaluate([u], xd, kappa)
        ctx.agg, ctx.e, ctx.u = agg, e, u
        ctx.rhs, ctx.gx = ev["rhs"][0], ev["gx"][0]
        ctx.de = penal * xd ** (penal - 1) * (1 - emin)
        return x.new_tensor(float(ev["g"][0]))

    @staticmethod
    def backward(ctx, gout):
        grid = ctx.agg.grid
        _set(grid, ctx.e)
        lam = grid.solve_many([ctx.rhs])[0]
        grad = ctx.gx - ctx.de * grid.element_product(lam, ctx.u)
        return (gout * torch.as_tensor(grad, dtype=gout.dtype, device=gout.device), None, None, None, None, None)


def compliance(x: torch.Tensor, grid, f, penal: float = 3.0, emin: float = 1e-9) -> torch.Tensor:
    return Compliance.apply(x, grid, np.asarray(f, float), float(penal), float(emin))


def stress_pnorm(x: torch.Tensor, agg, f, penal: float = 3.0, emin: float = 1e-9, kappa: float = 1.0):
    return StressPNormObjective.apply(x, agg, np.asarray(f, float), float(penal), float(emin), float(kappa))


class Projection(torch.autograd.Function):
    """Heaviside projection (simp.py:1443-1450) with its analytic derivative,
    on the GPU (lp_project): float64 tensors, CUDA (device pointers) or CPU."""

    @staticmethod
    def forward(ctx, xt, beta, eta):
        import ctypes as C

        from . import _lib

        x = xt.detach().to(torch.float64).contiguous()
        out = torch.empty_like(x)
        dout = torch.empty_like(x)
        dev = 1 if x.is_cuda else 0
        if dev:
            torch.cuda.synchronize()
        _lib.check(_lib.lib().lp_project(x.numel(), C.c_void_p(x.data_ptr()), float(beta), float(eta),
                                         C.c_void_p(out.data_ptr()), C.c_void_p(dout.data_ptr()), dev), "project")
        ctx.save_for_backward(dout)
        return out.to(xt.dtype)

    @staticmethod
    def backward(ctx, gout):
        (dout,) = ctx.saved_tensors
        return gout * dout.to(gout.dtype), None, None


def project(xt: torch.Tensor, beta: float, eta: float = 0.5) -> torch.Tensor:
    return Projection.apply(xt, float(beta), float(eta))
