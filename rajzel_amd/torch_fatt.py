"""First-arrival traveltimes as a differentiable torch operation.

``traveltimes(vp, survey)`` returns the receiver times of every shot
(nshots, ntr) for a velocity model ``vp`` (ny, nx) on the GPU.  Its backward
is the reference's adjoint-state gradient: the incoming gradient dL/dTmod is
the adjoint source in place of the residual Tmod - Tobs (calc_FA_res.c:42),
scaled by the receiver slope as calc_FA_res.c:52-69, propagated by the
Leung-Qian sweeps (sweep_adj.c) and assembled as lambda*S^3 with the frame
copy (ass_grad.c:23, cp_grad_frame.c).  For L = 0.5*||Tmod - Tobs||^2 the
backward is therefore bit for bit the gradient grad_obj() returns.

That gradient is a continuous-adjoint approximation of dL/dVp: for surface
acquisition it agrees with central finite differences up to a positive,
smoothly varying factor (tests/test_gpu_torch.py checks the correlation),
which is what the reference's descent direction (descent.c:22) relies on.

The forward keeps the converged traveltime fields of every shot on the
device; when they do not fit the context's HBM budget the backward
recomputes them batch by batch.  A Solver holds the fields of its last
forward only: a backward whose forward was followed by another one on the
same Solver uploads its own slowness again, and the adjoint recomputes the
fields that belong to it.

torch is imported before the HIP library so that both use one HIP runtime
(torch ships its own copy of libamdhip64.so.7).
"""
from __future__ import annotations

import torch

from . import native


class Solver:
    """A native context for one survey (model grid + acquisition) bound to
    torch's current HIP stream."""

    def __init__(self, survey, device=0, rows_per_lane=0, hbm_budget=0):
        if native._lib is not None and not native._loaded_after_torch:
            raise RuntimeError("librajzel_amd.so was loaded before torch; import rajzel_amd.torch_fatt "
                               "(or torch) first so that both share one HIP runtime")
        model = dict(nx=survey.nx, ny=survey.ny, dh=survey.dh, time=survey.time, ttnorm=survey.ttnorm,
                     eps_adj=survey.eps_adj, src_ix=survey.src_ix, src_iy=survey.src_iy,
                     rec_ix=survey.rec_ix, rec_iy=survey.rec_iy)
        self.shape = (survey.ny, survey.nx)
        self.nshots, self.ntr = survey.nshots, survey.ntr
        self.device = torch.device("cuda", device)
        self.ctx = native.Context([model], device=device, rows_per_lane=rows_per_lane, hbm_budget=hbm_budget)
        self._forwards = 0     # forwards run so far
        self._current = 0      # the forward whose slowness the context holds

    def _bind_stream(self):
        self.ctx.set_stream(torch.cuda.current_stream(self.device).cuda_stream)

    def close(self):
        self.ctx.close()


def slowness(vp: torch.Tensor) -> torch.Tensor:
    """S = 1.0/Vp evaluated in double, stored as float (calc_S.c:18-24)."""
    return (1.0 / vp.double()).float()


class _Traveltimes(torch.autograd.Function):
    @staticmethod
    def forward(fctx, vp, solver):
        if vp.dtype != torch.float32 or vp.shape != solver.shape or vp.device != solver.device:
            raise ValueError(f"vp must be float32 {solver.shape} on {solver.device}")
        solver._bind_stream()
        S = slowness(vp).contiguous()
        solver.ctx.set_slowness(0, S.data_ptr())
        solver.ctx.forward()
        solver._forwards += 1
        solver._current = solver._forwards
        tmod = torch.empty((solver.nshots, solver.ntr), dtype=torch.float32, device=vp.device)
        solver.ctx.tmod(0, device_ptr=tmod.data_ptr())
        fctx.solver = solver
        fctx.S = S
        fctx.token = solver._current
        return tmod

    @staticmethod
    def backward(fctx, grad_out):
        solver = fctx.solver
        solver._bind_stream()
        if solver._current != fctx.token:
            # another forward ran since: restore this one's slowness, so that
            # the adjoint recomputes this forward's fields
            solver.ctx.set_slowness(0, fctx.S.data_ptr())
            solver._current = fctx.token
        g = grad_out.contiguous().float()
        solver.ctx.set_residuals(0, g.data_ptr())
        solver.ctx.adjoint()
        grad = torch.empty(solver.shape, dtype=torch.float32, device=g.device)
        solver.ctx.grad(0, device_ptr=grad.data_ptr())
        return grad, None


def traveltimes(vp: torch.Tensor, solver: Solver) -> torch.Tensor:
    """Receiver first-arrival times (nshots, ntr) of ``vp`` (ny, nx)."""
    return _Traveltimes.apply(vp, solver)


def misfit(vp: torch.Tensor, solver: Solver, tobs: torch.Tensor) -> torch.Tensor:
    """L2 traveltime misfit 0.5*sum((Tmod - Tobs)^2) over picks >= 0
    (grad_obj.c:61, calc_FA_res.c:39-47); its gradient is grad_obj's."""
    t = traveltimes(vp, solver)
    r = torch.where(tobs >= 0, t - tobs, torch.zeros_like(t))
    return 0.5 * (r.double() ** 2).sum()
