"""torch.autograd bindings of the differentiable step (module name after the
reference's Python package, pypsum/, which binds MAS file I/O only; the
reference has no derivative code).

    PicFieldEnergy.apply(stepper, steps, ckpt, x0, y0, vx0, vy0, ...) -> L
        L = field energy after `steps` steps from the given state (one
        species: D position tensors then D velocity tensors; several species
        follow one another).  forward = checkpointed run on the HIP engine
        (checkpoint every `ckpt` steps); backward = the hand-written adjoint
        kernels with per-segment recompute.
    ModeDisplace.apply(xi, amps, lo, L) -> x
        x = xi + sum_k amps[k-1] L/(2 pi k) sin(2 pi k (xi - lo)/L), wrapped
        into [lo, lo+L); backward projects dL/dx on the modes (HIP kernels).

All tensors live on the GPU; there is no CPU path.
"""
import ctypes as C

import torch

from ._lib import PSUM_F32, PSUM_F64, check, lib


def _stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


class PicFieldEnergy(torch.autograd.Function):
    @staticmethod
    def forward(ctx, stepper, steps, ckpt, *state):
        D = stepper.dim
        ns = len(state) // (2 * D)
        if ns * 2 * D != len(state) or ns != len(stepper.n):
            raise ValueError("PicFieldEnergy: expected D positions + D velocities per species")
        torch.cuda.current_stream().synchronize()
        for s in range(ns):
            chunk = state[2 * D * s: 2 * D * (s + 1)]
            stepper.load(s, [t.detach().contiguous() for t in chunk[:D]],
                         [t.detach().contiguous() for t in chunk[D:]])
        loss, ms = stepper.record(steps, ckpt)
        ctx.stepper = stepper
        ctx.ns = ns
        ctx.t_record_ms = ms
        return torch.tensor(loss, dtype=torch.float64, device=state[0].device)

    @staticmethod
    def backward(ctx, grad_loss):
        gx, gv, ms, _ = ctx.stepper.reverse(float(grad_loss))
        ctx.stepper.t_reverse_ms = ms
        grads = []
        for s in range(ctx.ns):
            grads += gx[s] + gv[s]
        return (None, None, None, *grads)


class ModeDisplace(torch.autograd.Function):
    @staticmethod
    def forward(ctx, xi, amps, lo, L):
        if xi.dtype not in (torch.float32, torch.float64) or not xi.is_cuda:
            raise ValueError("ModeDisplace: xi must be a float32/float64 CUDA tensor")
        xi = xi.contiguous()
        a = amps.detach().to(torch.float64).contiguous()
        x = torch.empty_like(xi)
        dt = PSUM_F64 if xi.dtype == torch.float64 else PSUM_F32
        check(lib().psum_hip_mode_displace(dt, C.c_void_p(xi.data_ptr()), C.c_void_p(x.data_ptr()),
                                           xi.numel(), C.c_void_p(a.data_ptr()), a.numel(),
                                           float(lo), float(L), _stream()), "mode_displace")
        ctx.save_for_backward(xi)
        ctx.nm, ctx.lo, ctx.L, ctx.adtype = a.numel(), float(lo), float(L), amps.dtype
        return x

    @staticmethod
    def backward(ctx, gx):
        (xi,) = ctx.saved_tensors
        gx = gx.to(xi.dtype).contiguous()
        out = torch.empty(ctx.nm, dtype=torch.float64, device=xi.device)
        nb = lib().psum_hip_mode_project_workspace(xi.numel(), ctx.nm)
        work = torch.empty(max(1, nb // 8), dtype=torch.float64, device=xi.device)
        dt = PSUM_F64 if xi.dtype == torch.float64 else PSUM_F32
        check(lib().psum_hip_mode_project(dt, C.c_void_p(xi.data_ptr()), C.c_void_p(gx.data_ptr()),
                                          xi.numel(), ctx.nm, ctx.lo, ctx.L,
                                          C.c_void_p(out.data_ptr()), C.c_void_p(work.data_ptr()),
                                          _stream()), "mode_project")
        return None, out.to(ctx.adtype), None, None
