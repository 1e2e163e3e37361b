"""Forward-mode derivative of the time loop: a torch.autograd.Function whose
forward runs the HIP step (smaug_step) and whose jvp runs the tangent-linear
run (smaug_tl_*), so torch's forward-mode AD (torch.autograd.forward_ad: dual
tensors) differentiates

    L(w0, bg) = sum_{n=1..nsteps} loss_fn(w_n, bg)

in any direction of the initial evolved state w0 (np*N values, field-major,
ghost cells included), of the static background bg (nb*N values: e_b,
rho_b, b_b) and of the driver amplitude drv[0] (a 0-d tensor whose value is
the engine's own: the amplitude's tangent enters through
smaug_tl_driver_amplitude; no driver amplitude map may be set).  loss_fn is a torch function of one state's evolved fields and
of the background (flat float64 tensors on w0's device).  The
hyperdiffusion coefficients, the wave-speed maxima and dt are held at their
values along the trajectory, as in the reverse mode (smaug_amd.adjoint).  In
a mixed-precision engine the run's stencils use FP32 value/tangent pairs.

States pass through host memory (smaug_set_state / smaug_get_state and the
tangent-linear run's host arrays) on the engine's own stream; one jvp costs
one forward run of the loop and one tangent-linear run.
"""
from __future__ import annotations

from typing import Callable, Tuple

import numpy as np
import torch

from .engine import Engine

LossFn = Callable[[torch.Tensor, torch.Tensor], torch.Tensor]


def line_vz2_loss_bg(eng: Engine, i_line: int) -> LossFn:
    """sum over the interior cells of the plane (line in 2D) of dim-0 index
    i_line of v_z^2 = (m_0 / (rho + rho_b))^2, as a function of the evolved
    fields and the background."""
    n = eng.rp.n
    nd = eng.nd
    shape = tuple(n[:nd][::-1])
    inner = tuple(slice(2, n[d] - 2) for d in range(nd - 1, 0, -1)) + (i_line,)

    def fn(w: torch.Tensor, bg: torch.Tensor) -> torch.Tensor:
        v = w.reshape((2 + 2 * nd,) + shape)
        rhob = bg.reshape((2 + nd,) + shape)[1]
        vz = v[1][inner] / (v[0][inner] + rhob[inner])
        return (vz * vz).sum()

    return fn


def _host(t: torch.Tensor) -> np.ndarray:
    return t.detach().to("cpu", torch.float64).contiguous().numpy()


def _run(eng: Engine, w0: np.ndarray, bg: np.ndarray, clock0, nsteps: int, loss_fn: LossFn, dev,
         dw0=None, dbg=None, damp: float = 0.0) -> Tuple[torch.Tensor, torch.Tensor]:
    """(L, dL): the forward loop from (w0, bg), and with a direction the
    tangent-linear run alongside it."""
    k = eng.npert
    eng.set_state(np.concatenate([w0.reshape(k, -1), bg.reshape(-1, eng.N)]))
    eng.fill_ghosts()  # as at creation: the ghost cells are images of the interior
    eng.set_clock(*clock0)
    bgt = torch.from_numpy(bg.reshape(-1)).to(dev)
    loss = torch.zeros((), dtype=torch.float64, device=dev)
    dloss = torch.zeros((), dtype=torch.float64, device=dev)
    if dw0 is None:
        for _ in range(nsteps):
            eng.step(1)
            w = torch.from_numpy(eng.state()[:k].reshape(-1)).to(dev)
            loss = loss + loss_fn(w, bgt)
        return loss, dloss
    eng.tl_begin(dw0.reshape(k, -1), dbg=None if dbg is None else dbg.reshape(-1, eng.N))
    if damp:
        eng.tl_driver_amplitude(damp)
    dbgt = torch.zeros_like(bgt) if dbg is None else torch.from_numpy(dbg.reshape(-1)).to(dev)
    for _ in range(nsteps):
        eng.tl_step(1)
        wv, dw = eng.tl_get()
        w = torch.from_numpy(wv.reshape(-1)).to(dev)
        d = torch.from_numpy(dw.reshape(-1)).to(dev)
        # the loss function's own derivative, by reverse mode on its small
        # graph (forward mode cannot nest inside this jvp)
        with torch.enable_grad():
            x = w.requires_grad_(True)
            y = bgt.clone().requires_grad_(True)
            val = loss_fn(x, y)
            gx, gy = torch.autograd.grad(val, (x, y), allow_unused=True)
        tan = torch.zeros((), dtype=torch.float64, device=dev)
        if gx is not None:
            tan = tan + (gx * d).sum()
        if gy is not None:
            tan = tan + (gy * dbgt).sum()
        loss = loss + val.detach()
        dloss = dloss + tan
    eng.tl_end()
    return loss, dloss


class _StateLoss(torch.autograd.Function):
    @staticmethod
    def forward(w0, bg, amp, eng, nsteps, loss_fn, clock0):
        loss, _ = _run(eng, _host(w0), _host(bg), clock0, nsteps, loss_fn, w0.device)
        return loss

    @staticmethod
    def setup_context(ctx, inputs, output):
        w0, bg, _, eng, nsteps, loss_fn, clock0 = inputs
        ctx.eng, ctx.nsteps, ctx.loss_fn, ctx.clock0 = eng, nsteps, loss_fn, clock0
        ctx.w0, ctx.bg = w0.detach(), bg.detach()

    @staticmethod
    def jvp(ctx, dw0, dbg, damp, *_):
        eng = ctx.eng
        k = eng.npert
        dw = np.zeros((k, eng.N)) if dw0 is None else _host(dw0)
        db = None if dbg is None else _host(dbg)
        da = 0.0 if damp is None else float(damp)
        _, dloss = _run(eng, _host(ctx.w0), _host(ctx.bg), ctx.clock0, ctx.nsteps, ctx.loss_fn, ctx.w0.device,
                        dw0=dw, dbg=db, damp=da)
        return dloss


def state_loss(eng: Engine, w0: torch.Tensor, bg: torch.Tensor, nsteps: int, loss_fn: LossFn,
               clock0=None, amplitude: torch.Tensor = None) -> torch.Tensor:
    """L(w0, bg, A) = sum_{n=1..nsteps} loss_fn(w_n, bg) from the initial
    evolved state w0 (np*N float64) and background bg (nb*N float64), with
    forward-mode derivatives in both and in the driver amplitude A (a 0-d
    float64 tensor equal to the engine's drv[0]; default: that value with no
    tangent): pass dual tensors (torch.autograd.forward_ad.make_dual) and
    read the tangent of the result.  The functorch transforms
    (torch.func.jvp) are not supported: the engine reads its inputs'
    storage.  clock0 = (it, qt) of w0 (default: the engine's clock at
    entry)."""
    if w0.dtype != torch.float64 or bg.dtype != torch.float64:
        raise TypeError("w0 and bg must be float64")
    if w0.numel() != eng.npert * eng.N or bg.numel() != (eng.nvar - eng.npert) * eng.N:
        raise ValueError("w0 needs np*N values and bg nb*N values")
    a0 = float(eng.rp.drv[0])
    if amplitude is None:
        amplitude = torch.tensor(a0, dtype=torch.float64, device=w0.device)
    elif amplitude.numel() != 1 or float(amplitude.detach()) != a0:
        raise ValueError(f"the amplitude must be one value equal to the engine's driver amplitude {a0}")
    if clock0 is None:
        it, _, qt, _ = eng.clock()
        clock0 = (it, qt)
    return _StateLoss.apply(w0.reshape(-1), bg.reshape(-1), amplitude.reshape(()), eng, int(nsteps), loss_fn,
                            tuple(clock0))
