"""Differentiable time loop: torch.autograd.Function over the HIP step and
its hand-written adjoint (smaug_adjoint_step), with checkpoint-every-K
recomputation.

The loss is a sum over the states after each step,
    L = sum_{n=1..nsteps} loss_fn(w_n),
where loss_fn is any torch function of the evolved fields of one state
(a flat float64 device tensor of np*N values, np = 2 + 2*ndim, field-major,
ghost cells included; it should read interior cells only).  Controls are
the driver amplitude map (one value per (j, k) column) and the initial
evolved state.  The hyperdiffusion coefficients and wave-speed maxima are
held at their forward values in the derivative.

Memory: one checkpoint (np*N doubles) per K steps plus, during the
backward pass, one segment of K step tapes (4*np*N doubles each for RK4).
"""
from __future__ import annotations

import math
import time
from typing import Callable, Dict, Optional

import torch

from .engine import Engine

LossFn = Callable[[torch.Tensor], torch.Tensor]


def line_vz2_loss(eng: Engine, rhob: torch.Tensor, i_line: int) -> LossFn:
    """sum over the interior cells of the plane (line in 2D) of dim-0 index
    i_line of v_z^2 = (m_0 / (rho + rho_b))^2."""
    n = eng.rp.n
    nd = eng.nd
    shape = (2 + 2 * nd,) + tuple(n[:nd][::-1])  # (np, [n2,] n1, n0)
    inner = tuple(slice(2, n[d] - 2) for d in range(nd - 1, 0, -1))
    rb = rhob.reshape(shape[1:])[inner + (i_line,)]

    def fn(w: torch.Tensor) -> torch.Tensor:
        v = w.reshape(shape)
        rho = v[0][inner + (i_line,)]
        m0 = v[1][inner + (i_line,)]
        vz = m0 / (rho + rb)
        return (vz * vz).sum()

    return fn


def checkpoint_every(nsteps: int) -> int:
    """sqrt(N) schedule: segments of about sqrt(nsteps) steps."""
    return max(1, int(math.ceil(math.sqrt(nsteps))))


class _DriverLoss(torch.autograd.Function):
    @staticmethod
    def forward(ctx, amap: torch.Tensor, w0: torch.Tensor, eng: Engine, nsteps: int, every: int,
                loss_fn: LossFn, clock0, stats: Optional[Dict]):
        F = eng.npert * eng.N
        dev = amap.device
        t0 = time.perf_counter()
        if eng.rp.driver:
            eng.set_driver_map(device_ptr=amap.data_ptr())
        eng.load_pert(w0.data_ptr())
        eng.set_clock(*clock0)
        ckpts, clocks = [], []
        ws = torch.empty(F, dtype=torch.float64, device=dev)
        loss = torch.zeros((), dtype=torch.float64, device=dev)
        it, qt = clock0
        for n in range(nsteps):
            if n % every == 0:
                ck = torch.empty(F, dtype=torch.float64, device=dev)
                eng.save_pert(ck.data_ptr())
                ckpts.append(ck)
                clocks.append((it, qt))
            eng.step(1)
            it, qt = it + 1, qt + eng.rp.dt
            eng.save_pert(ws.data_ptr())
            loss = loss + loss_fn(ws)
        wN = torch.empty(F, dtype=torch.float64, device=dev)
        eng.save_pert(wN.data_ptr())
        if stats is not None:
            torch.cuda.current_stream().synchronize()
            stats["forward_s"] = time.perf_counter() - t0
        ctx.eng, ctx.nsteps, ctx.every, ctx.loss_fn, ctx.stats = eng, nsteps, every, loss_fn, stats
        ctx.ckpts, ctx.clocks = ckpts, clocks
        ctx.save_for_backward(amap, wN)
        return loss

    @staticmethod
    def backward(ctx, gout: torch.Tensor):
        eng, nsteps, every, loss_fn = ctx.eng, ctx.nsteps, ctx.every, ctx.loss_fn
        amap, wN = ctx.saved_tensors
        F = eng.npert * eng.N
        dev = amap.device
        t0 = time.perf_counter()
        st = torch.cuda.current_stream()  # the forward's stream
        if st.cuda_stream == 0:
            torch.cuda.synchronize()
        eng.set_stream(st.cuda_stream)
        if eng.rp.driver:
            eng.set_driver_map(device_ptr=amap.data_ptr())
        T = eng.tape_doubles()
        tapes = torch.empty((min(every, nsteps), T), dtype=torch.float64, device=dev)
        lam = torch.zeros(F, dtype=torch.float64, device=dev)
        gmap = torch.zeros_like(amap)
        wend = torch.empty(F, dtype=torch.float64, device=dev)

        def seed(w: torch.Tensor):
            with torch.enable_grad():
                x = w.detach().requires_grad_(True)
                (g,) = torch.autograd.grad(loss_fn(x), x)
            lam.add_(g * gout)

        nseg = len(ctx.ckpts)
        for seg in range(nseg - 1, -1, -1):
            start = seg * every
            end = min(start + every, nsteps)
            eng.load_pert(ctx.ckpts[seg].data_ptr())
            eng.set_clock(*ctx.clocks[seg])
            for k in range(end - start):
                eng.step_taped(tapes[k].data_ptr())
            eng.save_pert(wend.data_ptr())
            for k in range(end - start - 1, -1, -1):
                wn = wend if k == end - start - 1 else tapes[k + 1, :F]
                seed(wn)  # d loss_fn(w_n) / d w_n, w_n the state after this step
                eng.adjoint_step(tapes[k].data_ptr(), wn.data_ptr(), lam.data_ptr(), gmap.data_ptr())
        if st.cuda_stream == 0:
            eng.sync()
        if ctx.stats is not None:
            torch.cuda.current_stream().synchronize()
            ctx.stats["backward_s"] = time.perf_counter() - t0
            ctx.stats["segments"] = nseg
            ctx.stats["tape_bytes"] = int(tapes.numel()) * 8
        return gmap, lam, None, None, None, None, None, None


def _hip_runtimes() -> set:
    """Paths of the HIP runtime libraries mapped into this process."""
    try:
        with open("/proc/self/maps") as fh:
            return {ln.split()[-1] for ln in fh if "libamdhip64" in ln}
    except OSError:
        return set()


def _check_shared_runtime():
    """The engine and torch must use one HIP runtime for a torch stream
    handle to mean anything to the engine.  That holds when torch is
    imported before the engine library is first loaded (the library then
    binds to torch's libamdhip64 by soname); loaded the other way round the
    process holds two runtimes."""
    libs = _hip_runtimes()
    if len(libs) > 1:
        raise RuntimeError("two HIP runtimes are loaded (%s): import torch before creating the first "
                           "smaug_amd Engine so both use one runtime" % ", ".join(sorted(libs)))


def _engine_stream(eng: Engine) -> "torch.cuda.Stream":
    """A torch stream (default priority) shared by the engine and the torch
    work of the loss; kept on the engine so it outlives every queued call."""
    st = getattr(eng, "_torch_stream", None)
    if st is None:
        st = torch.cuda.Stream()
        eng._torch_stream = st
    return st


def driver_loss(eng: Engine, amap: torch.Tensor, w0: torch.Tensor, nsteps: int, loss_fn: LossFn,
                every: Optional[int] = None, clock0=None, stats: Optional[Dict] = None) -> torch.Tensor:
    """L(amap, w0) = sum_{n=1..nsteps} loss_fn(w_n), differentiable in the
    driver amplitude map amap (n1*n2 float64 device tensor; its gradient is
    zero when the engine has no driver) and the initial evolved state w0
    (np*N float64 device tensor, ghost cells filled).

    The engine and the torch work of the loss share one stream, so no host
    synchronisation happens per step.  clock0 = (it, qt) of w0 (default: the
    engine's clock at entry).  every = checkpoint interval (default
    sqrt(nsteps)).  stats, if given, receives forward/backward wall times."""
    if amap.dtype != torch.float64 or w0.dtype != torch.float64:
        raise TypeError("amap and w0 must be float64")
    if not (amap.is_cuda and w0.is_cuda):
        raise TypeError("amap and w0 must be device tensors")
    if amap.numel() != eng.N // eng.rp.n[0] or w0.numel() != eng.npert * eng.N:
        raise ValueError("amap needs n1*n2 values and w0 np*N values")
    amap = amap.contiguous()
    w0 = w0.contiguous()
    if clock0 is None:
        it, _, qt, _ = eng.clock()
        clock0 = (it, qt)
    every = int(every or checkpoint_every(nsteps))
    _check_shared_runtime()
    cur = torch.cuda.current_stream()
    side = _engine_stream(eng)
    side.wait_stream(cur)
    with torch.cuda.stream(side):
        eng.set_stream(side.cuda_stream)
        out = _DriverLoss.apply(amap, w0, eng, int(nsteps), every, loss_fn, tuple(clock0), stats)
    cur.wait_stream(side)
    out.record_stream(cur)
    return out
