"""torch.autograd binding of the batched eikonal solve.

`eikonal_picks(slowness, batch)` returns the batch's receiver picks as a
differentiable function of the (nx, nz) slowness tensor: forward is
seisfat_batch_run_record, backward the adjoint-state vector-Jacobian product
(seisfat_batch_vjp_device), both on torch's current stream.  Isotropic
batches only.  A traveltime-misfit tomography step is then

    picks = eikonal_picks(s, batch)
    loss = 0.5 * ((picks - obs) ** 2).sum()
    loss.backward()          # s.grad = dJ/ds (adjoint state, fp64 accumulated)
"""
from __future__ import annotations

import torch

from .eikonal import Batch


class _DeviceArray:
    """__cuda_array_interface__ view of n float32 values at a device pointer."""

    def __init__(self, ptr: int, n: int):
        self.__cuda_array_interface__ = {"shape": (n,), "typestr": "<f4", "data": (int(ptr or 0), False),
                                         "version": 2}


class EikonalPicks(torch.autograd.Function):
    @staticmethod
    def forward(ctx, slowness: torch.Tensor, batch: Batch):
        if not slowness.is_cuda:
            raise ValueError("slowness must be a CUDA tensor")
        eik = batch.eik
        if tuple(slowness.shape) != (eik.nx, eik.nz):
            raise ValueError(f"slowness must have shape (nx, nz) = {(eik.nx, eik.nz)}")
        stream = torch.cuda.current_stream(slowness.device).cuda_stream
        eik.set_slowness(slowness.detach())
        batch.run_record(stream)
        if batch.nrec == 0:
            picks = torch.empty(0, dtype=torch.float32, device=slowness.device)
        else:
            view = torch.as_tensor(_DeviceArray(batch.device_picks_ptr(), batch.nrec), device=slowness.device)
            picks = view.clone()
        ctx.batch = batch
        ctx.record = batch.records
        ctx.dtype = slowness.dtype
        return picks

    @staticmethod
    def backward(ctx, grad_picks: torch.Tensor):
        b = ctx.batch
        if b.records != ctx.record:
            # the batch holds one recorded trajectory: the one of its latest
            # forward, which is not this graph's
            raise RuntimeError("eikonal_picks backward: the batch ran another forward (record "
                               f"{b.records}) after this graph's (record {ctx.record}); its adjoint state "
                               "no longer describes this graph's slowness. Use one batch per live graph, or "
                               "call backward before the next forward.")
        w = grad_picks.detach().to(torch.float32).contiguous()
        g = torch.empty((b.eik.nx, b.eik.nz), dtype=torch.float64, device=w.device)
        b.vjp_device(w.data_ptr() if w.numel() else 0, g.data_ptr(), torch.cuda.current_stream(w.device).cuda_stream)
        return g.to(ctx.dtype), None


def eikonal_picks(slowness: torch.Tensor, batch: Batch) -> torch.Tensor:
    return EikonalPicks.apply(slowness, batch)
