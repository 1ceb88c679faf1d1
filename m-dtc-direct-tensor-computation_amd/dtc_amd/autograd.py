"""torch.autograd binding of the effective stiffness of voxel domains.

    C = effective_stiffness(modulus, domain)      # (6, 6) float64, differentiable in modulus

Forward runs the whole GPU path (mesh, 24 load-case solves, force projection,
6x6 fit: f-src/mod_struct_calcmat.f90:252-473) through the C ABI; backward calls
dtc_batch_grad, which evaluates dL/dE_v = sum_j h_j,v^T Khat u_j,v from the
forward displacements (the problem is self-adjoint, so no second solve).
PyTorch only carries the tensors: the modulus is handed to the library as a host
array and the gradient comes back as one.
"""
from __future__ import annotations

import dataclasses
from typing import Sequence

import numpy as np
import torch

from . import DTC_OK, DTC_ERR_NO_CONVERGENCE, Batch, Domain


class _EffectiveStiffness(torch.autograd.Function):
    @staticmethod
    def forward(ctx, modulus: torch.Tensor, domain: Domain, kw: dict):
        m = modulus.detach().to("cpu", torch.float64).contiguous().numpy()
        d = dataclasses.replace(domain, modulus=m)
        b = Batch([d], **kw)
        r = b.run()[0]
        if r["status"] not in (DTC_OK, DTC_ERR_NO_CONVERGENCE):
            b.close()
            raise RuntimeError(f"effective_stiffness: domain status {r['status']}")
        ctx.batch = b
        ctx.device = modulus.device
        ctx.dtype = modulus.dtype
        return torch.as_tensor(r["C"], dtype=torch.float64, device=modulus.device)

    @staticmethod
    def backward(ctx, gC: torch.Tensor):
        g = ctx.batch.grad(0, gC.detach().to("cpu", torch.float64).numpy())
        ctx.batch.close()
        return torch.as_tensor(g, device=ctx.device, dtype=ctx.dtype), None, None


def effective_stiffness(modulus: torch.Tensor, domain: Domain, **kw) -> torch.Tensor:
    """6x6 effective stiffness C (Voigt xx,yy,zz,xy,xz,yz) of one domain whose voxels
    carry the moduli `modulus` (shape of domain.phi; voxels below llimit are void and
    get zero gradient).  Keyword arguments are the solve options of dtc_amd.opts()."""
    kw = dict(kw)
    kw.setdefault("rotation", False)
    if tuple(modulus.shape) != tuple(domain.phi.shape):
        raise ValueError(f"modulus shape {tuple(modulus.shape)} != voxel shape {domain.phi.shape}")
    return _EffectiveStiffness.apply(modulus, domain, kw)


def effective_stiffness_batch(moduli: Sequence[torch.Tensor], domains: Sequence[Domain], **kw) -> torch.Tensor:
    """Stack of C for several domains (one forward batch per domain pair list)."""
    return torch.stack([effective_stiffness(m, d, **kw) for m, d in zip(moduli, domains)])
