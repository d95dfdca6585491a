"""Differentiable float64 restatement of the per-voxel forward step in torch.

Used by the adjoint tests as the autograd reference: the same model as the HIP
forward kernel (strata_fdtd_amd/csrc/adjoint.hip) written with out-of-place
torch ops in the style of the reference's torch solver (core/solver_gpu.py:
652-713: slice differences, masked updates, decay multiplies), so that
torch.autograd can differentiate the probe traces with respect to the design
region's density and bulk modulus.
"""

from __future__ import annotations

import numpy as np
import torch
import torch.nn.functional as F


def forward(pb, rho_box, K_box, air: np.ndarray, decays, wave, dtype=torch.float64, device="cuda"):
    nx, ny, nz = pb.shape
    dt = pb.courant * (pb.resolution / (pb.c * np.sqrt(3)))
    cv0 = -dt / (pb.rho * pb.resolution)
    cp0 = -pb.rho * pb.c**2 * dt / pb.resolution
    lo, hi = pb.box_lo, pb.box_hi
    ones = torch.ones(pb.shape, dtype=dtype, device=device)
    # the kernels hold r = rho/rho0 and kappa = K/K0 as float32 (formed in float32);
    # the forward takes those rounded values while the derivative stays the exact
    # 1/rho0 (1/K0) in `dtype` (autograd through the float32 division would round
    # the gradient itself to float32)
    r64 = rho_box.to(dtype) / pb.rho
    k64 = K_box.to(dtype) / pb.K0
    r_box = r64 + ((rho_box.float() / pb.rho).to(dtype) - r64).detach()
    k_box = k64 + ((K_box.float() / pb.K0).to(dtype) - k64).detach()
    pad = (lo[2], nz - hi[2], lo[1], ny - hi[1], lo[0], nx - hi[0])
    r = F.pad(r_box - 1.0, pad) + ones
    kap = F.pad(k_box - 1.0, pad) + ones
    airt = torch.from_numpy(air.astype(bool)).to(device)
    ox = airt[:-1] & airt[1:]
    oy = airt[:, :-1] & airt[:, 1:]
    oz = airt[:, :, :-1] & airt[:, :, 1:]
    cvx = cv0 / ((r[:-1] + r[1:]) * 0.5)
    cvy = cv0 / ((r[:, :-1] + r[:, 1:]) * 0.5)
    cvz = cv0 / ((r[:, :, :-1] + r[:, :, 1:]) * 0.5)
    dX, dY, dZ = (torch.from_numpy(np.asarray(d, dtype=np.float64)).to(device=device, dtype=dtype) for d in decays)
    p = torch.zeros(pb.shape, dtype=dtype, device=device)
    vx = torch.zeros_like(p)
    vy = torch.zeros_like(p)
    vz = torch.zeros_like(p)
    src = tuple(int(v) for v in pb.source.position)
    traces = []
    for n in range(pb.steps):
        vxn = torch.cat([torch.where(ox, vx[:-1] + cvx * (p[1:] - p[:-1]), torch.zeros_like(cvx)), vx[-1:]], 0)
        vyn = torch.cat([torch.where(oy, vy[:, :-1] + cvy * (p[:, 1:] - p[:, :-1]), torch.zeros_like(cvy)),
                         vy[:, -1:]], 1)
        vzn = torch.cat([torch.where(oz, vz[:, :, :-1] + cvz * (p[:, :, 1:] - p[:, :, :-1]),
                                     torch.zeros_like(cvz)), vz[:, :, -1:]], 2)
        dvx = vxn - F.pad(vxn[:-1], (0, 0, 0, 0, 1, 0))
        dvy = vyn - F.pad(vyn[:, :-1], (0, 0, 1, 0))
        dvz = vzn - F.pad(vzn[:, :, :-1], (1, 0))
        pn = torch.where(airt, p + (cp0 * kap) * ((dvx + dvy) + dvz), torch.zeros_like(p))
        pn = ((pn * dX[:, None, None]) * dY[None, :, None]) * dZ[None, None, :]
        if air[src]:
            bump = torch.zeros_like(pn)
            bump[src] = float(wave[n])
            pn = pn + bump
        vx = vxn * dX[:, None, None]
        vy = vyn * dY[None, :, None]
        vz = vzn * dZ[None, None, :]
        p = pn
        traces.append(torch.stack([p[tuple(q)] for q in pb.probes]))
    return torch.stack(traces)
