"""Describe a configured ``strata_fdtd_amd.FDTDSolver`` as an oracle Problem.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``). Used where a workload is
built through the product API (bench.py's CPU baseline, full-size parity
tests): geometry, material ids and PML decay tables are read from the solver
(they are pinned separately against ``oracle.tables``); coefficients, source
waveforms and ADE pole coefficients are re-derived here with the oracle's own
host formulas.
"""

from __future__ import annotations

import numpy as np

from . import cpu, membranes as M, tables as T


def problem_from_solver(solver, n_steps: int) -> cpu.Problem:
    grid = solver.grid
    nx, ny, nz = solver.shape
    dt = solver.dt  # courant * min_dx / (c sqrt 3), pinned against tables.time_step by the tests
    if grid.is_uniform:
        cv, cp = T.uniform_coefficients(dt, solver.dx, solver.rho, solver.c)
        inv_face = inv_cell = (None, None, None)
    else:
        cv, cp = T.base_coefficients(dt, solver.rho, solver.c)
        st = T.stencil_arrays(grid.x_coords, grid.y_coords, grid.z_coords)
        inv_face = (st["inv_dx_face"], st["inv_dy_face"], st["inv_dz_face"])
        inv_cell = (st["inv_dx_cell"], st["inv_dy_cell"], st["inv_dz_cell"])
    decay = tuple(d.detach().cpu().numpy() for d in solver._decay)
    times = T.accumulated_times(dt, n_steps)
    srcs = []
    for s in solver._sources:
        if s.source_type == "membrane":
            wf = s.waveform
            w = T.waveform_table(times, wf.frequency, wf.bandwidth, wf.amplitude)
            srcs.append((2, M.injection_field(s), 0, w, M.weights_of(s, grid)))
            continue
        w = T.waveform_table(times, s.frequency, s.bandwidth, s.amplitude)
        if s.source_type == "point":
            srcs.append((0, int(np.ravel_multi_index(tuple(int(v) for v in s.position), solver.shape)), 0, w))
        else:
            srcs.append((1, int(s.position["axis"]), int(s.position["index"]), w))
    probes = [int(np.ravel_multi_index(p.position, solver.shape)) for p in solver._probes.values()]
    poles = []
    rho_inf = np.zeros(256, np.float32)
    K_inf = np.zeros(256, np.float32)
    for mid, mat in solver._materials.items():
        rho_inf[mid] = mat.rho_inf
        K_inf[mid] = mat.K_inf
        for pole in mat.poles:
            tgt = 0 if pole.target == "density" else 1
            if tgt == 0:
                corr = T.native_density_correction(dt, mat.rho_inf, solver.dx)
            else:
                corr = T.native_modulus_correction(dt, mat.K_inf)
            if pole.is_debye:
                a, b = T.debye_coefficients(pole.tau, pole.delta_chi, dt)
                poles.append(cpu.Pole(mid, tgt, 0, a, b, 0.0, corr))
            else:
                a, b, d = T.lorentz_coefficients(pole.omega_0, pole.gamma, pole.delta_chi, dt)
                poles.append(cpu.Pole(mid, tgt, 1, a, b, d, corr))
    return cpu.Problem(
        shape=(nx, ny, nz), cv=cv, cp=cp, air=np.ascontiguousarray(solver.geometry, dtype=np.uint8),
        decay=decay, nonuniform=not grid.is_uniform, inv_face=inv_face, inv_cell=inv_cell,
        sources=srcs, probes=probes, poles=poles,
        material=np.ascontiguousarray(solver._material_id) if poles else None,
        ade_inv_dx=1.0 / solver.dx, rho_inf=rho_inf, K_inf=K_inf, dt=dt)
