"""CPU oracle for the MI355X FDTD backend — TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline``
leg may import this package, and only as the checker (or as the timed CPU
baseline). The product package ``strata_fdtd_amd`` never imports it.

Contents
--------
``cpu``          ctypes bindings for the C++17/OpenMP restatement of the
                 reference native step (``fdtd_oracle.cpp``) and for the
                 reference's own native kernels compiled from source into
                 ``oracle/_ref`` (``ref_harness.cpp``; built only where the
                 reference checkout exists).
``tables``       independent numpy restatement of the host-side precompute:
                 time step, coefficients, PML profiles/decays, nonuniform
                 spacing arrays, Gaussian-pulse waveform and grid-index rules.
``from_solver``  describes a configured product solver as an oracle problem
                 (full-size checks and the CPU baseline).
``microphones``  the reference's microphone recording (trilinear gathers,
                 polar patterns).
``membranes``    membrane-source weights over the whole grid (circular Bessel
                 and rectangular sine modes).

The float64 twin of the step (``cpu.run(..., impl="oracle64")``, the same
C++ template) is the "error vs FP64" reference. Voxel masks are pinned
directly by masks the reference's own sdf.py wrote (tests/golden/voxel_masks.npz),
so the oracle has no voxeliser of its own.

Parity status of each function is recorded in its docstring ("pinned by ...").
"""
