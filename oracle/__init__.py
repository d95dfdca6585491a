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
``numpy_step``   whole-array numpy restatement of one step (python ordering)
                 for tiny grids, in float32 or float64.
``voxel``        brute-force per-point SDF evaluation used to pin the host
                 voxelizer on small grids.
``microphones``  the reference's microphone recording (trilinear gathers,
                 polar patterns).
``membranes``    membrane-source weights over the whole grid (circular Bessel
                 and rectangular sine modes).

Parity status of each function is recorded in its docstring ("pinned by ...").
"""
