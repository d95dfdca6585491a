"""Write boundary-plane fixtures from the reference's own boundary code.

    python tests/golden/make_boundary_golden.py

Run in a container that holds the reference checkout (at /root/reference, or
STRATA_REF_SRC). The reference's first-order Mur ABC and radiation impedance
(boundaries/_boundaries.py ABCFirstOrder :420-527, RadiationImpedance
:529-792) exist only in python there (no native kernel), so their arithmetic
is pinned by vectors they write: each case builds boundary objects on a small
stand-in solver (shape, c, dt, dx and the four fields; ``h5py``, needed only
by the reference's HDF5 writer, replaced by an empty module), then for 5
steps sets p to a seeded random float32 field, calls every boundary's
apply_velocity, then every boundary's apply_pressure (the reference's
order, core/solver.py:2044-2047), and records p. ``boundary_planes.npz``
holds per case the shape, the boundary list (JSON), the input fields and the
outputs. Data only; the reference is not needed to read it.
"""

from __future__ import annotations

import json
import os
import sys
import types

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF_SRC = os.environ.get("STRATA_REF_SRC", "/root/reference/src")
OUT = os.path.join(HERE, "boundary_planes.npz")
STEPS = 5

# (name, shape, [boundary spec]); a spec is ("abc", axis) or
# ("rad", axis, side, reflection_coeff, pipe_radius)
CASES = [
    ("abc_all", (6, 7, 8), [("abc", "all")]),
    ("abc_yx", (6, 7, 8), [("abc", ["y", "x"])]),
    ("abc_z", (5, 6, 9), [("abc", "z")]),
    ("abc_all_thin", (2, 3, 4), [("abc", "all")]),
    ("rad_x_high", (7, 5, 6), [("rad", "x", "high", 0.7, None)]),
    ("rad_y_low_pipe", (6, 7, 5), [("rad", "y", "low", None, 0.01)]),
    ("rad_z_ends", (5, 6, 8), [("rad", "z", "low", 0.0, None), ("rad", "z", "high", 1.0, None)]),
    ("abc_then_rad", (6, 6, 7), [("abc", "x"), ("rad", "y", "high", 0.3, None)]),
    ("rad_then_abc", (6, 6, 7), [("rad", "x", "low", 0.45, None), ("abc", "all")]),
]


class _Solver:
    """The attributes the reference's boundaries read."""

    def __init__(self, shape):
        self.shape = shape
        self.c = 343.0
        self.dx = 1e-3
        self.dt = 0.95 * self.dx / (self.c * np.sqrt(3.0))
        self.p, self.vx, self.vy, self.vz = (np.zeros(shape, np.float32) for _ in range(4))


def reference_classes():
    sys.modules.setdefault("h5py", types.ModuleType("h5py"))
    sys.path.insert(0, REF_SRC)
    from strata_fdtd.boundaries import _boundaries as B
    assert B.__file__.startswith(REF_SRC), B.__file__
    return B.ABCFirstOrder, B.RadiationImpedance


def make(spec, ABC, Rad):
    if spec[0] == "abc":
        ax = spec[1]
        return ABC(axis=tuple(ax) if isinstance(ax, list) else ax)
    _, axis, side, refl, radius = spec
    return Rad(axis=axis, side=side, reflection_coeff=refl, pipe_radius=radius)


def main() -> None:
    ABC, Rad = reference_classes()
    out = {}
    rng = np.random.default_rng(808)
    for name, shape, specs in CASES:
        sol = _Solver(shape)
        bnds = [make(s, ABC, Rad) for s in specs]
        for b in bnds:
            b.initialize(sol)
        ins, outs = [], []
        for _ in range(STEPS):
            p = rng.standard_normal(shape).astype(np.float32)
            ins.append(p.copy())
            sol.p = p
            for b in bnds:
                b.apply_velocity(sol)
            for b in bnds:
                b.apply_pressure(sol)
            outs.append(sol.p.copy())
        out[f"{name}__shape"] = np.array(shape, np.int64)
        out[f"{name}__specs"] = np.array(json.dumps(specs))
        out[f"{name}__in"] = np.stack(ins)
        out[f"{name}__out"] = np.stack(outs)
        out[f"{name}__dt"] = np.array(sol.dt)
        print(f"{name}: {shape}, {len(specs)} boundaries, changed cells "
              f"{int((np.stack(ins) != np.stack(outs)).sum())}", flush=True)
    out["names"] = np.array([c[0] for c in CASES])
    np.savez_compressed(OUT, **out)
    print(f"wrote {OUT} ({os.path.getsize(OUT)} bytes)")


if __name__ == "__main__":
    main()
