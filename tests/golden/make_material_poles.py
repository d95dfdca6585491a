"""Write the poles of the reference's material library (materials/library.py,
materials/porous.py, materials/solids.py) as computed by the reference.

    python tests/golden/make_material_poles.py

Run in a container that holds the reference checkout (at /root/reference, or
STRATA_REF_SRC); ``h5py`` (needed only by the reference's HDF5 writer) is
replaced by an empty module. For every library material: rho_inf, K_inf and
each pole's target, type, delta_chi and tau, in the material's order, stored
in ``material_poles.npz`` as JSON (floats in their shortest round-trip form).
"""

from __future__ import annotations

import json
import os
import sys
import types

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF_SRC = os.environ.get("STRATA_REF_SRC", "/root/reference/src")


def main() -> None:
    sys.modules.setdefault("h5py", types.ModuleType("h5py"))
    sys.path.insert(0, REF_SRC)
    from strata_fdtd.materials import library as L
    assert L.__file__.startswith(REF_SRC), L.__file__
    out = {}
    for cat, mats in L.MATERIALS.items():
        for name, m in mats.items():
            out[name] = dict(category=cat, rho_inf=float(m.rho_inf), K_inf=float(m.K_inf),
                             poles=[[p.target, p.pole_type.value, float(p.delta_chi),
                                     float(p.tau) if p.tau is not None else None] for p in m.poles])
    path = os.path.join(HERE, "material_poles.npz")
    np.savez_compressed(path, json=np.array(json.dumps(out)))
    print(f"{len(out)} materials, {sum(len(v['poles']) for v in out.values())} poles; wrote {path}")


if __name__ == "__main__":
    main()
