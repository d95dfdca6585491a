"""Write golden vectors from the reference's own native kernels.

    python tests/golden/make_golden.py

Runs the reference's C++ kernels (``_kernels/src/*.cpp`` compiled from source
into ``oracle/_ref`` by ``oracle/Makefile``; needs the reference checkout at
build time) on small seeded scenarios, composed in the reference solver's
python step ordering, and stores inputs and outputs as ``.npz`` fixtures next
to this script. The fixtures are data only: seeds and sizes that rebuild the
scenario through ``tests/helpers.py``, the probe traces and the final fields.
"""

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

from helpers import oracle_problem, seeded_scenario  # noqa: E402
from oracle import cpu  # noqa: E402

# name -> (seed, shape, scenario kwargs, steps)
CASES = {
    "plain_pml": (21, (16, 18, 20), dict(geometry=False, pml=3), 40),
    "rigid_plane_source": (22, (16, 18, 20), dict(geometry=True, pml=3, plane=True), 40),
    "ade_debye_lorentz": (23, (18, 20, 22), dict(geometry=True, pml=3, ade=True), 40),
    "nonuniform_rigid": (24, (16, 18, 20), dict(geometry=True, pml=3, nonuniform=True), 40),
    "nonuniform_ade": (25, (18, 20, 22), dict(geometry=True, pml=3, ade=True, nonuniform=True), 40),
}


def main() -> None:
    cpu.build(ref=True)
    if not cpu.have_ref():
        raise SystemExit("oracle/_ref is not built (the reference checkout is needed)")
    for name, (seed, shape, kw, steps) in CASES.items():
        sc = seeded_scenario(seed, shape=shape, **kw)
        state, trace = cpu.run(oracle_problem(sc, steps), steps, impl="ref")
        path = os.path.join(HERE, f"{name}.npz")
        np.savez_compressed(path, seed=seed, shape=np.array(shape), steps=steps,
                            kwargs=np.array(repr(sorted(kw.items()))), trace=trace,
                            **{k: state[k] for k in ("p", "vx", "vy", "vz")})
        print(f"wrote {path}")


if __name__ == "__main__":
    main()
