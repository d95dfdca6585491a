"""Boundary conditions: PML absorbing layers and the rigid-wall protocol.

``PML`` keeps the reference's interface (boundaries/_boundaries.py:104-417:
``PML(depth, axis, max_sigma, order)``, ``initialize(solver)``,
``apply_velocity/apply_pressure(solver)``, ``reset()``,
``get_interior_slice()``). On ``backend="mi355x"`` the damping itself runs
inside the fused step kernel; this class only builds the three float32 decay
tables ``exp(-sigma*dt)`` exactly as the native ``initialize_pml`` does
(_kernels/src/pml.cpp:13-45) and hands them to the solver. Interior entries are
exactly 1.0f, so the kernel's multiply is an identity there: damping is
confined to the slabs.
"""

from __future__ import annotations

import ctypes
import ctypes.util
from typing import Literal

import numpy as np
from numpy.typing import NDArray


_libm = None


def _expf(x: np.ndarray) -> np.ndarray:
    """C's expf on each float32 value (the function std::exp(float) calls)."""
    global _libm
    if _libm is None:
        _libm = ctypes.CDLL(ctypes.util.find_library("m") or "libm.so.6")
        _libm.expf.restype = ctypes.c_float
        _libm.expf.argtypes = [ctypes.c_float]
    return np.array([_libm.expf(float(v)) for v in np.asarray(x, dtype=np.float32)], dtype=np.float32)


class RigidBoundary:
    """No-op boundary: rigid walls come from the geometry mask (boundaries/_boundaries.py:75-101)."""

    def initialize(self, solver) -> None:
        pass

    def apply_velocity(self, solver) -> None:
        pass

    def apply_pressure(self, solver) -> None:
        pass

    def reset(self) -> None:
        pass


class PML:
    """Polynomially graded absorbing layer (boundaries/_boundaries.py:104-417)."""

    def __init__(self, depth: int = 10,
                 axis: Literal["all", "x", "y", "z"] | tuple[str, ...] = "all",
                 max_sigma: float | None = None, order: int = 3):
        self.depth = depth
        self.order = order
        if axis == "all":
            self.axes = ("x", "y", "z")
        elif isinstance(axis, str):
            self.axes = (axis,)
        else:
            self.axes = tuple(axis)
        self._max_sigma = max_sigma
        self._initialized = False
        self._sigma_x: NDArray[np.float32] | None = None
        self._sigma_y: NDArray[np.float32] | None = None
        self._sigma_z: NDArray[np.float32] | None = None
        self._solver = None

    # ---- profile construction (host, float64 -> float32) -------------------
    def initialize(self, solver) -> None:
        """Build sigma profiles for the solver's grid (boundaries/_boundaries.py:166-226)."""
        self._solver = solver
        nx, ny, nz = solver.shape
        grid = solver.grid
        if self._max_sigma is None:
            r_target = 1e-6
            d = self.depth * grid.min_spacing
            self._max_sigma = -(self.order + 1) * solver.c * np.log(r_target) / (2 * d)
        if grid.is_uniform:
            self._sigma_x = self._create_profile(nx, solver.dx) if "x" in self.axes else None
            self._sigma_y = self._create_profile(ny, solver.dx) if "y" in self.axes else None
            self._sigma_z = self._create_profile(nz, solver.dx) if "z" in self.axes else None
        else:
            self._sigma_x = (self._create_profile_nonuniform(grid.x_coords, grid.dx)
                             if "x" in self.axes else None)
            self._sigma_y = (self._create_profile_nonuniform(grid.y_coords, grid.dy)
                             if "y" in self.axes else None)
            self._sigma_z = (self._create_profile_nonuniform(grid.z_coords, grid.dz)
                             if "z" in self.axes else None)
        self._initialized = True

    def _create_profile(self, n: int, dx: float) -> NDArray[np.float32]:
        """sigma_max * (dist/d)**order on both slabs, float64 values stored as float32 (:228-255)."""
        sigma = np.zeros(n, dtype=np.float32)
        d = self.depth
        for i in range(d):
            sigma[i] = self._max_sigma * (((d - i) / d) ** self.order)
        for i in range(n - d, n):
            sigma[i] = self._max_sigma * (((i - (n - d - 1)) / d) ** self.order)
        return sigma

    def _create_profile_nonuniform(self, coords, cell_spacing) -> NDArray[np.float32]:
        """Profile from physical distance to the PML interface (:257-315)."""
        n = len(coords)
        sigma = np.zeros(n, dtype=np.float32)
        d = self.depth
        if d >= n // 2:
            d = max(1, n // 4)
        left_thickness = float(np.sum(cell_spacing[:d]))
        right_thickness = float(np.sum(cell_spacing[-d:]))
        left_interface = coords[d] - cell_spacing[d] / 2 if d < n else coords[-1]
        right_interface = (coords[n - d - 1] + cell_spacing[n - d - 1] / 2) if n - d > 0 else coords[0]
        for i in range(d):
            dist = left_interface - coords[i]
            nd = min(1.0, dist / left_thickness) if left_thickness > 0 else 0.0
            sigma[i] = self._max_sigma * (nd ** self.order)
        for i in range(n - d, n):
            dist = coords[i] - right_interface
            nd = min(1.0, dist / right_thickness) if right_thickness > 0 else 0.0
            sigma[i] = self._max_sigma * (nd ** self.order)
        return sigma

    def decay_tables(self, dt: float, shape) -> tuple[NDArray[np.float32], ...]:
        """float32 exp(-sigma*dt) per axis; an axis without PML gets ones.

        Formed as the native initialize_pml does (_kernels/src/pml.cpp:13-45):
        float sigma times float dt, then std::exp on float (C's expf). Interior
        cells have sigma = 0, so their decay is exactly 1.0f.
        """
        out = []
        dt32 = np.float32(dt)
        for sig, n in ((self._sigma_x, shape[0]), (self._sigma_y, shape[1]), (self._sigma_z, shape[2])):
            if sig is None:
                out.append(np.ones(n, dtype=np.float32))
            else:
                out.append(_expf((-sig) * dt32))
        return tuple(out)

    # ---- boundary protocol -------------------------------------------------
    def apply_velocity(self, solver) -> None:
        """Fused into the step kernel on the MI355X backend (nothing to do here)."""

    def apply_pressure(self, solver) -> None:
        """Fused into the step kernel on the MI355X backend (nothing to do here)."""

    def reset(self) -> None:
        """The PML holds no time-dependent state."""

    def get_interior_slice(self) -> tuple[slice, slice, slice]:
        if self._solver is None:
            raise RuntimeError("PML not initialized")
        d = self.depth
        nx, ny, nz = self._solver.shape
        return (slice(d, nx - d) if "x" in self.axes else slice(None),
                slice(d, ny - d) if "y" in self.axes else slice(None),
                slice(d, nz - d) if "z" in self.axes else slice(None))

    @property
    def is_initialized(self) -> bool:
        return self._initialized


def _mur_coefficient(solver) -> float:
    """(c dt - dx) / (c dt + dx), in float64 from the solver's c, dt and dx."""
    return (solver.c * solver.dt - solver.dx) / (solver.c * solver.dt + solver.dx)


class ABCFirstOrder:
    """First-order Mur absorbing boundary on both faces of the chosen axes
    (boundaries/_boundaries.py:420-527): after the step's pressure update,
    p[0] = prev + k (p[1] - p[0]) and prev = p[1], likewise p[N-1] from
    p[N-2], with k = (c dt - dx) / (c dt + dx) and prev the inner plane of the
    previous step; axes in the order x, y, z. On ``backend="mi355x"`` the
    planes are updated on the device after each sweep (``k_boundary_plane``)."""

    def __init__(self, axis: Literal["all", "x", "y", "z"] | tuple[str, ...] = "all"):
        if axis == "all":
            self.axes = ("x", "y", "z")
        elif isinstance(axis, str):
            self.axes = (axis,)
        else:
            self.axes = tuple(axis)
        self._solver = None
        self._coeff = 0.0

    def initialize(self, solver) -> None:
        self._solver = solver
        self._coeff = _mur_coefficient(solver)

    def planes(self) -> list[tuple]:
        """(kind 0, axis, side, coeff, refl, omr) per axis, in x, y, z order."""
        return [(0, ax, 0, self._coeff, 0.0, 0.0) for ax, name in enumerate("xyz") if name in self.axes]

    def apply_velocity(self, solver) -> None:
        """No velocity modification (a Mur boundary acts on the pressure)."""

    def apply_pressure(self, solver) -> None:
        """Applied on the device after each sweep on the MI355X backend."""

    def reset(self) -> None:
        if self._solver is not None:
            self._solver._reset_boundary_state()


class RadiationImpedance:
    """Partially reflecting open end on one face (boundaries/_boundaries.py:
    529-792): after the step's pressure update, with the inner plane p_in, the
    face plane p_b and the previous step's inner plane prev,
    abc = prev + m (p_in - p_b), p_b = R p_in + (1 - R) abc, prev = p_in, where
    m is the Mur coefficient. R is the given reflection coefficient, or for a
    pipe of radius a: |(z - 1)/(z + 1)| with z = (k a)^2 / 4 at k = 2 pi 1 kHz / c,
    capped at 0.95."""

    def __init__(self, axis: Literal["x", "y", "z"], side: Literal["low", "high"],
                 reflection_coeff: float | None = None, pipe_radius: float | None = None):
        self.axis = axis
        self.side = side
        if reflection_coeff is not None:
            if not 0.0 <= reflection_coeff <= 1.0:
                raise ValueError("reflection_coeff must be in [0, 1]")
            self._reflection_coeff = reflection_coeff
            self._frequency_dependent = False
        elif pipe_radius is not None:
            if pipe_radius <= 0:
                raise ValueError("pipe_radius must be positive")
            self._pipe_radius = pipe_radius
            self._frequency_dependent = True
            self._reflection_coeff = None
        else:
            raise ValueError("Must specify either reflection_coeff or pipe_radius")
        self._solver = None
        self._initialized = False
        self._coeff_a = 0.0
        self._coeff_b = 0.0
        self._mur = 0.0

    def initialize(self, solver) -> None:
        self._solver = solver
        mur = _mur_coefficient(solver)
        if self._frequency_dependent:
            ka = (2 * np.pi * 1000 / solver.c) * self._pipe_radius
            z = ka**2 / 4
            R = min(abs((z - 1) / (z + 1)), 0.95)
        else:
            R = self._reflection_coeff
        self._coeff_a = (1 - R) * mur
        self._coeff_b = R
        self._mur = mur
        self._initialized = True

    def planes(self) -> list[tuple]:
        ax = "xyz".index(self.axis)
        R = self._coeff_b
        return [(1, ax, 0 if self.side == "low" else 1, self._mur, R, 1 - R)]

    def apply_velocity(self, solver) -> None:
        """The reference keeps the face's velocity without using it: nothing to do."""

    def apply_pressure(self, solver) -> None:
        """Applied on the device after each sweep on the MI355X backend."""

    def reset(self) -> None:
        if self._solver is not None:
            self._solver._reset_boundary_state()

    @property
    def reflection_coefficient(self) -> float:
        return self._coeff_b

    @property
    def is_initialized(self) -> bool:
        return self._initialized
