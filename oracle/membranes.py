"""Membrane-source weights, restated over the whole grid — TEST INFRASTRUCTURE
ONLY (see ``oracle/__init__.py``).

Follows core/solver.py of the reference: the membrane-local coordinates of
every cell centre from a full 3D meshgrid (``_grid_to_membrane_coords``
:297-336, ``_grid_to_rectangular_coords`` :631-666), the mask on the plane
nearest the centre (``get_injection_mask`` :499-530, :700-731), and the mode
shape of the masked cells normalised by their maximum (``mode_shape``
:451-497, :668-698; ``get_injection_weights`` :532-556, :733-757).

The product evaluates only the membrane's plane; this whole-grid restatement
is what its weights are compared with, bit for bit.

Pinned by closed forms in tests/test_membranes_host.py: the (0,1) circular mode
is 1 at a cell on the centre and J0(alpha r/a)/J0(alpha r0/a) elsewhere,
(1,1) rectangular cells are |sin(pi x') sin(pi y')| over the largest, every
weight lies in [0, 1] with 0 off the patch, and a (m, n) circular mode's
radial nodes sit at the zeros of J_m.
"""

from __future__ import annotations

import numpy as np
from scipy.special import j0, jn_zeros, jv

_ZEROS = {(0, 1): 2.4048255576957727, (0, 2): 5.5200781102863115, (1, 1): 3.8317059702075125,
          (1, 2): 7.0155866698156190, (2, 1): 5.1356223018406826, (2, 2): 8.4172441403998649}


def _mesh(coords):
    return np.meshgrid(coords[0], coords[1], coords[2], indexing="ij")


def _plane_mask(shape, inside3d, axis: int, plane: int) -> np.ndarray:
    mask = np.zeros(shape, dtype=bool)
    sl = [slice(None)] * 3
    sl[axis] = plane
    mask[tuple(sl)] = inside3d[tuple(sl)]
    return mask


def _norm(shape):
    shape = np.abs(shape)
    mx = np.max(shape)
    return shape / mx if mx > 1e-10 else np.zeros_like(shape)


def circular_weights(coords, center, radius: float, normal: str, mode=(0, 1)) -> np.ndarray:
    """float64 weights over the grid of a circular membrane."""
    x, y, z = _mesh(coords)
    cx, cy, cz = center
    axis = "xyz".index(normal)
    if normal == "z":
        d1, d2 = x - cx, y - cy
    elif normal == "y":
        d1, d2 = x - cx, z - cz
    else:
        d1, d2 = y - cy, z - cz
    r = np.sqrt(d1**2 + d2**2)
    theta = np.arctan2(d2, d1)
    plane = int(np.argmin(np.abs(coords[axis] - center[axis])))
    mask = _plane_mask(x.shape, r <= radius, axis, plane)
    m, n = mode
    alpha = _ZEROS.get((m, n)) or float(jn_zeros(m, n)[-1])
    rho = r[mask] / radius
    shape = j0(alpha * rho) if m == 0 else jv(m, alpha * rho) * np.cos(m * theta[mask])
    shape = np.where(rho <= 1.0, shape, 0.0)
    w = np.zeros(x.shape, dtype=np.float64)
    w[mask] = _norm(shape)
    return w


def rectangular_weights(coords, center, size, normal: str, mode=(1, 1)) -> np.ndarray:
    """float64 weights over the grid of a rectangular membrane."""
    x, y, z = _mesh(coords)
    cx, cy, cz = center
    axis = "xyz".index(normal)
    if normal == "z":
        u, v = x - cx, y - cy
    elif normal == "y":
        u, v = x - cx, z - cz
    else:
        u, v = y - cy, z - cz
    plane = int(np.argmin(np.abs(coords[axis] - center[axis])))
    lx, ly = size
    mask = _plane_mask(x.shape, (np.abs(u) <= lx / 2) & (np.abs(v) <= ly / 2), axis, plane)
    m, n = mode
    um, vm = u[mask], v[mask]
    shape = np.sin(m * np.pi * (um / lx + 0.5)) * np.sin(n * np.pi * (vm / ly + 0.5))
    shape = np.where((np.abs(um) <= lx / 2) & (np.abs(vm) <= ly / 2), shape, 0.0)
    w = np.zeros(x.shape, dtype=np.float64)
    w[mask] = _norm(shape)
    return w


def weights_of(source, grid) -> np.ndarray:
    """Weights of a product membrane source object, from this restatement."""
    coords = (grid.x_coords, grid.y_coords, grid.z_coords)
    if hasattr(source, "radius"):
        return circular_weights(coords, source.center, source.radius, source.normal_axis, tuple(source.mode))
    return rectangular_weights(coords, source.center, source.size, source.normal_axis, tuple(source.mode))


def injection_field(source) -> int:
    return 0 if source.injection_type == "pressure" else 1 + "xyz".index(source.normal_axis)
