"""Paths for swept-volume geometry (geometry/paths.py): a 2D logarithmic or
Archimedean spiral, a 3D helix, and the distance from points to a polyline.
The sampled paths and the distances are formed with the reference's numpy
expressions, so the swept primitives built on them (SpiralHorn, HelicalTube)
voxelise to the reference's masks bit for bit.
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import Literal

import numpy as np
from numpy.typing import NDArray


@dataclass
class SpiralPath:
    """r = a exp(b theta) (logarithmic) or r = a + b theta (Archimedean) from
    inner_radius to outer_radius over `turns` turns about `center`
    (geometry/paths.py:30-134)."""

    center: NDArray[np.floating] | tuple[float, float]
    inner_radius: float
    outer_radius: float
    turns: float
    spiral_type: Literal["logarithmic", "archimedean"] = "logarithmic"

    def __post_init__(self) -> None:
        self.center = np.asarray(self.center, dtype=np.float64)
        if self.center.shape != (2,):
            raise ValueError(f"center must be 2D point, got shape {self.center.shape}")
        if self.inner_radius <= 0:
            raise ValueError(f"inner_radius must be positive, got {self.inner_radius}")
        if self.outer_radius <= self.inner_radius:
            raise ValueError(f"outer_radius ({self.outer_radius}) must be > inner_radius ({self.inner_radius})")
        if self.turns <= 0:
            raise ValueError(f"turns must be positive, got {self.turns}")
        if self.spiral_type not in ("logarithmic", "archimedean"):
            raise ValueError(f"spiral_type must be 'logarithmic' or 'archimedean', got '{self.spiral_type}'")

    def sample_points(self, n_points: int = 100) -> NDArray[np.floating]:
        if n_points < 2:
            raise ValueError(f"n_points must be >= 2, got {n_points}")
        theta = np.linspace(0, 2 * np.pi * self.turns, n_points)
        if self.spiral_type == "logarithmic":
            k = np.log(self.outer_radius / self.inner_radius) / (2 * np.pi * self.turns)
            r = self.inner_radius * np.exp(k * theta)
        else:
            k = (self.outer_radius - self.inner_radius) / (2 * np.pi * self.turns)
            r = self.inner_radius + k * theta
        return np.stack([self.center[0] + r * np.cos(theta), self.center[1] + r * np.sin(theta)], axis=1)

    @property
    def total_length(self) -> float:
        """Length of the 1000-point polyline through the spiral."""
        pts = self.sample_points(1000)
        return float(np.sum(np.linalg.norm(np.diff(pts, axis=0), axis=1)))


@dataclass
class HelixPath:
    """Helix of `radius` about the vertical axis through `center`, rising
    `pitch` per turn (geometry/paths.py:137-223)."""

    center: NDArray[np.floating] | tuple[float, float, float]
    radius: float
    pitch: float
    turns: float

    def __post_init__(self) -> None:
        self.center = np.asarray(self.center, dtype=np.float64)
        if self.center.shape != (3,):
            raise ValueError(f"center must be 3D point, got shape {self.center.shape}")
        for name in ("radius", "pitch", "turns"):
            if getattr(self, name) <= 0:
                raise ValueError(f"{name} must be positive, got {getattr(self, name)}")

    def sample_points(self, n_points: int = 100) -> NDArray[np.floating]:
        if n_points < 2:
            raise ValueError(f"n_points must be >= 2, got {n_points}")
        theta = np.linspace(0, 2 * np.pi * self.turns, n_points)
        return np.stack([self.center[0] + self.radius * np.cos(theta),
                         self.center[1] + self.radius * np.sin(theta),
                         self.center[2] + (theta / (2 * np.pi)) * self.pitch], axis=1)

    @property
    def total_length(self) -> float:
        return float(self.turns * np.sqrt((2 * np.pi * self.radius) ** 2 + self.pitch**2))

    @property
    def total_height(self) -> float:
        return float(self.turns * self.pitch)


def distance_to_polyline(points, polyline) -> NDArray[np.floating]:
    """Distance from each point to the nearest of the polyline's segments
    (geometry/paths.py:226-293): per segment the projection parameter
    (p - a) . s / |s|^2 clipped to [0, 1] (a degenerate segment: the distance
    to its start), the minimum over segments in order."""
    points = np.asarray(points, dtype=np.float64)
    polyline = np.asarray(polyline, dtype=np.float64)
    if points.shape[1] != polyline.shape[1]:
        raise ValueError(f"Dimension mismatch: points are {points.shape[1]}D, polyline is {polyline.shape[1]}D")
    if polyline.shape[0] < 2:
        raise ValueError(f"Polyline must have >= 2 vertices, got {polyline.shape[0]}")
    best = np.full(points.shape[0], np.inf, dtype=np.float64)
    for a, b in zip(polyline[:-1], polyline[1:]):
        seg = b - a
        l2 = np.dot(seg, seg)
        if l2 < 1e-14:
            d = np.linalg.norm(points - a, axis=1)
        else:
            t = np.clip(np.dot(points - a, seg) / l2, 0, 1)
            d = np.linalg.norm(points - (a + t[:, np.newaxis] * seg), axis=1)
        best = np.minimum(best, d)
    return best


def distance_to_helix(points, center, radius: float, pitch: float, turns: float) -> NDArray[np.floating]:
    """Distance to a helix sampled at max(100, 50 turns) points
    (geometry/paths.py:296-344)."""
    points = np.asarray(points, dtype=np.float64)
    center = np.asarray(center, dtype=np.float64)
    if points.ndim != 2 or points.shape[1] != 3:
        raise ValueError(f"points must be Nx3 array, got shape {points.shape}")
    if center.shape != (3,):
        raise ValueError(f"center must be 3D point, got shape {center.shape}")
    n = max(100, int(50 * turns))
    return distance_to_polyline(points, HelixPath(center=center, radius=radius, pitch=pitch, turns=turns)
                                .sample_points(n))


def nearest_sample(points, samples, chunk: int = 8192) -> NDArray[np.intp]:
    """Index of the nearest path sample per point (first one on ties), by the
    Euclidean norm of sample - point, as np.argmin of np.linalg.norm over
    the samples gives it point by point."""
    points = np.asarray(points, dtype=np.float64)
    out = np.empty(points.shape[0], dtype=np.intp)
    for c0 in range(0, points.shape[0], chunk):
        diff = samples[None, :, :] - points[c0:c0 + chunk, None, :]
        out[c0:c0 + chunk] = np.argmin(np.linalg.norm(diff, axis=2), axis=1)
    return out
