"""strata_fdtd_amd — MI355X (gfx950) backend for strata-fdtd's 3D staggered
pressure-velocity FDTD time step.

Drop-in surface (the reference's names and semantics, ``backend="mi355x"``):
``FDTDSolver``, ``UniformGrid``, ``NonuniformGrid``, ``PML``, ``GaussianPulse``,
membrane sources (``CircularMembraneSource``, ``RectangularMembraneSource``),
``Microphone``,
SDF geometry (``Box``, ``Sphere``, ``Cylinder``, ``Cone``, ``Union``, ``Intersection``,
``Difference``, their smooth forms, ``Translate``/``Scale``/``Rotate``,
``HelmholtzResonator``, ``MaterializedGeometry``) and ADE
materials (``Pole``, ``PoleType``, ``SimpleMaterial``, ``AIR_20C``).
Beyond the reference: ``serving`` (continuous batching of many solves) and
``adjoint`` (gradients w.r.t. per-voxel density and bulk modulus).
"""

from .boundaries import PML, ABCFirstOrder, RadiationImpedance, RigidBoundary
from .geometry import (Box, Cone, CSGNode, Cylinder, Difference, HelmholtzResonator, Horn, Intersection,
                       MaterializedGeometry, Rotate, Scale, SDFPrimitive, SmoothDifference, SmoothIntersection,
                       SmoothUnion, Sphere, SpiralHorn, HelicalTube, Transform, Translate, Union, conical_horn, exponential_horn,
                       hyperbolic_horn, rotation_matrix, tractrix_horn)
from .hdf5_in import HDF5ResultReader
from .grid import NonuniformGrid, UniformGrid, explicit_core_stretch_coords
from .materials import AIR_15C, AIR_20C, AIR_25C, WATER_20C, AcousticMaterial, Pole, PoleType, SimpleMaterial
from .membranes import CircularMembraneSource, MembraneSource, RectangularMembraneSource
from .microphones import POLAR_PATTERNS, Microphone
from .solver import FDTDSolver, has_gpu_support, has_mi355x_backend, has_native_kernels
from .sources import GaussianPulse, Probe
from .waveforms import AudioFileWaveform
from .material_models import DampingMaterial, PorousMaterial, ViscoelasticSolid, get_material, list_materials

__version__ = "0.1.0"

__all__ = [
    "FDTDSolver", "has_mi355x_backend", "has_native_kernels", "has_gpu_support", "UniformGrid", "NonuniformGrid", "explicit_core_stretch_coords",
    "PML", "RigidBoundary", "ABCFirstOrder", "RadiationImpedance", "GaussianPulse", "Probe", "SDFPrimitive", "Box", "Sphere", "Cylinder", "Cone",
    "Union", "Intersection", "Difference", "CSGNode", "SmoothUnion", "SmoothIntersection", "SmoothDifference",
    "Transform", "Translate", "Scale", "Rotate", "rotation_matrix", "Horn", "conical_horn", "exponential_horn",
    "hyperbolic_horn", "tractrix_horn", "SpiralHorn", "HelicalTube", "HelmholtzResonator", "MaterializedGeometry", "Pole",
    "PoleType", "SimpleMaterial", "AcousticMaterial", "AIR_20C", "AIR_25C", "AIR_15C", "WATER_20C", "Microphone",
    "POLAR_PATTERNS", "HDF5ResultReader", "AudioFileWaveform", "PorousMaterial", "ViscoelasticSolid",
    "DampingMaterial", "get_material", "list_materials", "MembraneSource", "CircularMembraneSource", "RectangularMembraneSource",
]
