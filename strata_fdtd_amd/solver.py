"""FDTDSolver with the MI355X backend.

Same constructor, methods and semantics as the reference's ``FDTDSolver``
(core/solver.py:1442-3443) for the staggered pressure-velocity step, with the
compute on one MI355X: ``backend="mi355x"`` (also what ``"auto"`` selects).

Device state (torch tensors on ``cuda:<device>``): two ping-pong sets of
{p, vx, vy, vz} float32 fields, the cell-flag byte, 1D PML decay and spacing
tables, the source waveform table (float64, one column per step), the probe
trace buffer, a device step counter and a health word. ``step()``/``run()``
enqueue the fused step kernel through the C ABI; nothing is copied back until
the caller reads fields or probe data.

Host views: ``solver.p`` (and vx, vy, vz) return numpy arrays synchronised
from the device. If a caller writes into such an array, the write is uploaded
before the next step (the arrays handed out are treated as possibly modified),
so reference-style code like ``solver.p[50, 50, 50] = 1.0`` keeps working.
"""

from __future__ import annotations

import warnings
from typing import Callable, Literal

import numpy as np
from numpy.typing import NDArray

from . import _native
from .boundaries import PML, ABCFirstOrder, RadiationImpedance
from .grid import NonuniformGrid, UniformGrid
from .materials import density_correction_coeff, modulus_correction_coeff
from .sources import GaussianPulse, Probe, accumulated_times, waveform_table
from .tuning import normalize_tiling, pick_tiling

_FIELDS = ("p", "vx", "vy", "vz")


def _torch():
    import torch
    return torch


def has_mi355x_backend() -> bool:
    """True if the HIP library is built and a GPU is visible."""
    try:
        _native.load()
        return _torch().cuda.is_available()
    except (ImportError, OSError):
        return False


def has_native_kernels() -> bool:
    """The reference's availability check (core/solver.py:95-101) for this
    backend: True if the compiled HIP kernel library loads."""
    try:
        _native.load()
        return True
    except OSError:
        return False


def has_gpu_support() -> bool:
    """The reference's GPU check (core/solver_gpu.py:101-107) for this
    backend: True if the kernel library loads and a GPU is visible."""
    return has_mi355x_backend()


def _fast_waveform(source, times: NDArray[np.float64], dt: float) -> NDArray[np.float64]:
    """Vectorised sampling when it reproduces the per-step calls exactly."""
    if isinstance(source, GaussianPulse):
        full = source.waveform(times, dt)
        m = min(len(times), 257)
        if np.array_equal(full[:m], waveform_table(source, times[:m], dt)):
            return full
    return waveform_table(source, times, dt)


# point and x/y-plane sources the sweep adds itself (FDTD_MAX_SOURCES of the C ABI)
SWEEP_SOURCES = 8


def _table_source(source, cells: int) -> bool:
    """Whether a source is added from per-cell tables after the sweep rather
    than inside it: membranes, and plane sources normal to z. A z plane meets
    every strip of one 64-wide k segment, where the sweep's per-cell source
    test costs more than a pass over the plane's cells. The tables hold 32-bit
    cell indices, so on grids of 2^31 cells or more z planes stay in the
    sweep."""
    if source.source_type == "membrane":
        return True
    return source.source_type == "plane" and int(source.position["axis"]) == 2 and cells < 2**31


class FDTDSolver:
    """3D acoustic FDTD solver (staggered Yee grid) on one MI355X GPU.

    Args mirror core/solver.py:1507-1518; ``backend`` accepts "auto" or
    "mi355x" ("native"/"gpu"/"python" are the reference's CPU/MPS backends and
    are not provided by this package). ``device`` selects the GPU ordinal.
    """

    def __init__(self, shape=None, resolution=None, grid=None, c: float = 343.0, rho: float = 1.2,
                 courant: float = 0.95,
                 backend: Literal["auto", "mi355x"] = "auto",
                 warn_energy_drift: bool = False, energy_drift_threshold: float = 0.01,
                 device: int = 0, tiling: tuple[int, int] | None = None,
                 voxelizer: Literal["auto", "device", "host"] = "auto"):
        if grid is not None:
            self._grid = grid
            self.shape = tuple(grid.shape)
        elif shape is not None and resolution is not None:
            self._grid = UniformGrid(shape=tuple(shape), resolution=resolution)
            self.shape = self._grid.shape
        else:
            raise ValueError("Must provide either 'grid' or both 'shape' and 'resolution'")
        if backend not in ("auto", "mi355x"):
            raise ValueError(f"backend={backend!r} is not provided by strata_fdtd_amd "
                             "(use 'mi355x' or 'auto')")
        torch = _torch()
        _native.load()  # raise early if the library is missing
        if not torch.cuda.is_available():
            raise RuntimeError("backend='mi355x' needs a visible MI355X GPU (torch.cuda is unavailable)")

        self.dx = self._grid.min_spacing
        self.c = c
        self.rho = rho
        self.backend = "mi355x"
        min_spacing = self._grid.min_spacing
        self.dt = courant * (min_spacing / (self.c * np.sqrt(3)))
        if self._grid.is_uniform:
            self._coeff_p = -self.rho * self.c**2 * self.dt / self.dx
            self._coeff_v = -self.dt / (self.rho * self.dx)
        else:
            self._coeff_p_base = -self.rho * self.c**2 * self.dt
            self._coeff_v_base = -self.dt / self.rho
            self._spacing_arrays = self._grid.get_spacing_arrays_for_stencil()

        self._device = torch.device(f"cuda:{device}")
        self._dev_index = device
        nx, ny, nz = self.shape
        dev = self._device
        f32 = torch.float32
        self._f = [[torch.zeros(self.shape, dtype=f32, device=dev) for _ in _FIELDS] for _ in range(2)]
        self._flags = torch.empty(self.shape, dtype=torch.uint8, device=dev)
        self._decay = [torch.ones(n, dtype=f32, device=dev) for n in self.shape]
        self._step_dev = torch.zeros(1, dtype=torch.int64, device=dev)
        self._health = torch.zeros(1, dtype=torch.int32, device=dev)
        if not self._grid.is_uniform:
            sa = self._spacing_arrays
            self._tables = {k: torch.from_numpy(np.ascontiguousarray(v)).to(dev) for k, v in sa.items()}

        self.geometry = np.ones(self.shape, dtype=bool)
        self._material_id = np.zeros(self.shape, dtype=np.uint8)
        self._materials: dict = {}
        self._sources: list = []
        self._probes: dict[str, Probe] = {}
        self._probe_start: dict[str, int] = {}
        self._microphones: dict = {}
        self._mic_start: dict[str, int] = {}   # first step recorded on the device
        self._mic_synced: dict[str, int] = {}  # steps already moved into the Microphone
        self._step_times: list[float] = []     # solver time at which each step was taken
        self._boundaries: list = []
        self._bnd_state = None                 # device float32 planes of Mur / radiation boundaries
        self._bnd_planes: list = []
        self._step_count = 0
        self._time = 0.0
        self._snapshots: list = []
        self._velocity_snapshots: list = []
        self._snapshot_interval: int | None = None
        self._snapshot_velocity = False
        self._warn_energy_drift = warn_energy_drift
        self._energy_drift_threshold = energy_drift_threshold
        self._energy_history: list[tuple[int, float, float]] = []
        self._track_energy = False
        self._energy_sample_interval = 1
        # energy samples enqueued on the device and not yet read back:
        # (step count, time) per row of _energy_dev (float64 [cap][2] sums)
        self._energy_pending: list[tuple[int, float]] = []
        self._energy_dev = None
        self._d2h_pending = False  # snapshot copies in flight to pinned host memory

        # host mirrors of the current field set
        self._mirror: list[np.ndarray | None] = [None] * 4
        self._mirror_valid = False
        self._exported: set[int] = set()

        # device tables with growable capacity
        self._cap = 0
        self._wave = None
        self._mwave = None
        self._trace = None
        self._probe_order: list[str] = []
        self._mic_out = None
        self._mic_order: list[str] = []
        self._mic_f64 = np.zeros((0, 4), dtype=np.int32)
        self._dirty_geometry = True
        self._dirty_sources = True
        self._dirty_ade = True
        self._ade_tensors: dict = {}

        stream = torch.cuda.current_stream(dev).cuda_stream
        h = _native._vp()
        _native.call("fdtd_ctx_create", device, stream, _native.C.byref(h))
        self._ctx = h.value
        self._stream = torch.cuda.current_stream(dev)
        self._tiling = normalize_tiling(tiling) if tiling is not None else None
        if voxelizer not in ("auto", "device", "host"):
            raise ValueError(f"voxelizer must be 'auto', 'device' or 'host', got {voxelizer!r}")
        self._voxelizer = voxelizer
        self._bind_static()

    def __del__(self):
        ctx = getattr(self, "_ctx", None)
        if ctx:
            try:
                _native.load().fdtd_ctx_destroy(ctx)
            except Exception:
                pass
            self._ctx = None

    # ------------------------------------------------------------------ props
    time = property(lambda self: self._time)
    step_count = property(lambda self: self._step_count)
    grid = property(lambda self: self._grid)
    using_native = property(lambda self: False)
    using_gpu = property(lambda self: True)
    has_materials = property(lambda self: len(self._materials) > 0)
    material_count = property(lambda self: len(self._materials))

    @property
    def _cur(self) -> int:
        return self._step_count & 1

    def device_fields(self) -> dict:
        """Current device tensors {p, vx, vy, vz} (read them after a sync)."""
        return dict(zip(_FIELDS, self._f[self._cur]))

    def _host_field(self, idx: int) -> np.ndarray:
        if not self._mirror_valid:
            torch = _torch()
            torch.cuda.synchronize(self._device)
            self._mirror = [t.cpu().numpy() for t in self._f[self._cur]]
            self._mirror_valid = True
        self._exported.add(idx)
        return self._mirror[idx]

    def _set_field(self, idx: int, value) -> None:
        arr = np.ascontiguousarray(np.asarray(value, dtype=np.float32))
        if arr.shape != self.shape:
            raise ValueError(f"field shape {arr.shape} != solver shape {self.shape}")
        self._flush_mirrors()
        self._f[self._cur][idx].copy_(_torch().from_numpy(arr))
        self._mirror_valid = False

    p = property(lambda self: self._host_field(0), lambda self, v: self._set_field(0, v))
    vx = property(lambda self: self._host_field(1), lambda self, v: self._set_field(1, v))
    vy = property(lambda self: self._host_field(2), lambda self, v: self._set_field(2, v))
    vz = property(lambda self: self._host_field(3), lambda self, v: self._set_field(3, v))

    def _flush_mirrors(self) -> None:
        """Upload host arrays handed out since the last device step."""
        if self._exported and self._mirror_valid:
            torch = _torch()
            for idx in sorted(self._exported):
                self._f[self._cur][idx].copy_(torch.from_numpy(self._mirror[idx]))
        self._exported.clear()
        self._mirror_valid = False

    # ---------------------------------------------------------------- set-up
    def _bind_static(self) -> None:
        nx, ny, nz = self.shape
        ctx = self._ctx
        if self._grid.is_uniform:
            _native.call("fdtd_ctx_set_grid", ctx, nx, ny, nz, 0, float(np.float32(self._coeff_v)),
                         float(np.float32(self._coeff_p)), None, None, None, None, None, None)
        else:
            t = self._tables
            _native.call("fdtd_ctx_set_grid", ctx, nx, ny, nz, 1, float(np.float32(self._coeff_v_base)),
                         float(np.float32(self._coeff_p_base)),
                         *[_native.ptr(t[f"inv_d{a}_face"]) for a in "xyz"],
                         *[_native.ptr(t[f"inv_d{a}_cell"]) for a in "xyz"])
        _native.call("fdtd_ctx_set_fields", ctx, *[_native.ptr(x) for x in self._f[0]],
                     *[_native.ptr(x) for x in self._f[1]])
        _native.call("fdtd_ctx_set_flags", ctx, _native.ptr(self._flags))
        _native.call("fdtd_ctx_set_pml", ctx, *[_native.ptr(d) for d in self._decay])
        _native.call("fdtd_ctx_set_counters", ctx, _native.ptr(self._step_dev), _native.ptr(self._health))
        tiling = self._tiling or pick_tiling(self.shape, not self._grid.is_uniform, False)
        _native.call("fdtd_ctx_set_tiling", ctx, *tiling)

    def set_tiling(self, rows: int, planes: int, stack: int = 1) -> None:
        """Launch shape of the fused kernel: rows per wave, planes per wave and
        waves per j-stack (halo rows shared through LDS)."""
        self._tiling = normalize_tiling((rows, planes, stack))
        _native.call("fdtd_ctx_set_tiling", self._ctx, *self._tiling)

    @property
    def tiling(self) -> tuple[int, int, int]:
        return self._tiling or pick_tiling(self.shape, not self._grid.is_uniform, bool(self._ade_tensors))

    def set_geometry(self, geometry) -> None:
        """Bool mask (True = air), an SDF primitive, or a MaterializedGeometry
        (core/solver.py:1682-1780)."""
        from .geometry import MaterializedGeometry, SDFPrimitive
        if isinstance(geometry, MaterializedGeometry):
            mask, ids = self._voxelize(geometry)
            if mask.shape != self.shape:
                raise ValueError(f"Geometry shape {mask.shape} doesn't match solver shape {self.shape}")
            self.geometry = mask
            for mid, material in geometry.get_material_table().items():
                if mid == 0:
                    continue
                if mid not in self._materials:
                    self.register_material(material, material_id=mid)
                region = ids == mid
                if np.any(region):
                    self.set_material_region(region, material_id=mid)
        elif isinstance(geometry, SDFPrimitive):
            mask, _ = self._voxelize(geometry)
            if mask.shape != self.shape:
                raise ValueError(f"Geometry shape {mask.shape} doesn't match solver shape {self.shape}")
            self.geometry = mask
            self._material_id.fill(0)
            self._dirty_ade = True
        else:
            mask = np.asarray(geometry)
            if mask.shape != self.shape:
                raise ValueError(f"Geometry shape {mask.shape} doesn't match solver shape {self.shape}")
            self.geometry = mask.astype(bool)
        self._dirty_geometry = True
        self._dirty_sources = True   # table sources cover the air cells of their patch, plane or point

    def _voxelize(self, geometry):
        """(air mask bool, material ids uint8 or None) of an SDF tree or a
        MaterializedGeometry: on the GPU when the tree is one the device
        voxeliser evaluates exactly (voxelizer="auto" or "device"), else with
        the host voxeliser. Both give the reference's mask bit for bit."""
        from .geometry import MaterializedGeometry
        from . import voxelize_device as VD
        tree = geometry.geometry if isinstance(geometry, MaterializedGeometry) else geometry
        use_device = self._voxelizer == "device" or (self._voxelizer == "auto" and VD.supports(tree) and all(
            VD.supports(r) for r, _m, _i in getattr(geometry, "_material_regions", [])))
        if use_device:
            air, ids = VD.voxelize_device(geometry, self.grid, device=self._device, stream=self._stream)
            self._stream.synchronize()
            mask = air.cpu().numpy().astype(bool)
            if isinstance(geometry, MaterializedGeometry):
                ids_h = ids.cpu().numpy() if ids is not None else np.zeros(self.shape, dtype=np.uint8)
                return mask, ids_h
            return mask, None
        if isinstance(geometry, MaterializedGeometry):
            return geometry.voxelize_with_materials(self.grid)
        return geometry.voxelize(self.grid), None

    def _position_to_index(self, pos, what: str):
        if isinstance(pos, tuple) and len(pos) == 3:
            if any(isinstance(v, float) and v < max(self.shape) for v in pos):
                gp = tuple(int(round(v / self.dx)) for v in pos)
                for ax, (ix, dim) in enumerate(zip(gp, self.shape)):
                    if not 0 <= ix < dim:
                        raise ValueError(f"{what} {'xyz'[ax]} position {pos[ax]:.4f}m "
                                         f"(index {ix}) is outside grid (0-{dim-1})")
                return gp
        return pos

    def add_source(self, source) -> None:
        """Point, plane or membrane source (core/solver.py:1782-1841)."""
        if getattr(source, "source_type", "point") == "membrane":
            extent = self._grid.physical_extent()
            for ax, name in enumerate("xyz"):
                v = source.center[ax]
                if not 0 <= v <= extent[ax]:
                    raise ValueError(f"Membrane center {name}={v:.4f}m is outside grid (0-{extent[ax]:.4f}m)")
            self._sources.append(source)
            self._dirty_sources = True
            return
        if source.source_type == "point":
            source.position = self._position_to_index(source.position, "Source")
        self._sources.append(source)
        self._dirty_sources = True

    def add_probe(self, name: str, position) -> None:
        """Pressure probe at a grid index or physical position (core/solver.py:1843-1885)."""
        if name in self._probes:
            raise ValueError(f"Probe '{name}' already exists")
        position = self._position_to_index(position, "Probe")
        self._probes[name] = Probe(name=name, position=tuple(int(v) for v in position))
        self._probe_start[name] = self._step_count
        self._dirty_sources = True

    def add_microphone(self, position, name: str | None = None, pattern="omni", direction=None, up=None):
        """Virtual microphone at a physical position (metres), recorded on the
        device after every step (core/solver.py:1887-1971)."""
        from .microphones import Microphone
        mic = position if isinstance(position, Microphone) else Microphone(
            position=position, name=name, pattern=pattern, direction=direction, up=up)
        mic_name = mic.name or f"mic_{len(self._microphones)}"
        mic.name = mic_name
        if mic_name in self._microphones:
            raise ValueError(f"Microphone '{mic_name}' already exists")
        mic._initialize(self)
        self._sync_microphones()
        self._microphones[mic_name] = mic
        self._mic_start[mic_name] = self._step_count
        self._mic_synced[mic_name] = self._step_count
        mic._solver = self
        self._dirty_sources = True
        return mic

    @property
    def microphones(self) -> dict:
        return self._microphones

    def _sync_microphones(self) -> None:
        """Move device-recorded microphone samples into the Microphone objects."""
        if not self._microphones or self._mic_out is None:
            return
        pending = [n for n in self._mic_order if self._mic_synced.get(n, 0) < self._step_count]
        if not pending:
            return
        lo = min(self._mic_synced[n] for n in pending)
        rec = self._mic_out[lo:self._step_count].cpu().numpy()
        for name in pending:
            mic = self._microphones[name]
            q = self._mic_order.index(name)
            a = self._mic_synced[name]
            rows = rec[a - lo:, q, :]
            mic._data.extend(mic._combine(rows, self._mic_f64[q]))
            mic._times.extend(self._step_times[a:self._step_count])
            self._mic_synced[name] = self._step_count

    def add_boundary(self, boundary) -> None:
        """Boundaries act in the order they are added (core/solver.py:1963-1990,
        2044-2047). The PML damps inside the step kernel; Mur ABC and radiation
        impedance planes are updated after it, so a PML must come before them."""
        boundary.initialize(self)
        if isinstance(boundary, PML):
            if any(isinstance(b, PML) for b in self._boundaries):
                raise NotImplementedError("more than one PML per solver")
            if self._bnd_planes:
                raise NotImplementedError("a PML added after a Mur or radiation boundary (its damping runs "
                                          "inside the step, before the boundary planes)")
            torch = _torch()
            for ax, tab in enumerate(boundary.decay_tables(self.dt, self.shape)):
                self._decay[ax].copy_(torch.from_numpy(np.ascontiguousarray(tab)))
        elif isinstance(boundary, (ABCFirstOrder, RadiationImpedance)):
            self._boundaries.append(boundary)
            self._setup_boundary_planes()
            return
        elif type(boundary).__name__ != "RigidBoundary":
            raise NotImplementedError(f"{type(boundary).__name__} is not implemented on the mi355x backend")
        self._boundaries.append(boundary)

    def _setup_boundary_planes(self) -> None:
        """Hand the Mur / radiation planes to the context, in list order, with
        a zeroed device state (each entry's previous-step planes); every source
        moves to the per-cell tables, which run after the planes."""
        import ctypes as C
        torch = _torch()
        planes = [pl for b in self._boundaries if isinstance(b, (ABCFirstOrder, RadiationImpedance))
                  for pl in b.planes()]
        if len(planes) > 16:
            raise NotImplementedError(f"{len(planes)} boundary planes (at most 16)")
        offs, total = [], 0
        for kind, axis, *_ in planes:
            if self.shape[axis] < 2:
                raise ValueError(f"a boundary plane on axis {'xyz'[axis]} needs at least 2 cells")
            offs.append(total)
            total += int(np.prod(self.shape)) // self.shape[axis] * (2 if kind == 0 else 1)
        self._stream.synchronize()  # steps in flight may still use the previous state
        with torch.cuda.stream(self._stream):
            self._bnd_state = torch.zeros(max(total, 1), dtype=torch.float32, device=self._device)
        arr = (_native.BoundaryPlane * max(len(planes), 1))()
        for q, ((kind, axis, side, coeff, refl, omr), off) in enumerate(zip(planes, offs)):
            arr[q] = _native.BoundaryPlane(kind, axis, side, 0, coeff, refl, omr, off)
        _native.call("fdtd_ctx_set_boundary_planes", self._ctx, len(planes), C.cast(arr, C.c_void_p),
                     _native.ptr(self._bnd_state), max(total, 1))
        self._bnd_planes = planes
        self._dirty_sources = True

    def _reset_boundary_state(self) -> None:
        if self._bnd_state is not None:
            with _torch().cuda.stream(self._stream):
                self._bnd_state.zero_()

    def enable_snapshots(self, interval: int, capture_velocity: bool = False) -> None:
        self._snapshot_interval = interval
        self._snapshot_velocity = capture_velocity

    # ---- materials (core/solver.py:2810-2942) --------------------------------
    def register_material(self, material, material_id: int | None = None) -> int:
        if material_id is None:
            used = set(self._materials)
            for cand in range(1, 256):
                if cand not in used:
                    material_id = cand
                    break
            else:
                raise ValueError("Maximum of 255 materials reached")
        if not 1 <= material_id <= 255:
            raise ValueError("material_id must be in range 1-255 (0 is reserved for air)")
        if material_id in self._materials:
            raise ValueError(f"Material ID {material_id} is already registered")
        self._materials[material_id] = material
        self._dirty_ade = True
        return material_id

    def set_material_region(self, mask, material_id: int) -> None:
        if mask.shape != self.shape:
            raise ValueError(f"Mask shape {mask.shape} doesn't match solver shape {self.shape}")
        if material_id != 0 and material_id not in self._materials:
            raise ValueError(f"Material ID {material_id} not registered. Use register_material() first.")
        self._material_id[mask] = material_id
        self._dirty_ade = True

    def set_material_box(self, material_id, x_range, y_range, z_range) -> None:
        mask = np.zeros(self.shape, dtype=bool)
        mask[x_range[0]:x_range[1], y_range[0]:y_range[1], z_range[0]:z_range[1]] = True
        self.set_material_region(mask, material_id)

    def get_material_at(self, position):
        mid = self._material_id[position]
        return None if mid == 0 else self._materials.get(mid)

    # ------------------------------------------------------ device preparation
    def _upload_geometry(self) -> None:
        torch = _torch()
        air = torch.from_numpy(np.ascontiguousarray(self.geometry, dtype=np.uint8)).to(self._device)
        mat = torch.from_numpy(np.ascontiguousarray(self._material_id)).to(self._device)
        nx, ny, nz = self.shape
        _native.call("fdtd_build_flags", _native.ptr(air), _native.ptr(mat), _native.ptr(self._flags),
                     nx, ny, nz, self._stream.cuda_stream)
        self._air_dev = air
        self._mat_dev = mat
        self._dirty_geometry = False

    def _pole_list(self):
        poles = []
        for mid, material in self._materials.items():
            for pole in material.poles:
                poles.append((mid, material, pole))
        return poles

    def _setup_ade(self) -> None:
        torch = _torch()
        poles = self._pole_list()
        if not poles:
            _native.call("fdtd_ctx_set_ade", self._ctx, 0, *([None] * 12), 0.0)
            self._ade_tensors = {}
            self._dirty_ade = False
            return
        mids = sorted({m for m, _, _ in poles})
        sel = np.isin(self._material_id, mids)
        if not np.any(sel):
            _native.call("fdtd_ctx_set_ade", self._ctx, 0, *([None] * 12), 0.0)
            self._ade_tensors = {}
            self._dirty_ade = False
            return
        idx = np.nonzero(sel)
        lo = np.array([int(a.min()) for a in idx], dtype=np.int32)
        hi = np.array([int(a.max()) + 1 for a in idx], dtype=np.int32)
        box_n = int(np.prod(hi - lo))
        n = len(poles)
        # the native correction is called with the scalar 1/dx (min spacing) on
        # every grid, nonuniform ones included (core/solver.py:3295-3305)
        inv_dx = 1.0 / self.dx
        mat_id = np.array([m for m, _, _ in poles], dtype=np.int32)
        target = np.array([0 if p.target == "density" else 1 for _, _, p in poles], dtype=np.int32)
        lorentz = np.array([0 if p.is_debye else 1 for _, _, p in poles], dtype=np.int32)
        a = np.zeros(n, np.float32)
        b = np.zeros(n, np.float32)
        d = np.zeros(n, np.float32)
        corr = np.zeros(n, np.float32)
        for q, (_m, material, pole) in enumerate(poles):
            co = pole.fdtd_coefficients(self.dt)
            a[q], b[q] = co[0], co[1]
            if not pole.is_debye:
                d[q] = co[2]
            if pole.target == "density":
                corr[q] = density_correction_coeff(self.dt, material.rho_inf, inv_dx)
            else:
                corr[q] = modulus_correction_coeff(self.dt, material.K_inf)
        J = torch.zeros((n, box_n), dtype=torch.float32, device=self._device)
        Jp = torch.zeros((n, box_n), dtype=torch.float32, device=self._device)
        mat_dev = torch.from_numpy(np.ascontiguousarray(self._material_id)).to(self._device)
        keep = dict(J=J, Jp=Jp, mat=mat_dev, mat_id=mat_id, target=target, lorentz=lorentz, a=a, b=b,
                    d=d, corr=corr, lo=lo, hi=hi)
        cp = lambda arr: arr.ctypes.data  # noqa: E731 - host arrays for the C ABI
        _native.call("fdtd_ctx_set_ade", self._ctx, n, cp(mat_id), cp(target), cp(lorentz), cp(a), cp(b),
                     cp(d), cp(corr), _native.ptr(mat_dev), cp(lo), cp(hi), _native.ptr(J), _native.ptr(Jp),
                     float(np.float32(1.0 / self.dx)))
        self._ade_tensors = keep
        self._dirty_ade = False
        tiling = self._tiling or pick_tiling(self.shape, not self._grid.is_uniform, True)
        _native.call("fdtd_ctx_set_tiling", self._ctx, *tiling)

    def _ensure_capacity(self, total_steps: int) -> None:
        """Waveform and trace tables must cover steps [0, total_steps)."""
        torch = _torch()
        if total_steps <= self._cap and not self._dirty_sources:
            return
        cap = self._cap
        if total_steps > cap:
            cap = max(total_steps, 2 * cap, 1024)
        # The sweep adds the sources listed before the first table source; that
        # one and every later source are added after the sweep from per-cell
        # tables, each cell's entries in list order. Every cell thus receives
        # its sources in the order of the list (core/solver.py:2386-2433).
        n_cells = int(np.prod(self.shape))
        first = next((q for q, s in enumerate(self._sources) if _table_source(s, n_cells)), len(self._sources))
        if self._bnd_planes and self._sources:
            # boundary planes come between the step and its sources: every
            # source is added after them, from the tables (32-bit cell indices)
            if n_cells >= 2**31:
                raise NotImplementedError("sources with Mur / radiation boundaries on 2^31 cells or more")
            first = 0
        if first > SWEEP_SOURCES:
            # the sweep adds at most SWEEP_SOURCES sources itself; the rest of
            # the list goes to the tables (32-bit cell indices)
            if n_cells >= 2**31:
                raise ValueError(f"{first} point/plane sources on a grid of 2^31 cells or more "
                                 f"(at most {SWEEP_SOURCES} there)")
            first = SWEEP_SOURCES
        srcs = self._sources[:first]
        table = self._sources[first:]
        times = accumulated_times(self.dt, cap) if self._sources else None
        if srcs:
            wave = np.stack([_fast_waveform(s, times, self.dt) for s in srcs]).astype(np.float64)
            self._wave = torch.from_numpy(np.ascontiguousarray(wave)).to(self._device)
        else:
            self._wave = None
        if table:
            # a membrane samples its own waveform object; point and plane
            # sources are their own waveform
            wave = np.stack([_fast_waveform(m.waveform if m.source_type == "membrane" else m, times, self.dt)
                             for m in table]).astype(np.float64)
            self._mwave = torch.from_numpy(np.ascontiguousarray(wave)).to(self._device)
        else:
            self._mwave = None
        names = list(self._probes)
        new_trace = torch.zeros((cap, max(len(names), 1)), dtype=torch.float32, device=self._device)
        if self._trace is not None and self._step_count:
            n_old = min(self._step_count, self._trace.shape[0])
            for q_new, name in enumerate(names):
                if name in self._probe_order:
                    q_old = self._probe_order.index(name)
                    new_trace[:n_old, q_new] = self._trace[:n_old, q_old]
        self._trace = new_trace
        self._probe_order = names
        self._sync_microphones()
        mics = list(self._microphones)
        new_mic = None
        if mics:
            new_mic = torch.zeros((cap, len(mics), 4), dtype=torch.float64, device=self._device)
            if self._mic_out is not None and self._step_count:
                n_old = min(self._step_count, self._mic_out.shape[0])
                for q_new, name in enumerate(mics):
                    if name in self._mic_order:
                        new_mic[:n_old, q_new] = self._mic_out[:n_old, self._mic_order.index(name)]
            tabs = [self._microphones[n].gather_table(self.shape) for n in mics]
            idx = np.ascontiguousarray(np.stack([t[0] for t in tabs]), dtype=np.int32)
            wts = np.ascontiguousarray(np.stack([t[1] for t in tabs]), dtype=np.float64)
            self._mic_f64 = np.ascontiguousarray(np.stack([t[2] for t in tabs]), dtype=np.int32)
            _native.call("fdtd_ctx_set_microphones", self._ctx, len(mics), idx.ctypes.data, wts.ctypes.data,
                         self._mic_f64.ctypes.data, _native.ptr(new_mic), cap)
        else:
            _native.call("fdtd_ctx_set_microphones", self._ctx, 0, None, None, None, None, 0)
        self._mic_out = new_mic
        self._mic_order = mics
        kind, a, b = [], [], []
        for s in srcs:
            if s.source_type == "point":
                i, j, k = (int(v) for v in s.position)
                kind.append(0)
                a.append(np.ravel_multi_index((i, j, k), self.shape))
                b.append(0)
            else:
                kind.append(1)
                a.append(int(s.position["axis"]))
                b.append(int(s.position["index"]))
        kind = np.array(kind or [0], dtype=np.int32)
        a = np.array(a or [0], dtype=np.int64)
        b = np.array(b or [0], dtype=np.int64)
        _native.call("fdtd_ctx_set_sources", self._ctx, len(srcs), kind.ctypes.data, a.ctypes.data,
                     b.ctypes.data, _native.ptr(self._wave), cap)
        flat = np.array([np.ravel_multi_index(self._probes[n].position, self.shape) for n in names] or [0],
                        dtype=np.int64)
        _native.call("fdtd_ctx_set_probes", self._ctx, len(names), flat.ctypes.data,
                     _native.ptr(self._trace), cap)
        self._set_table_sources(table, names, cap)
        self._cap = cap
        self._dirty_sources = False

    def _table_cells(self, src, air: NDArray[np.bool_]) -> tuple[NDArray, NDArray, int]:
        """(flat cells, float64 weights, field) of one table source: the air
        cells of a membrane's patch with their mode weights, of a plane, or
        the point's cell, the last two with weight 1.0 (the product w * 1.0 is
        w exactly, so the add equals the sweep's)."""
        if src.source_type == "membrane":
            if src._cached_weights is None:
                src._cached_weights = src.get_injection_weights(self._grid)
                src._cached_mask = src._cached_weights > 0
            idx = np.flatnonzero(src._cached_mask.ravel() & air.ravel())
            return idx, src._cached_weights.ravel()[idx], src.injection_field()
        if src.source_type == "plane":
            ax, index = int(src.position["axis"]), int(src.position["index"])
            if not (0 <= ax <= 2 and 0 <= index < self.shape[ax]):
                raise ValueError(f"plane source: axis {ax} index {index} outside the grid {self.shape}")
            plane = np.zeros(self.shape, dtype=bool)
            sl = [slice(None)] * 3
            sl[ax] = index
            plane[tuple(sl)] = True
            idx = np.flatnonzero(plane.ravel() & air.ravel())
        else:
            flat = int(np.ravel_multi_index(tuple(int(v) for v in src.position), self.shape))
            idx = np.array([flat] if air.ravel()[flat] else [], dtype=np.int64)
        return idx, np.ones(idx.size, dtype=np.float64), 0

    def _set_table_sources(self, table: list, probe_names: list, cap: int) -> None:
        """Per-cell tables of the sources added after the sweep (membranes,
        z-normal planes, and every source listed after the first of them):
        their air cells, each cell once, with its entries (source, field,
        weight) in list order and the probes recorded on it after them."""
        air = np.asarray(self.geometry, dtype=bool)
        cells, order, rows, fields, weights = [], [], [], [], []
        for q, src in enumerate(table):
            idx, w, field = self._table_cells(src, air)
            cells.append(idx.astype(np.int64))
            order.append(np.full(idx.size, q, dtype=np.int64))
            rows.append(np.full(idx.size, q, dtype=np.int32))
            fields.append(np.full(idx.size, field, dtype=np.int32))
            weights.append(w)
        cells = np.concatenate(cells) if cells else np.zeros(0, np.int64)
        if cells.size == 0:
            _native.call("fdtd_ctx_set_membranes", self._ctx, 0, None, None, 0, None, None, None, None, 0, None,
                         0, None, 0)
            return
        order = np.concatenate(order)
        perm = np.lexsort((order, cells))          # by cell, then list order
        cells = cells[perm]
        rows = np.concatenate(rows)[perm]
        fields = np.concatenate(fields)[perm]
        weights = np.ascontiguousarray(np.concatenate(weights)[perm], dtype=np.float64)
        ucells, first = np.unique(cells, return_index=True)
        estart = np.append(first, cells.size).astype(np.int32)
        # probes re-recorded after the injection on cells whose pressure changes
        pflat = [(np.ravel_multi_index(self._probes[n].position, self.shape), q) for q, n in enumerate(probe_names)]
        pstart, pidx = [0], []
        for e, c in enumerate(ucells):
            if pflat and np.any(fields[estart[e]:estart[e + 1]] == 0):
                pidx += [q for f, q in pflat if f == c]
            pstart.append(len(pidx))
        ucells = np.ascontiguousarray(ucells, dtype=np.int64)
        rows = np.ascontiguousarray(rows, dtype=np.int32)
        fields = np.ascontiguousarray(fields, dtype=np.int32)
        pstart = np.array(pstart, dtype=np.int32)
        pidx = np.array(pidx or [0], dtype=np.int32)
        _native.call("fdtd_ctx_set_membranes", self._ctx, int(ucells.size), ucells.ctypes.data, estart.ctypes.data,
                     int(cells.size), rows.ctypes.data, fields.ctypes.data, weights.ctypes.data,
                     pstart.ctypes.data, int(pstart[-1]), pidx.ctypes.data, len(table),
                     _native.ptr(self._mwave), cap)

    def _prepare(self, n_steps: int) -> None:
        if self._dirty_geometry:
            self._upload_geometry()
        if self._dirty_ade:
            self._setup_ade()
        self._ensure_capacity(self._step_count + n_steps)
        self._flush_mirrors()

    def _advance_host(self, n: int) -> None:
        for _ in range(n):
            self._step_times.append(self._time)
            self._time += self.dt
        self._step_count += n
        self._mirror_valid = False

    # ------------------------------------------------------------------ steps
    def step(self) -> None:
        """Advance one time step (core/solver.py:2003-2077)."""
        self._prepare(1)
        n_before, t_before = self._step_count, self._time
        _native.call("fdtd_step_n", self._ctx, 1)
        self._advance_host(1)
        self._after_step_hooks(n_before, t_before)

    def _after_step_hooks(self, n_before: int, t_before: float) -> None:
        if self._snapshot_interval and n_before % self._snapshot_interval == 0:
            self._snapshots.append((t_before, self._snapshot_p()))
            if self._snapshot_velocity:
                self._velocity_snapshots.append((t_before, *self._centred_velocity()))
        if self._track_energy and self._step_count % self._energy_sample_interval == 0:
            self._queue_energy()

    def _hooks_fire(self, n_before: int) -> bool:
        """Whether the step taken at count n_before snapshots or samples energy."""
        if self._snapshot_interval and n_before % self._snapshot_interval == 0:
            return True
        return bool(self._track_energy and (n_before + 1) % self._energy_sample_interval == 0)

    def _to_pinned(self, dev_tensor) -> np.ndarray:
        """Copy a device tensor to new pinned host memory on the solver's
        stream without waiting; the array is valid after _sync_d2h()."""
        torch = _torch()
        host = torch.empty(dev_tensor.shape, dtype=dev_tensor.dtype, pin_memory=True)
        with torch.cuda.stream(self._stream):
            host.copy_(dev_tensor, non_blocking=True)
        self._d2h_pending = True
        return host.numpy()

    def _sync_d2h(self) -> None:
        if self._d2h_pending:
            self._stream.synchronize()
            self._d2h_pending = False

    def _snapshot_p(self) -> np.ndarray:
        """The current pressure, copied asynchronously to pinned host memory."""
        self._flush_mirrors()
        return self._to_pinned(self._f[self._cur][0])

    def time_sweeps(self, n: int) -> tuple[float, float]:
        """Advance n (1..4096) steps eagerly with HIP events around every fused
        sweep launch; returns (summed sweep ms, whole-window ms)."""
        import ctypes as C
        self._prepare(n)
        ms_sweep, ms_total = C.c_float(), C.c_float()
        _native.call("fdtd_time_sweeps", self._ctx, n, C.byref(ms_sweep), C.byref(ms_total))
        self._advance_host(n)
        return float(ms_sweep.value), float(ms_total.value)

    def stream_ceiling(self, reps: int = 20, order: str = "linear") -> float:
        """Milliseconds per pass of a streaming kernel with the sweep's bytes
        per cell and no stencil, on this solve's buffers (state unchanged):
        the practical bandwidth ceiling the sweep is measured against.
        order "linear": 16-byte accesses in memory order; "march": the
        sweep's own strips (one lane per k cell, its rows x planes)."""
        import ctypes as C
        self._prepare(1)
        ms = C.c_float()
        fn = {"linear": "fdtd_stream_ceiling", "march": "fdtd_stream_ceiling_march"}[order]
        _native.call(fn, self._ctx, reps, C.byref(ms))
        return float(ms.value)

    def step_n(self, n: int, graph_chunk: int = 0) -> None:
        """Advance n steps on the device without host round trips.

        graph_chunk = 0 (the default) enqueues the steps' kernels one by one;
        graph_chunk = m replays a captured hipGraph of m steps. Measured on
        MI355X, the graph replay costs about 3.5 us more per step than the
        same kernels enqueued back to back (100^3: 12.9 against 9.6 us per
        step; 512^3: 901 against 863 us), so graphs are only worth it where
        the host cannot keep up with the launches."""
        if n <= 0:
            return
        self._prepare(n)
        done = 0
        if graph_chunk and n >= 2 * graph_chunk:
            reps = n // graph_chunk
            _native.call("fdtd_graph_run", self._ctx, graph_chunk, reps)
            done = reps * graph_chunk
        if n - done:
            _native.call("fdtd_step_n", self._ctx, n - done)
        self._advance_host(n)

    def run(self, duration: float, progress: bool = False, track_energy: bool = False,
            energy_sample_interval: int = 1, output_file: str | None = None,
            script_content: str | None = None, callback: Callable[[int], None] | None = None,
            snapshot_interval: int | None = None) -> None:
        """Run ceil(duration/dt) steps (core/solver.py:2520-2606).

        output_file: an HDF5 file in the reference's result layout
        (io/hdf5.py), written when the run ends (also when it ends with an
        exception), with the pressure field after every step whose index in
        this run is a multiple of snapshot_interval."""
        import time as _time_mod
        n_steps = int(np.ceil(duration / self.dt))
        self._track_energy = track_energy
        self._energy_sample_interval = energy_sample_interval
        if track_energy and not self._energy_history and not self._energy_pending:
            self._queue_energy()
        h5_snaps = [] if output_file and snapshot_interval else None
        started = _time_mod.time()
        bar = None
        if progress:
            from tqdm import tqdm
            bar = tqdm(total=n_steps, desc="FDTD simulation")
        try:
            step = 0
            while step < n_steps:
                if not callback:
                    # steps at which nothing is recorded go to the device as one
                    # run (with a progress bar, runs of at most 256 steps, each
                    # waited for before the bar moves)
                    j = step
                    limit = min(n_steps, step + 256) if bar is not None else n_steps
                    while j < limit and not self._hooks_fire(self._step_count + j - step) and not (
                            h5_snaps is not None and j % snapshot_interval == 0):
                        j += 1
                    if j > step:
                        self.step_n(j - step)
                        if bar is not None:
                            self._stream.synchronize()
                            bar.update(j - step)
                        step = j
                        continue
                t_before = self._time
                self.step()
                if h5_snaps is not None and step % snapshot_interval == 0:
                    h5_snaps.append((t_before, self._snapshot_p()))
                if callback:
                    callback(self._step_count - 1)
                if bar is not None:
                    bar.update(1)
                step += 1
        finally:
            if bar is not None:
                bar.close()
            self._sync_d2h()
            if output_file:
                from datetime import datetime, timezone

                from . import hdf5_out
                tree = hdf5_out.solver_tree(self, h5_snaps, snapshot_interval, _time_mod.time() - started,
                                            script_content, datetime.now(timezone.utc).isoformat())
                hdf5_out.write(str(output_file), tree)
        self.check_health()
        self._resolve_energy()
        if self._warn_energy_drift and len(self._energy_history) >= 2:
            rep = self.energy_report()
            if abs(rep["energy_change_percent"]) > self._energy_drift_threshold * 100:
                warnings.warn(f"Energy drift detected: {rep['energy_change_percent']:.2f}% change "
                              f"(threshold: {self._energy_drift_threshold * 100:.1f}%). "
                              f"Status: {rep['conservation_status']}", UserWarning, stacklevel=2)

    def synchronize(self) -> None:
        _torch().cuda.synchronize(self._device)

    def check_health(self) -> int:
        """Read the device health word: bit0 non-finite probe sample, bit1 table overrun."""
        h = int(self._health.item())
        if h & 2:
            raise RuntimeError("device step ran past the source/probe tables (internal error)")
        if h & 1:
            warnings.warn("non-finite probe sample detected (simulation unstable?)", RuntimeWarning,
                          stacklevel=2)
        return h

    # ---------------------------------------------------------------- results
    def get_probe_data(self, name: str | None = None) -> dict[str, NDArray[np.floating]]:
        if name is not None and name not in self._probes:
            raise KeyError(f"Probe '{name}' not found")
        names = [name] if name is not None else list(self._probes)
        if self._trace is None or self._step_count == 0 or self._probe_order != list(self._probes):
            return {n: np.zeros(0, dtype=np.float32) for n in names}
        tr = self._trace[: self._step_count].cpu().numpy()
        out = {}
        for n in names:
            q = self._probe_order.index(n)
            out[n] = np.ascontiguousarray(tr[self._probe_start.get(n, 0):, q], dtype=np.float32)
        return out

    def get_snapshots(self):
        self._sync_d2h()
        return self._snapshots

    def get_velocity_snapshots(self):
        self._sync_d2h()
        return self._velocity_snapshots

    def _centred_velocity(self):
        """Cell-centred vx, vy, vz of the current state (core/solver.py:
        2647-2687), formed on the device and copied asynchronously to pinned
        host memory."""
        torch = _torch()
        self._flush_mirrors()
        with torch.cuda.stream(self._stream):
            outs = [torch.empty(self.shape, dtype=torch.float32, device=self._device) for _ in range(3)]
        _native.call("fdtd_ctx_centred_velocity", self._ctx, *(_native.ptr(t) for t in outs))
        return [self._to_pinned(t) for t in outs]

    def _enqueue_energy(self, out) -> None:
        """Enqueue the energy sums of the current state into the device
        float64 pair `out` (fdtd_ctx_energy), after the steps enqueued so far."""
        self._flush_mirrors()
        if self._dirty_geometry:
            self._upload_geometry()
        _native.call("fdtd_ctx_energy", self._ctx, _native.ptr(out))

    def _energy_from_sums(self, sp: float, sv: float) -> float:
        dV = self.dx**3
        pe = 0.5 * sp / (self.rho * self.c**2) * dV
        ke = 0.5 * self.rho * sv * dV
        return pe + ke

    def compute_energy(self) -> float:
        """1/2 sum_air p^2/(rho c^2) dV + 1/2 rho sum_air |v|^2 dV (core/solver.py:2689-2706),
        reduced on the device in float64 by one pass over the fields."""
        torch = _torch()
        with torch.cuda.stream(self._stream):
            out = torch.empty(2, dtype=torch.float64, device=self._device)
        self._enqueue_energy(out)
        self._stream.synchronize()
        sp, sv = out.cpu().tolist()
        return self._energy_from_sums(sp, sv)

    def _queue_energy(self) -> None:
        """Record an energy sample of the current state without waiting: the
        sums go to a device history read back when the history is read."""
        torch = _torch()
        n = len(self._energy_pending)
        with torch.cuda.stream(self._stream):
            if self._energy_dev is None or self._energy_dev.shape[0] <= n:
                grown = torch.empty((max(256, 2 * (n + 1)), 2), dtype=torch.float64, device=self._device)
                if n:
                    grown[:n].copy_(self._energy_dev[:n])
                self._energy_dev = grown
        self._enqueue_energy(self._energy_dev[n])
        self._energy_pending.append((self._step_count, self._time))

    def _resolve_energy(self) -> None:
        if not self._energy_pending:
            return
        n = len(self._energy_pending)
        self._stream.synchronize()
        sums = self._energy_dev[:n].cpu().numpy()
        for (count, t), (sp, sv) in zip(self._energy_pending, sums):
            self._energy_history.append((count, t, self._energy_from_sums(float(sp), float(sv))))
        self._energy_pending.clear()

    def get_energy_history(self):
        self._resolve_energy()
        return self._energy_history.copy()

    def energy_report(self) -> dict:
        self._resolve_energy()
        if not self._energy_history:
            raise ValueError("No energy history recorded. Call run() with track_energy=True first.")
        e = np.array([x for _, _, x in self._energy_history])
        e0, e1 = e[0], e[-1]
        if e0 == 0:
            pct = 0.0 if e1 == 0 else float("inf")
        else:
            pct = (e1 - e0) / e0 * 100
        status = "stable" if abs(pct) <= 1.0 else ("growing" if pct > 0 else "decaying")
        return {"initial_energy": float(e0), "final_energy": float(e1), "max_energy": float(e.max()),
                "min_energy": float(e.min()), "energy_change_percent": pct,
                "conservation_status": status, "n_samples": len(self._energy_history)}

    def reset(self) -> None:
        """Zero fields, clock, probes, energy history and ADE state (core/solver.py:2781-2800)."""
        for s in self._f:
            for t in s:
                t.zero_()
        self._step_dev.zero_()
        self._health.zero_()
        if self._trace is not None:
            self._trace.zero_()
        for key in ("J", "Jp"):
            if key in self._ade_tensors:
                self._ade_tensors[key].zero_()
        self._step_count = 0
        self._time = 0.0
        self._sync_d2h()
        self._snapshots.clear()
        self._velocity_snapshots.clear()
        self._energy_history.clear()
        self._energy_pending.clear()
        self._track_energy = False
        self._exported.clear()
        self._mirror_valid = False
        for name in self._probe_start:
            self._probe_start[name] = 0
        self._step_times.clear()
        if self._mic_out is not None:
            self._mic_out.zero_()
        for name, mic in self._microphones.items():
            mic._data.clear()
            mic._times.clear()
            self._mic_start[name] = 0
            self._mic_synced[name] = 0
        for b in self._boundaries:
            b.reset()

    def get_sample_rate(self) -> float:
        return 1.0 / self.dt

    def get_frequency_response(self, probe_name: str, n_fft: int | None = None):
        data = self.get_probe_data(probe_name)[probe_name]
        if n_fft is None:
            n_fft = int(2 ** np.ceil(np.log2(len(data))))
        spectrum = np.fft.rfft(data, n=n_fft)
        return np.fft.rfftfreq(n_fft, self.dt), np.abs(spectrum)


__all__ = ["FDTDSolver", "has_mi355x_backend", "NonuniformGrid", "UniformGrid"]
