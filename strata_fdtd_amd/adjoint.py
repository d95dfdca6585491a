"""Probe-loss gradients with respect to per-voxel density and bulk modulus.

``DesignProblem`` describes a solve whose design region (a box of cells)
carries per-voxel density rho and bulk modulus K; outside the box the medium is
the background air. ``probe_traces(rho, K, problem)`` is a
``torch.autograd.Function``: its forward runs the per-voxel forward step on
the GPU (checkpointing the state every ``checkpoint`` steps) and returns the
probe traces; its backward runs the hand-written time-reversed adjoint kernel
segment by segment — restoring a checkpoint, recomputing the segment's states,
then stepping the adjoint backwards through them — and accumulates dL/drho and
dL/dK with Kahan-compensated summation over time. Any loss built from the
traces in torch gets its gradient through this Function.

Model (no counterpart in the reference, which has scalar rho and c):
face coefficient cv0 / mean(r_lo, r_hi) with r = rho/rho0, cell coefficient
cp0 * K/K0; with rho = rho0 and K = K0 everywhere it is the reference step.
Float64 mode (``dtype=torch.float64``) runs the same kernels in double for
finite-difference checks.
"""

from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _native
from .boundaries import PML
from .grid import UniformGrid
from .sources import GaussianPulse, accumulated_times

_vp = C.c_void_p


def _lib():
    return _native.load()


def _ptr(t):
    return None if t is None else t.data_ptr()


@dataclass
class DesignProblem:
    shape: tuple[int, int, int]
    resolution: float = 1e-3
    c: float = 343.0
    rho: float = 1.204
    pml_depth: int = 12
    source: GaussianPulse | None = None
    probes: list = field(default_factory=list)
    steps: int = 2000
    box_lo: tuple[int, int, int] = (0, 0, 0)
    box_hi: tuple[int, int, int] = (0, 0, 0)
    checkpoint: int = 100
    geometry: np.ndarray | None = None        # bool air mask (None: all air)
    courant: float = 0.95

    @property
    def box_shape(self) -> tuple[int, int, int]:
        return tuple(h - l for l, h in zip(self.box_lo, self.box_hi))

    @property
    def K0(self) -> float:
        return self.rho * self.c**2


class _Device:
    """Static device data of a DesignProblem (flags, decays, probes, waveform)."""

    def __init__(self, pb: DesignProblem, dtype, device):
        import torch
        self.pb = pb
        self.dtype = dtype
        self.dev = device
        nx, ny, nz = pb.shape
        self.dt = pb.courant * (pb.resolution / (pb.c * np.sqrt(3)))
        self.cv0 = -self.dt / (pb.rho * pb.resolution)
        self.cp0 = -pb.rho * pb.c**2 * self.dt / pb.resolution
        if dtype == torch.float32:
            self.cv0, self.cp0 = float(np.float32(self.cv0)), float(np.float32(self.cp0))
        air = np.ones(pb.shape, np.uint8) if pb.geometry is None else pb.geometry.astype(np.uint8)
        air_d = torch.from_numpy(air).to(device)
        self.flags = torch.empty(pb.shape, dtype=torch.uint8, device=device)
        _native.call("fdtd_build_flags", _ptr(air_d), None, _ptr(self.flags), nx, ny, nz,
                     torch.cuda.current_stream(device).cuda_stream)
        self.air = air_d

        class _S:  # PML.initialize reads these attributes
            pass
        s = _S()
        s.grid = UniformGrid(pb.shape, pb.resolution)
        s.shape, s.c, s.dx, s.dt = pb.shape, pb.c, pb.resolution, self.dt
        pml = PML(depth=pb.pml_depth)
        if pb.pml_depth:
            pml.initialize(s)
            tabs = pml.decay_tables(self.dt, pb.shape)
        else:
            tabs = [np.ones(n, np.float32) for n in pb.shape]
        self.dec = [torch.from_numpy(np.ascontiguousarray(t)).to(device) for t in tabs]
        self.probe_cells = torch.tensor([np.ravel_multi_index(p, pb.shape) for p in pb.probes] or [0],
                                        dtype=torch.int64, device=device)
        self.n_probe = len(pb.probes)
        src = pb.source
        self.src_cell = int(np.ravel_multi_index(tuple(src.position), pb.shape)) if src is not None else -1
        times = accumulated_times(self.dt, pb.steps)
        self.wave = np.ascontiguousarray(src.waveform(times, self.dt) if src is not None else np.zeros(pb.steps),
                                         dtype=np.float64)
        self.lo = np.array(pb.box_lo, dtype=np.int32)
        self.hi = np.array(pb.box_hi, dtype=np.int32)
        self.code = 8 if dtype == torch.float64 else 4
        self.stream = torch.cuda.current_stream(device).cuda_stream

    def zeros_state(self):
        import torch
        return [torch.zeros(self.pb.shape, dtype=self.dtype, device=self.dev) for _ in range(4)]

    def zeros_state_pair(self, pairs: int = 1):
        """Both ping-pong sets {p, vx, vy, vz} as views of one zeroed slab
        (`pairs` such pairs from the same slab). Eight separate allocations
        land wherever the allocator finds room, and at 192^3 (fields plus
        parameters just inside the 256 MiB Infinity Cache) that moved the fast
        forward between 91 and 111 ms for 2000 steps; one slab ran it in
        90-92 ms every time."""
        import torch
        n = int(np.prod(self.pb.shape))
        slab = torch.zeros(8 * pairs * n, dtype=self.dtype, device=self.dev)
        views = [slab[q * n:(q + 1) * n].view(self.pb.shape) for q in range(8 * pairs)]
        out = [[views[8 * m:8 * m + 4], views[8 * m + 4:8 * m + 8]] for m in range(pairs)]
        return out[0] if pairs == 1 else out

    def step(self, src_state, dst_state, n, r, kappa, trace_row):
        pb = self.pb
        rc = _lib().fdtd_param_step(self.code, *[_ptr(t) for t in src_state], *[_ptr(t) for t in dst_state],
                                    *pb.shape, _ptr(self.flags), *[_ptr(d) for d in self.dec],
                                    self.lo.ctypes.data, self.hi.ctypes.data, _ptr(r), _ptr(kappa),
                                    self.cv0, self.cp0, self.src_cell, float(self.wave[n]), self.n_probe,
                                    _ptr(self.probe_cells), _ptr(trace_row), self.stream)
        if rc != 0:
            raise _native.NativeError(f"fdtd_param_step failed ({rc})")

    def _common(self, r, kappa):
        pb = self.pb
        return (*pb.shape, _ptr(self.flags), *[_ptr(d) for d in self.dec], self.lo.ctypes.data,
                self.hi.ctypes.data, _ptr(r), _ptr(kappa), self.cv0, self.cp0)

    def run(self, states, step0, n_steps, pingpong, r, kappa, trace):
        """n_steps forward steps in one native call (see fdtd_param_run)."""
        if n_steps <= 0:
            return
        ptrs = [_ptr(t) for st in states for t in st]
        arr = (_vp * len(ptrs))(*ptrs)
        _native.call("fdtd_param_run", self.code, arr, len(states), step0, n_steps, 1 if pingpong else 0,
                     *self._common(r, kappa), self.src_cell, self.wave.ctypes.data, self.n_probe,
                     _ptr(self.probe_cells), _ptr(trace), self.stream)

    def adjoint_run(self, lam_pair, states, n_lo, n_hi, r, kappa, g, acc):
        """Adjoint steps n_hi-1 .. n_lo in one native call; returns the index of
        the buffer of `lam_pair` holding lambda_{n_lo}."""
        ptrs = [_ptr(t) for st in lam_pair for t in st]
        lam_arr = (_vp * 8)(*ptrs)
        sp = [_ptr(t) for st in states for t in st]
        st_arr = (_vp * max(len(sp), 1))(*sp) if sp else (_vp * 1)()
        _native.call("fdtd_adjoint_run", self.code, lam_arr, st_arr, n_lo, n_hi, *self._common(r, kappa),
                     self.n_probe, _ptr(self.probe_cells), _ptr(g), *[_ptr(a) for a in acc], self.stream)
        return (n_hi - n_lo) & 1

    def adjoint(self, lam, lam_out, state, r, kappa, g_row, acc):
        pb = self.pb
        rc = _lib().fdtd_adjoint_step(self.code, *[_ptr(t) for t in lam], *[_ptr(t) for t in lam_out],
                                      *[_ptr(t) for t in state], *pb.shape, _ptr(self.flags),
                                      *[_ptr(d) for d in self.dec], self.lo.ctypes.data, self.hi.ctypes.data,
                                      _ptr(r), _ptr(kappa), self.cv0, self.cp0, self.n_probe,
                                      _ptr(self.probe_cells), _ptr(g_row), *[_ptr(a) for a in acc],
                                      self.stream)
        if rc != 0:
            raise _native.NativeError(f"fdtd_adjoint_step failed ({rc})")


class _Fast:
    """Float32 march kernels of the design solve (csrc/design.hip): forward
    sweep with optional box history, adjoint sweep reading that history.
    PLANES / ADJ_PLANES: planes per strip of the forward (2-row strips) and
    of the adjoint (4-row strips), the fastest measured at 192^3."""

    PLANES = 2
    ADJ_PLANES = 4

    def __init__(self, dv: _Device):
        import torch
        self.dv = dv
        pb = dv.pb
        self.wave_dev = torch.from_numpy(np.ascontiguousarray(dv.wave, dtype=np.float64)).to(dv.dev)
        src = pb.source
        self.src = (C.c_int32 * 3)(*[int(v) for v in src.position]) if src is not None else None
        self.probe_ijk = torch.tensor(np.array(pb.probes or [(0, 0, 0)], dtype=np.int32).reshape(-1, 3),
                                      device=dv.dev)
        self.dec = (_vp * 3)(*[_ptr(d) for d in dv.dec])
        lo, hi = np.array(pb.box_lo), np.array(pb.box_hi)
        self.ecells = int(np.prod(hi - lo + 2))
        self.bcells = int(np.prod(hi - lo))

    @staticmethod
    def _sets(sets):
        return (_vp * 8)(*[_ptr(t) for st in sets for t in st])

    @staticmethod
    def pack(r, kappa):
        """(r, kappa) interleaved per box voxel, as the kernels read them (one
        8-byte load per voxel); None when both are absent (air)."""
        import torch
        if r is None and kappa is None:
            return None
        one = torch.ones_like(r if r is not None else kappa)
        return torch.stack((r if r is not None else one, kappa if kappa is not None else one), dim=-1).contiguous()

    def _common(self, rk):
        dv, pb = self.dv, self.dv.pb
        return (_ptr(dv.flags), self.dec, *pb.shape, dv.cv0, dv.cp0, dv.lo.ctypes.data, dv.hi.ctypes.data,
                _ptr(rk))

    def history(self, steps):
        import torch
        return (torch.empty(steps * self.ecells, dtype=torch.float32, device=self.dv.dev),
                torch.empty(steps * self.bcells, dtype=torch.float32, device=self.dv.dev))

    def forward(self, sets, rk, trace, step0, n_steps, hist=None, hist0=0):
        hp, hd = hist if hist is not None else (None, None)
        hist_steps = (hp.numel() // self.ecells) if hp is not None else 0
        rows = trace.shape[0] if self.dv.n_probe > 0 else self.wave_dev.numel()
        _native.call("fdtd_design_forward", self._sets(sets), *self._common(rk), self.src,
                     _ptr(self.wave_dev), self.dv.n_probe, _ptr(self.probe_ijk), _ptr(trace),
                     min(rows, self.wave_dev.numel()), step0, n_steps, _ptr(hp), _ptr(hd), hist0, hist_steps,
                     self.PLANES, self.dv.stream)

    def adjoint(self, lam, rk, g, n_lo, n_hi, hist, hist0, acc):
        """acc: float32 [box][4], Kahan (sum, compensation) of dL/dr, then dL/dkappa."""
        import torch
        hp, hd = hist
        if getattr(self, "_faces", None) is None:
            n = _native.load().fdtd_design_face_count(self.dv.lo.ctypes.data, self.dv.hi.ctypes.data)
            self._faces = torch.empty(n, dtype=torch.float32, device=self.dv.dev)
            self._pb_box = torch.empty(2 * self.bcells, dtype=torch.float32, device=self.dv.dev)
        _native.call("fdtd_design_adjoint", self._sets(lam), *self._common(rk), self.dv.n_probe,
                     _ptr(self.probe_ijk), _ptr(g), g.shape[0], n_lo, n_hi, _ptr(hp), _ptr(hd), hist0,
                     hp.numel() // self.ecells, _ptr(acc), _ptr(self._faces), _ptr(self._pb_box), self.ADJ_PLANES,
                     self.dv.stream)


def history_bytes(pb: DesignProblem, steps: int | None = None) -> int:
    """HBM bytes of the float32 forward history (p on the box grown by one
    cell, div on the box) for `steps` steps."""
    lo, hi = np.array(pb.box_lo), np.array(pb.box_hi)
    return 4 * int(np.prod(hi - lo + 2) + np.prod(hi - lo)) * (pb.steps if steps is None else steps)


def _pick_history(pb: DesignProblem, history: str, device) -> str:
    if history != "auto":
        return history
    import torch
    free, _total = torch.cuda.mem_get_info(device)
    return "all" if history_bytes(pb) <= 0.5 * free else "checkpoint"


def _forward_fast(pb, r, kappa, dv, history: str):
    import torch
    fast = _Fast(dv)
    rk = fast.pack(r, kappa)
    rec = None
    if history == "checkpoint":
        # the recomputation's fields come from the same slab as the forward's,
        # so every call sees the same placement (one slab, not two)
        sets, rec = dv.zeros_state_pair(pairs=2)
    else:
        sets = dv.zeros_state_pair()
    traces = torch.zeros((pb.steps, max(dv.n_probe, 1)), dtype=torch.float32, device=dv.dev)
    ckpts, hist = [], None
    if history == "all":
        hist = fast.history(pb.steps)
        fast.forward(sets, rk, traces, 0, pb.steps, hist, 0)
    elif history == "checkpoint":
        K = pb.checkpoint
        for n0 in range(0, pb.steps, K):
            ckpts.append([t.clone() for t in sets[n0 & 1]])
            fast.forward(sets, rk, traces, n0, min(K, pb.steps - n0))
    else:
        fast.forward(sets, rk, traces, 0, pb.steps)
    return traces[:, :dv.n_probe], ("fast", dv, rk, fast, history, hist, ckpts, (sets, rec))


def _adjoint_fast(saved, grad_traces):
    import torch
    _, dv, rk, fast, history, hist, ckpts, (sets, rec) = saved
    pb = dv.pb
    # The forward's field slab is free once its history (or checkpoints) is
    # kept: reused, the adjoint and the next forward see the same buffer
    # placement every call (which the cache-resident sweeps are sensitive to).
    for t in sets[0] + sets[1]:
        t.zero_()
    lam = sets
    acc = torch.zeros(tuple(pb.box_shape) + (4,), dtype=torch.float32, device=dv.dev)
    g = torch.zeros((pb.steps, max(dv.n_probe, 1)), dtype=torch.float32, device=dv.dev)
    g[:, :dv.n_probe] = grad_traces.to(torch.float32)
    if history == "all":
        fast.adjoint(lam, rk, g, 0, pb.steps, hist, 0, acc)
    elif history == "checkpoint":
        K = pb.checkpoint
        seg_hist = fast.history(K)
        scratch = torch.zeros_like(g)
        for seg in reversed(range(len(ckpts))):
            n0 = seg * K
            n1 = min(pb.steps, n0 + K)
            for t, c in zip(rec[n0 & 1], ckpts[seg]):
                t.copy_(c)
            fast.forward(rec, rk, scratch, n0, n1 - n0, seg_hist, n0)
            fast.adjoint(lam, rk, g, n0, n1, seg_hist, n0, acc)
    else:
        raise RuntimeError("forward ran without history or checkpoints")
    # Kahan: the compensation holds minus the low-order part the sums lost
    return acc[..., 0] - acc[..., 1], acc[..., 2] - acc[..., 3]


def design_ratio(x, ref: float, device="cuda"):
    """The per-voxel ratio the kernels read: x / ref formed in float32 on the
    device (rho/rho0 or K/K0). Finite-difference checks use it to find the
    perturbation that actually reaches the kernels."""
    import torch
    return (x.detach().to(device=device, dtype=torch.float32) / ref).contiguous()


def forward_traces(pb: DesignProblem, rho=None, K=None, dtype=None, device="cuda",
                   keep_checkpoints: bool = False, impl: str = "auto", history: str = "auto"):
    """Run the per-voxel forward solve; returns (traces [steps, n_probe], saved
    state for the adjoint). impl "fast" (float32 march kernels) or "cell" (one
    thread per cell, float32 or float64); "auto" picks "fast" for float32.
    history (fast, with keep_checkpoints): "all" keeps the box history of every
    step in HBM, "checkpoint" keeps the state every `pb.checkpoint` steps and
    recomputes each segment's history in the backward pass; "auto" takes "all"
    when it fits in half the free HBM."""
    import torch
    dtype = dtype or torch.float32
    dev = torch.device(device)
    dv = _Device(pb, dtype, dev)
    r = design_ratio(rho, pb.rho, device) if rho is not None else None
    kappa = design_ratio(K, pb.K0, device) if K is not None else None
    if impl == "auto":
        impl = "fast" if dtype == torch.float32 else "cell"
    if impl == "fast":
        if dtype != torch.float32:
            raise ValueError("the fast design kernels are float32")
        mode = _pick_history(pb, history, dev) if keep_checkpoints else "none"
        return _forward_fast(pb, r, kappa, dv, mode)
    a, b = dv.zeros_state(), dv.zeros_state()
    traces = torch.zeros((pb.steps, max(dv.n_probe, 1)), dtype=dtype, device=device)
    ckpts = []
    K = pb.checkpoint if keep_checkpoints else pb.steps
    for n0 in range(0, pb.steps, K):
        if keep_checkpoints:
            ckpts.append([t.clone() for t in a])
        m = min(K, pb.steps - n0)
        dv.run([a, b], n0, m, True, r, kappa, traces)
        if m & 1:
            a, b = b, a
    return traces[:, :dv.n_probe], ("cell", dv, r, kappa, ckpts)


def adjoint_gradients(saved, grad_traces):
    """Backward pass: returns (dL/dr, dL/dkappa) on the design box."""
    import torch
    if saved[0] == "fast":
        return _adjoint_fast(saved, grad_traces)
    _, dv, r, kappa, ckpts = saved
    pb = dv.pb
    K = pb.checkpoint
    lam, lam_out = dv.zeros_state(), dv.zeros_state()
    acc = [torch.zeros(pb.box_shape, dtype=dv.dtype, device=dv.dev) for _ in range(4)]
    g = torch.zeros((pb.steps, max(dv.n_probe, 1)), dtype=dv.dtype, device=dv.dev)
    g[:, :dv.n_probe] = grad_traces.to(dv.dtype)
    slots = [dv.zeros_state() for _ in range(K)]
    scratch = torch.zeros((pb.steps, max(dv.n_probe, 1)), dtype=dv.dtype, device=dv.dev)
    n_seg = (pb.steps + K - 1) // K
    for seg in reversed(range(n_seg)):
        n0 = seg * K
        n1 = min(pb.steps, n0 + K)
        for t, c in zip(slots[0], ckpts[seg]):
            t.copy_(c)
        dv.run(slots[:n1 - n0], n0, n1 - n0 - 1, False, r, kappa, scratch)
        if dv.adjoint_run([lam, lam_out], slots[:n1 - n0], n0, n1, r, kappa, g, acc):
            lam, lam_out = lam_out, lam
    # Kahan: the compensation holds minus the low-order part the sums lost
    return acc[0] - acc[1], acc[2] - acc[3]


class _ProbeTraces:
    """Namespace holder so the Function is built lazily (torch import)."""

    fn = None


def probe_traces(rho, K, problem: DesignProblem, impl: str = "auto", history: str = "auto"):
    """Differentiable probe traces [steps, n_probe] as a function of the design
    region's density and bulk modulus (torch tensors of the box shape). impl and
    history as in forward_traces."""
    import torch

    if _ProbeTraces.fn is None:
        class ProbeTraces(torch.autograd.Function):
            @staticmethod
            def forward(ctx, rho_, K_, pb, impl_, history_):
                traces, saved = forward_traces(pb, rho_, K_, dtype=rho_.dtype if rho_.dtype == torch.float64
                                               else torch.float32, device=rho_.device,
                                               keep_checkpoints=True, impl=impl_, history=history_)
                ctx.saved = saved
                ctx.pb = pb
                return traces

            @staticmethod
            def backward(ctx, grad_traces):
                gr, gk = adjoint_gradients(ctx.saved, grad_traces.contiguous())
                pb = ctx.pb
                ctx.saved = None
                return (gr / pb.rho).to(grad_traces.dtype), (gk / pb.K0).to(grad_traces.dtype), None, None, None

        _ProbeTraces.fn = ProbeTraces
    return _ProbeTraces.fn.apply(rho, K, problem, impl, history)


def design_problem_192(n: int = 192, steps: int = 2000, seed: int = 5, pml: int = 12,
                       box: int = 128, checkpoint: int = 100) -> tuple[DesignProblem, object, object, object]:
    """The adjoint workload: n^3 at 1 mm, air, PML 12, a box^3 central design
    region of per-voxel rho and K = air * (1 + 0.05 N(0,1)) (seeded), 1 kHz
    pulse, 4 probes, `steps` steps, L2 loss to seeded target traces.
    Returns (problem, rho, K, target) as torch tensors on cuda."""
    import torch
    rng = np.random.default_rng(seed)
    lo = ((n - box) // 2,) * 3
    hi = tuple(v + box for v in lo)
    c0 = n // 2
    src = GaussianPulse(position=(c0 - box // 4, c0, c0), frequency=1000.0)
    probes = [tuple(int(v) for v in rng.integers(lo[0] + 4, hi[0] - 4, 3)) for _ in range(4)]
    pb = DesignProblem(shape=(n, n, n), pml_depth=pml, source=src, probes=probes, steps=steps,
                       box_lo=lo, box_hi=hi, checkpoint=checkpoint)
    rho = torch.tensor(pb.rho * (1 + 0.05 * rng.standard_normal((box,) * 3)), dtype=torch.float32,
                       device="cuda")
    K = torch.tensor(pb.K0 * (1 + 0.05 * rng.standard_normal((box,) * 3)), dtype=torch.float32,
                     device="cuda")
    target = torch.tensor(0.8 * rng.standard_normal((steps, 4)) * 1e-3, dtype=torch.float32, device="cuda")
    return pb, rho, K, target
