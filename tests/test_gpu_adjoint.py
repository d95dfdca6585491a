"""Adjoint gradients w.r.t. per-voxel density and bulk modulus.

Pins: (1) with rho = rho0 and K = K0 the per-voxel forward equals the plain
solver bit for bit; (2) in float64 the hand-written adjoint equals torch
autograd through a differentiable restatement of the same forward; (3) in
float64 it matches central finite differences; (4) the float32 adjoint (Kahan
accumulation over time) stays close to float64; (5) the checkpoint interval
does not change the result.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _problem(steps=40, shape=(16, 14, 18), checkpoint=10, geometry=False):
    from strata_fdtd_amd import GaussianPulse
    from strata_fdtd_amd.adjoint import DesignProblem
    air = None
    if geometry:
        air = np.ones(shape, bool)
        air[3:5, 2:4, 12:15] = False
    return DesignProblem(shape=shape, pml_depth=3, source=GaussianPulse(position=(4, 7, 6), frequency=9000.0),
                         probes=[(11, 6, 9), (8, 9, 12), (5, 5, 5)], steps=steps, box_lo=(5, 4, 5),
                         box_hi=(11, 10, 12), checkpoint=checkpoint, geometry=air)


def _params(pb, dtype, seed=3):
    import torch
    rng = np.random.default_rng(seed)
    rho = torch.tensor(pb.rho * (1 + 0.05 * rng.standard_normal(pb.box_shape)), dtype=dtype, device="cuda")
    K = torch.tensor(pb.K0 * (1 + 0.05 * rng.standard_normal(pb.box_shape)), dtype=dtype, device="cuda")
    target = torch.tensor(1e-3 * rng.standard_normal((pb.steps, len(pb.probes))), dtype=dtype, device="cuda")
    return rho, K, target


def _grad(pb, rho, K, target):
    from strata_fdtd_amd.adjoint import probe_traces
    rho = rho.clone().requires_grad_(True)
    K = K.clone().requires_grad_(True)
    tr = probe_traces(rho, K, pb)
    loss = ((tr - target) ** 2).sum()
    loss.backward()
    return loss.item(), rho.grad.detach(), K.grad.detach()


def test_uniform_parameters_reduce_to_plain_solver():
    import torch
    from helpers import Scenario, make_solver
    from strata_fdtd_amd.adjoint import forward_traces
    pb = _problem(steps=50, geometry=True)
    rho = torch.full(pb.box_shape, pb.rho, device="cuda")
    K = torch.full(pb.box_shape, pb.K0, device="cuda")
    tr, _ = forward_traces(pb, rho, K)
    sc = Scenario(shape=pb.shape, pml_depth=3, air=pb.geometry, sources=[("point", pb.source.position, 9000.0)],
                  probes=pb.probes)
    s = make_solver(sc)
    s.step_n(pb.steps)
    ref = s.get_probe_data()
    for q in range(len(pb.probes)):
        assert np.array_equal(tr[:, q].cpu().numpy(), ref[f"probe{q}"])


def test_fp64_adjoint_equals_torch_autograd():
    import torch
    import torch_reference
    from strata_fdtd_amd.adjoint import _Device
    pb = _problem(geometry=True)
    rho, K, target = _params(pb, torch.float64)
    _, g_rho, g_K = _grad(pb, rho, K, target)
    dv = _Device(pb, torch.float64, torch.device("cuda"))
    decays = [d.cpu().numpy() for d in dv.dec]
    rho_t = rho.clone().requires_grad_(True)
    K_t = K.clone().requires_grad_(True)
    tr = torch_reference.forward(pb, rho_t, K_t, pb.geometry, decays, dv.wave)
    loss = ((tr - target) ** 2).sum()
    loss.backward()
    for a, b in ((g_rho, rho_t.grad), (g_K, K_t.grad)):
        err = (a - b).norm() / b.norm()
        print(f"fp64 adjoint vs autograd: rel {err.item():.3e}")
        assert err.item() < 1e-9


def test_fp64_adjoint_equals_torch_autograd_64cubed():
    """The adjoint workload shrunk to 64^3 (PML 12, a central 32^3 design box
    of air + 5% seeded noise, 1 kHz pulse, 4 probes), 240 steps: float64
    adjoint vs torch autograd through the differentiable restatement."""
    import torch
    import torch_reference
    from strata_fdtd_amd import GaussianPulse
    from strata_fdtd_amd.adjoint import DesignProblem, _Device
    rng = np.random.default_rng(5)
    lo, hi = (16, 16, 16), (48, 48, 48)
    pb = DesignProblem(shape=(64, 64, 64), pml_depth=12, source=GaussianPulse(position=(24, 32, 32), frequency=1000.0),
                       probes=[tuple(int(v) for v in rng.integers(20, 44, 3)) for _ in range(4)], steps=240,
                       box_lo=lo, box_hi=hi, checkpoint=20)
    print(f"[adjoint] autograd 64^3, probes {pb.probes}", flush=True)
    rho, K, target = _params(pb, torch.float64, seed=5)
    _, g_rho, g_K = _grad(pb, rho, K, target)
    dv = _Device(pb, torch.float64, torch.device("cuda"))
    decays = [d.cpu().numpy() for d in dv.dec]
    rho_t = rho.clone().requires_grad_(True)
    K_t = K.clone().requires_grad_(True)
    tr = torch_reference.forward(pb, rho_t, K_t, np.ones(pb.shape, bool), decays, dv.wave)
    ((tr - target) ** 2).sum().backward()
    for a, b in ((g_rho, rho_t.grad), (g_K, K_t.grad)):
        err = (a - b).norm() / b.norm()
        print(f"fp64 adjoint vs autograd at 64^3: rel {err.item():.3e}")
        assert b.norm() > 0 and err.item() < 1e-9


def test_fp64_adjoint_matches_central_differences():
    """Central differences of the float64 loss in the ratio the kernels read
    (r = rho/rho0, kappa = K/K0, formed in float32): the step is the difference
    of the two float32 ratios that reach the kernels, and the adjoint's dL/drho
    times rho0 is dL/dr. Truncation error ~h^2 = 1e-8; bound 1e-5 relative."""
    import torch
    from strata_fdtd_amd.adjoint import design_ratio, forward_traces
    pb = _problem()
    rho, K, target = _params(pb, torch.float64)
    _, g_rho, g_K = _grad(pb, rho, K, target)
    rng = np.random.default_rng(8)

    def loss(r, k):
        tr, _ = forward_traces(pb, r, k, dtype=torch.float64)
        return float(((tr - target) ** 2).sum())

    for which, base, grad, ref in (("rho", rho, g_rho, pb.rho), ("K", K, g_K, pb.K0)):
        for _ in range(4):
            idx = tuple(int(rng.integers(0, n)) for n in pb.box_shape)
            h = 1e-4 * float(base[idx])
            plus, minus = base.clone(), base.clone()
            plus[idx] += h
            minus[idx] -= h
            dr = float(design_ratio(plus, ref)[idx].double() - design_ratio(minus, ref)[idx].double())
            lp = loss(plus if which == "rho" else rho, plus if which == "K" else K)
            lm = loss(minus if which == "rho" else rho, minus if which == "K" else K)
            fd = (lp - lm) / dr
            g = float(grad[idx]) * ref
            print(f"{which}{idx}: adjoint dL/dratio {g:.9e} fd {fd:.9e} rel {abs(g - fd) / abs(fd):.2e}")
            assert abs(g - fd) <= 1e-5 * abs(fd) + 1e-12 * abs(lp)


def test_fp32_adjoint_close_to_fp64():
    import torch
    pb = _problem(steps=60)
    rho64, K64, t64 = _params(pb, torch.float64)
    _, gr64, gk64 = _grad(pb, rho64, K64, t64)
    _, gr32, gk32 = _grad(pb, rho64.float(), K64.float(), t64.float())
    for a, b in ((gr32.double(), gr64), (gk32.double(), gk64)):
        err = ((a - b).norm() / b.norm()).item()
        print(f"fp32 vs fp64 gradient: rel {err:.3e}")
        assert err < 2e-3


def test_checkpoint_interval_does_not_change_gradient():
    import torch
    pb7 = _problem(checkpoint=7)
    pb40 = _problem(checkpoint=40)
    rho, K, target = _params(pb7, torch.float32)
    _, a_r, a_k = _grad(pb7, rho, K, target)
    _, b_r, b_k = _grad(pb40, rho, K, target)
    assert torch.equal(a_r, b_r) and torch.equal(a_k, b_k)


def test_fast_forward_equals_cell_kernels():
    """The float32 march forward performs the per-cell kernel's operations in
    the same order: identical traces, with and without recorded history."""
    import torch
    from strata_fdtd_amd.adjoint import forward_traces
    pb = _problem(steps=45, geometry=True)
    rho, K, _ = _params(pb, torch.float32)
    a, _ = forward_traces(pb, rho, K, impl="cell")
    for hist in ("all", "checkpoint"):
        b, _ = forward_traces(pb, rho, K, impl="fast", keep_checkpoints=True, history=hist)
        assert torch.equal(a, b), hist
    c, _ = forward_traces(pb, rho, K, impl="fast")
    assert torch.equal(a, c)


@pytest.mark.parametrize("history", ["all", "checkpoint"])
def test_fast_gradient_equals_cell_kernels(history):
    """The float32 march adjoint (forward state from the recorded box history)
    gives the per-cell adjoint's gradients bit for bit."""
    import torch
    from strata_fdtd_amd.adjoint import probe_traces
    pb = _problem(steps=45, checkpoint=10, geometry=True)
    rho, K, target = _params(pb, torch.float32)
    out = {}
    for impl in ("cell", "fast"):
        r = rho.clone().requires_grad_(True)
        k = K.clone().requires_grad_(True)
        tr = probe_traces(r, k, pb, impl=impl, history=history)
        ((tr - target) ** 2).sum().backward()
        out[impl] = (tr.detach(), r.grad, k.grad)
    for x, y in zip(out["cell"], out["fast"]):
        assert torch.equal(x, y), history
    assert out["fast"][1].abs().max() > 0 and out["fast"][2].abs().max() > 0


def test_fast_equals_cell_kernels_on_many_strips():
    """A grid of many strips: three 64-wide k segments of which the box meets
    one, a partial last row strip inside the box's row range, several plane
    strips; probes inside and outside the box. The strips that miss the box
    (constant coefficients) and the box strips (split finer, per-voxel
    coefficients, gradient) together give the per-cell kernels' traces and
    gradients bit for bit, also with rho and K absent (= air)."""
    import torch
    from strata_fdtd_amd import GaussianPulse
    from strata_fdtd_amd.adjoint import DesignProblem, forward_traces, probe_traces
    shape = (40, 37, 150)
    air = np.ones(shape, bool)
    air[12:15, 20:30, 80:84] = False
    pb = DesignProblem(shape=shape, pml_depth=3, source=GaussianPulse(position=(20, 18, 60), frequency=9000.0),
                       probes=[(20, 30, 100), (33, 5, 20), (10, 35, 140), (29, 8, 71)], steps=45,
                       box_lo=(9, 7, 70), box_hi=(30, 36, 110), checkpoint=10, geometry=air)
    print(f"[adjoint] many strips {shape}", flush=True)
    a, _ = forward_traces(pb, None, None, impl="cell")
    b, _ = forward_traces(pb, None, None, impl="fast", keep_checkpoints=True, history="all")
    assert torch.equal(a, b)
    rho, K, target = _params(pb, torch.float32, seed=9)
    out = {}
    for impl in ("cell", "fast"):
        r = rho.clone().requires_grad_(True)
        k = K.clone().requires_grad_(True)
        tr = probe_traces(r, k, pb, impl=impl, history="all")
        ((tr - target) ** 2).sum().backward()
        out[impl] = (tr.detach(), r.grad, k.grad)
    for x, y in zip(out["cell"], out["fast"]):
        assert torch.equal(x, y)
    assert out["fast"][0].abs().max() > 0 and out["fast"][1].abs().max() > 0


@pytest.mark.slow
def test_adjoint_workload_192_vs_central_differences():
    """The adjoint workload at its own size: 192^3 at 1 mm, PML 12, a 128^3
    design box of rho and K = air * (1 + 0.05 N(0,1)), 1 kHz pulse, 4 probes,
    2000 steps, L2 loss to seeded targets, gradient with a checkpoint every
    100 steps and recomputation (float32 fast kernels, Kahan accumulation).
    Checked against float64 central differences with h = 1e-4 relative on 32
    seeded voxels (16 of rho, 16 of K), in the ratio the kernels read.

    Bounds: every voxel |g - fd| <= 5e-6 ||g|| (||g|| the L2 norm of the whole
    rho or K gradient; the float32 adjoint's rounding floor, measured at about
    4e-7 ||g||), and |g - fd| <= 0.02 |fd| + 5e-6 ||g|| per voxel."""
    import torch
    from strata_fdtd_amd.adjoint import design_problem_192, design_ratio, forward_traces, probe_traces
    pb, rho, K, target = design_problem_192()
    assert pb.checkpoint == 100
    print("[adjoint] 192^3 gradient (checkpoint every 100 steps)", flush=True)
    r = rho.clone().requires_grad_(True)
    k = K.clone().requires_grad_(True)
    tr = probe_traces(r, k, pb, history="checkpoint")
    ((tr - target) ** 2).sum().backward()
    g_rho, g_K = r.grad.double(), k.grad.double()
    norms = {"rho": float(g_rho.norm()), "K": float(g_K.norm())}
    assert norms["rho"] > 0 and norms["K"] > 0
    rng = np.random.default_rng(17)
    t64 = target.double()

    def loss64(rr, kk):
        out, _ = forward_traces(pb, rr, kk, dtype=torch.float64)
        return float(((out - t64) ** 2).sum())

    rel = []
    for v in range(32):
        which = "rho" if v % 2 == 0 else "K"
        base, grad, ref = (rho, g_rho, pb.rho) if which == "rho" else (K, g_K, pb.K0)
        idx = tuple(int(rng.integers(0, s)) for s in pb.box_shape)
        b64 = base.double()
        plus, minus = b64.clone(), b64.clone()
        plus[idx] *= 1 + 1e-4
        minus[idx] *= 1 - 1e-4
        # the step that reaches the kernels: the difference of the float32 ratios
        dr = float(design_ratio(plus, ref)[idx].double() - design_ratio(minus, ref)[idx].double())
        lp = loss64(plus if which == "rho" else rho.double(), plus if which == "K" else K.double())
        lm = loss64(minus if which == "rho" else rho.double(), minus if which == "K" else K.double())
        fd = (lp - lm) / (dr * ref)
        g = float(grad[idx])
        err = abs(g - fd)
        rel.append(err / abs(fd))
        print(f"{which}{idx}: adjoint {g:.6e} fd {fd:.6e} |g-fd|/|fd| {err / abs(fd):.2e} "
              f"|g-fd|/||g|| {err / norms[which]:.2e}", flush=True)
        assert err <= 5e-6 * norms[which], (which, idx)
        assert err <= 0.02 * abs(fd) + 5e-6 * norms[which], (which, idx)
    print(f"per-voxel relative error: median {np.median(rel):.2e}, max {np.max(rel):.2e}")
