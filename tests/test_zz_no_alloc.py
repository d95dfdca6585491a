"""fdtd_voxelize_static (the voxeliser entry point that allocates nothing)
against fdtd_voxelize with a caller workspace, through the C ABI.

Two calls in a row on one stream, the second with a smaller grid and a
smaller program set, must give the workspace entry point's air masks and
material ids bit for bit (the staging array keeps the first call's programs
and must not leak into the second). A program set larger than the fixed
staging array is refused on the host with an error, before any upload.
"""

import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _programs(geometry):
    from strata_fdtd_amd.voxelize_device import _regions_of, compile_programs
    tree, regions = _regions_of(geometry)
    return compile_programs(tree, regions)


def _call(ps, grid, static: bool, stream):
    import torch
    from strata_fdtd_amd import _native
    lib = _native.load()
    nx, ny, nz = grid.shape
    dev = torch.device("cuda", torch.cuda.current_device())
    coords = [torch.from_numpy(np.ascontiguousarray(c, dtype=np.float64)).to(dev)
              for c in (grid.x_coords, grid.y_coords, grid.z_coords)]
    # poison the outputs so an unwritten cell cannot pass
    air = torch.full((nx, ny, nz), 0xA5, dtype=torch.uint8, device=dev)
    ids = torch.full((nx, ny, nz), 0x5A, dtype=torch.uint8, device=dev) if ps.n_prog > 1 else None
    ops = np.ascontiguousarray(ps.ops)
    prm = ps.params.ctypes.data if len(ps.params) else None
    torch.cuda.synchronize()
    if static:
        _native.call("fdtd_voxelize_static", ops.ctypes.data, len(ops), ps.start.ctypes.data, ps.ids.ctypes.data,
                     ps.n_prog, prm, len(ps.params), *[_native.ptr(c) for c in coords], nx, ny, nz,
                     _native.ptr(air), _native.ptr(ids), stream.cuda_stream)
    else:
        need = int(lib.fdtd_voxelize_scratch_bytes(len(ops), ps.n_prog, len(ps.params)))
        scratch = torch.empty(max(need, 8), dtype=torch.uint8, device=dev)
        torch.cuda.synchronize()
        _native.call("fdtd_voxelize", ops.ctypes.data, len(ops), ps.start.ctypes.data, ps.ids.ctypes.data,
                     ps.n_prog, prm, len(ps.params), _native.ptr(scratch), scratch.numel(),
                     *[_native.ptr(c) for c in coords], nx, ny, nz, _native.ptr(air), _native.ptr(ids),
                     stream.cuda_stream)
    stream.synchronize()
    return air.cpu().numpy(), None if ids is None else ids.cpu().numpy()


def test_static_voxelize_equals_workspace_entry_point():
    import torch
    import strata_fdtd_amd as sfa
    from strata_fdtd_amd.workloads import enclosure_512_scene
    stream = torch.cuda.current_stream()
    big, *_ = enclosure_512_scene(96)
    small = sfa.MaterializedGeometry(sfa.Difference(sfa.Box(center=(0.02, 0.02, 0.02), size=(0.03, 0.028, 0.034)),
                                                    sfa.Sphere(center=(0.021, 0.019, 0.02), radius=0.009)))
    small.set_material(sfa.Box(center=(0.01, 0.02, 0.02), size=(0.012, 0.04, 0.04)),
                       sfa.SimpleMaterial(name="m", _rho=1.2, _c=340.0))
    cases = [(big, sfa.UniformGrid((96, 96, 96), 1e-3)),
             (small, sfa.UniformGrid((40, 36, 44), 1e-3))]
    sizes = [len(_programs(g).ops) for g, _ in cases]
    assert sizes[1] < sizes[0]
    for geo, grid in cases:
        print(f"[no-alloc] voxelize {grid.shape}", flush=True)
        ps = _programs(geo)
        a_ref, i_ref = _call(ps, grid, False, stream)
        a_st, i_st = _call(ps, grid, True, stream)
        assert set(np.unique(a_ref)) <= {0, 1}
        assert np.array_equal(a_st, a_ref)
        assert np.array_equal(i_st, i_ref)
        mask, ids = geo.voxelize_with_materials(grid)
        assert np.array_equal(a_st.astype(bool), mask) and np.array_equal(i_st, ids)


def test_static_voxelize_refuses_programs_beyond_its_array():
    import torch
    import strata_fdtd_amd as sfa
    from strata_fdtd_amd import _native
    lib = _native.load()
    boxes = [sfa.Box(center=(0.001 * q, 0.01, 0.01), size=(0.002, 0.002, 0.002)) for q in range(6000)]
    ps = _programs(sfa.Union(*boxes))
    need = int(lib.fdtd_voxelize_scratch_bytes(len(ps.ops), ps.n_prog, len(ps.params)))
    assert need > 256 * 1024
    dev = torch.device("cuda", torch.cuda.current_device())
    xc = torch.zeros(4, dtype=torch.float64, device=dev)
    air = torch.zeros((4, 4, 4), dtype=torch.uint8, device=dev)
    ops = np.ascontiguousarray(ps.ops)
    rc = lib.fdtd_voxelize_static(ops.ctypes.data, len(ops), ps.start.ctypes.data, ps.ids.ctypes.data, ps.n_prog,
                                  ps.params.ctypes.data, len(ps.params), xc.data_ptr(), xc.data_ptr(),
                                  xc.data_ptr(), 4, 4, 4, air.data_ptr(), None,
                                  C.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert rc < 0
    assert b"staging" in lib.fdtd_last_error()
